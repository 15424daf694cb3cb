// Batched dense complex LU with partial pivoting, the matching multi-RHS
// solve, complex GEMM, and the mixed-precision (complex64 factor, complex128
// refinement) solve, for gfx950.
//
// Replaces LAPACK zgetrf / zgetrs as reached through scipy.linalg.lu_factor
// and lu_solve (capytaine tools/block_circulant_matrices.py:361-385, called
// from bem/engines.py:326-338).  LAPACK conventions: column-major storage,
// P A = L U with L unit lower triangular, pivots chosen by the largest
// |Re|+|Im| (first index on ties), ipiv[i] = row swapped with row i (0-based
// here), info = 1 + index of the first exactly-zero pivot.
//
// Every kernel is templated on the real type R: R = double gives complex128
// on v_mfma_f64_16x16x4f64, R = float gives complex64 on
// v_mfma_f32_16x16x4f32.
//
// Right-looking blocked algorithm (block width 256 from 2048 rows, 128 from
// 1024, else 64; tunable), one launch per step for the whole batch:
//   1. panel: recursive halving down to 8- or 16-column leaves factored by
//      one work-group per matrix, the halves joined by trsm + MFMA GEMM;
//   2. row interchanges applied to the columns outside the panel;
//   3. U12 = L11^-1 A12 in place: L11's 64 x 64 diagonal blocks inverted in
//      one launch, each block row multiplied by its inverse on MFMA, the
//      rows below updated by the MFMA GEMM;
//   4. A22 -= L21 U12 with the MFMA complex GEMM, the next panel's columns
//      first so that, with a side stream, the next panel can be factored
//      while the rest is updated.
// The solve inverts the factor's 64 x 64 diagonal blocks once and runs one
// substitution launch per block and direction (MFMA products only).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <cmath>
#include <utility>
#include <vector>
#include "cpt_internal.hpp"

namespace cpt {
namespace {

typedef double dvec4 __attribute__((ext_vector_type(4)));
typedef float fvec4 __attribute__((ext_vector_type(4)));
typedef _Float16 hvec4 __attribute__((ext_vector_type(4)));
typedef float fvec4u __attribute__((ext_vector_type(4), aligned(8)));  // 16 bytes at 8-byte alignment

template <typename R>
struct cx {
    R re, im;
};

template <typename R>
__device__ __forceinline__ cx<R> zmul(cx<R> a, cx<R> b)
{
    return cx<R>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R>
__device__ __forceinline__ cx<R> zsub(cx<R> a, cx<R> b)
{
    return cx<R>{a.re - b.re, a.im - b.im};
}
template <typename R>
__device__ __forceinline__ cx<R> zrcp(cx<R> a)
{
    // Smith's reciprocal, no overflow for |a| up to the floating-point range.
    if (fabs(a.re) >= fabs(a.im)) {
        const R r = a.im / a.re, d = a.re + a.im * r;
        return cx<R>{R(1) / d, -r / d};
    } else {
        const R r = a.re / a.im, d = a.re * r + a.im;
        return cx<R>{r / d, R(-1) / d};
    }
}
// p[i] where ok, else zero, as a load on every lane (index 0 where !ok)
// and a select of values: a select of addresses would turn the access into
// a flat load between the buffer and a zero in private memory.
template <typename R>
__device__ __forceinline__ cx<R> ld_or_zero(const cx<R>* p, size_t i, bool ok)
{
    const cx<R> v = p[ok ? i : 0];
    return cx<R>{ok ? v.re : R(0), ok ? v.im : R(0)};
}

template <typename R>
__device__ __forceinline__ R cabs1(cx<R> a)
{
    return fabs(a.re) + fabs(a.im);
}

// MFMA 16 x 16 x 4 of each precision.  Operand lanes: A[i][k] at lane
// i + 16 k, B[k][j] at lane j + 16 k.  Accumulator register v of lane l holds
// D[row(l, v)][l % 16]: row 4 v + l/16 for FP64 (four groups of one row),
// 4 (l/16) + v for FP32 (one group of four rows).
template <typename R>
struct Mfma;
template <>
struct Mfma<double> {
    typedef dvec4 acc;
    static __device__ __forceinline__ acc op(double a, double b, acc c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * v + (lane >> 4); }
};
template <>
struct Mfma<float> {
    typedef fvec4 acc;
    static __device__ __forceinline__ acc op(float a, float b, acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

constexpr int PB = 64;  // widest triangle the panel and trsm kernels hold (one wavefront)

inline int round_up_host(int x, int m) { return (x + m - 1) / m * m; }

// The panel kernel's sub-block update loads L for groups of panel_qg rows
// and panel_jc columns per round trip: at most `cap` complex values per
// thread in flight; panel_jc divides SB and panel_qg divides RPT.
constexpr int panel_qg(int rpt, int cap) { return rpt < cap ? rpt : cap; }
constexpr int panel_jc(int rpt, int sb, int cap)
{
    const int jc = rpt >= cap ? 1 : cap / rpt;
    return jc < sb ? jc : sb;
}

// ---------------------------------------------------------------------------
// 1. Panel factorisation
//
// One work-group of NT threads per matrix.  Rows k0..N-1 of the panel are
// distributed round-robin (row t = tid + NT q, q < RPT); the panel columns
// are processed in sub-blocks of SB columns held in registers:
//   (a) load the sub-block;
//   (b) its rows above the diagonal block are U: solve with the unit lower
//       triangle L11 of the columns already factored (L11 staged in LDS, one
//       wavefront per column, the running solution held across the lanes);
//   (c) left-looking update of the rows below with L(rows, k0:c0) U;
//   (d) factor the SB columns: block-wide pivot search with the candidate
//       rows broadcast through LDS (one barrier per column), row
//       interchange, scaling and rank-1 updates in registers;
//   (e) store the sub-block, then apply its interchanges to the other
//       columns [sw_lo, sw_hi) in global memory (once per sub-block).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double lane_bcast(double v, int j)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), j);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), j);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float lane_bcast(float v, int j)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}

// Value of lane j (wave-uniform j) in every lane, through a scalar register.
// DPP moves within rows of 16 lanes: xor 1, xor 2 (quad permutations),
// then the mirrors of 8 and of 16 lanes.
constexpr int DPP_XOR1 = 0xB1, DPP_XOR2 = 0x4E, DPP_HALF_MIRROR = 0x141, DPP_MIRROR = 0x140;

template <int CTRL>
__device__ __forceinline__ int dpp_i(int v)
{
    return __builtin_amdgcn_update_dpp(v, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_r(float v)
{
    return __int_as_float(dpp_i<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_r(double v)
{
    const int lo = dpp_i<CTRL>(__double2loint(v)), hi = dpp_i<CTRL>(__double2hiint(v));
    return __hiloint2double(hi, lo);
}

// (best, idx) <- the largest best of the wavefront, the smallest idx among
// equal values (LAPACK's first index on ties); result uniform.  Lanes
// without a candidate hold best = -1, idx = INT_MAX.
template <typename R, int CTRL>
__device__ __forceinline__ void argmax_step(R& best, int& idx)
{
    const R ob = dpp_r<CTRL>(best);
    const int oi = dpp_i<CTRL>(idx);
    if (ob > best || (ob == best && oi < idx)) {
        best = ob;
        idx = oi;
    }
}
template <typename R>
__device__ __forceinline__ void wave_argmax(R& best, int& idx)
{
    argmax_step<R, DPP_XOR1>(best, idx);
    argmax_step<R, DPP_XOR2>(best, idx);
    argmax_step<R, DPP_HALF_MIRROR>(best, idx);
    argmax_step<R, DPP_MIRROR>(best, idx);
    // every lane of a row of 16 holds the row's result: combine the four rows
    R b = lane_bcast(best, 0);
    int i = __builtin_amdgcn_readlane(idx, 0);
#pragma unroll
    for (int r = 16; r < 64; r += 16) {
        const R rb = lane_bcast(best, r);
        const int ri = __builtin_amdgcn_readlane(idx, r);
        if (rb > b || (rb == b && ri < i)) {
            b = rb;
            i = ri;
        }
    }
    best = b;
    idx = i;
}

template <typename R, int NT, int RPT, int SB>
__global__ void __launch_bounds__(NT) panel_kernel(cx<R>* __restrict__ Abase, int64_t lda, int64_t strideA, int N,
                                                   int k0, int kb, int sw_lo, int sw_hi,
                                                   int* __restrict__ ipiv_base, int64_t strideP,
                                                   int* __restrict__ info_base)
{
    typedef cx<R> C;
    C* A = Abase + (size_t)blockIdx.x * strideA;
    int* ipiv = ipiv_base + (size_t)blockIdx.x * strideP;
    const int M = N - k0;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    constexpr int NW = NT / 64;
    const C zero{R(0), R(0)};

    __shared__ C sL[PB * PB];  // L11, column-major (i + PB j)
    __shared__ C sU[PB * SB];
    __shared__ C sCand[2][NW][SB];  // per column parity: each wavefront's candidate row
    __shared__ C sDiag[2][SB];      // and the diagonal row
    __shared__ R sBest[2][NW];
    __shared__ int sIdx[2][NW];
    __shared__ int sPiv[SB];

    C v[RPT][SB];

    for (int c0 = k0; c0 < k0 + kb; c0 += SB) {
        const int sbw = min(SB, k0 + kb - c0);
        const int t0 = c0 - k0;
        // (a)
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int t = tid + NT * q;
#pragma unroll
            for (int s = 0; s < SB; ++s) {
                if (t < M && s < sbw) v[q][s] = A[(k0 + t) + (size_t)(c0 + s) * lda];
                else v[q][s].re = v[q][s].im = R(0);
            }
        }
        if (t0 > 0) {
            // (b)
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int t = tid + NT * q;
                if (t < t0)
#pragma unroll
                    for (int s = 0; s < SB; ++s) sU[t * SB + s] = v[q][s];
            }
            for (int e = tid; e < t0 * t0; e += NT) {
                const int i = e % t0, j = e / t0;
                sL[i + PB * j] = A[(k0 + i) + (size_t)(k0 + j) * lda];
            }
            __syncthreads();
            // column sweep, one wavefront per column (NW may be below SB):
            // x_j broadcast from lane j, lanes below update
            for (int s = wave; s < sbw; s += NW) {
                C x = ld_or_zero(sU, lane * SB + s, lane < t0);
                for (int j = 0; j < t0 - 1; ++j) {
                    const C xj{lane_bcast(x.re, j), lane_bcast(x.im, j)};
                    if (lane > j && lane < t0) x = zsub(x, zmul(sL[lane + PB * j], xj));
                }
                if (lane < t0) sU[lane * SB + s] = x;
            }
            __syncthreads();
            // (c) the rows above t0 take the solved block; the rows below
            // subtract L's columns [k0, k0 + t0) times it, JC columns of QG
            // rows loaded at once (one round trip per group of rows and JC
            // columns; t0 is a multiple of SB), j ascending as in a
            // column-by-column update
            constexpr int CAP = sizeof(R) == 4 ? 16 : 8;
            constexpr int QG = panel_qg(RPT, CAP);
            constexpr int JC = panel_jc(RPT, SB, CAP);
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int t = tid + NT * q;
                if (t < t0)
#pragma unroll
                    for (int s = 0; s < SB; ++s) v[q][s] = sU[t * SB + s];
            }
#pragma unroll
            for (int g = 0; g < RPT; g += QG) {
                for (int j0 = 0; j0 < t0; j0 += JC) {
                    C l[QG][JC];
#pragma unroll
                    for (int qq = 0; qq < QG; ++qq) {
                        const int t = tid + NT * (g + qq);
#pragma unroll
                        for (int u = 0; u < JC; ++u) l[qq][u] = zero;
                        if (t >= t0 && t < M) {
                            const C* Lrow = A + (k0 + t) + (size_t)(k0 + j0) * lda;
#pragma unroll
                            for (int u = 0; u < JC; ++u) l[qq][u] = Lrow[(size_t)u * lda];
                        }
                    }
#pragma unroll
                    for (int qq = 0; qq < QG; ++qq) {
                        const int t = tid + NT * (g + qq);
                        if (t >= t0 && t < M)
#pragma unroll
                            for (int u = 0; u < JC; ++u)
#pragma unroll
                                for (int s = 0; s < SB; ++s)
                                    v[g + qq][s] = zsub(v[g + qq][s], zmul(l[qq][u], sU[(j0 + u) * SB + s]));
                    }
                }
            }
        }
        // (d) one barrier per column: every wavefront writes its best
        // candidate (|Re|+|Im|, row, and that row's SB values) and the owner
        // of the diagonal row writes that row, each into the half of a
        // double buffer of this column's parity (a wavefront can only write
        // a half again after the next column's barrier, which every reader
        // of this column has passed); after the barrier every thread picks
        // the same winner and reads the pivot row directly
#pragma unroll
        for (int s = 0; s < SB; ++s) {
            if (s < sbw) {
                const int td = t0 + s;
                const int buf = s & 1;
                R best = R(-1);
                int bidx = INT_MAX;
#pragma unroll
                for (int q = 0; q < RPT; ++q) {
                    const int t = tid + NT * q;
                    if (t >= td && t < M) {
                        const R a = cabs1(v[q][s]);
                        if (a > best) {
                            best = a;
                            bidx = t;
                        }
                    }
                }
                wave_argmax(best, bidx);
                if (lane == 0) {
                    sBest[buf][wave] = best;
                    sIdx[buf][wave] = bidx;
                }
#pragma unroll
                for (int q = 0; q < RPT; ++q) {
                    const int t = tid + NT * q;
                    if (t == bidx)
#pragma unroll
                        for (int s2 = 0; s2 < SB; ++s2) sCand[buf][wave][s2] = v[q][s2];
                    if (t == td)
#pragma unroll
                        for (int s2 = 0; s2 < SB; ++s2) sDiag[buf][s2] = v[q][s2];
                }
                __syncthreads();
                // every thread combines the wave maxima (same order, same result)
                R b = sBest[buf][0];
                int p = sIdx[buf][0], wb = 0;
#pragma unroll
                for (int w = 1; w < NW; ++w) {
                    const R bw = sBest[buf][w];
                    const int iw = sIdx[buf][w];
                    if (bw > b || (bw == b && iw < p)) {
                        b = bw;
                        p = iw;
                        wb = w;
                    }
                }
                const bool pivoted = b > R(0);
                if (!pivoted) p = td;  // zero column: no interchange, report singularity
                if (tid == 0) {
                    if (!pivoted && info_base[blockIdx.x] == 0) info_base[blockIdx.x] = c0 + s + 1;
                    ipiv[c0 + s] = k0 + p;
                    sPiv[s] = p;
                }
                // the pivot row becomes row td, the old row td goes to row p
                const C* prow = pivoted ? sCand[buf][wb] : sDiag[buf];
#pragma unroll
                for (int q = 0; q < RPT; ++q) {
                    const int t = tid + NT * q;
                    if (p != td && t == td)
#pragma unroll
                        for (int s2 = 0; s2 < SB; ++s2) v[q][s2] = prow[s2];
                    if (p != td && t == p)
#pragma unroll
                        for (int s2 = 0; s2 < SB; ++s2) v[q][s2] = sDiag[buf][s2];
                }
                const C piv = prow[s];
                const bool nonzero = (piv.re != R(0) || piv.im != R(0));
                const C r = nonzero ? zrcp(piv) : C{R(1), R(0)};
#pragma unroll
                for (int q = 0; q < RPT; ++q) {
                    const int t = tid + NT * q;
                    if (t > td && t < M) {
                        const C l = nonzero ? zmul(v[q][s], r) : v[q][s];
                        v[q][s] = l;
#pragma unroll
                        for (int s2 = s + 1; s2 < SB; ++s2) v[q][s2] = zsub(v[q][s2], zmul(l, prow[s2]));
                    }
                }
            }
        }
        __syncthreads();  // sPiv of the sub-block complete
        // (e)
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int t = tid + NT * q;
            if (t < M)
#pragma unroll
                for (int s = 0; s < SB; ++s)
                    if (s < sbw) A[(k0 + t) + (size_t)(c0 + s) * lda] = v[q][s];
        }
        // the sub-block's interchanges on the other columns [sw_lo, sw_hi):
        // one thread per (column, involved row) pair finds, walking the
        // interchanges backwards, which original row ends in that row, loads
        // it, and after a barrier stores it -- one round trip per chunk of
        // whole columns (columns are independent, so chunks need no order)
        {
            constexpr int PPT = 4;  // pairs per thread and chunk
            const int nslot = 2 * sbw;
            const int ncol = sw_hi - sw_lo;
            const int cols_per_chunk = (NT * PPT) / nslot;
            for (int cb = 0; cb < ncol; cb += cols_per_chunk) {
                const int npairs = min(cols_per_chunk, ncol - cb) * nslot;
                // values as scalars and destinations as offsets (-1: none),
                // all in registers
                R vre[PPT], vim[PPT];
                int64_t dst[PPT];
#pragma unroll
                for (int u = 0; u < PPT; ++u) {
                    const int pair = tid + NT * u;
                    vre[u] = vim[u] = R(0);
                    dst[u] = -1;
                    if (pair < npairs) {
                        const int col = sw_lo + cb + pair / nslot, j = pair % nslot;
                        if (col < c0 || col >= c0 + sbw) {
                            const int r = (j & 1) ? sPiv[j >> 1] : t0 + (j >> 1);
                            int x = r;
                            for (int s = sbw - 1; s >= 0; --s) {
                                if (x == t0 + s) x = sPiv[s];
                                else if (x == sPiv[s]) x = t0 + s;
                            }
                            const int64_t colo = k0 + (int64_t)col * lda;
                            const C val = A[colo + x];
                            vre[u] = val.re;
                            vim[u] = val.im;
                            dst[u] = colo + r;
                        }
                    }
                }
                __syncthreads();
#pragma unroll
                for (int u = 0; u < PPT; ++u)
                    if (dst[u] >= 0) A[dst[u]] = C{vre[u], vim[u]};
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// 1b. Panel leaf over several work-groups per matrix.
//
// P work-groups of 512 threads per matrix each own a slab of 512 or 1024
// rows of the leaf, one or two rows per thread with all W leaf columns in
// registers: the leaf is loaded once and stored once.  Row interchanges move no data:
// each register slot carries the (leaf-relative) row it currently holds, an
// interchange of rows j and p swaps two labels, and the final store writes
// each slot to its row.
//
// Per column, one barrier inside the work-group: every wave finds its best
// candidate (largest |Re|+|Im|, first row on ties) with a DPP reduction and
// stages the candidate's row, value and reciprocal in LDS (two buffers, by
// column parity, so that the next column's staging never meets a slow
// wave's reads); after the barrier every wave combines the staged entries
// itself (same inputs, same result), so no second barrier hands the winner
// around.  With P > 1 the work-group's winner is published -- value, row,
// reciprocal and the row's W entries -- in an exchange record, and the
// whole work-group reads all P records of its matrix in one round of loads
// into LDS; one more barrier tells every wave that each of its words showed
// the column's sequence number.  Each 64-bit word of a record holds 32 bits
// of data and, in its high half, the sequence number: a record is complete
// when every word shows the current sequence, so no fence is needed
// (agent-scope atomic loads and stores, coherent across the XCDs).  Records
// alternate between two parities: a work-group writes the record of column
// j + 2 only after it has read every record of column j + 1, which the other
// work-groups write only after reading those of column j.  Sequence numbers
// are unique within a factorisation, whose start clears the records.  The
// wait is bounded (spin polls; a negative spin gives up at once, a test
// hook): a work-group that gives up sets the matrix's info to -1 and
// returns, so that every work-group ends.  Register indices stay static:
// the update of column j writes each slot's result one slot down, so the
// pivot column is always slot 0 and the finished column moves to the end.
// ---------------------------------------------------------------------------
constexpr int MW_PMAX = 32;    // work-groups per matrix
constexpr int MW_ROWW = 64;    // words of one row of W entries (2 W for float pairs, 4 W for double pairs)
constexpr int MW_REC = 72;     // words per record: header, candidate row
constexpr int MW_XWORDS = 2 * MW_PMAX * MW_REC;  // words per matrix: two parities of MW_PMAX records
constexpr int MW_LEAF_BUDGET = 256;  // work-groups of one leaf launch over the batch

template <typename R>
constexpr int mw_width() { return sizeof(R) == 4 ? 32 : 16; }

// Header of a record, in 32-bit data words: best (1 or 2 words) at 0, row at
// 2, reciprocal of the pivot (re, im) from 4; the row's entries follow.
template <typename R>
constexpr int mw_hdr() { return 4 + 2 * (int)(sizeof(R) / 4); }

__device__ __forceinline__ void mw_store(uint64_t* p, uint32_t seq, uint32_t bits)
{
    __hip_atomic_store(p, ((uint64_t)seq << 32) | bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t mw_load(const uint64_t* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One step of the argmax with a DPP move restricted to the rows in ROWMASK;
// lanes of the other rows compare with themselves and keep their values.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ int dpp_keep(int v)
{
    return __builtin_amdgcn_update_dpp(v, v, CTRL, ROWMASK, 0xF, false);
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_keep(float v)
{
    return __int_as_float(dpp_keep<CTRL, ROWMASK>(__float_as_int(v)));
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_keep(double v)
{
    const int lo = dpp_keep<CTRL, ROWMASK>(__double2loint(v)), hi = dpp_keep<CTRL, ROWMASK>(__double2hiint(v));
    return __hiloint2double(hi, lo);
}
// Full-row moves (every source lane valid) need no old value.
template <int CTRL>
__device__ __forceinline__ int dpp_any(int v)
{
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_any(float v)
{
    return __int_as_float(dpp_any<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_any(double v)
{
    const int lo = dpp_any<CTRL>(__double2loint(v)), hi = dpp_any<CTRL>(__double2hiint(v));
    return __hiloint2double(hi, lo);
}
template <typename R>
__device__ __forceinline__ void argmax_take(R& best, int& idx, R ob, int oi)
{
    // all three comparisons evaluated (no short circuit, no branches)
    const bool take = (ob > best) | ((ob == best) & (oi < idx));
    best = take ? ob : best;
    idx = take ? oi : idx;
}
template <typename R, int CTRL, int ROWMASK>
__device__ __forceinline__ void argmax_keep(R& best, int& idx)
{
    if constexpr (ROWMASK == 0xF)
        argmax_take(best, idx, dpp_any<CTRL>(best), dpp_any<CTRL>(idx));
    else
        argmax_take(best, idx, dpp_keep<CTRL, ROWMASK>(best), dpp_keep<CTRL, ROWMASK>(idx));
}
constexpr int DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143;

// (best, idx) over the lanes of row 0 (lanes 0..15), uniform: the largest
// best, the smallest idx among equal values.
template <typename R>
__device__ __forceinline__ void row0_argmax(R& best, int& idx)
{
    argmax_keep<R, DPP_XOR1, 0xF>(best, idx);
    argmax_keep<R, DPP_XOR2, 0xF>(best, idx);
    argmax_keep<R, DPP_HALF_MIRROR, 0xF>(best, idx);
    argmax_keep<R, DPP_MIRROR, 0xF>(best, idx);
    best = lane_bcast(best, 0);
    idx = __builtin_amdgcn_readlane(idx, 0);
}
// The same over the whole wavefront: rows reduced by the mirrors, then row
// 0 into row 1 and row 2 into row 3 (broadcast of lane 15 of each row), then
// rows 0-1 into row 3 (broadcast of lane 31); lane 63 holds the result.
template <typename R>
__device__ __forceinline__ void wave_argmax_bcast(R& best, int& idx)
{
    argmax_keep<R, DPP_XOR1, 0xF>(best, idx);
    argmax_keep<R, DPP_XOR2, 0xF>(best, idx);
    argmax_keep<R, DPP_HALF_MIRROR, 0xF>(best, idx);
    argmax_keep<R, DPP_MIRROR, 0xF>(best, idx);
    argmax_keep<R, DPP_ROW_BCAST15, 0xA>(best, idx);
    argmax_keep<R, DPP_ROW_BCAST31, 0xC>(best, idx);
    best = lane_bcast(best, 63);
    idx = __builtin_amdgcn_readlane(idx, 63);
}

__device__ __forceinline__ float rfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double rfma(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Word k of a real as 32-bit data (k < sizeof(R) / 4).
__device__ __forceinline__ uint32_t real_word(float v, int) { return __float_as_uint(v); }
__device__ __forceinline__ uint32_t real_word(double v, int k)
{
    return (uint32_t)(k ? __double2hiint(v) : __double2loint(v));
}

template <typename R, int W, int MW_NT, int RPT>
__global__ void __launch_bounds__(MW_NT) panel_mw_kernel(cx<R>* __restrict__ Abase, int64_t lda, int64_t strideA,
                                                         int N, int k0, int kb, int sw_lo, int sw_hi,
                                                         int* __restrict__ ipiv_base, int64_t strideP,
                                                         int* __restrict__ info_base, uint64_t* __restrict__ xbase,
                                                         int P, uint32_t seq0, int spin)
{
    typedef cx<R> C;
    constexpr int NV = sizeof(R) / 4;         // 32-bit words per real
    constexpr int HDR = mw_hdr<R>();
    constexpr int PER = HDR + 2 * NV * W;     // words read per record
    constexpr int WPT = (MW_PMAX * PER + MW_NT - 1) / MW_NT;  // record words per thread
    static_assert(2 * NV * W <= MW_ROWW && PER <= MW_REC, "leaf row does not fit a record");
    static_assert(HDR % (2 * NV) == 0 && (MW_REC * 4) % 16 == 0, "record data misaligned for its reals");
    constexpr int NW = MW_NT / 64;
    static_assert(NW <= 16, "the wave entries are combined within one row of 16 lanes");
    constexpr int Rw = MW_NT * RPT;  // RPT rows per thread, each with the leaf's W entries in registers
    const int g = blockIdx.x, b = blockIdx.y;
    C* A = Abase + (size_t)blockIdx.y * strideA;
    int* ipiv = ipiv_base + (size_t)b * strideP;
    uint64_t* X = xbase + (size_t)b * MW_XWORDS;
    const int M = N - k0;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int row0 = g * Rw;
    const int rend = min(M, row0 + Rw);

    __shared__ C sCand[2][NW][W];  // each wave's candidate row, finished columns as zeros
    __shared__ C sRcp[2][NW];      // its reciprocal
    __shared__ R sBest[2][NW];
    __shared__ int sIdx[2][NW];
    __shared__ int sOk[2][NW];     // exchange polls: the wave saw every word of its share
    __shared__ C sZero[W];         // the pivot row of a column without pivot
    __shared__ int sPiv[W];
    __shared__ int sSrc[2 * W], sDst[2 * W];
    // the records of one column as read: the data halves of their words,
    // record r from word r * MW_REC (16-byte aligned)
    __shared__ __align__(16) uint32_t sXd[MW_PMAX * MW_REC + 1];

    if (tid < W) sZero[tid] = C{R(0), R(0)};

    C v[RPT][W];
    int lab[RPT];  // leaf row held by slot q (INT_MAX: no row)
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int t = row0 + tid + MW_NT * q;
        lab[q] = t < rend ? t : INT_MAX;
#pragma unroll
        for (int s = 0; s < W; ++s) {
            if (t < rend && s < kb) v[q][s] = A[(k0 + t) + (size_t)(k0 + s) * lda];
            else v[q][s] = C{R(0), R(0)};
        }
    }

    // after j steps, register slot s holds leaf column (s + j) mod W.  The
    // loop is not unrolled: the code of one step fits the instruction cache.
    bool quit = false;
    int jend = kb;  // set to the current step when giving up: one exit only
#pragma unroll 1
    for (int j = 0; j < jend; ++j) {
        const int buf = j & 1;
        const uint32_t seq = seq0 + (uint32_t)j;
        // entries of the pivot row beyond the active columns (the finished
        // ones, at the end of the slots) enter the update as zeros
        const int active = W - j;
        {
            R best = R(-1);
            int bidx = INT_MAX;
            C cb = v[0][0];  // the thread's best entry
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const R a = cabs1(v[q][0]);
                const bool take = (lab[q] != INT_MAX) & (lab[q] >= j) & ((a > best) | ((a == best) & (lab[q] < bidx)));
                best = take ? a : best;
                bidx = take ? lab[q] : bidx;
                cb = C{take ? v[q][0].re : cb.re, take ? v[q][0].im : cb.im};
            }
            // the reciprocal of the thread's best entry, computed beside the
            // reduction: the winner's is staged with its row
            const C myrcp = (cb.re != R(0) || cb.im != R(0)) ? zrcp(cb) : C{R(1), R(0)};
            wave_argmax_bcast(best, bidx);
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                if (lab[q] == bidx) {
#pragma unroll
                    for (int s = 0; s < W; ++s) sCand[buf][wave][s] = v[q][s];
                    sRcp[buf][wave] = myrcp;
                }
            }
            // after the owner's stores (one wave's LDS operations complete
            // in order)
            if (lane >= active && lane < W) sCand[buf][wave][lane] = C{R(0), R(0)};
            if (lane == 0) {
                sBest[buf][wave] = best;
                sIdx[buf][wave] = bidx;
            }
        }
        __syncthreads();
        // the work-group's candidate, combined by every wave
        const int wl = lane < NW ? lane : 0;  // indices stay inside the arrays
        R bg = lane < NW ? sBest[buf][wl] : R(-1);
        int pg = lane < NW ? sIdx[buf][wl] : INT_MAX;
        const int mine = pg;
        row0_argmax(bg, pg);
        const unsigned long long whit = __ballot(lane < NW && mine == pg);
        const int wg = whit ? __builtin_ctzll(whit) : 0;
        R wbest = bg;
        int pwin = pg;
        C rcp = sRcp[buf][wg];
        const C* prow = &sCand[buf][wg][0];
        int ab = 0;  // uniform: every wave reads the same flags
        if (P > 1) {
            uint64_t* recs = X + (size_t)buf * MW_PMAX * MW_REC;
            if (wave == 0) {
                // this work-group's record of the column
                uint64_t* rec = recs + (size_t)g * MW_REC;
                if (lane < 2 * NV * W) {
                    const int e = lane / (2 * NV), k = lane % (2 * NV);
                    const C c = prow[e];
                    mw_store(rec + HDR + lane, seq, real_word(k < NV ? c.re : c.im, k % NV));
                }
                if (lane < HDR) {
                    uint32_t bits = 0;
                    if (lane < NV) bits = real_word(bg, lane);
                    else if (lane == 2) bits = (uint32_t)pg;
                    else if (lane >= 4) {
                        const int k = lane - 4;
                        bits = real_word(k < NV ? rcp.re : rcp.im, k % NV);
                    }
                    mw_store(rec + lane, seq, bits);
                }
            }
            // all P records in one round of loads by the whole work-group
            // (word w: record w / PER, its word w % PER), repeated until
            // every word shows this column's sequence
            const int nw = P * PER;
            for (int it = 0;; ++it) {
                // every load issued before the first wait: words past the
                // records read word 0 again, and are not kept
                uint64_t x[WPT];
                int at[WPT];
#pragma unroll
                for (int u = 0; u < WPT; ++u) {
                    const int w = tid + MW_NT * u;
                    const int r = w / PER, o = w - r * PER;
                    at[u] = w < nw ? r * MW_REC + o : -1;
                    x[u] = mw_load(recs + max(at[u], 0));
                }
                bool ok = spin >= 0;
#pragma unroll
                for (int u = 0; u < WPT; ++u) {
                    ok = ok & ((at[u] < 0) | ((uint32_t)(x[u] >> 32) == seq));
                    sXd[at[u] >= 0 ? at[u] : MW_PMAX * MW_REC] = (uint32_t)x[u];  // past the end: discarded
                }
                const bool wok = __all(ok);
                if (lane == 0) sOk[it & 1][wave] = wok;
                __syncthreads();
                if (__all(lane >= NW || sOk[it & 1][wl])) break;
                if (spin < 0 || it >= spin) {
                    ab = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (!ab) {
                // the winner over the records: rows are distinct, the
                // lowest record holding the winning row is its owner
                R rb = R(-1);
                int ri = INT_MAX;
                if (lane < P) {
                    rb = *reinterpret_cast<const R*>(&sXd[lane * MW_REC]);
                    ri = (int)sXd[lane * MW_REC + 2];
                }
                const int rmine = ri;
                wave_argmax_bcast(rb, ri);
                const unsigned long long hit = __ballot(lane < P && rmine == ri);
                const int rr = hit ? __builtin_ctzll(hit) : 0;
                wbest = rb;
                pwin = ri;
                rcp = *reinterpret_cast<const C*>(&sXd[rr * MW_REC + 4]);
                prow = reinterpret_cast<const C*>(&sXd[rr * MW_REC + HDR]);
            }
        }
        if (ab) {
            // the columns left unfinished get no interchange, so that the
            // kernels after this one find valid rows in ipiv
            if (g == 0)
                for (int jj = j + tid; jj < kb; jj += MW_NT) ipiv[k0 + jj] = k0 + jj;
            if (tid == 0) info_base[b] = -1;
            quit = true;
            jend = j;
        }
        // no nonzero candidate (a zero column, or one of NaNs, which no
        // comparison selects): no interchange, singularity reported; after
        // giving up nothing more is written
        const bool pivoted = !quit && wbest > R(0);
        const int p = pivoted ? pwin : j;
        if (tid == 0 && !quit) {
            sPiv[j] = p;
            if (g == 0) {
                if (!pivoted && info_base[b] == 0) info_base[b] = k0 + j + 1;
                ipiv[k0 + j] = k0 + p;
            }
        }
        // rows j and p trade places: their labels
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            if (lab[q] == p) lab[q] = j;
            else if (lab[q] == j) lab[q] = p;
        }
        if (!pivoted) {
            prow = sZero;
            rcp = C{R(1), R(0)};
        }
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            // rows below the pivot take l and the rank-1 update, the others
            // a zero multiplier (no branch: one version of the registers
            // across the loop); each result goes one slot down, the finished
            // column to the last slot
            // (selects of values: a select of the two operands' addresses
            // would keep v in memory)
            const bool below = lab[q] != INT_MAX && lab[q] > j;
            const C cur = v[q][0];
            const C lq = pivoted ? zmul(cur, rcp) : cur;
            const C l{below ? lq.re : R(0), below ? lq.im : R(0)};
            const C fin{below ? lq.re : cur.re, below ? lq.im : cur.im};
#pragma unroll
            for (int s = 1; s < W; ++s) {
                // v - l p in four fused operations
                const C pv = prow[s], x = v[q][s];
                v[q][s - 1] = C{rfma(-l.re, pv.re, rfma(l.im, pv.im, x.re)),
                                rfma(-l.re, pv.im, rfma(-l.im, pv.re, x.im))};
            }
            v[q][W - 1] = fin;
        }
    }
    if (quit) return;
    // slot s holds column (s + kb) mod W
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        if (lab[q] != INT_MAX)
#pragma unroll
            for (int s = 0; s < W; ++s) {
                const int c = (s + kb) % W;
                if (c < kb) A[(k0 + lab[q]) + (size_t)(k0 + c) * lda] = v[q][s];
            }
    }
    __syncthreads();  // sPiv complete
    // the leaf's interchanges on the panel's other columns [sw_lo, sw_hi),
    // column c to work-group (c - sw_lo) mod P: the rows involved (j and its
    // pivot, for each column j) and, walking the interchanges backwards,
    // the original row each ends with, once in LDS; then every (column,
    // row) pair is loaded, and after a barrier stored
    {
        const int nslot = 2 * kb;
        if (tid < nslot) {
            const int r = (tid & 1) ? sPiv[tid >> 1] : (tid >> 1);
            int x = r;
            for (int s = kb - 1; s >= 0; --s) {
                if (x == s) x = sPiv[s];
                else if (x == sPiv[s]) x = s;
            }
            sDst[tid] = r;
            sSrc[tid] = x;
        }
        __syncthreads();
        const int ncol = sw_hi - sw_lo;
        const int mycols = ncol > g ? (ncol - g + P - 1) / P : 0;
        const int npairs = mycols * nslot;
        constexpr int PPT = 8;  // pairs per thread in flight
        // whole columns per chunk: a column's loads all precede its stores
        const int chunk = nslot ? (MW_NT * PPT / nslot) * nslot : 1;
        for (int p0 = 0; p0 < npairs; p0 += chunk) {
            C val[PPT];
            int64_t dst[PPT];
#pragma unroll
            for (int u = 0; u < PPT; ++u) {
                const int pr = p0 + tid + MW_NT * u;
                val[u] = C{R(0), R(0)};
                dst[u] = -1;
                if (tid + MW_NT * u < chunk && pr < npairs) {
                    const int col = sw_lo + g + (pr / nslot) * P, js = pr % nslot;
                    if (col < k0 || col >= k0 + kb) {
                        const int64_t colo = k0 + (int64_t)col * lda;
                        val[u] = A[colo + sSrc[js]];
                        dst[u] = colo + sDst[js];
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < PPT; ++u)
                if (dst[u] >= 0) A[dst[u]] = val[u];
        }
    }
}

__global__ void zero_words_kernel(uint64_t* __restrict__ p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// 2. Row interchanges ipiv[r0 .. r1) applied to the columns [0, c_lo) and
//    [c_hi, ncols) of a column-major array (LAPACK zlaswp); with `swapped`,
//    only for the matrices whose flag is set.
// ---------------------------------------------------------------------------
template <typename R>
__global__ void laswp_kernel(cx<R>* __restrict__ Abase, int64_t lda, int64_t strideA, int ncols, int c_lo, int c_hi,
                             const int* __restrict__ ipiv_base, int64_t strideP, int r0, int r1,
                             const int* __restrict__ swapped)
{
    // a speculative panel that passed its check interchanged no rows
    if (swapped && !swapped[blockIdx.y]) return;
    __shared__ int sp[256];
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const bool mine = col < ncols && !(col >= c_lo && col < c_hi);
    cx<R>* A = Abase + (size_t)blockIdx.y * strideA + (size_t)col * lda;
    const int* ipiv = ipiv_base + (size_t)blockIdx.y * strideP;
    // the pivots in chunks of 256 through LDS (one load per thread), then
    // every column walks them in order
    for (int i0 = r0; i0 < r1; i0 += 256) {
        const int n = min(256, r1 - i0);
        __syncthreads();
        if ((int)threadIdx.x < n) sp[threadIdx.x] = ipiv[i0 + threadIdx.x];
        __syncthreads();
        if (!mine) continue;
        for (int u = 0; u < n; ++u) {
            const int i = i0 + u, p = sp[u];
            if (p != i) {
                const cx<R> x = A[i], y = A[p];
                A[i] = y;
                A[p] = x;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// 3. Triangular solves with a kb x kb diagonal block T (kb <= PB = 64).
//    stage_triangle / solve_triangle: lane i holds x_i of TRSM_ILP columns
//    at once; x_j is broadcast with v_readlane (j is wave-uniform), T(i, j)
//    comes from LDS, where the triangle is staged once per work-group (loads
//    batched four per thread).  UPPER = false: T unit lower triangular;
//    UPPER = true: T upper, non-unit (the inversion of 3b).  trsm_kernel:
//    L^-1 B in place on the kb rows of B, one wavefront per group of
//    columns, the triangle reused for TRSM_WAVES x TRSM_CPW columns.
// ---------------------------------------------------------------------------
constexpr int TRSM_WAVES = 4, TRSM_CPW = 8, TRSM_ILP = 4;

// The kb x kb triangle of T (column-major, leading dimension ldt) into LDS
// (sT[i + PB j]), loads batched four per thread; for UPPER the diagonal is
// replaced by its reciprocals.  Ends with a barrier.
template <typename R, bool UPPER, int NT>
__device__ __forceinline__ void stage_triangle(const cx<R>* __restrict__ T, int64_t ldt, int kb, cx<R>* sT)
{
    typedef cx<R> C;
    const C zero{R(0), R(0)};
    const int tot = kb * kb;
    for (int e0 = threadIdx.x; e0 < tot; e0 += 4 * NT) {
        C v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = e0 + u * NT;
            const int i = e % kb, j = e / kb;
            v[u] = ld_or_zero(T, i + (size_t)j * ldt, e < tot && (UPPER ? (i <= j) : (i > j)));
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = e0 + u * NT;
            if (e < tot) sT[(e % kb) + PB * (e / kb)] = v[u];
        }
    }
    __syncthreads();
    if (UPPER) {
        for (int i = threadIdx.x; i < kb; i += NT) sT[i + PB * i] = zrcp(sT[i + PB * i]);
        __syncthreads();
    }
}

// x <- T^-1 x for TRSM_ILP columns held by one wavefront (lane i holds row
// i; rows >= kb are zero and stay zero); T from stage_triangle.
template <typename R, bool UPPER>
__device__ __forceinline__ void solve_triangle(const cx<R>* sT, int kb, int lane, cx<R> (&x)[TRSM_ILP])
{
    typedef cx<R> C;
    if (!UPPER) {
        for (int j = 0; j < kb - 1; ++j) {
            const C t = sT[lane + PB * j];
            const bool below = lane > j && lane < kb;
#pragma unroll
            for (int q = 0; q < TRSM_ILP; ++q) {
                const C y{lane_bcast(x[q].re, j), lane_bcast(x[q].im, j)};
                if (below) x[q] = zsub(x[q], zmul(t, y));
            }
        }
    } else {
        for (int j = kb - 1; j >= 0; --j) {
            const C d = sT[j + PB * j];  // reciprocal of the diagonal
#pragma unroll
            for (int q = 0; q < TRSM_ILP; ++q)
                if (lane == j) x[q] = zmul(x[q], d);
            const C t = sT[lane + PB * j];
            const bool above = lane < j;
#pragma unroll
            for (int q = 0; q < TRSM_ILP; ++q) {
                const C y{lane_bcast(x[q].re, j), lane_bcast(x[q].im, j)};
                if (above) x[q] = zsub(x[q], zmul(t, y));
            }
        }
    }
}

template <typename R>
__global__ void __launch_bounds__(64 * TRSM_WAVES) trsm_kernel(const cx<R>* __restrict__ Tbase, int64_t ldt,
                                                               int64_t strideT, int kb, cx<R>* __restrict__ Bbase,
                                                               int64_t ldb, int64_t strideB, int ncols)
{
    typedef cx<R> C;
    __shared__ C sT[PB * PB];
    const C zero{R(0), R(0)};
    stage_triangle<R, false, 64 * TRSM_WAVES>(Tbase + (size_t)blockIdx.y * strideT, ldt, kb, sT);
    const int lane = threadIdx.x & 63;
    const int cbase = (blockIdx.x * TRSM_WAVES + (threadIdx.x >> 6)) * TRSM_CPW;
    C* Bm = Bbase + (size_t)blockIdx.y * strideB;
    for (int c = 0; c < TRSM_CPW; c += TRSM_ILP) {
        const int col0 = cbase + c;
        if (col0 >= ncols) break;
        C x[TRSM_ILP];
#pragma unroll
        for (int q = 0; q < TRSM_ILP; ++q)
            x[q] = ld_or_zero(Bm, lane + (size_t)(col0 + q) * ldb, col0 + q < ncols && lane < kb);
        solve_triangle<R, false>(sT, kb, lane, x);
        if (lane < kb) {
#pragma unroll
            for (int q = 0; q < TRSM_ILP; ++q)
                if (col0 + q < ncols) Bm[lane + (size_t)(col0 + q) * ldb] = x[q];
        }
    }
}

// ---------------------------------------------------------------------------
// 3b. Inverses of the PB x PB diagonal blocks of an LU factor, for the
//     substitution below: inv(L_kk) (unit lower, its strict lower part kept)
//     and inv(U_kk) (upper, diagonal included) share one PB x PB block of D
//     the way L and U share the factor, so D may be the factor itself (dstep
//     = 1: block k at D(k0, k0)) or a strip of PB columns (dstep = 0: block k
//     at D(k0, 0)).  One work-group per (block, matrix); the columns of the
//     identity are solved TRSM_ILP at a time by the wavefront triangle solve.
//     In place, each triangle is staged in LDS before any of its entries is
//     overwritten, and the two triangles do not overlap.
// ---------------------------------------------------------------------------
template <typename R, bool UPPER>
__device__ __forceinline__ void invert_triangle(const cx<R>* sT, int kb, int lane, int wave, cx<R>* D, int64_t ldd)
{
    typedef cx<R> C;
    for (int g = 0; g < PB / (4 * TRSM_ILP); ++g) {
        const int col0 = (4 * g + wave) * TRSM_ILP;
        C x[TRSM_ILP];
#pragma unroll
        for (int q = 0; q < TRSM_ILP; ++q) x[q] = C{R(lane == col0 + q ? 1 : 0), R(0)};
        solve_triangle<R, UPPER>(sT, kb, lane, x);
#pragma unroll
        for (int q = 0; q < TRSM_ILP; ++q) {
            const int col = col0 + q;
            if (col < kb && lane < kb && (UPPER ? lane <= col : lane > col)) D[lane + (size_t)col * ldd] = x[q];
        }
    }
}

template <typename R, bool WITH_U>
__global__ void __launch_bounds__(256) invert_diag_kernel(const cx<R>* __restrict__ LUbase, int64_t lda,
                                                          int64_t strideA, int N, cx<R>* Dbase, int64_t ldd,
                                                          int64_t strideD, int dstep)
{
    typedef cx<R> C;
    __shared__ C sT[PB * PB];
    const int k0 = blockIdx.x * PB, kb = min(PB, N - k0);
    const C* T = LUbase + (size_t)blockIdx.y * strideA + k0 + (size_t)k0 * lda;
    C* D = Dbase + (size_t)blockIdx.y * strideD + k0 + (size_t)(dstep * k0) * ldd;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    stage_triangle<R, false, 256>(T, lda, kb, sT);
    invert_triangle<R, false>(sT, kb, lane, wave, D, ldd);
    if (!WITH_U) return;
    __syncthreads();  // every wavefront is done with L before U replaces it
    stage_triangle<R, true, 256>(T, lda, kb, sT);
    invert_triangle<R, true>(sT, kb, lane, wave, D, ldd);
}

// ---------------------------------------------------------------------------
// 3d. B_j <- inv(L_jj) B_j in place for a wide B (kb <= PB rows, inverse
//     from invert_diag_kernel): work-group of four wavefronts over
//     APPLY_COLS columns of one matrix, 16 columns at a time.  Each group of
//     columns is read from memory once, coalesced (thread t: row t % 64),
//     into an LDS buffer while the previous group is multiplied
//     (double-buffered, one barrier per group); wavefront w keeps the
//     inverse's rows 16 w .. +16 as MFMA operands and writes those rows of
//     the result, in place: a group is already in LDS when it is written.
// ---------------------------------------------------------------------------
constexpr int APPLY_COLS = 128;

template <typename R>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
apply_inverse_kernel(const cx<R>* __restrict__ Dbase, int64_t ldd, int64_t strideD, int kb, cx<R>* Bbase, int64_t ldb,
                     int64_t strideB, int ncols)
{
    typedef cx<R> C;
    typedef typename Mfma<R>::acc acc;
    __shared__ C sB[2][PB][17];  // [buffer][row k][column], padded against bank conflicts
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const C* D = Dbase + (size_t)blockIdx.y * strideD;
    C* B = Bbase + (size_t)blockIdx.y * strideB;
    const int i = 16 * wave + li;
    C a[PB / 4];
#pragma unroll
    for (int s = 0; s < PB / 4; ++s) {
        const int k = 4 * s + lk;
        const bool in = i < kb && k < kb;
        a[s] = ld_or_zero(D, i + (size_t)k * ldd, in && i > k);
        if (in && i == k) a[s] = C{R(1), R(0)};
    }
    const int cbeg = blockIdx.x * APPLY_COLS, cend = min(ncols, cbeg + APPLY_COLS);
    // thread t stages row t % 64 of the columns t / 64 + 4 u, u < 4
    const int sr = threadIdx.x & 63, sc = threadIdx.x >> 6;
    C g[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int col = cbeg + sc + 4 * u;
        g[u] = ld_or_zero(B, sr + (size_t)col * ldb, sr < kb && col < cend);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) sB[0][sr][sc + 4 * u] = g[u];
    __syncthreads();
    int buf = 0;
    for (int c0 = cbeg; c0 < cend; c0 += 16) {
        const bool more = c0 + 16 < cend;
        if (more) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int col = c0 + 16 + sc + 4 * u;
                g[u] = ld_or_zero(B, sr + (size_t)col * ldb, sr < kb && col < cend);
            }
        }
        acc xr{0, 0, 0, 0}, xi{0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < PB / 4; ++s) {
            const C bv = sB[buf][4 * s + lk][li];
            xr = Mfma<R>::op(a[s].re, bv.re, xr);
            xi = Mfma<R>::op(a[s].re, bv.im, xi);
            xr = Mfma<R>::op(-a[s].im, bv.im, xr);
            xi = Mfma<R>::op(a[s].im, bv.re, xi);
        }
        const int col = c0 + li;
        if (col < cend) {
            C* Bc = B + (size_t)col * ldb;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = 16 * wave + Mfma<R>::row(lane, v);
                if (r < kb) Bc[r] = C{xr[v], xi[v]};
            }
        }
        if (more) {
#pragma unroll
            for (int u = 0; u < 4; ++u) sB[buf ^ 1][sr][sc + 4 * u] = g[u];
        }
        __syncthreads();  // the next group is staged; this buffer is free again
        buf ^= 1;
    }
}

// ---------------------------------------------------------------------------
// 3c. Substitution of the solve (getrs) for a few right-hand sides, one
//     launch per diagonal block: every work-group forms x_k = inv(T_kk) b_k
//     for its right-hand sides (SUB_RC at a time) with MFMA -- wavefront w
//     the rows 16 w .. +16 -- and, with UPDATE, subtracts T(rows, k) x_k
//     from its slab of sub_slab rows of [row_lo, row_hi) on MFMA: the rows
//     below the block for L, above it for U.  The block's own rows of B are
//     not written by an UPDATE launch, so every work-group reads the same
//     b_k; the last launch (UPDATE = false, block = blockIdx.x) writes x_k
//     for every block in place.
// ---------------------------------------------------------------------------
constexpr int SUB_RC = 16;
// 16-row tiles per wavefront of an UPDATE launch (all of their operands in
// registers at once) and rows per work-group
template <typename R>
constexpr int sub_tiles() { return sizeof(R) == 8 ? 1 : 2; }
template <typename R>
constexpr int sub_slab() { return 64 * sub_tiles<R>(); }

// x_k(rows 16 wave .., cols c0 + (lane & 15)) = inv(T_kk) b_k, accumulator
// layout (row Mfma<R>::row(lane, v) of the tile); inverse entries outside
// the triangle read as zero, the unit diagonal of inv(L) as one.
template <typename R, bool UPPER>
__device__ __forceinline__ void diag_product(const cx<R>* Dk, int64_t ldd, const cx<R>* Bk, int64_t ldb, int kb,
                                             int c0, int nrhs, int wave, int lane, typename Mfma<R>::acc& xr,
                                             typename Mfma<R>::acc& xi)
{
    typedef cx<R> C;
    const int li = lane & 15, lk = lane >> 4;
    const int i = 16 * wave + li, col = c0 + li;
    xr = typename Mfma<R>::acc{0, 0, 0, 0};
    xi = typename Mfma<R>::acc{0, 0, 0, 0};
    // all loads of a batch in flight at once: the whole block for complex64,
    // half of it for complex128 (register budget)
    constexpr int DS = sizeof(R) == 8 ? PB / 8 : PB / 4;
#pragma unroll 1
    for (int s0 = 0; s0 < PB / 4; s0 += DS) {
        C a[DS], b[DS];
#pragma unroll
        for (int s = 0; s < DS; ++s) {
            const int k = 4 * (s0 + s) + lk;
            const bool in = i < kb && k < kb;
            a[s] = ld_or_zero(Dk, i + (size_t)k * ldd, in && (UPPER ? i <= k : i > k));
            if (!UPPER && in && i == k) a[s] = C{R(1), R(0)};
            b[s] = ld_or_zero(Bk, k + (size_t)col * ldb, k < kb && col < nrhs);
        }
#pragma unroll
        for (int s = 0; s < DS; ++s) {
            xr = Mfma<R>::op(a[s].re, b[s].re, xr);
            xi = Mfma<R>::op(a[s].re, b[s].im, xi);
            xr = Mfma<R>::op(-a[s].im, b[s].im, xr);
            xi = Mfma<R>::op(a[s].im, b[s].re, xi);
        }
    }
}

template <typename R, bool UPPER, bool UPDATE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
substitution_kernel(const cx<R>* __restrict__ LUbase, int64_t lda, int64_t strideA, int N, int kblk,
                    const cx<R>* __restrict__ Dbase, int64_t ldd, int64_t strideD, int dstep, cx<R>* __restrict__ Bbase,
                    int64_t ldb, int64_t strideB, int nrhs, int row_lo, int row_hi)
{
    typedef cx<R> C;
    typedef typename Mfma<R>::acc acc;
    __shared__ C sX[PB][SUB_RC];  // x_k(k, col)
    const int k0 = (UPDATE ? kblk : (int)blockIdx.x) * PB;
    const int kb = min(PB, N - k0);
    const C* LU = LUbase + (size_t)blockIdx.y * strideA;
    const C* Dk = Dbase + (size_t)blockIdx.y * strideD + k0 + (size_t)(dstep * k0) * ldd;
    C* B = Bbase + (size_t)blockIdx.y * strideB;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    // UPDATE: wavefront w takes the rows wrow .. +16 TW of the slab as TW
    // 16 x 16 tiles; all of its T(rows, k) operands (lane l: row
    // 16 t + (l & 15), k 4 s + (l >> 4)) are loaded before x_k is formed, so
    // that their latency overlaps that of the diagonal product
    constexpr int TW = sub_tiles<R>();
    const int wrow = row_lo + blockIdx.x * sub_slab<R>() + 16 * TW * wave;
    const C* Tk = LU + (size_t)k0 * lda;
    C ta[UPDATE ? PB / 4 : 1][TW];
    if (UPDATE) {
#pragma unroll
        for (int s = 0; s < PB / 4; ++s)
#pragma unroll
            for (int t = 0; t < TW; ++t) {
                const int k = 4 * s + lk, row = wrow + 16 * t + li;
                ta[s][t] = ld_or_zero(Tk, row + (size_t)k * lda, k < kb && row < row_hi);
            }
    }
    for (int c0 = 0; c0 < nrhs; c0 += SUB_RC) {
        acc xr, xi;
        diag_product<R, UPPER>(Dk, ldd, B + k0, ldb, kb, c0, nrhs, wave, lane, xr, xi);
        if (!UPDATE) {
            // every wavefront has consumed b_k (its MFMAs read it) before any
            // row of x_k replaces it
            __syncthreads();
            if (c0 + li < nrhs) {
                C* Bc = B + k0 + (size_t)(c0 + li) * ldb;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int r = 16 * wave + Mfma<R>::row(lane, v);
                    if (r < kb) Bc[r] = C{xr[v], xi[v]};
                }
            }
            continue;
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) sX[16 * wave + Mfma<R>::row(lane, v)][li] = C{xr[v], xi[v]};
        __syncthreads();
        // B(rows, cols) -= T(rows, k) x_k on MFMA (x_k(k 4 s + (l >> 4),
        // col l & 15) from LDS)
        acc cr[TW], ci[TW];
#pragma unroll
        for (int t = 0; t < TW; ++t) {
            cr[t] = acc{0, 0, 0, 0};
            ci[t] = acc{0, 0, 0, 0};
        }
#pragma unroll
        for (int s = 0; s < PB / 4; ++s) {
            const int k = 4 * s + lk;
            const C xb = ld_or_zero(&sX[0][0], (size_t)k * SUB_RC + li, k < kb);
#pragma unroll
            for (int t = 0; t < TW; ++t) {
                const C a = ta[s][t];
                cr[t] = Mfma<R>::op(a.re, xb.re, cr[t]);
                ci[t] = Mfma<R>::op(a.re, xb.im, ci[t]);
                cr[t] = Mfma<R>::op(-a.im, xb.im, cr[t]);
                ci[t] = Mfma<R>::op(a.im, xb.re, ci[t]);
            }
        }
        if (c0 + li < nrhs) {
            C* Bc = B + (size_t)(c0 + li) * ldb;
#pragma unroll
            for (int t = 0; t < TW; ++t)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = wrow + 16 * t + Mfma<R>::row(lane, v);
                    if (row < row_hi) {
                        Bc[row].re -= cr[t][v];
                        Bc[row].im -= ci[t][v];
                    }
                }
        }
        __syncthreads();  // sX is rewritten by the next group of columns
    }
}

template <typename R>
int launch_trsm(const cx<R>* T, int64_t ldt, int64_t sT, int kb, cx<R>* B, int64_t ldb, int64_t sB, int ncols,
                int batch, hipStream_t stream)
{
    constexpr int per_block = TRSM_WAVES * TRSM_CPW;
    hipLaunchKernelGGL((trsm_kernel<R>), dim3((ncols + per_block - 1) / per_block, batch),
                       dim3(64 * TRSM_WAVES), 0, stream, T, ldt, sT, kb, B, ldb, sB, ncols);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// 4. C (M x N) = alpha * A (M x K) * B (K x N) + (beta_zero ? 0 : C), complex,
//    column-major, batched by blockIdx.z.  Work-group tile 64 x 64, four
//    wavefronts of 32 x 32, each a 2 x 2 grid of 16 x 16 MFMA tiles, operand
//    fragments read straight from global memory; the complex product is four
//    real MFMAs per tile pair (re*re - im*im, re*im + im*re).  Used for small
//    M (forces, diagonal blocks).
// ---------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(256) gemm_small_kernel(int M, int N, int K, R alpha, const cx<R>* __restrict__ Abase,
                                                         int64_t lda, int64_t strideA, const cx<R>* __restrict__ Bbase,
                                                         int64_t ldb, int64_t strideB, cx<R>* __restrict__ Cbase,
                                                         int64_t ldc, int64_t strideC, int beta_zero)
{
    typedef typename Mfma<R>::acc acc;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row0 = blockIdx.x * 64 + (wave & 1) * 32;
    const int col0 = blockIdx.y * 64 + (wave >> 1) * 32;
    const cx<R>* A = Abase + (size_t)blockIdx.z * strideA;
    const cx<R>* B = Bbase + (size_t)blockIdx.z * strideB;
    cx<R>* C = Cbase + (size_t)blockIdx.z * strideC;
    const int li = lane & 15, lk = lane >> 4;
    acc cr[2][2], ci[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            cr[a][b] = acc{0, 0, 0, 0};
            ci[a][b] = acc{0, 0, 0, 0};
        }
    for (int k = 0; k < K; k += 4) {
        const int kk = k + lk;
        R ar[2], ai[2], br[2], bi[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int row = row0 + 16 * t + li, col = col0 + 16 * t + li;
            const cx<R> a = (row < M && kk < K) ? A[row + (size_t)kk * lda] : cx<R>{R(0), R(0)};
            const cx<R> b = (col < N && kk < K) ? B[kk + (size_t)col * ldb] : cx<R>{R(0), R(0)};
            ar[t] = a.re;
            ai[t] = a.im;
            br[t] = b.re;
            bi[t] = b.im;
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                cr[a][b] = Mfma<R>::op(ar[a], br[b], cr[a][b]);
                cr[a][b] = Mfma<R>::op(-ai[a], bi[b], cr[a][b]);
                ci[a][b] = Mfma<R>::op(ar[a], bi[b], ci[a][b]);
                ci[a][b] = Mfma<R>::op(ai[a], br[b], ci[a][b]);
            }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = row0 + 16 * a + Mfma<R>::row(lane, v);
                const int col = col0 + 16 * b + li;
                if (row < M && col < N) {
                    cx<R>* c = C + row + (size_t)col * ldc;
                    R re = beta_zero ? R(0) : c->re;
                    R im = beta_zero ? R(0) : c->im;
                    re += alpha * cr[a][b][v];
                    im += alpha * ci[a][b][v];
                    c->re = re;
                    c->im = im;
                }
            }
}

// ---------------------------------------------------------------------------
// 4c. Skinny product for few right-hand sides: C (M x N) = alpha A B +
//     (beta_zero ? 0 : C) with N <= 2 SKN.  Work-group of four wavefronts on
//     a 64-row x 16-column tile of C; the wavefronts split K in four, each
//     streams its quarter of the 64-row slab of A (the next 16 columns in
//     flight while the current ones feed the MFMAs), and the four partial
//     tiles are summed through LDS.  The stream of A from HBM bounds it.
// ---------------------------------------------------------------------------
constexpr int SKM = 64, SKN = 16, SKK = 16;

// C = alpha A B (+ C) for a product with few rows (M <= 16) and a long K
// (the generalised forces W^T Phi: 6 dof x nrhs over the panels): one
// work-group per (column, matrix), each thread summing every row over
// k = tid, tid + 256, ..., then a reduction across the work-group.
constexpr int FEW_ROWS = 16;

template <typename R>
__global__ void __launch_bounds__(256) gemm_fewrows_kernel(int M, int K, R alpha, const cx<R>* __restrict__ Abase,
                                                           int64_t lda, int64_t strideA, const cx<R>* __restrict__ Bbase,
                                                           int64_t ldb, int64_t strideB, cx<R>* __restrict__ Cbase,
                                                           int64_t ldc, int64_t strideC, int beta_zero)
{
    typedef cx<R> C;
    __shared__ C red[4][FEW_ROWS];
    const int col = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const C* A = Abase + (size_t)blockIdx.y * strideA;
    const C* B = Bbase + (size_t)blockIdx.y * strideB + (size_t)col * ldb;
    C* Cc = Cbase + (size_t)blockIdx.y * strideC + (size_t)col * ldc;
    C acc[FEW_ROWS];
#pragma unroll
    for (int i = 0; i < FEW_ROWS; ++i) acc[i] = C{R(0), R(0)};
    for (int k = tid; k < K; k += 256) {
        const C x = B[k];
#pragma unroll
        for (int i = 0; i < FEW_ROWS; ++i)
            if (i < M) {
                const C a = A[i + (size_t)k * lda];
                acc[i].re += a.re * x.re - a.im * x.im;
                acc[i].im += a.re * x.im + a.im * x.re;
            }
    }
#pragma unroll
    for (int i = 0; i < FEW_ROWS; ++i) {
        R re = acc[i].re, im = acc[i].im;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            re += __shfl_xor(re, off);
            im += __shfl_xor(im, off);
        }
        if (lane == 0) red[wave][i] = C{re, im};
    }
    __syncthreads();
    if (tid < M) {
        R re = R(0), im = R(0);
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            re += red[w][tid].re;
            im += red[w][tid].im;
        }
        C* c = Cc + tid;
        c->re = (beta_zero ? R(0) : c->re) + alpha * re;
        c->im = (beta_zero ? R(0) : c->im) + alpha * im;
    }
}

template <typename R>
__global__ void __launch_bounds__(256) gemm_skinny_kernel(int M, int N, int K, R alpha,
                                                          const cx<R>* __restrict__ Abase, int64_t lda,
                                                          int64_t strideA, const cx<R>* __restrict__ Bbase,
                                                          int64_t ldb, int64_t strideB, cx<R>* __restrict__ Cbase,
                                                          int64_t ldc, int64_t strideC, int beta_zero)
{
    typedef cx<R> Cx;
    typedef typename Mfma<R>::acc acc;
    __shared__ Cx sC[4][SKM * SKN];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int row0 = blockIdx.x * SKM, col0 = blockIdx.y * SKN;
    const Cx* A = Abase + (size_t)blockIdx.z * strideA;
    const Cx* B = Bbase + (size_t)blockIdx.z * strideB;
    Cx* C = Cbase + (size_t)blockIdx.z * strideC;
    const int nchunks = (K + SKK - 1) / SKK;
    const int per = (nchunks + 3) / 4;
    const int kbeg = min(K, wave * per * SKK);
    const int kend = min(K, kbeg + per * SKK);
    const int bcol = col0 + li;
    const bool b_ok = bcol < N;
    acc cr[4], ci[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        cr[a] = acc{0, 0, 0, 0};
        ci[a] = acc{0, 0, 0, 0};
    }
    Cx ra[4][4], rb[4];
#define CPT_SKINNY_LOAD(kc_)                                                                  \
    do {                                                                                      \
        _Pragma("unroll") for (int ks = 0; ks < 4; ++ks)                                      \
        {                                                                                     \
            const int k = (kc_) + 4 * ks + lk;                                                \
            _Pragma("unroll") for (int a = 0; a < 4; ++a)                                     \
            {                                                                                 \
                const int row = row0 + 16 * a + li;                                           \
                if (k < kend && row < M) ra[ks][a] = A[row + (size_t)k * lda];                \
                else ra[ks][a].re = ra[ks][a].im = R(0);                                      \
            }                                                                                 \
            if (k < kend && b_ok) rb[ks] = B[k + (size_t)bcol * ldb];                         \
            else rb[ks].re = rb[ks].im = R(0);                                                \
        }                                                                                     \
    } while (0)

    if (kbeg < kend) CPT_SKINNY_LOAD(kbeg);
    for (int kc = kbeg; kc < kend; kc += SKK) {
        Cx ca[4][4], cb[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            cb[ks] = rb[ks];
#pragma unroll
            for (int a = 0; a < 4; ++a) ca[ks][a] = ra[ks][a];
        }
        if (kc + SKK < kend) CPT_SKINNY_LOAD(kc + SKK);
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                cr[a] = Mfma<R>::op(ca[ks][a].re, cb[ks].re, cr[a]);
                ci[a] = Mfma<R>::op(ca[ks][a].re, cb[ks].im, ci[a]);
                cr[a] = Mfma<R>::op(-ca[ks][a].im, cb[ks].im, cr[a]);
                ci[a] = Mfma<R>::op(ca[ks][a].im, cb[ks].re, ci[a]);
            }
    }
#undef CPT_SKINNY_LOAD
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = 16 * a + Mfma<R>::row(lane, v);
            sC[wave][r + SKM * li] = Cx{cr[a][v], ci[a][v]};
        }
    __syncthreads();
    for (int e = threadIdx.x; e < SKM * SKN; e += 256) {
        const int r = e % SKM, c = e / SKM;
        const int row = row0 + r, col = col0 + c;
        if (row < M && col < N) {
            R re = (sC[0][e].re + sC[1][e].re) + (sC[2][e].re + sC[3][e].re);
            R im = (sC[0][e].im + sC[1][e].im) + (sC[2][e].im + sC[3][e].im);
            Cx* p = C + row + (size_t)col * ldc;
            const R c_re = beta_zero ? R(0) : p->re;
            const R c_im = beta_zero ? R(0) : p->im;
            p->re = c_re + alpha * re;
            p->im = c_im + alpha * im;
        }
    }
}

// ---------------------------------------------------------------------------
// 4b. Same product, LDS-tiled for large M: work-group tile 128 x TBN (TBN =
//     64, or 128 for complex64), K staged in chunks of 16 through LDS (A as [k][row], B as
//     [col][k] padded to 17 so that both the stores and the MFMA fragment
//     reads are free of bank conflicts), the next chunk prefetched into
//     registers while the current one feeds the MFMAs.  Wavefront w computes
//     rows 64 (w & 1) .. +64 and columns (TBN/2) (w >> 1) .. +TBN/2: a
//     4 x TBN/32 grid of 16 x 16 tiles.  The wider tile doubles the MFMA work
//     per barrier and per A fragment (used for complex64, whose MFMA takes
//     half the cycles of complex128's).
// ---------------------------------------------------------------------------
constexpr int TBM = 128, TKC = 16;

// GAUSS: three real products per complex one, P1 = Ar Br, P2 = Ai Bi,
// P3 = (Ar + Ai)(Br + Bi), Re = P1 - P2, Im = P3 - P1 - P2 (6 M N K real
// MFMA flops for the 8 M N K of the product; one more accumulator set).
template <typename R, int TBN, bool GAUSS>
__device__ __forceinline__ void gemm_tiled_body(int M, int N, int K, R alpha, const cx<R>* __restrict__ Abase,
                                                int64_t lda, int64_t strideA, const cx<R>* __restrict__ Bbase,
                                                int64_t ldb, int64_t strideB, cx<R>* __restrict__ Cbase, int64_t ldc,
                                                int64_t strideC, int beta_zero)
{
    typedef cx<R> Cx;
    typedef typename Mfma<R>::acc acc;
    constexpr int FB = TBN / 32;        // 16-column tiles per wavefront
    constexpr int NB_LOAD = TBN / 16;   // B values per thread per chunk
    __shared__ Cx sA[TKC][TBM];
    __shared__ Cx sB[TBN][TKC + 1];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int wm = wave & 1, wn = wave >> 1;
    const int row0 = blockIdx.x * TBM, col0 = blockIdx.y * TBN;
    const Cx* A = Abase + (size_t)blockIdx.z * strideA;
    const Cx* B = Bbase + (size_t)blockIdx.z * strideB;
    Cx* C = Cbase + (size_t)blockIdx.z * strideC;

    Cx ra[8], rb[NB_LOAD];
    const int a_row = row0 + (tid & 127);
    const bool a_ok = a_row < M;
    acc cr[4][FB], ci[4][FB], cs[GAUSS ? 4 : 1][GAUSS ? FB : 1];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < FB; ++b) {
            cr[a][b] = acc{0, 0, 0, 0};
            ci[a][b] = acc{0, 0, 0, 0};
            if (GAUSS) cs[GAUSS ? a : 0][GAUSS ? b : 0] = acc{0, 0, 0, 0};
        }

#define CPT_LOAD_CHUNK(kc_)                                                                   \
    do {                                                                                      \
        _Pragma("unroll") for (int i = 0; i < 8; ++i)                                         \
        {                                                                                     \
            const int k = (kc_) + (tid >> 7) + 2 * i;                                         \
            if (a_ok && k < K) ra[i] = A[a_row + (size_t)k * lda];                            \
            else ra[i].re = ra[i].im = R(0);                                                  \
        }                                                                                     \
        _Pragma("unroll") for (int i = 0; i < NB_LOAD; ++i)                                   \
        {                                                                                     \
            const int idx = tid + 256 * i;                                                    \
            const int col = col0 + (idx >> 4), k = (kc_) + (idx & 15);                        \
            if (col < N && k < K) rb[i] = B[k + (size_t)col * ldb];                           \
            else rb[i].re = rb[i].im = R(0);                                                  \
        }                                                                                     \
    } while (0)

    CPT_LOAD_CHUNK(0);
    for (int kc = 0; kc < K; kc += TKC) {
#pragma unroll
        for (int i = 0; i < 8; ++i) sA[(tid >> 7) + 2 * i][tid & 127] = ra[i];
#pragma unroll
        for (int i = 0; i < NB_LOAD; ++i) {
            const int idx = tid + 256 * i;
            sB[idx >> 4][idx & 15] = rb[i];
        }
        __syncthreads();
        if (kc + TKC < K) CPT_LOAD_CHUNK(kc + TKC);
#pragma unroll
        for (int ks = 0; ks < TKC; ks += 4) {
            R ar[4], ai[4], br[FB], bi[FB];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const Cx v = sA[ks + lk][wm * 64 + a * 16 + li];
                ar[a] = v.re;
                ai[a] = v.im;
            }
#pragma unroll
            for (int b = 0; b < FB; ++b) {
                const Cx v = sB[wn * (TBN / 2) + b * 16 + li][ks + lk];
                br[b] = v.re;
                bi[b] = v.im;
            }
            if (GAUSS) {
                // cr: P1, ci: P2, cs: P3
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < FB; ++b) {
                        cr[a][b] = Mfma<R>::op(ar[a], br[b], cr[a][b]);
                        ci[a][b] = Mfma<R>::op(ai[a], bi[b], ci[a][b]);
                        cs[GAUSS ? a : 0][GAUSS ? b : 0] =
                            Mfma<R>::op(ar[a] + ai[a], br[b] + bi[b], cs[GAUSS ? a : 0][GAUSS ? b : 0]);
                    }
            } else {
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < FB; ++b) {
                        cr[a][b] = Mfma<R>::op(ar[a], br[b], cr[a][b]);
                        ci[a][b] = Mfma<R>::op(ar[a], bi[b], ci[a][b]);
                    }
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < FB; ++b) {
                        cr[a][b] = Mfma<R>::op(-ai[a], bi[b], cr[a][b]);
                        ci[a][b] = Mfma<R>::op(ai[a], br[b], ci[a][b]);
                    }
            }
        }
        __syncthreads();
    }
#undef CPT_LOAD_CHUNK
    if (GAUSS) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < FB; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const R p1 = cr[a][b][v], p2 = ci[a][b][v];
                    cr[a][b][v] = p1 - p2;
                    ci[a][b][v] = cs[GAUSS ? a : 0][GAUSS ? b : 0][v] - p1 - p2;
                }
    }
    if constexpr (sizeof(R) == 4) {
        // complex64: a lane holds rows 4 lk .. 4 lk + 3 of one column of a
        // 16 x 16 tile, 32 contiguous bytes of C, as two 16-byte vectors
        // (8-byte alignment asked); tiles cut by the last row one row at a time
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < FB; ++b) {
                const int rowb = row0 + wm * 64 + 16 * a + 4 * lk;
                const int col = col0 + wn * (TBN / 2) + 16 * b + li;
                if (col >= N || rowb >= M) continue;
                Cx* c = C + rowb + (size_t)col * ldc;
                if (rowb + 3 < M) {
                    fvec4u* cv = reinterpret_cast<fvec4u*>(c);
                    fvec4u lo = {alpha * cr[a][b][0], alpha * ci[a][b][0], alpha * cr[a][b][1], alpha * ci[a][b][1]};
                    fvec4u hi = {alpha * cr[a][b][2], alpha * ci[a][b][2], alpha * cr[a][b][3], alpha * ci[a][b][3]};
                    if (!beta_zero) {
                        lo += cv[0];
                        hi += cv[1];
                    }
                    cv[0] = lo;
                    cv[1] = hi;
                } else {
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        if (rowb + v < M) {
                            c[v].re = (beta_zero ? R(0) : c[v].re) + alpha * cr[a][b][v];
                            c[v].im = (beta_zero ? R(0) : c[v].im) + alpha * ci[a][b][v];
                        }
                }
            }
    } else {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < FB; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = row0 + wm * 64 + 16 * a + Mfma<R>::row(lane, v);
                    const int col = col0 + wn * (TBN / 2) + 16 * b + li;
                    if (row < M && col < N) {
                        Cx* c = C + row + (size_t)col * ldc;
                        R re = beta_zero ? R(0) : c->re;
                        R im = beta_zero ? R(0) : c->im;
                        re += alpha * cr[a][b][v];
                        im += alpha * ci[a][b][v];
                        c->re = re;
                        c->im = im;
                    }
                }
    }
}

template <typename R, int TBN, bool GAUSS = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
gemm_tiled_kernel(int M, int N, int K, R alpha, const cx<R>* __restrict__ Abase, int64_t lda, int64_t strideA,
                  const cx<R>* __restrict__ Bbase, int64_t ldb, int64_t strideB, cx<R>* __restrict__ Cbase,
                  int64_t ldc, int64_t strideC, int beta_zero)
{
    gemm_tiled_body<R, TBN, GAUSS>(M, N, K, alpha, Abase, lda, strideA, Bbase, ldb, strideB, Cbase, ldc, strideC,
                                   beta_zero);
}

// The Gauss form with one wavefront per SIMD: its three accumulator sets fit
// the registers of a lane without spilling where two wavefronts would not.
template <typename R, int TBN>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1)))
gemm_tiled_gauss_w1_kernel(int M, int N, int K, R alpha, const cx<R>* __restrict__ Abase, int64_t lda,
                           int64_t strideA, const cx<R>* __restrict__ Bbase, int64_t ldb, int64_t strideB,
                           cx<R>* __restrict__ Cbase, int64_t ldc, int64_t strideC, int beta_zero)
{
    gemm_tiled_body<R, TBN, true>(M, N, K, alpha, Abase, lda, strideA, Bbase, ldb, strideB, Cbase, ldc, strideC,
                                  beta_zero);
}


// Complex64 tile product with FP16 inputs and FP32 accumulation, four-
// product form (Re += Ar Br - Ai Bi, Im += Ar Bi + Ai Br; -Ai by a sign flip
// in registers), 128 x 64 tiles, k in chunks of KC = 16.  A and B are
// converted once per work-group as they are staged: two FP16 planes (Re,
// Im) per operand, k contiguous, so a lane reads the KC / 4 k values of an
// MFMA operand (row li of A or column li of B, k = KC/4 lk and on) with one
// LDS read.  A and B fill the instruction's k slots the
// same way, so the product sums over every k of the chunk.  Staging: a
// thread loads KC / 2 consecutive k of one row of A (lanes on consecutive
// rows: coalesced) and KC / 4 consecutive k of one column of B.
template <int KC>
struct H16Op;
template <>
struct H16Op<16> {
    typedef hvec4 v;
    static constexpr int KP = 20;  // k stride of a plane row in halfs: 40 bytes, 8-byte aligned
    static __device__ __forceinline__ fvec4 mma(v a, v b, fvec4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
    }
};
// (a 32-deep chunk on v_mfma_f32_16x16x32_f16 needs 208 VGPRs and 46 KB of
// LDS, two wavefronts per SIMD instead of three, and measured 3 % slower)

template <int KC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
gemm_h16_kernel(int M, int N, int K, float alpha, const cx<float>* __restrict__ Abase, int64_t lda, int64_t strideA,
                const cx<float>* __restrict__ Bbase, int64_t ldb, int64_t strideB, cx<float>* __restrict__ Cbase,
                int64_t ldc, int64_t strideC, int beta_zero)
{
    typedef cx<float> Cx;
    typedef typename H16Op<KC>::v hv;
    constexpr int TBN = 64, FB = TBN / 32, KP = H16Op<KC>::KP;
    constexpr int AK = KC / 2, BK = KC / 4, LK = KC / 4;  // k per thread of A, of B; k per lane of an operand
    static_assert(TBM == 128 && KP >= KC, "staging maps 256 threads onto 128 x KC of A and KC x 64 of B");
    __shared__ __attribute__((aligned(16))) _Float16 sA[2][TBM * KP];
    __shared__ __attribute__((aligned(16))) _Float16 sB[2][TBN * KP];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int wm = wave & 1, wn = wave >> 1;
    const int row0 = blockIdx.x * TBM, col0 = blockIdx.y * TBN;
    const Cx* A = Abase + (size_t)blockIdx.z * strideA;
    const Cx* B = Bbase + (size_t)blockIdx.z * strideB;
    Cx* C = Cbase + (size_t)blockIdx.z * strideC;

    // staging roles: A row ar_ (k = AK akh and on); B column bc_ (k = BK bkq and on)
    const int ar_ = tid & 127, akh = tid >> 7;
    const int bc_ = tid >> 2, bkq = tid & 3;
    const bool a_ok = row0 + ar_ < M, b_ok = col0 + bc_ < N;
    const Cx* Arow = A + (row0 + ar_);
    const Cx* Bcol = B + (size_t)(col0 + bc_) * ldb;
    Cx ra[AK], rb[BK];

    fvec4 cr[4][FB], ci[4][FB];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < FB; ++b) {
            cr[a][b] = fvec4{0, 0, 0, 0};
            ci[a][b] = fvec4{0, 0, 0, 0};
        }

    auto load = [&](int kc) {
#pragma unroll
        for (int i = 0; i < AK; ++i) {
            const int k = kc + AK * akh + i;
            if (a_ok && k < K) ra[i] = Arow[(size_t)k * lda];
            else ra[i].re = ra[i].im = 0.0f;
        }
#pragma unroll
        for (int j = 0; j < BK; ++j) {
            const int k = kc + BK * bkq + j;
            if (b_ok && k < K) rb[j] = Bcol[k];
            else rb[j].re = rb[j].im = 0.0f;
        }
    };

    load(0);
    for (int kc = 0; kc < K; kc += KC) {
#pragma unroll
        for (int q = 0; q < AK / LK; ++q) {
            hv hr, hi;
#pragma unroll
            for (int j = 0; j < LK; ++j) {
                const Cx v = ra[LK * q + j];
                hr[j] = (_Float16)v.re;
                hi[j] = (_Float16)v.im;
            }
            const int off = ar_ * KP + AK * akh + LK * q;
            *reinterpret_cast<hv*>(&sA[0][off]) = hr;
            *reinterpret_cast<hv*>(&sA[1][off]) = hi;
        }
#pragma unroll
        for (int q = 0; q < BK / LK; ++q) {
            hv hr, hi;
#pragma unroll
            for (int j = 0; j < LK; ++j) {
                const Cx v = rb[LK * q + j];
                hr[j] = (_Float16)v.re;
                hi[j] = (_Float16)v.im;
            }
            const int off = bc_ * KP + BK * bkq + LK * q;
            *reinterpret_cast<hv*>(&sB[0][off]) = hr;
            *reinterpret_cast<hv*>(&sB[1][off]) = hi;
        }
        __syncthreads();
        if (kc + KC < K) load(kc + KC);
        hv yr[FB], yi[FB];
#pragma unroll
        for (int b = 0; b < FB; ++b) {
            const int off = (wn * (TBN / 2) + b * 16 + li) * KP + LK * lk;
            yr[b] = *reinterpret_cast<const hv*>(&sB[0][off]);
            yi[b] = *reinterpret_cast<const hv*>(&sB[1][off]);
        }
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int off = (wm * 64 + a * 16 + li) * KP + LK * lk;
            const hv xr = *reinterpret_cast<const hv*>(&sA[0][off]);
            const hv xi = *reinterpret_cast<const hv*>(&sA[1][off]);
            const hv xn = -xi;
#pragma unroll
            for (int b = 0; b < FB; ++b) {
                cr[a][b] = H16Op<KC>::mma(xr, yr[b], cr[a][b]);
                ci[a][b] = H16Op<KC>::mma(xr, yi[b], ci[a][b]);
                cr[a][b] = H16Op<KC>::mma(xn, yi[b], cr[a][b]);
                ci[a][b] = H16Op<KC>::mma(xi, yr[b], ci[a][b]);
            }
        }
        __syncthreads();
    }
    // a lane holds rows 4 lk .. 4 lk + 3 of one column of each 16 x 16 tile
    // (Mfma<float>::row): 32 contiguous bytes of C, read and written as two
    // 16-byte vectors (8-byte alignment asked: ldc and offsets may be odd);
    // tiles cut by the last row take one row at a time
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < FB; ++b) {
            const int rowb = row0 + wm * 64 + 16 * a + 4 * lk;
            const int col = col0 + wn * (TBN / 2) + 16 * b + li;
            if (col >= N || rowb >= M) continue;
            float re[4], im[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                re[v] = alpha * cr[a][b][v];
                im[v] = alpha * ci[a][b][v];
            }
            Cx* c = C + rowb + (size_t)col * ldc;
            if (rowb + 3 < M) {
                fvec4u* cv = reinterpret_cast<fvec4u*>(c);
                fvec4u lo = {re[0], im[0], re[1], im[1]}, hi = {re[2], im[2], re[3], im[3]};
                if (!beta_zero) {
                    lo += cv[0];
                    hi += cv[1];
                }
                cv[0] = lo;
                cv[1] = hi;
            } else {
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    if (rowb + v < M) {
                        c[v].re = (beta_zero ? 0.0f : c[v].re) + re[v];
                        c[v].im = (beta_zero ? 0.0f : c[v].im) + im[v];
                    }
            }
        }
}

// ---------------------------------------------------------------------------
// 1c. Speculative panel: the factorisation with the pivots it would take
//     known in advance, checked against LAPACK's pivot rule.
//
// On the BEM influence matrices (second-kind operators, K = I/2 + a compact
// part) partial pivoting interchanges no rows: at every column j of the
// elimination the diagonal entry has the largest |Re|+|Im| of the column
// (on the configured bodies the largest entry below it stays under 0.35 of
// it).  A panel is then factored with the diagonal as pivots: no pivot
// search, and the rows below each leaf's leading w x w block are
// independent (each solves l U11 = a by itself).
// Every entry a_ij^(j) the elimination meets below a pivot u_jj is compared
// with it exactly as the pivot search compares candidates: cabs1(a_ij^(j)) >
// cabs1(u_jj), or a zero or non-number u_jj, means the search would have
// taken another row.  A matrix whose panel fails the check has the panel
// restored from a copy taken before the panel step and factored again with
// partial pivoting by the repair kernel (LAPACK's rule column by column,
// ties to the first row, zero columns reported in info), so the factor is
// LAPACK's P A = L U in both cases, up to rounding.
// ---------------------------------------------------------------------------
constexpr int SPEC_KB_MAX = 16 * PB;  // widest panel (block width) the copy holds

// columns of the panel copy per matrix
inline int spec_bak_cols(int N) { return N < SPEC_KB_MAX ? N : SPEC_KB_MAX; }

// Leaf width of the speculative panels: 32 columns, so that a thread holds
// one row of a leaf in registers (32 complex values) with straight-line code
// that stays in the instruction cache.
template <typename R>
constexpr int spec_width()
{
    return 32;
}

// copy of the panel columns [k0, k0 + kb), rows [k0, N), leading dimension
// N - k0, per matrix (four rows per thread, 1024 per work-group and
// column); clears the matrix's interchange flag of this panel
template <typename R>
__global__ void __launch_bounds__(256) spec_backup_kernel(const cx<R>* __restrict__ Abase, int64_t lda,
                                                          int64_t strideA, int N, int k0, int kb,
                                                          cx<R>* __restrict__ bak, int64_t strideBak,
                                                          int* __restrict__ swapped)
{
    const int M = N - k0;
    const int j = blockIdx.y;
    const size_t b = blockIdx.z;
    if (blockIdx.x == 0 && j == 0 && threadIdx.x == 0) swapped[b] = 0;
    const cx<R>* src = Abase + b * strideA + k0 + (size_t)(k0 + j) * lda;
    cx<R>* dst = bak + b * strideBak + (size_t)j * M;
    cx<R> v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = blockIdx.x * 1024 + q * 256 + threadIdx.x;
        if (i < M) v[q] = src[i];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int i = blockIdx.x * 1024 + q * 256 + threadIdx.x;
        if (i < M) dst[i] = v[q];
    }
}

// Leading w x w block (w <= spec_width) of the leaf at column c0, factored
// without interchanges by one wavefront per matrix in LDS: lane (i, h) owns
// row i's columns of parity h (rows and columns beyond w padded with the
// identity, which leaves the leading block's factor unchanged).  At column j
// every lane forms l_ij from the column and the pivot (both final since the
// last barrier), the owner of column j stores it apart, and each lane
// updates its own columns beyond j with row j (which it updated itself):
// one barrier per column.  Each entry a_ij^(j) met below the pivot is
// checked; the block is written back with ipiv = identity.
template <typename R>
__global__ void __launch_bounds__(64) spec_top_kernel(cx<R>* __restrict__ Abase, int64_t lda, int64_t strideA,
                                                      int c0, int w, int* __restrict__ ipiv_base, int64_t strideP,
                                                      int* __restrict__ flags)
{
    typedef cx<R> C;
    constexpr int SW = spec_width<R>();
    constexpr int LD = SW + 1;  // padded row stride: lanes of one column hit different banks
    __shared__ C sT[SW * LD];   // sT[i * LD + k]: the working block, row-major
    __shared__ C sL[SW * LD];   // the multipliers l_ij
    const int b = blockIdx.x;
    if (flags[b]) return;  // the panel is factored again anyway
    C* A = Abase + (size_t)b * strideA + c0 + (size_t)c0 * lda;
    const int lane = threadIdx.x, i = lane & (SW - 1), h = lane >> 5;
    for (int k = h; k < SW; k += 2) sT[i * LD + k] = (i < w && k < w) ? A[i + (size_t)k * lda] : C{R(i == k ? 1 : 0), R(0)};
    __syncthreads();
    bool bad = false;
    for (int j = 0; j < SW; ++j) {
        const C d = sT[j * LD + j];
        const R ad = cabs1(d);
        const C a = sT[i * LD + j];
        const bool below = i > j;
        bad = bad || !(ad > R(0)) || (below && cabs1(a) > ad);
        const C l = zmul(a, zrcp(d));
        if (below && h == (j & 1)) sL[i * LD + j] = l;
        if (below)
            for (int k = j + 1 + ((j + 1 + h) & 1); k < SW; k += 2)
                sT[i * LD + k] = zsub(sT[i * LD + k], zmul(l, sT[j * LD + k]));
        __syncthreads();
    }
    if (__ballot(bad)) {
        if (lane == 0) flags[b] = 1;
        return;
    }
    if (i < w) {
        for (int k = h; k < w; k += 2) A[i + (size_t)k * lda] = k < i ? sL[i * LD + k] : sT[i * LD + k];
        if (h == 0) ipiv_base[(size_t)b * strideP + c0 + i] = c0 + i;
    }
}

// Rows [c0 + w, N) of the leaf's columns: l U11 = a solved row by row
// (forward substitution over the columns, U11 in LDS with the reciprocals of
// its diagonal), a thread per row holding it in registers; each entry
// a_rc^(c) met before its division is checked against u_cc.
template <typename R>
__global__ void __launch_bounds__(256) spec_rows_kernel(cx<R>* __restrict__ Abase, int64_t lda, int64_t strideA,
                                                        int N, int c0, int w, int* __restrict__ flags)
{
    typedef cx<R> C;
    constexpr int SW = spec_width<R>();
    __shared__ C sU[SW * SW];
    __shared__ R sAd[SW];
    const int b = blockIdx.y;
    if (flags[b]) return;
    C* A = Abase + (size_t)b * strideA;
    const int tid = threadIdx.x;
#pragma unroll
    for (int e0 = 0; e0 < SW * SW; e0 += 256) {
        const int e = e0 + tid, i = e % SW, k = e / SW;
        if (k < w && i <= k) {
            const C u = A[(c0 + i) + (size_t)(c0 + k) * lda];
            if (i == k) {
                sU[e] = zrcp(u);
                sAd[k] = cabs1(u);
            } else {
                sU[e] = u;
            }
        }
    }
    __syncthreads();
    const int r = c0 + w + blockIdx.x * 256 + tid;
    if (r >= N) return;
    C x[SW];
#pragma unroll
    for (int c = 0; c < SW; ++c)
        if (c < w) x[c] = A[r + (size_t)(c0 + c) * lda];
    bool bad = false;
#pragma unroll
    for (int c = 0; c < SW; ++c) {
        if (c < w) {
            C t = x[c];
#pragma unroll
            for (int k = 0; k < c; ++k) t = zsub(t, zmul(x[k], sU[k + SW * c]));
            bad = bad || cabs1(t) > sAd[c];
            x[c] = zmul(t, sU[c + SW * c]);
        }
    }
#pragma unroll
    for (int c = 0; c < SW; ++c)
        if (c < w) A[r + (size_t)(c0 + c) * lda] = x[c];
    if (bad) atomicOr(flags + b, 1);
}

// Partial pivoting of the panel columns [k0, k0 + kb), rows [k0, N), of the
// matrices whose check failed: the panel back from the copy, then LAPACK's
// rule column by column (largest |Re|+|Im|, first row on ties; a column
// without a nonzero candidate keeps its row and sets info), interchanges
// within the panel, the scaled column and the rank-1 update of the panel's
// remaining columns.  One work-group per matrix, global memory: the rare
// path.  Clears the matrix's flag, marks an interchange for the rows
// outside the panel, and counts the repair.
template <typename R>
__global__ void __launch_bounds__(512) spec_repair_kernel(cx<R>* __restrict__ Abase, int64_t lda, int64_t strideA,
                                                          int N, int k0, int kb, const cx<R>* __restrict__ bak,
                                                          int64_t strideBak, int* __restrict__ ipiv_base,
                                                          int64_t strideP, int* __restrict__ info_base,
                                                          int* __restrict__ flags, int* __restrict__ swapped,
                                                          int* __restrict__ repairs)
{
    typedef cx<R> C;
    constexpr int NT = 512, NW = NT / 64;
    __shared__ R sBest[NW];
    __shared__ int sIdx[NW];
    const int b = blockIdx.x;
    if (!flags[b]) return;
    const int M = N - k0;
    C* P = Abase + (size_t)b * strideA + k0 + (size_t)k0 * lda;
    const C* S = bak + (size_t)b * strideBak;
    int* ipiv = ipiv_base + (size_t)b * strideP;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (size_t e = tid; e < (size_t)M * kb; e += NT) P[(e % M) + (e / M) * lda] = S[e];
    __syncthreads();
    for (int j = 0; j < kb; ++j) {
        R best = R(-1);
        int bidx = INT_MAX;
        for (int t = j + tid; t < M; t += NT) {
            const R a = cabs1(P[t + (size_t)j * lda]);
            if (a > best) {
                best = a;
                bidx = t;
            }
        }
        wave_argmax(best, bidx);
        if (lane == 0) {
            sBest[wave] = best;
            sIdx[wave] = bidx;
        }
        __syncthreads();
        R bb = sBest[0];
        int p = sIdx[0];
#pragma unroll
        for (int v = 1; v < NW; ++v)
            if (sBest[v] > bb || (sBest[v] == bb && sIdx[v] < p)) {
                bb = sBest[v];
                p = sIdx[v];
            }
        const bool pivoted = bb > R(0);
        if (!pivoted) p = j;
        if (tid == 0) {
            if (!pivoted && info_base[b] == 0) info_base[b] = k0 + j + 1;
            ipiv[k0 + j] = k0 + p;
            if (p != j) swapped[b] = 1;
        }
        if (p != j)
            for (int c = tid; c < kb; c += NT) {
                const C x = P[j + (size_t)c * lda];
                P[j + (size_t)c * lda] = P[p + (size_t)c * lda];
                P[p + (size_t)c * lda] = x;
            }
        __syncthreads();
        const C piv = P[j + (size_t)j * lda];
        const bool nonzero = piv.re != R(0) || piv.im != R(0);
        const C rr = nonzero ? zrcp(piv) : C{R(1), R(0)};
        for (int t = j + 1 + tid; t < M; t += NT) {
            C l = P[t + (size_t)j * lda];
            if (nonzero) l = zmul(l, rr);
            P[t + (size_t)j * lda] = l;
            for (int c = j + 1; c < kb; ++c)
                P[t + (size_t)c * lda] = zsub(P[t + (size_t)c * lda], zmul(l, P[j + (size_t)c * lda]));
        }
        __syncthreads();
    }
    if (tid == 0) {
        flags[b] = 0;
        if (repairs) atomicAdd(repairs, 1);
    }
}


// ---------------------------------------------------------------------------
// 5c. Packed FP16 tile product for the complex64 trailing updates.
//
// C += alpha A B with A (M x K) and B (K x N) complex64 in the four-product
// form (Cr += Ar Br - Ai Bi, Ci += Ar Bi + Ai Br), FP16 inputs and FP32
// accumulation, with the operands packed to FP16 once per product instead
// of in every tile: pack_a_h16_kernel and pack_b_h16_kernel write two
// planes (re, im) with k contiguous per row of A and per column of B (zero
// beyond M, N and K, K padded to 32), so that the tile kernel stages them to
// LDS with plain 16-byte loads and feeds v_mfma_f32_16x16x32_f16 (eight k
// per lane); -Ai is formed in registers by a sign flip.  Two planes and two
// accumulator sets (against three of each in the three-product form) halve
// the LDS bytes per matrix-core instruction and fit two wavefronts per SIMD,
// so that one wave's LDS and global waits overlap the other's products.
// Tiles of 128 x 128 per work-group of four wavefronts (64 x 64 each), the
// 32-deep chunks double-buffered in LDS (80 KiB: two work-groups per CU).
// The batch is spread so that every work-group of one matrix runs on the
// same XCD (work-group ids go round-robin over the 8 XCDs), and the
// matrix's packed A and B are read into one L2 only.
// ---------------------------------------------------------------------------
typedef _Float16 hvec8 __attribute__((ext_vector_type(8)));
constexpr int HP_BM = 128, HP_BK = 32, HP_LDK = HP_BK + 8;  // LDS row stride in halves (80 bytes)
constexpr int HP_GROUP = 4;   // row panels per group of tiles
constexpr int HP_PLANES = 2;  // re, im

// A (M x K, column-major) -> Ap: plane p, row m at Ap + (p Mp + m) Kp, k
// contiguous.  64 rows x 32 k per work-group, transposed through LDS.
__global__ void __launch_bounds__(256) pack_a_h16_kernel(const cx<float>* __restrict__ Abase, int64_t lda,
                                                         int64_t strideA, int M, int K, _Float16* __restrict__ Apbase,
                                                         int Mp, int Kp, int64_t strideAp)
{
    __shared__ cx<float> t[32][65];
    const int m0 = blockIdx.x * 64, k0 = blockIdx.y * 32, tid = threadIdx.x;
    const cx<float>* A = Abase + (size_t)blockIdx.z * strideA;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int k = (tid >> 6) + 4 * q, m = tid & 63;
        cx<float> v{0.0f, 0.0f};
        if (m0 + m < M && k0 + k < K) v = A[(m0 + m) + (size_t)(k0 + k) * lda];
        t[k][m] = v;
    }
    __syncthreads();
    const int m = tid >> 2, seg = tid & 3;
    hvec8 re, im;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const cx<float> v = t[8 * seg + j][m];
        re[j] = (_Float16)v.re;
        im[j] = (_Float16)v.im;
    }
    const size_t plane = (size_t)Mp * Kp;
    _Float16* P = Apbase + (size_t)blockIdx.z * strideAp + (size_t)(m0 + m) * Kp + k0 + 8 * seg;
    *reinterpret_cast<hvec8*>(P) = re;
    *reinterpret_cast<hvec8*>(P + plane) = im;
}

// B (K x N, column-major) -> Bp: plane p, column n at Bp + (p Np + n) Kp, k
// contiguous (B's own order): eight k per thread.
__global__ void __launch_bounds__(256) pack_b_h16_kernel(const cx<float>* __restrict__ Bbase, int64_t ldb,
                                                         int64_t strideB, int K, int N, _Float16* __restrict__ Bpbase,
                                                         int Np, int Kp, int64_t strideBp)
{
    const int segs = Kp / 8;
    const int id = blockIdx.x * 256 + threadIdx.x;
    if (id >= Np * segs) return;
    const int n = id / segs, seg = id % segs;
    const cx<float>* B = Bbase + (size_t)blockIdx.y * strideB + (size_t)n * ldb;
    hvec8 re, im;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int k = 8 * seg + j;
        cx<float> v{0.0f, 0.0f};
        if (n < N && k < K) v = B[k];
        re[j] = (_Float16)v.re;
        im[j] = (_Float16)v.im;
    }
    const size_t plane = (size_t)Np * Kp;
    _Float16* P = Bpbase + (size_t)blockIdx.y * strideBp + (size_t)n * Kp + 8 * seg;
    *reinterpret_cast<hvec8*>(P) = re;
    *reinterpret_cast<hvec8*>(P + plane) = im;
}

// C (M x N of the columns [c_lo, c_lo + N) of the packed B) += alpha A B;
// one-dimensional grid of Mt Nt batch work-groups, Mt = ceil(M / 128),
// Nt = ceil(N / 128).  Two LDS stages: chunk i is multiplied from one while
// chunk i + 1, loaded into registers during chunk i - 1, is stored to the
// other, and chunk i + 2 is loaded: one barrier per chunk.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
gemm_h16p_kernel(int M, int N, int Kp, float alpha, const _Float16* __restrict__ Apbase, int Mp, int64_t strideAp,
                 const _Float16* __restrict__ Bpbase, int Np, int c_lo, int64_t strideBp, cx<float>* __restrict__ Cbase,
                 int64_t ldc, int64_t strideC, int batch)
{
    typedef cx<float> Cx;
    constexpr int STAGE = HP_PLANES * HP_BM * HP_LDK;  // halves of one operand's stage
    __shared__ __attribute__((aligned(16))) _Float16 sA[2 * STAGE];
    __shared__ __attribute__((aligned(16))) _Float16 sB[2 * STAGE];
    const int Mt = (M + HP_BM - 1) / HP_BM, Nt = (N + HP_BM - 1) / HP_BM, T = Mt * Nt;
    // work-group -> (matrix, tile): all tiles of a matrix on one XCD when the
    // batch is a multiple of 8
    int z, t;
    const int L = blockIdx.x;
    if (batch % 8 == 0) {
        const int x = L % 8, j = L / 8;
        z = x + 8 * (j / T);
        t = j % T;
    } else {
        z = L / T;
        t = L % T;
    }
    // tiles in groups of HP_GROUP row panels, column panels outer within a
    // group: the work-groups of an XCD in flight at once share a few A and B
    // panels in its L2
    const int gsz = HP_GROUP * Nt, g = t / gsz, first = g * HP_GROUP;
    const int grows = min(HP_GROUP, Mt - first), tg = t % gsz;
    const int row0 = (first + tg % grows) * HP_BM, col0 = (tg / grows) * HP_BM;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int wm = wave & 1, wn = wave >> 1;
    const size_t planeA = (size_t)Mp * Kp, planeB = (size_t)Np * Kp;
    const _Float16* Ap = Apbase + (size_t)z * strideAp + (size_t)row0 * Kp;
    const _Float16* Bp = Bpbase + (size_t)z * strideBp + (size_t)(c_lo + col0) * Kp;
    Cx* C = Cbase + (size_t)z * strideC;
    // staging role: rows (or columns) sr and sr + 64, k = 8 sseg .. + 8
    const int sr = tid >> 2, sseg = tid & 3;
    hvec8 ra[HP_PLANES][2], rb[HP_PLANES][2];
    auto load = [&](int kc) {
#pragma unroll
        for (int p = 0; p < HP_PLANES; ++p)
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                ra[p][i] = *reinterpret_cast<const hvec8*>(Ap + p * planeA + (size_t)(sr + 64 * i) * Kp + kc + 8 * sseg);
                rb[p][i] = *reinterpret_cast<const hvec8*>(Bp + p * planeB + (size_t)(sr + 64 * i) * Kp + kc + 8 * sseg);
            }
    };
    auto store = [&](int st) {
#pragma unroll
        for (int p = 0; p < HP_PLANES; ++p)
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int off = st * STAGE + p * HP_BM * HP_LDK + (sr + 64 * i) * HP_LDK + 8 * sseg;
                *reinterpret_cast<hvec8*>(&sA[off]) = ra[p][i];
                *reinterpret_cast<hvec8*>(&sB[off]) = rb[p][i];
            }
    };
    fvec4 cr[4][4], ci[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            cr[a][b] = fvec4{0, 0, 0, 0};
            ci[a][b] = fvec4{0, 0, 0, 0};
        }
    const int nk = Kp / HP_BK;
    load(0);
    store(0);
    if (nk > 1) load(HP_BK);
    for (int c = 0; c < nk; ++c) {
        // stage c & 1 holds chunk c; the registers hold chunk c + 1
        __syncthreads();
        const int st = c & 1;
        hvec8 yr[4], yi[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int off = st * STAGE + (wn * 64 + b * 16 + li) * HP_LDK + 8 * lk;
            yr[b] = *reinterpret_cast<const hvec8*>(&sB[off]);
            yi[b] = *reinterpret_cast<const hvec8*>(&sB[off + HP_BM * HP_LDK]);
        }
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const int off = st * STAGE + (wm * 64 + a * 16 + li) * HP_LDK + 8 * lk;
            const hvec8 xr = *reinterpret_cast<const hvec8*>(&sA[off]);
            const hvec8 xi = *reinterpret_cast<const hvec8*>(&sA[off + HP_BM * HP_LDK]);
            const hvec8 xn = -xi;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                cr[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xr, yr[b], cr[a][b], 0, 0, 0);
                ci[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xr, yi[b], ci[a][b], 0, 0, 0);
                cr[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xn, yi[b], cr[a][b], 0, 0, 0);
                ci[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xi, yr[b], ci[a][b], 0, 0, 0);
            }
        }
        // chunk c + 1 into the other stage (its last readers, of chunk c - 1,
        // passed this iteration's barrier), then load chunk c + 2
        if (c + 1 < nk) {
            store(st ^ 1);
            if (c + 2 < nk) load((c + 2) * HP_BK);
        }
    }
    // a lane holds rows 4 lk .. 4 lk + 3 of one column of each 16 x 16 block:
    // 32 contiguous bytes of C, read and written as two 16-byte vectors
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int rowb = row0 + wm * 64 + 16 * a + 4 * lk;
            const int col = col0 + wn * 64 + 16 * b + li;
            if (col >= N || rowb >= M) continue;
            float re[4], im[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                re[v] = alpha * cr[a][b][v];
                im[v] = alpha * ci[a][b][v];
            }
            Cx* cp = C + rowb + (size_t)col * ldc;
            if (rowb + 3 < M) {
                fvec4u* cv = reinterpret_cast<fvec4u*>(cp);
                fvec4u lo = {re[0], im[0], re[1], im[1]}, hi = {re[2], im[2], re[3], im[3]};
                lo += cv[0];
                hi += cv[1];
                cv[0] = lo;
                cv[1] = hi;
            } else {
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    if (rowb + v < M) {
                        cp[v].re += re[v];
                        cp[v].im += im[v];
                    }
            }
        }
}

// Halves of the packed operands of products with M, N <= n and K <= kmax:
// A and B, two planes each (B with one tile of columns to spare), per
// matrix.
inline size_t h16p_pack_halves(int n, int kmax)
{
    const size_t np = (size_t)round_up_host(n, HP_BM) + HP_BM, kp = (size_t)round_up_host(kmax, HP_BK);
    return 2 * HP_PLANES * np * kp;
}

// C += alpha A B (complex64, column-major, batched) through the packed
// operands in `pack` (h16p_pack_halves(max(M, N), K) halves per matrix).
// The columns [0, n_first) of C and B are multiplied first, then the rest,
// with `between` called in between (the look-ahead's fork).
template <typename F>
int gemm_h16p(int M, int N, int K, float alpha, const cx<float>* A, int64_t lda, int64_t sA, const cx<float>* B,
              int64_t ldb, int64_t sB, cx<float>* C, int64_t ldc, int64_t sC, int batch, _Float16* pack,
              int64_t sPack, int n_first, F between, hipStream_t stream)
{
    if (M <= 0 || N <= 0 || K <= 0 || batch <= 0) return 0;
    // B's packed columns reach one tile beyond N: the second launch starts
    // at a column that need not be a multiple of 128
    const int Mp = round_up_host(M, HP_BM), Np = round_up_host(N, HP_BM) + HP_BM, Kp = round_up_host(K, HP_BK);
    _Float16* Ap = pack;
    _Float16* Bp = pack + (size_t)HP_PLANES * Mp * Kp;
    hipLaunchKernelGGL(pack_a_h16_kernel, dim3(Mp / 64, Kp / 32, batch), dim3(256), 0, stream, A, lda, sA, M, K, Ap,
                       Mp, Kp, sPack);
    const int segs = Kp / 8;
    hipLaunchKernelGGL(pack_b_h16_kernel, dim3((Np * segs + 255) / 256, batch), dim3(256), 0, stream, B, ldb, sB, K, N,
                       Bp, Np, Kp, sPack);
    int e;
    if ((e = (int)hipGetLastError())) return e;
    auto launch = [&](int c_lo, int n) {
        if (n <= 0) return;
        // algorithmic bytes: the packed operands (two FP16 planes) read
        // once, C read and written
        TimedLaunch timed(stream, TK_CGEMM_PACKED, 8.0 * M * n * K * batch,
                          (4.0 * ((double)M * K + (double)K * n) + 16.0 * M * n) * batch);
        const int Mt = (M + HP_BM - 1) / HP_BM, Nt = (n + HP_BM - 1) / HP_BM;
        hipLaunchKernelGGL(gemm_h16p_kernel, dim3(Mt * Nt * batch), dim3(256), 0, stream, M, n, Kp, alpha,
                           (const _Float16*)Ap, Mp, sPack, (const _Float16*)Bp, Np, c_lo, sPack, C + (size_t)c_lo * ldc,
                           ldc, sC, batch);
    };
    const int n1 = n_first > 0 && n_first < N ? n_first : N;
    launch(0, n1);
    if ((e = between())) return e;
    launch(n1, N - n1);
    return (int)hipGetLastError();
}

template <typename R, int NT, int RPT, int SB>
void launch_panel(cx<R>* A, int64_t lda, int64_t sA, int N, int k0, int kb, int sw_lo, int sw_hi, int* ipiv,
                  int64_t sP, int* info, int batch, hipStream_t stream)
{
    hipLaunchKernelGGL((panel_kernel<R, NT, RPT, SB>), dim3(batch), dim3(NT), 0, stream, A, lda, sA, N, k0, kb, sw_lo,
                       sw_hi, ipiv, sP, info);
}

// Panel settings of one factorisation: single-work-group leaf width and
// threads, and for the multi-work-group leaves the exchange records (null:
// single-work-group leaves only), the work-groups one launch may take over
// the batch, the spin bound of their waits and the next sequence number.
constexpr int PANEL_SINGLE_MAX_ROWS = 8192;  // rows of the one-work-group leaf kernels

struct PanelCfg {
    int leaf = 16, nt = 512;
    bool mw = true;  // multi-work-group leaves where they fit
    uint64_t* xbuf = nullptr;
    int budget = MW_LEAF_BUDGET;
    int spin = 1 << 20;
    uint32_t seq = 1;
    // speculative panels (1c): the panel copy, the per-matrix check flags
    // and the context's repair counter (may be null)
    bool spec = false;
    void* spec_bak = nullptr;
    int* spec_flags = nullptr;
    int* spec_swapped = nullptr;  // per matrix: the panel interchanged rows (repair)
    int* spec_repairs = nullptr;
};

// Rows per work-group (512 threads with one row each, or two: 512 or 1024
// rows), work-groups per matrix and matrices per launch of a
// multi-work-group leaf of M rows: 512 rows when they hold the leaf, else
// 1024 (fewer work-groups, fewer exchanges to wait for), up to 32
// work-groups per matrix; one launch of at most `budget` work-groups when
// the batch fits, else the batch in chunks of budget / P matrices; nt = 0
// when M exceeds 32 x 1024 rows.
inline void mw_shape(int M, int batch, const PanelCfg& pc, int& P, int& nt, int& per_launch)
{
    nt = M <= 512 ? 512 : M <= 1024 * MW_PMAX ? 1024 : 0;
    P = nt ? (M + nt - 1) / nt : 0;
    per_launch = P ? std::max(1, std::min(batch, pc.budget / P)) : 0;
}

template <typename R, int NT, int RPT>
void launch_panel_mw(cx<R>* A, int64_t lda, int64_t sA, int N, int k0, int kb, int sw_lo, int sw_hi, int* ipiv,
                     int64_t sP, int* info, int batch, int P, const PanelCfg& pc, uint32_t seq, hipStream_t stream)
{
    hipLaunchKernelGGL((panel_mw_kernel<R, mw_width<R>(), NT, RPT>), dim3(P, batch), dim3(NT), 0, stream, A, lda, sA, N,
                       k0, kb, sw_lo, sw_hi, ipiv, sP, info, pc.xbuf, P, seq, pc.spin);
}

// Factor columns [k0, k0 + kb), kb <= mw_width, of rows [k0, N) over
// several work-groups per matrix; sets `done` when the leaf fits them.
template <typename R>
int panel_leaf_mw(cx<R>* A, int64_t lda, int64_t sA, int N, int k0, int kb, int sw_lo, int sw_hi, int* ipiv,
                  int64_t sP, int* info, int batch, PanelCfg& pc, bool& done, hipStream_t stream)
{
    done = false;
    if (!pc.xbuf || kb > mw_width<R>()) return 0;
    int P, nt, per;
    mw_shape(N - k0, batch, pc, P, nt, per);
    if (!nt) return 0;
    const uint32_t seq = pc.seq;
    pc.seq += 64;
    PanelCfg part = pc;
    for (int b0 = 0; b0 < batch; b0 += per) {
        const int nb = std::min(per, batch - b0);
        cx<R>* Ab = A + (size_t)b0 * sA;
        int* pb = ipiv + (size_t)b0 * sP;
        int* ib = info + b0;
        part.xbuf = pc.xbuf + (size_t)b0 * MW_XWORDS;
        if (nt == 512) launch_panel_mw<R, 512, 1>(Ab, lda, sA, N, k0, kb, sw_lo, sw_hi, pb, sP, ib, nb, P, part, seq, stream);
        else launch_panel_mw<R, 512, 2>(Ab, lda, sA, N, k0, kb, sw_lo, sw_hi, pb, sP, ib, nb, P, part, seq, stream);
    }
    done = true;
    return (int)hipGetLastError();
}

// Factor columns [k0, k0 + kb) of rows [k0, N) with one work-group per
// matrix; interchanges are applied to the columns [sw_lo, sw_hi).
template <typename R>
int panel_leaf(cx<R>* A, int64_t lda, int64_t sA, int N, int k0, int kb, int sw_lo, int sw_hi, int* ipiv, int64_t sP,
               int* info, int batch, int nt, hipStream_t stream)
{
    // 512 threads (two wavefronts per SIMD, up to 256 VGPRs each) or 256
    // (one wavefront per SIMD, up to 512); the register-resident sub-block
    // holds RPT rows x SB columns per thread.  A 256-thread work-group fits
    // on a CU beside one work-group of the tile product (the tuning option
    // CPT_TUNE_PANEL_THREADS).
    const int M = N - k0;
    // every instantiation launched here keeps its rows in registers without
    // scratch: 512 threads up to 2048 rows (complex128) or 8192 (complex64),
    // 256 threads up to 8192 rows; taller leaves are factored by several
    // work-groups per matrix
    constexpr bool dbl = sizeof(R) == 8;
    if (nt == 256 || M > (dbl ? 2048 : 8192)) {
        if (M <= 256) launch_panel<R, 256, 1, 8>(A, lda, sA, N, k0, kb, sw_lo, sw_hi, ipiv, sP, info, batch, stream);
        else if (M <= 512) launch_panel<R, 256, 2, 8>(A, lda, sA, N, k0, kb, sw_lo, sw_hi, ipiv, sP, info, batch, stream);
        else if (M <= 1024) launch_panel<R, 256, 4, 8>(A, lda, sA, N, k0, kb, sw_lo, sw_hi, ipiv, sP, info, batch, stream);
        else if (M <= 2048) launch_panel<R, 256, 8, 4>(A, lda, sA, N, k0, kb, sw_lo, sw_hi, ipiv, sP, info, batch, stream);
        else if (M <= 4096) launch_panel<R, 256, 16, 2>(A, lda, sA, N, k0, kb, sw_lo, sw_hi, ipiv, sP, info, batch, stream);
        else if (M <= 8192) launch_panel<R, 256, 32, 1>(A, lda, sA, N, k0, kb, sw_lo, sw_hi, ipiv, sP, info, batch, stream);
        else return (int)hipErrorInvalidValue;  // taller leaves go to the multi-work-group kernel
        return (int)hipGetLastError();
    }
    if (M <= 512) launch_panel<R, 512, 1, 8>(A, lda, sA, N, k0, kb, sw_lo, sw_hi, ipiv, sP, info, batch, stream);
    else if (M <= 1024) launch_panel<R, 512, 2, 8>(A, lda, sA, N, k0, kb, sw_lo, sw_hi, ipiv, sP, info, batch, stream);
    else if (M <= 2048) launch_panel<R, 512, 4, 8>(A, lda, sA, N, k0, kb, sw_lo, sw_hi, ipiv, sP, info, batch, stream);
    else if constexpr (!dbl) {
        if (M <= 4096) launch_panel<R, 512, 8, 4>(A, lda, sA, N, k0, kb, sw_lo, sw_hi, ipiv, sP, info, batch, stream);
        else launch_panel<R, 512, 16, 2>(A, lda, sA, N, k0, kb, sw_lo, sw_hi, ipiv, sP, info, batch, stream);
    }
    return (int)hipGetLastError();
}

template <typename R>
int gemm(int M, int N, int K, R alpha, const cx<R>* A, int64_t lda, int64_t sA, const cx<R>* B, int64_t ldb,
         int64_t sB, cx<R>* C, int64_t ldc, int64_t sC, int beta_zero, int batch, hipStream_t stream);

// Inverted diagonal blocks of the factor LU into D (dstep = 1: D is the
// factor itself, its diagonal blocks overwritten; 0: a strip of PB columns
// per matrix, leading dimension ldd >= N).
template <typename R>
int invert_diag(const cx<R>* LU, int N, int64_t lda, int64_t strideA, cx<R>* D, int64_t ldd, int64_t strideD,
                int dstep, int batch, hipStream_t stream, bool with_u = true)
{
    if (N <= 0 || batch <= 0) return 0;
    if (with_u)
        hipLaunchKernelGGL((invert_diag_kernel<R, true>), dim3((N + PB - 1) / PB, batch), dim3(256), 0, stream, LU,
                           lda, strideA, N, D, ldd, strideD, dstep);
    else
        hipLaunchKernelGGL((invert_diag_kernel<R, false>), dim3((N + PB - 1) / PB, batch), dim3(256), 0, stream, LU,
                           lda, strideA, N, D, ldd, strideD, dstep);
    return (int)hipGetLastError();
}

template <typename R>
int trsm_lower(const cx<R>* L, int64_t lda, int64_t sA, int kb, cx<R>* B, int64_t ldb, int64_t sB, int ncols,
               int batch, hipStream_t stream, cx<R>* inv = nullptr);
template <typename R>
int trsm_rec(const cx<R>* L, int64_t lda, int64_t sA, int kb, cx<R>* B, int64_t ldb, int64_t sB, int ncols, int batch,
             hipStream_t stream, cx<R>* inv);

// Recursive factorisation of columns [c0, c0 + w) of rows [c0, N) inside the
// panel [sw_lo, sw_hi): factor the left half, solve its unit-lower triangle
// against the top of the right half (trsm), update the rest of the right
// half with an MFMA product spread over the whole device, factor the right
// half.  The one-work-group leaves stream only `leaf` columns.
template <typename R>
int panel_rec(cx<R>* A, int64_t lda, int64_t sA, int N, int c0, int w, int sw_lo, int sw_hi, int* ipiv, int64_t sP,
              int* info, int batch, PanelCfg& pc, hipStream_t stream)
{
    const int M = N - c0;
    int e;
    // multi-work-group leaves where enabled, and always above the 8192 rows
    // the one-work-group kernel keeps in registers
    if (w <= mw_width<R>() && pc.xbuf && (pc.mw || M > PANEL_SINGLE_MAX_ROWS)) {
        bool done;
        if ((e = panel_leaf_mw<R>(A, lda, sA, N, c0, w, sw_lo, sw_hi, ipiv, sP, info, batch, pc, done, stream)) || done)
            return e;
    }
    // the one-work-group leaf kernel holds at most PB columns (its LDS
    // triangle) and 8192 rows; taller panels split down to the other's width
    if (w <= PB && (w <= pc.leaf || M <= 2 * pc.leaf) && M <= PANEL_SINGLE_MAX_ROWS)
        return panel_leaf<R>(A, lda, sA, N, c0, w, sw_lo, sw_hi, ipiv, sP, info, batch, pc.nt, stream);
    const int h = w / 2;
    if ((e = panel_rec<R>(A, lda, sA, N, c0, h, sw_lo, sw_hi, ipiv, sP, info, batch, pc, stream))) return e;
    cx<R>* A11 = A + c0 + (size_t)c0 * lda;
    cx<R>* A12 = A + c0 + (size_t)(c0 + h) * lda;
    cx<R>* A21 = A + (c0 + h) + (size_t)c0 * lda;
    cx<R>* A22 = A + (c0 + h) + (size_t)(c0 + h) * lda;
    if ((e = trsm_lower<R>(A11, lda, sA, h, A12, lda, sA, w - h, batch, stream))) return e;
    if (M > h && (e = gemm<R>(M - h, w - h, h, R(-1), A21, lda, sA, A12, lda, sA, A22, lda, sA, 0, batch, stream)))
        return e;
    return panel_rec<R>(A, lda, sA, N, c0 + h, w - h, sw_lo, sw_hi, ipiv, sP, info, batch, pc, stream);
}

// Speculative factorisation (1c) of columns [c0, c0 + w) of rows [c0, N):
// leaves of at most spec_width columns (leading block by one wavefront,
// rows below by forward substitution), joined as in panel_rec by trsm +
// MFMA GEMM.
template <typename R>
int panel_rec_spec(cx<R>* A, int64_t lda, int64_t sA, int N, int c0, int w, int* ipiv, int64_t sP, int batch,
                   const PanelCfg& pc, hipStream_t stream)
{
    int e;
    if (w <= spec_width<R>()) {
        hipLaunchKernelGGL(spec_top_kernel<R>, dim3(batch), dim3(64), 0, stream, A, lda, sA, c0, w, ipiv, sP,
                           pc.spec_flags);
        const int rows = N - c0 - w;
        if (rows > 0)
            hipLaunchKernelGGL(spec_rows_kernel<R>, dim3((rows + 255) / 256, batch), dim3(256), 0, stream, A, lda, sA,
                               N, c0, w, pc.spec_flags);
        return (int)hipGetLastError();
    }
    const int h = w / 2;
    const int M = N - c0;
    if ((e = panel_rec_spec<R>(A, lda, sA, N, c0, h, ipiv, sP, batch, pc, stream))) return e;
    cx<R>* A11 = A + c0 + (size_t)c0 * lda;
    cx<R>* A12 = A + c0 + (size_t)(c0 + h) * lda;
    cx<R>* A21 = A + (c0 + h) + (size_t)c0 * lda;
    cx<R>* A22 = A + (c0 + h) + (size_t)(c0 + h) * lda;
    if ((e = trsm_rec<R>(A11, lda, sA, h, A12, lda, sA, w - h, batch, stream, nullptr))) return e;
    if (M > h && (e = gemm<R>(M - h, w - h, h, R(-1), A21, lda, sA, A12, lda, sA, A22, lda, sA, 0, batch, stream)))
        return e;
    return panel_rec_spec<R>(A, lda, sA, N, c0 + h, w - h, ipiv, sP, batch, pc, stream);
}

// LU with partial pivoting of the panel columns [k0, k0 + kb), rows [k0, N);
// interchanges applied within the panel columns only.  Speculative panels:
// copy, factorisation with the diagonal pivots and its check, then the
// repair of the matrices whose check failed.
template <typename R>
int panel_dispatch(cx<R>* A, int64_t lda, int64_t sA, int N, int k0, int kb, int* ipiv, int64_t sP, int* info,
                   int batch, PanelCfg& pc, hipStream_t stream)
{
    const int M = N - k0;
    // complex LU of an M x kb panel: 4 (M kb^2 - kb^3 / 3) real flops
    TimedLaunch timed(stream, sizeof(R) == 8 ? TK_PANEL_Z : TK_PANEL_C,
                      4.0 * ((double)M * kb * kb - (double)kb * kb * kb / 3.0) * batch);
    if (pc.spec && kb <= SPEC_KB_MAX) {
        cx<R>* bak = (cx<R>*)pc.spec_bak;
        const int64_t sBak = (int64_t)N * spec_bak_cols(N);
        hipLaunchKernelGGL(spec_backup_kernel<R>, dim3((M + 1023) / 1024, kb, batch), dim3(256), 0, stream,
                           (const cx<R>*)A, lda, sA, N, k0, kb, bak, sBak, pc.spec_swapped);
        int e;
        if ((e = (int)hipGetLastError())) return e;
        if ((e = panel_rec_spec<R>(A, lda, sA, N, k0, kb, ipiv, sP, batch, pc, stream))) return e;
        hipLaunchKernelGGL(spec_repair_kernel<R>, dim3(batch), dim3(512), 0, stream, A, lda, sA, N, k0, kb,
                           (const cx<R>*)bak, sBak, ipiv, sP, info, pc.spec_flags, pc.spec_swapped,
                           pc.spec_repairs);
        return (int)hipGetLastError();
    }
    return panel_rec<R>(A, lda, sA, N, k0, kb, k0, k0 + kb, ipiv, sP, info, batch, pc, stream);
}

template <typename R>
int gemm(int M, int N, int K, R alpha, const cx<R>* A, int64_t lda, int64_t sA, const cx<R>* B, int64_t ldb,
         int64_t sB, cx<R>* C, int64_t ldc, int64_t sC, int beta_zero, int batch, hipStream_t stream)
{
    if (M <= 0 || N <= 0 || batch <= 0) return 0;
    if (M <= FEW_ROWS && K >= 256) {
        hipLaunchKernelGGL(gemm_fewrows_kernel<R>, dim3(N, batch), dim3(256), 0, stream, M, K, alpha, A, lda, sA, B,
                           ldb, sB, C, ldc, sC, beta_zero);
        return (int)hipGetLastError();
    }
    if (N <= 2 * SKN && M >= SKM && K >= 4 * SKK) {
        dim3 grid((M + SKM - 1) / SKM, (N + SKN - 1) / SKN, batch);
        const double nbytes =
            (double)sizeof(cx<R>) * ((double)M * K + (double)K * N + (beta_zero ? 1.0 : 2.0) * M * N) * batch;
        TimedLaunch timed(stream, sizeof(R) == 8 ? TK_SKINNY_Z : TK_SKINNY_C, nbytes, nbytes);
        hipLaunchKernelGGL(gemm_skinny_kernel<R>, grid, dim3(256), 0, stream, M, N, K, alpha, A, lda, sA, B, ldb, sB,
                           C, ldc, sC, beta_zero);
    } else if (M >= 96 && K >= 16) {
        TimedLaunch timed(stream, sizeof(R) == 8 ? TK_ZGEMM : TK_CGEMM, 8.0 * M * N * K * batch,
                          (double)sizeof(cx<R>) * ((double)M * K + (double)K * N + (beta_zero ? 1.0 : 2.0) * M * N) *
                              batch);
        // complex64: 128 x 128 tiles for wide products when tuned so
        // (complex128 would need more than the 512 registers of a lane)
        if constexpr (sizeof(R) == 4) {
            if (active_gemm_tiles == 1 && N > 64) {
                dim3 grid((M + TBM - 1) / TBM, (N + 127) / 128, batch);
                hipLaunchKernelGGL((gemm_tiled_kernel<R, 128>), grid, dim3(256), 0, stream, M, N, K, alpha, A, lda,
                                   sA, B, ldb, sB, C, ldc, sC, beta_zero);
                return (int)hipGetLastError();
            }
            if (active_gemm_tiles == 2 && N > 64) {
                dim3 grid((M + TBM - 1) / TBM, (N + 127) / 128, batch);
                hipLaunchKernelGGL((gemm_tiled_gauss_w1_kernel<R, 128>), grid, dim3(256), 0, stream, M, N, K,
                                   alpha, A, lda, sA, B, ldb, sB, C, ldc, sC, beta_zero);
                return (int)hipGetLastError();
            }
            if (active_gemm_tiles == 3) {
                dim3 grid((M + TBM - 1) / TBM, (N + 63) / 64, batch);
                hipLaunchKernelGGL((gemm_tiled_kernel<R, 64, true>), grid, dim3(256), 0, stream, M, N, K, alpha, A,
                                   lda, sA, B, ldb, sB, C, ldc, sC, beta_zero);
                return (int)hipGetLastError();
            }
            if (active_gemm_tiles == 4) {
                dim3 grid((M + TBM - 1) / TBM, (N + 63) / 64, batch);
                hipLaunchKernelGGL(gemm_h16_kernel<16>, grid, dim3(256), 0, stream, M, N, K, alpha, A, lda, sA, B, ldb,
                                   sB, C, ldc, sC, beta_zero);
                return (int)hipGetLastError();
            }
        }
        // complex128 keeps the four-product form: its Gauss form needs one
        // wavefront per SIMD to hold three accumulator sets, and loses
        {
            dim3 grid((M + TBM - 1) / TBM, (N + 63) / 64, batch);
            hipLaunchKernelGGL((gemm_tiled_kernel<R, 64>), grid, dim3(256), 0, stream, M, N, K, alpha, A, lda, sA, B,
                               ldb, sB, C, ldc, sC, beta_zero);
        }
    } else {
        dim3 grid((M + 63) / 64, (N + 63) / 64, batch);
        hipLaunchKernelGGL(gemm_small_kernel<R>, grid, dim3(256), 0, stream, M, N, K, alpha, A, lda, sA, B, ldb, sB, C,
                           ldc, sC, beta_zero);
    }
    return (int)hipGetLastError();
}

// L^-1 B in place for a unit lower kb x kb block L (any kb): blocked
// forward substitution over PB-row diagonal blocks, the blocks below updated
// by the product.  Each diagonal block is solved by the wavefront-wide
// triangle kernel, or, given `inv` (room for kb x PB per matrix), by the
// MFMA product with its inverse, all of L's diagonal blocks inverted first
// in one launch.
template <typename R>
int trsm_lower(const cx<R>* L, int64_t lda, int64_t sA, int kb, cx<R>* B, int64_t ldb, int64_t sB, int ncols,
               int batch, hipStream_t stream, cx<R>* inv)
{
    int e;
    const int64_t sInv = (int64_t)kb * PB;
    if (inv && (e = invert_diag<R>(L, kb, lda, sA, inv, kb, sInv, 0, batch, stream, false))) return e;
    for (int j0 = 0; j0 < kb; j0 += PB) {
        const int jb = min(PB, kb - j0);
        if (inv) {
            hipLaunchKernelGGL(apply_inverse_kernel<R>, dim3((ncols + APPLY_COLS - 1) / APPLY_COLS, batch), dim3(256),
                               0, stream, inv + j0, (int64_t)kb, sInv, jb, B + j0, ldb, sB, ncols);
            if ((e = (int)hipGetLastError())) return e;
        } else if ((e = launch_trsm<R>(L + j0 + (size_t)j0 * lda, lda, sA, jb, B + j0, ldb, sB, ncols, batch,
                                              stream))) {
            return e;
        }
        const int rest = kb - j0 - jb;
        if (rest > 0 && (e = gemm<R>(rest, ncols, jb, R(-1), L + (j0 + jb) + (size_t)j0 * lda, lda, sA, B + j0, ldb,
                                     sB, B + j0 + jb, ldb, sB, 0, batch, stream)))
            return e;
    }
    return 0;
}

// L^-1 B for a unit lower kb x kb block L of any size: halves of L
// recursively (X1 = L11^-1 B1, B2 -= L21 X1 on MFMA with K = kb / 2, X2 =
// L22^-1 B2) down to blocks of 256, which trsm_lower solves (with `inv`, room
// for 256 x PB per matrix, or the wavefront triangle kernel): the products
// keep a depth that makes them compute-bound.
template <typename R>
int trsm_rec(const cx<R>* L, int64_t lda, int64_t sA, int kb, cx<R>* B, int64_t ldb, int64_t sB, int ncols, int batch,
             hipStream_t stream, cx<R>* inv)
{
    if (kb <= 4 * PB) return trsm_lower<R>(L, lda, sA, kb, B, ldb, sB, ncols, batch, stream, inv);
    const int h = ((kb / 2 + PB - 1) / PB) * PB;
    int e;
    if ((e = trsm_rec<R>(L, lda, sA, h, B, ldb, sB, ncols, batch, stream, inv))) return e;
    if ((e = gemm<R>(kb - h, ncols, h, R(-1), L + h, lda, sA, B, ldb, sB, B + h, ldb, sB, 0, batch, stream))) return e;
    return trsm_rec<R>(L + h + (size_t)h * lda, lda, sA, kb - h, B + h, ldb, sB, ncols, batch, stream, inv);
}

// Block width of the right-looking factorisation: 128 columns halve the
// trailing-matrix traffic per flop against 64 where the matrices are large
// (256 halves it again, at the price of wider panels: a tuning choice).
inline int getrf_block(int N, int tuned) { return tuned ? tuned : (N >= 2048 ? 4 * PB : N >= 1024 ? 2 * PB : PB); }

// Workspace of the factorisation (elements): the inverted diagonal blocks
// of L11 for U12 = L11^-1 A12, kb x PB per matrix for the widest block,
// then the exchange records of the multi-work-group panel leaves.
template <typename R>
size_t getrf_inv_elems(int batch)
{
    return (size_t)4 * PB * PB * (size_t)batch;
}
template <typename R>
size_t getrf_xbuf_elems(int batch)
{
    const size_t xbytes = sizeof(uint64_t) * (size_t)MW_XWORDS * batch;
    return (xbytes + sizeof(cx<R>) - 1) / sizeof(cx<R>);
}
// ... then for the speculative panels: the panel copy (N x min(N,
// SPEC_KB_MAX) per matrix), the check flags and the interchange flags (a
// 64-bit word each per matrix)
template <typename R>
size_t getrf_spec_elems(int N, int batch)
{
    const size_t flag_elems = (2 * sizeof(uint64_t) * (size_t)batch + sizeof(cx<R>) - 1) / sizeof(cx<R>);
    return (size_t)N * spec_bak_cols(N) * batch + flag_elems;
}
// ... and, for complex64, the packed FP16 operands of the trailing updates
template <typename R>
size_t getrf_pack_elems(int N, int batch)
{
    if (sizeof(R) != 4) return 0;
    return (h16p_pack_halves(N, spec_bak_cols(N)) * sizeof(_Float16) * batch + sizeof(cx<R>) - 1) / sizeof(cx<R>);
}
template <typename R>
size_t getrf_ws_elems(int N, int batch)
{
    return getrf_inv_elems<R>(batch) + getrf_xbuf_elems<R>(batch) + getrf_spec_elems<R>(N, batch) +
           getrf_pack_elems<R>(N, batch);
}

// Blocked right-looking LU with partial pivoting (LAPACK zgetrf semantics),
// all matrices of the batch at once.  Per block step k0 of width kb:
// interchanges of the panel applied to the other columns, U12 = L11^-1 A12
// in place, A22 -= A21 U12 on MFMA.  With a side stream, the next panel's
// columns are updated first and factored on the side stream while the rest
// of A22 is updated (look-ahead of depth one).
template <typename R>
int getrf_steps(cx<R>* A, int N, int64_t lda, int64_t strideA, int* ipiv, int64_t strideP, int* info, int batch,
                cx<R>* work, const LuSetup* aux, hipStream_t stream);

// `work`: getrf_ws_elems elements, or null (then U12 is solved by the
// wavefront triangle kernel instead of the inverted diagonal blocks)
template <typename R>
int getrf(cx<R>* A, int N, int64_t lda, int64_t strideA, int* ipiv, int64_t strideP, int* info, int batch,
          cx<R>* work, const LuSetup* aux, hipStream_t stream)
{
    if (N <= 0 || batch <= 0) return 0;
    return getrf_steps<R>(A, N, lda, strideA, ipiv, strideP, info, batch, work, aux, stream);
}

template <typename R>
int getrf_steps(cx<R>* A, int N, int64_t lda, int64_t strideA, int* ipiv, int64_t strideP, int* info, int batch,
                cx<R>* work, const LuSetup* aux, hipStream_t stream)
{
    int e;
    int nb = getrf_block(N, aux ? aux->block : 0);
    PanelCfg pc;
    pc.leaf = aux ? aux->leaf : 16;
    pc.nt = aux ? aux->panel_threads : 512;
    if (aux && aux->mw_budget > 0) pc.budget = aux->mw_budget;
    if (aux && aux->panel_spin != 0) pc.spin = aux->panel_spin;
    pc.mw = !aux || aux->panel_mw;
    if (work) {
        // the exchange records start cleared: sequence numbers are unique
        // within this factorisation only
        pc.xbuf = (uint64_t*)(work + getrf_inv_elems<R>(batch));
        const size_t nwords = (size_t)MW_XWORDS * batch;
        hipLaunchKernelGGL(zero_words_kernel, dim3((unsigned)((nwords + 255) / 256)), dim3(256), 0, stream, pc.xbuf,
                           nwords);
        if (aux && aux->speculate && nb <= SPEC_KB_MAX) {
            cx<R>* sp = work + getrf_inv_elems<R>(batch) + getrf_xbuf_elems<R>(batch);
            pc.spec = true;
            pc.spec_bak = sp;
            pc.spec_flags = (int*)(sp + (size_t)N * spec_bak_cols(N) * batch);
            pc.spec_swapped = pc.spec_flags + 2 * (size_t)batch;
            pc.spec_repairs = aux ? aux->spec_repairs : nullptr;
            hipLaunchKernelGGL(zero_words_kernel, dim3((unsigned)((2 * batch + 255) / 256)), dim3(256), 0, stream,
                               (uint64_t*)pc.spec_flags, 2 * (size_t)batch);
        }
    }
    // blocks wider than 256 columns: speculative panels only (the pivot
    // searches' trsm and leaves hold at most 256)
    if (!pc.spec && nb > 4 * PB) nb = 4 * PB;
    // complex64 with FP16-input products: packed operands for the trailing
    // updates (blocks up to the panel copy's width)
    _Float16* pack = nullptr;
    int64_t sPack = 0;
    if (sizeof(R) == 4 && work && active_gemm_tiles == 4 && nb <= spec_bak_cols(N)) {
        pack = (_Float16*)(work + getrf_inv_elems<R>(batch) + getrf_xbuf_elems<R>(batch) +
                           getrf_spec_elems<R>(N, batch));
        sPack = (int64_t)h16p_pack_halves(N, spec_bak_cols(N));
    }
    const bool ahead = aux && aux->side && aux->lookahead;
    if ((e = panel_dispatch<R>(A, lda, strideA, N, 0, min(nb, N), ipiv, strideP, info, batch, pc, stream))) return e;
    for (int k0 = 0; k0 < N; k0 += nb) {
        const int kb = min(nb, N - k0);
        // panel k0 is factored (on `stream`, or on the side stream joined below)
        hipLaunchKernelGGL(laswp_kernel<R>, dim3((N + 255) / 256, batch), dim3(256), 0, stream, A, lda, strideA, N, k0,
                           k0 + kb, ipiv, strideP, k0, k0 + kb,
                           pc.spec && kb <= SPEC_KB_MAX ? pc.spec_swapped : nullptr);
        if ((e = (int)hipGetLastError())) return e;
        const int rest = N - k0 - kb;
        if (rest <= 0) break;
        cx<R>* A11 = A + k0 + (size_t)k0 * lda;
        cx<R>* A12 = A + k0 + (size_t)(k0 + kb) * lda;
        cx<R>* A21 = A + (k0 + kb) + (size_t)k0 * lda;
        cx<R>* A22 = A + (k0 + kb) + (size_t)(k0 + kb) * lda;
        if ((e = trsm_rec<R>(A11, lda, strideA, kb, A12, lda, strideA, rest, batch, stream, work))) return e;
        const int kn = min(nb, rest);  // width of the next panel
        // a panel beside the trailing update gets one work-group per matrix
        // per leaf: the work-groups of a multi-work-group leaf must all be
        // resident at once, which a concurrent update of the whole device
        // delays until it drains; panels taller than one work-group's reach
        // are factored in order
        const bool side = ahead && rest <= PANEL_SINGLE_MAX_ROWS;
        auto fork = [&]() -> int {
            if (!side) return 0;
            // ... factor the next panel on the side stream while the rest of
            // the trailing matrix is updated on `stream` (disjoint columns)
            if (hipEventRecord(aux->ready, stream) != hipSuccess) return (int)hipErrorUnknown;
            if (hipStreamWaitEvent(aux->side, aux->ready, 0) != hipSuccess) return (int)hipErrorUnknown;
            PanelCfg one = pc;
            one.mw = false;
            int e2;
            if ((e2 = panel_dispatch<R>(A, lda, strideA, N, k0 + kb, kn, ipiv, strideP, info, batch, one, aux->side)))
                return e2;
            if (hipEventRecord(aux->done, aux->side) != hipSuccess) return (int)hipErrorUnknown;
            return 0;
        };
        if (pack) {
            // complex64 with FP16-input products: operands packed once for
            // both halves of the update (the next panel's columns first)
            if ((e = gemm_h16p(rest, rest, kb, -1.0f, (const cx<float>*)A21, lda, strideA, (const cx<float>*)A12, lda,
                               strideA, (cx<float>*)A22, lda, strideA, batch, pack, sPack, kn, fork, stream)))
                return e;
        } else {
            // trailing update of the next panel's columns first ...
            if ((e = gemm<R>(rest, kn, kb, R(-1), A21, lda, strideA, A12, lda, strideA, A22, lda, strideA, 0, batch,
                             stream)))
                return e;
            if ((e = fork())) return e;
            if (rest > kn && (e = gemm<R>(rest, rest - kn, kb, R(-1), A21, lda, strideA, A12 + (size_t)kn * lda, lda,
                                          strideA, A22 + (size_t)kn * lda, lda, strideA, 0, batch, stream)))
                return e;
        }
        if (side) {
            if (hipStreamWaitEvent(stream, aux->done, 0) != hipSuccess) return (int)hipErrorUnknown;
        } else if ((e = panel_dispatch<R>(A, lda, strideA, N, k0 + kb, kn, ipiv, strideP, info, batch, pc,
                                          stream))) {
            return e;
        }
    }
    return 0;
}

// Solve with the factor (LAPACK zgetrs): interchanges of B, then the
// substitutions with the inverted diagonal blocks D (from invert_diag).
template <typename R>
int getrs(const cx<R>* LU, int N, int64_t lda, int64_t strideA, const int* ipiv, int64_t strideP, const cx<R>* D,
          int64_t ldd, int64_t strideD, int dstep, cx<R>* B, int nrhs, int64_t ldb, int64_t strideB, int batch,
          hipStream_t stream)
{
    if (N <= 0 || nrhs <= 0 || batch <= 0) return 0;
    hipLaunchKernelGGL(laswp_kernel<R>, dim3((nrhs + 255) / 256, batch), dim3(256), 0, stream, B, ldb, strideB, nrhs, 0,
                       0, ipiv, strideP, 0, N, nullptr);
    // forward substitution with the unit lower factor: the rows below each
    // diagonal block, then every diagonal block in place
    const int nblk = (N + PB - 1) / PB;
    for (int k = 0; k + 1 < nblk; ++k) {
        const int lo = (k + 1) * PB;
        hipLaunchKernelGGL((substitution_kernel<R, false, true>), dim3((N - lo + sub_slab<R>() - 1) / sub_slab<R>(), batch),
                           dim3(256), 0, stream, LU, lda, strideA, N, k, D, ldd, strideD, dstep, B, ldb, strideB, nrhs,
                           lo, N);
    }
    hipLaunchKernelGGL((substitution_kernel<R, false, false>), dim3(nblk, batch), dim3(256), 0, stream, LU, lda,
                       strideA, N, 0, D, ldd, strideD, dstep, B, ldb, strideB, nrhs, 0, 0);
    // backward substitution with the upper factor: the rows above each
    // diagonal block, then every diagonal block in place
    for (int k = nblk - 1; k >= 1; --k) {
        const int hi = k * PB;
        hipLaunchKernelGGL((substitution_kernel<R, true, true>), dim3((hi + sub_slab<R>() - 1) / sub_slab<R>(), batch),
                           dim3(256), 0, stream, LU, lda, strideA, N, k, D, ldd, strideD, dstep, B, ldb, strideB, nrhs,
                           0, hi);
    }
    hipLaunchKernelGGL((substitution_kernel<R, true, false>), dim3(nblk, batch), dim3(256), 0, stream, LU, lda,
                       strideA, N, 0, D, ldd, strideD, dstep, B, ldb, strideB, nrhs, 0, 0);
    return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// Mixed precision: complex128 <-> complex64 conversions, x += d, and the
// per-column squared norms used by the refinement's stopping test.
// ---------------------------------------------------------------------------
__global__ void z2c_kernel(const cx<double>* __restrict__ src, int64_t lds, int64_t strideS, cx<float>* __restrict__ dst,
                           int64_t ldd, int64_t strideD, int rows, int cols)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
    if (r >= rows) return;
    const cx<double> v = src[(size_t)blockIdx.z * strideS + r + (size_t)c * lds];
    dst[(size_t)blockIdx.z * strideD + r + (size_t)c * ldd] = cx<float>{(float)v.re, (float)v.im};
}

__global__ void c2z_add_kernel(const cx<float>* __restrict__ src, int64_t lds, int64_t strideS,
                               cx<double>* __restrict__ dst, int64_t ldd, int64_t strideD, int rows, int cols,
                               int accumulate)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
    if (r >= rows) return;
    const cx<float> v = src[(size_t)blockIdx.z * strideS + r + (size_t)c * lds];
    cx<double>* d = dst + (size_t)blockIdx.z * strideD + r + (size_t)c * ldd;
    if (accumulate) {
        d->re += (double)v.re;
        d->im += (double)v.im;
    } else {
        d->re = (double)v.re;
        d->im = (double)v.im;
    }
}

__global__ void copy_z_kernel(const cx<double>* __restrict__ src, int64_t lds, int64_t strideS,
                              cx<double>* __restrict__ dst, int64_t ldd, int64_t strideD, int rows)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x, c = blockIdx.y;
    if (r >= rows) return;
    dst[(size_t)blockIdx.z * strideD + r + (size_t)c * ldd] = src[(size_t)blockIdx.z * strideS + r + (size_t)c * lds];
}

__global__ void zero_ints_kernel(int* __restrict__ p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// out[b * cols + c] = sum_r |a(r, c)|^2 for one column per work-group.
__global__ void colnorm2_kernel(const cx<double>* __restrict__ a, int64_t lda, int64_t strideA, int rows, int cols,
                                double* __restrict__ out)
{
    __shared__ double part[256];
    const int c = blockIdx.x, b = blockIdx.y;
    const cx<double>* col = a + (size_t)b * strideA + (size_t)c * lda;
    double s = 0.0;
    for (int r = threadIdx.x; r < rows; r += blockDim.x) s += col[r].re * col[r].re + col[r].im * col[r].im;
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[(size_t)b * cols + c] = part[0];
}

}  // namespace

// --------------------------- entry points ---------------------------------
int launch_zgemm(int M, int N, int K, double alpha, const cplx* A, int64_t lda, int64_t sA, const cplx* B,
                 int64_t ldb, int64_t sB, cplx* C, int64_t ldc, int64_t sC, int beta_zero, int batch,
                 hipStream_t stream)
{
    return gemm<double>(M, N, K, alpha, (const cx<double>*)A, lda, sA, (const cx<double>*)B, ldb, sB, (cx<double>*)C,
                        ldc, sC, beta_zero, batch, stream);
}

size_t getrf_workspace_bytes(int N, int batch) { return sizeof(cplx) * getrf_ws_elems<double>(N, batch); }

int launch_getrf(cplx* A, int N, int64_t lda, int64_t strideA, int* ipiv, int64_t strideP, int* info, int batch,
                 cplx* work, const LuSetup* aux, hipStream_t stream)
{
    return getrf<double>((cx<double>*)A, N, lda, strideA, ipiv, strideP, info, batch, (cx<double>*)work, aux, stream);
}

// Workspace of the solve: the inverted diagonal blocks, a strip of N x PB
// per matrix.
size_t getrs_workspace_bytes(int N, int batch) { return sizeof(cplx) * (size_t)N * PB * (size_t)batch; }

int launch_getrs(const cplx* LU, int N, int64_t lda, int64_t strideA, const int* ipiv, int64_t strideP, cplx* B,
                 int nrhs, int64_t ldb, int64_t strideB, int batch, cplx* work, hipStream_t stream)
{
    if (N <= 0 || nrhs <= 0 || batch <= 0) return 0;
    int e;
    cx<double>* D = (cx<double>*)work;
    const int64_t sD = (int64_t)N * PB;
    if ((e = invert_diag<double>((const cx<double>*)LU, N, lda, strideA, D, N, sD, 0, batch, stream))) return e;
    return getrs<double>((const cx<double>*)LU, N, lda, strideA, ipiv, strideP, D, N, sD, 0, (cx<double>*)B, nrhs, ldb,
                         strideB, batch, stream);
}

// complex64 counterparts (the single-precision Green function's matrices):
// the same kernels instantiated on float.
int launch_cgemm(int M, int N, int K, float alpha, const void* A, int64_t lda, int64_t sA, const void* B,
                 int64_t ldb, int64_t sB, void* C, int64_t ldc, int64_t sC, int beta_zero, int batch,
                 hipStream_t stream)
{
    return gemm<float>(M, N, K, alpha, (const cx<float>*)A, lda, sA, (const cx<float>*)B, ldb, sB, (cx<float>*)C, ldc,
                       sC, beta_zero, batch, stream);
}

size_t cgetrf_workspace_bytes(int N, int batch) { return sizeof(cx<float>) * getrf_ws_elems<float>(N, batch); }

int launch_cgetrf(void* A, int N, int64_t lda, int64_t strideA, int* ipiv, int64_t strideP, int* info, int batch,
                  void* work, const LuSetup* aux, hipStream_t stream)
{
    return getrf<float>((cx<float>*)A, N, lda, strideA, ipiv, strideP, info, batch, (cx<float>*)work, aux, stream);
}

size_t cgetrs_workspace_bytes(int N, int batch) { return sizeof(cx<float>) * (size_t)N * PB * (size_t)batch; }

int launch_cgetrs(const void* LU, int N, int64_t lda, int64_t strideA, const int* ipiv, int64_t strideP, void* B,
                  int nrhs, int64_t ldb, int64_t strideB, int batch, void* work, hipStream_t stream)
{
    if (N <= 0 || nrhs <= 0 || batch <= 0) return 0;
    int e;
    cx<float>* D = (cx<float>*)work;
    const int64_t sD = (int64_t)N * PB;
    if ((e = invert_diag<float>((const cx<float>*)LU, N, lda, strideA, D, N, sD, 0, batch, stream))) return e;
    return getrs<float>((const cx<float>*)LU, N, lda, strideA, ipiv, strideP, D, N, sD, 0, (cx<float>*)B, nrhs, ldb,
                        strideB, batch, stream);
}

// ---------------------------------------------------------------------------
// Mixed precision.  A complex64 factor of `batch` matrices of order n lives
// in one buffer: the LU factors (batch n^2, column-major, leading dimension
// n), the pivots (batch n) and info (batch), and the inverted 64 x 64
// diagonal blocks of L and U (a strip of n x PB per matrix), each part
// 256-byte aligned.  The refinement takes the factor and its own workspace:
// FP64 residuals and complex64 corrections (batch n nrhs each) and the
// squared column norms.
// ---------------------------------------------------------------------------
static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t mixed_factor_bytes(int n, int batch)
{
    const size_t b = (size_t)batch;
    return align256(sizeof(cx<float>) * b * n * n) + align256(sizeof(int) * (b * n + b)) +
           align256(sizeof(cx<float>) * b * n * PB);
}

MixedFactor mixed_factor_at(void* base, int n, int batch)
{
    const size_t b = (size_t)batch;
    char* p = (char*)base;
    MixedFactor F;
    F.K32 = p;
    p += align256(sizeof(cx<float>) * b * n * n);
    F.ipiv = (int*)p;
    F.info = F.ipiv + b * n;
    p += align256(sizeof(int) * (b * n + b));
    F.D = p;
    return F;
}

size_t mixed_factor_ws_bytes(int n, int batch) { return align256(sizeof(cx<float>) * getrf_ws_elems<float>(n, batch)); }

size_t mixed_refine_bytes(const MixedGroup* g, int ng)
{
    size_t bytes = 0, norms = 0;
    for (int i = 0; i < ng; ++i) {
        const size_t br = (size_t)g[i].batch * g[i].n * g[i].nrhs;
        bytes += align256(sizeof(cx<double>) * br) + align256(sizeof(cx<float>) * br);
        norms += (size_t)g[i].batch * g[i].nrhs;
    }
    return bytes + align256(2 * sizeof(double) * norms);
}

// Workspace of factor and refinement of groups factored inside one call:
// per group its factor and getrf workspace, then the refinement's.
size_t mixed_grouped_bytes(const MixedGroup* g, int ng)
{
    size_t bytes = 256;
    for (int i = 0; i < ng; ++i) bytes += mixed_factor_bytes(g[i].n, g[i].batch) + mixed_factor_ws_bytes(g[i].n, g[i].batch);
    return bytes + mixed_refine_bytes(g, ng);
}

size_t mixed_workspace_bytes(int n, int nrhs, int batch)
{
    MixedGroup g{};
    g.n = n;
    g.nrhs = nrhs;
    g.batch = batch;
    return mixed_grouped_bytes(&g, 1);
}

// Complex64 factor of K on stream s: K32 = K rounded, partial-pivoting LU
// on FP32 MFMA (or the product form of the calling thread), inverted
// diagonal blocks for the substitutions.  `ws`: mixed_factor_ws_bytes.
int mixed_factor(const cplx* K, int n, int64_t lda, int64_t strideK, int batch, const MixedFactor& F, void* ws,
                 const LuSetup* tune, hipStream_t s)
{
    if (n <= 0 || batch <= 0) return 0;
    typedef cx<float> c64;
    const int64_t sK32 = (int64_t)n * n;
    const dim3 blk(256), gK((n + 255) / 256, n, batch);
    hipLaunchKernelGGL(z2c_kernel, gK, blk, 0, s, (const cx<double>*)K, lda, strideK, (c64*)F.K32, (int64_t)n, sK32,
                       n, n);
    hipLaunchKernelGGL(zero_ints_kernel, dim3((batch + 255) / 256), blk, 0, s, F.info, batch);
    int e;
    if ((e = getrf<float>((c64*)F.K32, n, n, sK32, F.ipiv, n, F.info, batch, (c64*)ws, tune, s))) return e;
    return invert_diag<float>((const c64*)F.K32, n, n, sK32, (c64*)F.D, n, (int64_t)n * PB, 0, batch, s);
}

// Refinement of K X = B for groups of complex128 systems with their complex64
// factors, each group on its own stream: x0 = LU32 \ b, then sweeps
// r = b - K x (FP64 MFMA product against the unrounded K) and
// x += LU32 \ r until max_j ||r_j|| / ||b_j|| <= tol or max_iter sweeps, per
// group.  The host waits for all streams once per sweep to read the
// residual norms of every group at once; at the first wait it also reads
// each factor's info: info_out[i] = first nonzero info of group i's
// matrices (1 + column of an exactly-zero complex64 pivot), else 0.
int mixed_refine(const MixedGroup* g, const MixedFactor* F, int ng, void* work, int max_iter, double tol,
                 int* iterations, double* residual, int* info_out)
{
    typedef cx<float> c64;
    typedef cx<double> c128;
    struct State {
        c128* Rz;
        c64* D;
        double* nb2;
        double* nr2;
        size_t norm_off;
        bool active;
    };
    std::vector<State> st(ng);
    char* p = (char*)(((uintptr_t)work + 255) & ~(uintptr_t)255);
    size_t nnorm = 0;
    for (int i = 0; i < ng; ++i) {
        const size_t br = (size_t)g[i].batch * g[i].n * g[i].nrhs;
        st[i].Rz = (c128*)p;
        p += align256(sizeof(c128) * br);
        st[i].D = (c64*)p;
        p += align256(sizeof(c64) * br);
        st[i].norm_off = nnorm;
        nnorm += (size_t)g[i].batch * g[i].nrhs;
    }
    double* nb2_all = (double*)p;
    double* nr2_all = nb2_all + nnorm;
    std::vector<double> hb(nnorm), hr(nnorm);
    int e;
    const dim3 blk(256);
    for (int i = 0; i < ng; ++i) {
        const MixedGroup& G = g[i];
        State& S = st[i];
        S.nb2 = nb2_all + S.norm_off;
        S.nr2 = nr2_all + S.norm_off;
        S.active = G.n > 0 && G.batch > 0 && G.nrhs > 0;
        iterations[i] = 0;
        residual[i] = 0.0;
        info_out[i] = 0;
        if (!S.active) continue;
        const int n = G.n, r = G.nrhs, batch = G.batch;
        const int64_t sK32 = (int64_t)n * n, sRD = (int64_t)n * r;
        const dim3 gB((n + 255) / 256, r, batch);
        // x0 = LU32 \ b
        hipLaunchKernelGGL(z2c_kernel, gB, blk, 0, G.s, (const c128*)G.B, G.ldb, G.strideB, S.D, (int64_t)n, sRD, n,
                           r);
        if ((e = getrs<float>((const c64*)F[i].K32, n, n, sK32, F[i].ipiv, n, (const c64*)F[i].D, n, (int64_t)n * PB,
                              0, S.D, r, n, sRD, batch, G.s)))
            return e;
        hipLaunchKernelGGL(c2z_add_kernel, gB, blk, 0, G.s, S.D, (int64_t)n, sRD, (c128*)G.X, G.ldx, G.strideX, n, r,
                           0);
        hipLaunchKernelGGL(colnorm2_kernel, dim3(r, batch), blk, 0, G.s, (const c128*)G.B, G.ldb, G.strideB, n, r,
                           S.nb2);
        if ((e = (int)hipGetLastError())) return e;
    }
    bool first = true;
    for (;;) {
        bool any = false;
        for (int i = 0; i < ng; ++i) {
            if (!st[i].active) continue;
            any = true;
            const MixedGroup& G = g[i];
            State& S = st[i];
            const int n = G.n, r = G.nrhs, batch = G.batch;
            const int64_t sRD = (int64_t)n * r;
            const dim3 gB((n + 255) / 256, r, batch);
            // r = b - K x  (FP64)
            hipLaunchKernelGGL(copy_z_kernel, gB, blk, 0, G.s, (const c128*)G.B, G.ldb, G.strideB, S.Rz, (int64_t)n,
                               sRD, n);
            if ((e = gemm<double>(n, r, n, -1.0, (const c128*)G.K, G.lda, G.strideK, (const c128*)G.X, G.ldx,
                                  G.strideX, S.Rz, n, sRD, 0, batch, G.s)))
                return e;
            hipLaunchKernelGGL(colnorm2_kernel, dim3(r, batch), blk, 0, G.s, S.Rz, (int64_t)n, sRD, n, r, S.nr2);
            if ((e = (int)hipGetLastError())) return e;
        }
        if (!any) break;
        for (int i = 0; i < ng; ++i)
            if (st[i].active && hipStreamSynchronize(g[i].s) != hipSuccess) return (int)hipErrorUnknown;
        if (first) {
            if (hipMemcpy(hb.data(), nb2_all, sizeof(double) * nnorm, hipMemcpyDeviceToHost) != hipSuccess)
                return (int)hipErrorUnknown;
            for (int i = 0; i < ng; ++i) {
                if (!st[i].active) continue;
                std::vector<int> inf(g[i].batch);
                if (hipMemcpy(inf.data(), F[i].info, sizeof(int) * g[i].batch, hipMemcpyDeviceToHost) != hipSuccess)
                    return (int)hipErrorUnknown;
                for (int v : inf)
                    if (v != 0) {
                        info_out[i] = v;
                        break;
                    }
            }
            first = false;
        }
        if (hipMemcpy(hr.data(), nr2_all, sizeof(double) * nnorm, hipMemcpyDeviceToHost) != hipSuccess)
            return (int)hipErrorUnknown;
        for (int i = 0; i < ng; ++i) {
            State& S = st[i];
            if (!S.active) continue;
            const MixedGroup& G = g[i];
            double worst = 0.0;
            for (size_t j = S.norm_off; j < S.norm_off + (size_t)G.batch * G.nrhs; ++j) {
                const double rel = hb[j] > 0 ? std::sqrt(hr[j] / hb[j]) : std::sqrt(hr[j]);
                worst = std::isnan(rel) ? INFINITY : std::max(worst, rel);
            }
            residual[i] = worst;
            // a singular complex64 factor cannot correct anything
            if (worst <= tol || iterations[i] >= max_iter || info_out[i] != 0) {
                S.active = false;
                continue;
            }
            // x += LU32 \ r
            const int n = G.n, r = G.nrhs, batch = G.batch;
            const int64_t sK32 = (int64_t)n * n, sRD = (int64_t)n * r;
            const dim3 gB((n + 255) / 256, r, batch);
            hipLaunchKernelGGL(z2c_kernel, gB, blk, 0, G.s, S.Rz, (int64_t)n, sRD, S.D, (int64_t)n, sRD, n, r);
            if ((e = getrs<float>((const c64*)F[i].K32, n, n, sK32, F[i].ipiv, n, (const c64*)F[i].D, n,
                                  (int64_t)n * PB, 0, S.D, r, n, sRD, batch, G.s)))
                return e;
            hipLaunchKernelGGL(c2z_add_kernel, gB, blk, 0, G.s, S.D, (int64_t)n, sRD, (c128*)G.X, G.ldx, G.strideX,
                               n, r, 1);
            if ((e = (int)hipGetLastError())) return e;
            ++iterations[i];
        }
    }
    return 0;
}

// Factor and refine groups inside one call: the factors and getrf
// workspaces at the start of `work` (mixed_grouped_bytes), each group's
// factorisation on its own stream.
int launch_mixed_grouped(const MixedGroup* g, int ng, void* work, const LuSetup* tune, int max_iter, double tol,
                         int* iterations, double* residual, int* info_out)
{
    std::vector<MixedFactor> F(ng);
    char* p = (char*)(((uintptr_t)work + 255) & ~(uintptr_t)255);
    int e;
    for (int i = 0; i < ng; ++i) {
        F[i] = mixed_factor_at(p, g[i].n, g[i].batch);
        p += mixed_factor_bytes(g[i].n, g[i].batch);
        void* ws = p;
        p += mixed_factor_ws_bytes(g[i].n, g[i].batch);
        if (g[i].n > 0 && g[i].batch > 0 && g[i].nrhs > 0 &&
            (e = mixed_factor(g[i].K, g[i].n, g[i].lda, g[i].strideK, g[i].batch, F[i], ws, tune, g[i].s)))
            return e;
    }
    return mixed_refine(g, F.data(), ng, p, max_iter, tol, iterations, residual, info_out);
}

// One factorisation workspace per group, back to back in `work`
// (getrf_workspace_bytes of each group, in group order).
int launch_getrf_lanes(const LuGroupArgs* g, int ng, const LuSetup* tune, void* work)
{
    int e;
    char* p = (char*)work;
    for (int i = 0; i < ng; ++i) {
        if (g[i].n <= 0 || g[i].batch <= 0) continue;
        hipLaunchKernelGGL(zero_ints_kernel, dim3((g[i].batch + 255) / 256), dim3(256), 0, g[i].s, g[i].info,
                           g[i].batch);
        if ((e = getrf<double>((cx<double>*)g[i].A, g[i].n, g[i].lda, g[i].strideA, g[i].ipiv, g[i].strideP,
                               g[i].info, g[i].batch, (cx<double>*)p, tune, g[i].s)))
            return e;
        p += getrf_workspace_bytes(g[i].n, g[i].batch);
    }
    return 0;
}

// One strip of inverted diagonal blocks per group, back to back in `work`
// (getrs_workspace_bytes of each group, in group order).
int launch_getrs_lanes(const LuGroupArgs* g, int ng, void* work)
{
    int e;
    char* p = (char*)work;
    for (int i = 0; i < ng; ++i) {
        if (g[i].n <= 0 || g[i].batch <= 0 || g[i].nrhs <= 0) continue;
        if ((e = launch_getrs((const cplx*)g[i].A, g[i].n, g[i].lda, g[i].strideA, g[i].ipiv, g[i].strideP,
                              (cplx*)g[i].B, g[i].nrhs, g[i].ldb, g[i].strideB, g[i].batch, (cplx*)p, g[i].s)))
            return e;
        p += getrs_workspace_bytes(g[i].n, g[i].batch);
    }
    return 0;
}

}  // namespace cpt
