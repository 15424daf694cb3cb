// Batched dense complex128 LU with partial pivoting and the matching
// multi-right-hand-side solve, for gfx950.
//
// Replaces LAPACK zgetrf / zgetrs as reached through scipy.linalg.lu_factor
// and lu_solve (capytaine tools/block_circulant_matrices.py:361-385, called
// from bem/engines.py:326-338).  Same conventions as LAPACK: column-major
// storage, P A = L U with L unit lower triangular, pivots chosen by the
// largest |Re|+|Im| (first index on ties), ipiv[i] = row swapped with row i
// (0-based here), info = 1 + first zero pivot.
//
// Right-looking blocked algorithm, one launch per step for the whole batch:
//   1. panel: one work-group per matrix factors the NB-column panel; the
//      panel is processed in register-resident sub-blocks of SB columns
//      (left-looking inside the panel), with a block-wide pivot search;
//   2. row interchanges applied to the columns outside the panel;
//   3. U12 = L11^-1 A12 (unit lower triangular solve, L11 staged in LDS);
//   4. A22 -= L21 U12 with an FP64-MFMA complex GEMM (v_mfma_f64_16x16x4).
#include <hip/hip_runtime.h>
#include <climits>
#include "cpt_internal.hpp"

namespace cpt {
namespace {

typedef double dvec4 __attribute__((ext_vector_type(4)));

struct z2 {
    double re, im;
};

__device__ __forceinline__ z2 zmul(z2 a, z2 b) { return z2{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ z2 zsub(z2 a, z2 b) { return z2{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ z2 zrcp(z2 a)
{
    // Smith's reciprocal, no overflow for |a| up to the double range.
    if (fabs(a.re) >= fabs(a.im)) {
        const double r = a.im / a.re, d = a.re + a.im * r;
        return z2{1.0 / d, -r / d};
    } else {
        const double r = a.re / a.im, d = a.re * r + a.im;
        return z2{r / d, -1.0 / d};
    }
}
__device__ __forceinline__ double cabs1(z2 a) { return fabs(a.re) + fabs(a.im); }

__device__ __forceinline__ z2 ld(const cplx* p) { return z2{p->re, p->im}; }
__device__ __forceinline__ void st(cplx* p, z2 v)
{
    p->re = v.re;
    p->im = v.im;
}

constexpr int NB = 64;  // panel width

// ---------------------------------------------------------------------------
// 1. Panel factorisation
// ---------------------------------------------------------------------------
template <int NT, int RPT, int SB>
__global__ void __launch_bounds__(NT) panel_kernel(cplx* __restrict__ Abase, int64_t lda, int64_t strideA, int N,
                                                   int k0, int kb, int* __restrict__ ipiv_base, int64_t strideP,
                                                   int* __restrict__ info_base)
{
    cplx* A = Abase + (size_t)blockIdx.x * strideA;
    int* ipiv = ipiv_base + (size_t)blockIdx.x * strideP;
    const int M = N - k0;
    const int tid = threadIdx.x;
    constexpr int NW = NT / 64;

    __shared__ z2 sU[NB * SB];
    __shared__ z2 sRow[SB];
    __shared__ z2 sSwap[2][SB];
    __shared__ double sBest[NW];
    __shared__ int sIdx[NW];
    __shared__ int sPiv;

    z2 v[RPT][SB];

    for (int c0 = k0; c0 < k0 + kb; c0 += SB) {
        const int sbw = min(SB, k0 + kb - c0);
        const int t0 = c0 - k0;
        // (a) load the sub-block
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int t = tid + NT * q;
#pragma unroll
            for (int s = 0; s < SB; ++s)
                v[q][s] = (t < M && s < sbw) ? ld(A + (k0 + t) + (size_t)(c0 + s) * lda) : z2{0.0, 0.0};
        }
        if (t0 > 0) {
            // (b) U rows of the sub-block: solve with the unit lower L11 (t0 x t0)
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int t = tid + NT * q;
                if (t < t0)
#pragma unroll
                    for (int s = 0; s < SB; ++s) sU[t * SB + s] = v[q][s];
            }
            __syncthreads();
            if (tid < sbw) {
                const int s = tid;
                for (int i = 1; i < t0; ++i) {
                    z2 acc = sU[i * SB + s];
                    for (int j = 0; j < i; ++j)
                        acc = zsub(acc, zmul(ld(A + (k0 + i) + (size_t)(k0 + j) * lda), sU[j * SB + s]));
                    sU[i * SB + s] = acc;
                }
            }
            __syncthreads();
            // (c) left-looking update of the rows below
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const int t = tid + NT * q;
                if (t < t0) {
#pragma unroll
                    for (int s = 0; s < SB; ++s) v[q][s] = sU[t * SB + s];
                } else if (t < M) {
                    for (int j = 0; j < t0; ++j) {
                        const z2 l = ld(A + (k0 + t) + (size_t)(k0 + j) * lda);
#pragma unroll
                        for (int s = 0; s < SB; ++s) v[q][s] = zsub(v[q][s], zmul(l, sU[j * SB + s]));
                    }
                }
            }
        }
        // (d) factor the sub-block column by column
#pragma unroll
        for (int s = 0; s < SB; ++s) {
            if (s < sbw) {
                const int td = t0 + s;
                double best = -1.0;
                int bidx = INT_MAX;
#pragma unroll
                for (int q = 0; q < RPT; ++q) {
                    const int t = tid + NT * q;
                    if (t >= td && t < M) {
                        const double a = cabs1(v[q][s]);
                        if (a > best) {
                            best = a;
                            bidx = t;
                        }
                    }
                }
                for (int off = 32; off > 0; off >>= 1) {
                    const double ob = __shfl_xor(best, off);
                    const int oi = __shfl_xor(bidx, off);
                    if (ob > best || (ob == best && oi < bidx)) {
                        best = ob;
                        bidx = oi;
                    }
                }
                if ((tid & 63) == 0) {
                    sBest[tid >> 6] = best;
                    sIdx[tid >> 6] = bidx;
                }
                __syncthreads();
                if (tid == 0) {
                    double b = sBest[0];
                    int bi = sIdx[0];
                    for (int w = 1; w < NW; ++w)
                        if (sBest[w] > b || (sBest[w] == b && sIdx[w] < bi)) {
                            b = sBest[w];
                            bi = sIdx[w];
                        }
                    int p = bi;
                    if (!(b > 0.0)) {  // zero column: no interchange, report singularity
                        p = td;
                        if (info_base[blockIdx.x] == 0) info_base[blockIdx.x] = c0 + s + 1;
                    }
                    sPiv = p;
                    ipiv[c0 + s] = k0 + p;
                }
                __syncthreads();
                const int p = sPiv;
                if (p != td) {
#pragma unroll
                    for (int q = 0; q < RPT; ++q) {
                        const int t = tid + NT * q;
                        if (t == td)
#pragma unroll
                            for (int s2 = 0; s2 < SB; ++s2) sSwap[0][s2] = v[q][s2];
                        if (t == p)
#pragma unroll
                            for (int s2 = 0; s2 < SB; ++s2) sSwap[1][s2] = v[q][s2];
                    }
                    // interchange in the other columns of the panel (global memory)
                    for (int c = tid; c < kb; c += NT) {
                        const int col = k0 + c;
                        if (col < c0 || col >= c0 + sbw) {
                            cplx* a = A + (k0 + td) + (size_t)col * lda;
                            cplx* b = A + (k0 + p) + (size_t)col * lda;
                            const z2 x = ld(a), y = ld(b);
                            st(a, y);
                            st(b, x);
                        }
                    }
                    __syncthreads();
#pragma unroll
                    for (int q = 0; q < RPT; ++q) {
                        const int t = tid + NT * q;
                        if (t == td)
#pragma unroll
                            for (int s2 = 0; s2 < SB; ++s2) v[q][s2] = sSwap[1][s2];
                        if (t == p)
#pragma unroll
                            for (int s2 = 0; s2 < SB; ++s2) v[q][s2] = sSwap[0][s2];
                    }
                }
                // broadcast the pivot row
#pragma unroll
                for (int q = 0; q < RPT; ++q)
                    if (tid + NT * q == td)
#pragma unroll
                        for (int s2 = 0; s2 < SB; ++s2) sRow[s2] = v[q][s2];
                __syncthreads();
                const z2 piv = sRow[s];
                const bool nonzero = (piv.re != 0.0 || piv.im != 0.0);
                const z2 r = nonzero ? zrcp(piv) : z2{1.0, 0.0};
#pragma unroll
                for (int q = 0; q < RPT; ++q) {
                    const int t = tid + NT * q;
                    if (t > td && t < M) {
                        const z2 l = nonzero ? zmul(v[q][s], r) : v[q][s];
                        v[q][s] = l;
#pragma unroll
                        for (int s2 = s + 1; s2 < SB; ++s2) v[q][s2] = zsub(v[q][s2], zmul(l, sRow[s2]));
                    }
                }
                __syncthreads();
            }
        }
        // (e) store the sub-block
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int t = tid + NT * q;
            if (t < M)
#pragma unroll
                for (int s = 0; s < SB; ++s)
                    if (s < sbw) st(A + (k0 + t) + (size_t)(c0 + s) * lda, v[q][s]);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// 2. Row interchanges ipiv[r0 .. r1) applied to the columns [0, c_lo) and
//    [c_hi, ncols) of a column-major array (LAPACK zlaswp).
// ---------------------------------------------------------------------------
__global__ void laswp_kernel(cplx* __restrict__ Abase, int64_t lda, int64_t strideA, int ncols, int c_lo, int c_hi,
                             const int* __restrict__ ipiv_base, int64_t strideP, int r0, int r1)
{
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncols || (col >= c_lo && col < c_hi)) return;
    cplx* A = Abase + (size_t)blockIdx.y * strideA + (size_t)col * lda;
    const int* ipiv = ipiv_base + (size_t)blockIdx.y * strideP;
    for (int i = r0; i < r1; ++i) {
        const int p = ipiv[i];
        if (p != i) {
            const z2 x = ld(A + i), y = ld(A + p);
            st(A + i, y);
            st(A + p, x);
        }
    }
}

// ---------------------------------------------------------------------------
// 3. Triangular solve with a kb x kb diagonal block T (kb <= NB), in place on
//    the kb rows of B, one thread per column of B.
//    UPPER = false: T unit lower triangular; UPPER = true: T upper, non-unit.
// ---------------------------------------------------------------------------
template <bool UPPER>
__global__ void __launch_bounds__(256) trsm_kernel(const cplx* __restrict__ Tbase, int64_t ldt, int64_t strideT,
                                                   int kb, cplx* __restrict__ Bbase, int64_t ldb, int64_t strideB,
                                                   int ncols)
{
    __shared__ z2 sT[NB * NB];
    __shared__ z2 sD[NB];
    const cplx* T = Tbase + (size_t)blockIdx.y * strideT;
    for (int e = threadIdx.x; e < kb * kb; e += blockDim.x) {
        const int i = e % kb, j = e / kb;
        sT[i + NB * j] = ld(T + i + (size_t)j * ldt);
    }
    __syncthreads();
    if (UPPER) {
        for (int i = threadIdx.x; i < kb; i += blockDim.x) sD[i] = zrcp(sT[i + NB * i]);
        __syncthreads();
    }
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncols) return;
    cplx* B = Bbase + (size_t)blockIdx.y * strideB + (size_t)col * ldb;
    z2 x[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) x[i] = (i < kb) ? ld(B + i) : z2{0.0, 0.0};
    if (!UPPER) {
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (j < kb) {
                const z2 xj = x[j];
#pragma unroll
                for (int i = j + 1; i < NB; ++i)
                    if (i < kb) x[i] = zsub(x[i], zmul(sT[i + NB * j], xj));
            }
        }
    } else {
#pragma unroll
        for (int j = NB - 1; j >= 0; --j) {
            if (j < kb) {
                x[j] = zmul(x[j], sD[j]);
                const z2 xj = x[j];
#pragma unroll
                for (int i = 0; i < j; ++i) x[i] = zsub(x[i], zmul(sT[i + NB * j], xj));
            }
        }
    }
#pragma unroll
    for (int i = 0; i < NB; ++i)
        if (i < kb) st(B + i, x[i]);
}

// ---------------------------------------------------------------------------
// 4. C (M x N) += alpha * A (M x K) * B (K x N), complex128, column-major,
//    batched by blockIdx.z.  Work-group tile 64 x 64, four wavefronts of
//    32 x 32, each a 2 x 2 grid of 16 x 16 FP64 MFMA tiles; the complex
//    product is four real MFMAs per tile pair (re*re - im*im, re*im + im*re).
//    Operand fragments are read straight from global memory (L1/L2) in the
//    MFMA lane layout: A[i][k] at lane (i + 16 k), B[k][j] at lane (j + 16 k);
//    the accumulator holds D[4 (lane/16) + v][lane % 16], v = 0..3.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) zgemm_kernel(int M, int N, int K, double alpha, const cplx* __restrict__ Abase,
                                                    int64_t lda, int64_t strideA, const cplx* __restrict__ Bbase,
                                                    int64_t ldb, int64_t strideB, cplx* __restrict__ Cbase,
                                                    int64_t ldc, int64_t strideC, int beta_zero)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row0 = blockIdx.x * 64 + (wave & 1) * 32;
    const int col0 = blockIdx.y * 64 + (wave >> 1) * 32;
    const cplx* A = Abase + (size_t)blockIdx.z * strideA;
    const cplx* B = Bbase + (size_t)blockIdx.z * strideB;
    cplx* C = Cbase + (size_t)blockIdx.z * strideC;
    const int li = lane & 15, lk = lane >> 4;

    dvec4 cr[2][2], ci[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            cr[a][b] = dvec4{0.0, 0.0, 0.0, 0.0};
            ci[a][b] = dvec4{0.0, 0.0, 0.0, 0.0};
        }
    int arow[2], bcol[2];
    bool aok[2], bok[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        arow[t] = row0 + 16 * t + li;
        bcol[t] = col0 + 16 * t + li;
        aok[t] = arow[t] < M;
        bok[t] = bcol[t] < N;
    }
    for (int k = 0; k < K; k += 4) {
        const int kk = k + lk;
        const bool kok = kk < K;
        double ar[2], ai[2], br[2], bi[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            if (aok[t] && kok) {
                const cplx a = A[arow[t] + (size_t)kk * lda];
                ar[t] = a.re;
                ai[t] = a.im;
            } else {
                ar[t] = ai[t] = 0.0;
            }
            if (bok[t] && kok) {
                const cplx b = B[kk + (size_t)bcol[t] * ldb];
                br[t] = b.re;
                bi[t] = b.im;
            } else {
                br[t] = bi[t] = 0.0;
            }
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[b], cr[a][b], 0, 0, 0);
                cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai[a], bi[b], cr[a][b], 0, 0, 0);
                ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], bi[b], ci[a][b], 0, 0, 0);
                ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], br[b], ci[a][b], 0, 0, 0);
            }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = row0 + 16 * a + 4 * v + lk;
                const int col = col0 + 16 * b + li;
                if (row < M && col < N) {
                    cplx* c = C + row + (size_t)col * ldc;
                    double re = beta_zero ? 0.0 : c->re;
                    double im = beta_zero ? 0.0 : c->im;
                    re += alpha * cr[a][b][v];
                    im += alpha * ci[a][b][v];
                    c->re = re;
                    c->im = im;
                }
            }
}

template <int NT, int RPT, int SB>
void launch_panel(cplx* A, int64_t lda, int64_t sA, int N, int k0, int kb, int* ipiv, int64_t sP, int* info,
                  int batch, hipStream_t stream)
{
    hipLaunchKernelGGL((panel_kernel<NT, RPT, SB>), dim3(batch), dim3(NT), 0, stream, A, lda, sA, N, k0, kb, ipiv,
                       sP, info);
}

int panel_dispatch(cplx* A, int64_t lda, int64_t sA, int N, int k0, int kb, int* ipiv, int64_t sP, int* info,
                   int batch, hipStream_t stream)
{
    // 512 threads (two wavefronts per SIMD, up to 256 VGPRs each); the
    // register-resident sub-block holds RPT rows x SB columns per thread,
    // RPT * SB <= 48 complex values.
    const int M = N - k0;
    if (M <= 512) launch_panel<512, 1, 8>(A, lda, sA, N, k0, kb, ipiv, sP, info, batch, stream);
    else if (M <= 1024) launch_panel<512, 2, 8>(A, lda, sA, N, k0, kb, ipiv, sP, info, batch, stream);
    else if (M <= 2048) launch_panel<512, 4, 8>(A, lda, sA, N, k0, kb, ipiv, sP, info, batch, stream);
    else if (M <= 3072) launch_panel<512, 6, 8>(A, lda, sA, N, k0, kb, ipiv, sP, info, batch, stream);
    else if (M <= 4096) launch_panel<512, 8, 4>(A, lda, sA, N, k0, kb, ipiv, sP, info, batch, stream);
    else if (M <= 6144) launch_panel<512, 12, 4>(A, lda, sA, N, k0, kb, ipiv, sP, info, batch, stream);
    else if (M <= 8192) launch_panel<512, 16, 2>(A, lda, sA, N, k0, kb, ipiv, sP, info, batch, stream);
    else if (M <= 12288) launch_panel<512, 24, 2>(A, lda, sA, N, k0, kb, ipiv, sP, info, batch, stream);
    else if (M <= 16384) launch_panel<512, 32, 1>(A, lda, sA, N, k0, kb, ipiv, sP, info, batch, stream);
    else if (M <= 24576) launch_panel<512, 48, 1>(A, lda, sA, N, k0, kb, ipiv, sP, info, batch, stream);
    else return (int)hipErrorInvalidValue;
    return (int)hipGetLastError();
}

}  // namespace

int launch_zgemm(int M, int N, int K, double alpha, const cplx* A, int64_t lda, int64_t sA, const cplx* B,
                 int64_t ldb, int64_t sB, cplx* C, int64_t ldc, int64_t sC, int beta_zero, int batch,
                 hipStream_t stream)
{
    if (M <= 0 || N <= 0 || batch <= 0) return 0;
    dim3 grid((M + 63) / 64, (N + 63) / 64, batch);
    hipLaunchKernelGGL(zgemm_kernel, grid, dim3(256), 0, stream, M, N, K, alpha, A, lda, sA, B, ldb, sB, C, ldc, sC,
                       beta_zero);
    return (int)hipGetLastError();
}

int launch_getrf(cplx* A, int N, int64_t lda, int64_t strideA, int* ipiv, int64_t strideP, int* info, int batch,
                 hipStream_t stream)
{
    if (N <= 0 || batch <= 0) return 0;
    int e;
    for (int k0 = 0; k0 < N; k0 += NB) {
        const int kb = min(NB, N - k0);
        if ((e = panel_dispatch(A, lda, strideA, N, k0, kb, ipiv, strideP, info, batch, stream))) return e;
        const dim3 gs((N + 255) / 256, batch);
        hipLaunchKernelGGL(laswp_kernel, gs, dim3(256), 0, stream, A, lda, strideA, N, k0, k0 + kb, ipiv, strideP, k0,
                           k0 + kb);
        if ((e = (int)hipGetLastError())) return e;
        const int rest = N - k0 - kb;
        if (rest > 0) {
            cplx* A11 = A + k0 + (size_t)k0 * lda;
            cplx* A12 = A + k0 + (size_t)(k0 + kb) * lda;
            cplx* A21 = A + (k0 + kb) + (size_t)k0 * lda;
            cplx* A22 = A + (k0 + kb) + (size_t)(k0 + kb) * lda;
            hipLaunchKernelGGL(trsm_kernel<false>, dim3((rest + 255) / 256, batch), dim3(256), 0, stream, A11, lda,
                               strideA, kb, A12, lda, strideA, rest);
            if ((e = (int)hipGetLastError())) return e;
            if ((e = launch_zgemm(rest, rest, kb, -1.0, A21, lda, strideA, A12, lda, strideA, A22, lda, strideA, 0,
                                  batch, stream)))
                return e;
        }
    }
    return 0;
}

int launch_getrs(const cplx* LU, int N, int64_t lda, int64_t strideA, const int* ipiv, int64_t strideP, cplx* B,
                 int nrhs, int64_t ldb, int64_t strideB, int batch, hipStream_t stream)
{
    if (N <= 0 || nrhs <= 0 || batch <= 0) return 0;
    int e;
    hipLaunchKernelGGL(laswp_kernel, dim3((nrhs + 255) / 256, batch), dim3(256), 0, stream, B, ldb, strideB, nrhs, 0,
                       0, ipiv, strideP, 0, N);
    if ((e = (int)hipGetLastError())) return e;
    // forward substitution with the unit lower factor
    for (int k0 = 0; k0 < N; k0 += NB) {
        const int kb = min(NB, N - k0);
        hipLaunchKernelGGL(trsm_kernel<false>, dim3((nrhs + 255) / 256, batch), dim3(256), 0, stream,
                           LU + k0 + (size_t)k0 * lda, lda, strideA, kb, B + k0, ldb, strideB, nrhs);
        if ((e = (int)hipGetLastError())) return e;
        const int rest = N - k0 - kb;
        if (rest > 0 &&
            (e = launch_zgemm(rest, nrhs, kb, -1.0, LU + (k0 + kb) + (size_t)k0 * lda, lda, strideA, B + k0, ldb,
                              strideB, B + k0 + kb, ldb, strideB, 0, batch, stream)))
            return e;
    }
    // backward substitution with the upper factor
    const int nblk = (N + NB - 1) / NB;
    for (int blk = nblk - 1; blk >= 0; --blk) {
        const int k0 = blk * NB;
        const int kb = min(NB, N - k0);
        hipLaunchKernelGGL(trsm_kernel<true>, dim3((nrhs + 255) / 256, batch), dim3(256), 0, stream,
                           LU + k0 + (size_t)k0 * lda, lda, strideA, kb, B + k0, ldb, strideB, nrhs);
        if ((e = (int)hipGetLastError())) return e;
        if (k0 > 0 && (e = launch_zgemm(k0, nrhs, kb, -1.0, LU + (size_t)k0 * lda, lda, strideA, B + k0, ldb, strideB,
                                        B, ldb, strideB, 0, batch, stream)))
            return e;
    }
    return 0;
}

}  // namespace cpt
