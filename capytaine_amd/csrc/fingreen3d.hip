// FinGreen3D (the HAMS finite-depth Green function) on the device: the wave
// part of each (collocation point, panel) pair, integrated with the panel's
// quadrature points and added to S and K after the Rankine terms
// (integral_of_wave_part_fingreen3D, Green_wave.f90:545-616, called from
// build_matrices, matrices.f90:199-209).
//
// One thread per pair, with a grid stride.  The series of the near regions
// (R/h < 0.5) end in an epsilon-algorithm limit over up to 99 partial sums;
// those tables live in a workspace of FG_WS_DOUBLES doubles per thread,
// interleaved so that a wavefront's accesses to one entry are contiguous.
#include "cpt_internal.hpp"
#include "green.hpp"
#include "fingreen3d.hpp"

namespace cpt {

static_assert(FG_WS_DOUBLES == fg::WS_ARRAYS * fg::LIMES_LEN, "workspace layout");

namespace {

// The wave part of pair (I, J), integrated over the panel's quadrature
// points and added to S and K.  CLASS 0..2: every quadrature point of the
// pair lies in that series' region; 3: any (Linton's form, or mixed).
template <int CLASS>
__device__ __forceinline__ void pair_wave(const FgArgs& A, int I, int J, double* ws, size_t st)
{
    const double x0 = A.colloc[3 * (size_t)I], x1 = A.colloc[3 * (size_t)I + 1], x2 = A.colloc[3 * (size_t)I + 2];
    double gre = 0.0, gim = 0.0, nre[3] = {0.0, 0.0, 0.0}, nim[3] = {0.0, 0.0, 0.0};
    for (int q = 0; q < A.nq; ++q) {
        const double* qp = A.quad + 4 * ((size_t)J * A.nq + q);
        const double dx = x0 - qp[0], dy = x1 - qp[1], w = qp[3];
        const double r = sqrt(dx * dx + dy * dy);
        // the reference passes the panel's depth first for the adjoint
        // double layer and the collocation point's first otherwise
        const double zf = A.adjoint ? qp[2] : x2, zp = A.adjoint ? x2 : qp[2];
        // the factor tables of the zf and zp points
        const size_t pc = (size_t)I, pq = (size_t)A.n + (size_t)J * A.nq + q;
        const fg::TrigTable tr{A.root_tab, A.tab + (A.adjoint ? pq : pc), A.tab + (A.adjoint ? pc : pq), A.npts};
        const fg::Grn g = CLASS == 3
                              ? fg::point(r, zf, zp, A.v, A.roots, A.depth, tr, ws, st)
                              : fg::point_in<CLASS == 3 ? 0 : CLASS>(r, zf, zp, A.v, A.roots, A.depth, tr, ws, st);
        double drdx1 = 0.0, drdx2 = 0.0;
        if (fabs(r) > 16 * DBL_EPS) {
            drdx1 = A.adjoint ? dx / r : -dx / r;
            drdx2 = A.adjoint ? dy / r : -dy / r;
        }
        gre = gre + g.g[0].re * w;
        gim = gim + g.g[0].im * w;
        nre[0] = nre[0] + drdx1 * g.g[1].re * w;
        nim[0] = nim[0] + drdx1 * g.g[1].im * w;
        nre[1] = nre[1] + drdx2 * g.g[1].re * w;
        nim[1] = nim[1] + drdx2 * g.g[1].im * w;
        nre[2] = nre[2] + g.g[2].re * w;
        nim[2] = nim[2] + g.g[2].im * w;
    }
    cplx* s = A.S + (size_t)I + (size_t)A.ldS * J;
    s->re += M1_4PI * gre;
    s->im += M1_4PI * gim;
    if (A.nk == 1) {
        const double* dn = A.dpn + 3 * (size_t)(A.adjoint ? I : J);
        cplx* k = A.K + (size_t)I + (size_t)A.ldK * J;
        k->re += M1_4PI * (dn[0] * nre[0] + dn[1] * nre[1] + dn[2] * nre[2]);
        k->im += M1_4PI * (dn[0] * nim[0] + dn[1] * nim[1] + dn[2] * nim[2]);
    } else {
        cplx* k = A.K + 3 * ((size_t)I + (size_t)A.ldK * J);
        for (int c = 0; c < 3; ++c) {
            k[c].re += M1_4PI * nre[c];
            k[c].im += M1_4PI * nim[c];
        }
    }
}

// The per-root factors of the series (pm, ac, umt by term index).
__global__ void fingreen3d_root_table_kernel(const double* __restrict__ wvn, double H, double V, double* root)
{
    const int i = threadIdx.x;
    if (i >= fg::TAB_TERMS) return;
    const fg::TrigDirect d{0.0, 0.0, H, V, wvn};
    const bool used = i >= 2;
    root[i] = used ? d.pm(i) : 0.0;
    root[fg::TAB_TERMS + i] = used ? d.ac(i) : 0.0;
    root[2 * fg::TAB_TERMS + i] = used ? d.umt(i) : 0.0;
}

// The per-point factors cos/sin of k (z + h) and of umt (z + h) for every
// term, with the expressions the series use; points are the collocation
// points, then the quadrature points.
__global__ void __launch_bounds__(256) fingreen3d_point_table_kernel(const double* __restrict__ colloc, int n,
                                                                     const double* __restrict__ quad,
                                                                     const double* __restrict__ wvn, double H,
                                                                     double* tab, size_t npts)
{
    const size_t total = npts * fg::TAB_TERMS;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (size_t)gridDim.x * blockDim.x) {
        const size_t p = idx % npts;
        const int i = (int)(idx / npts);
        if (i < 2) continue;
        const double z = p < (size_t)n ? colloc[3 * p + 2] : quad[4 * (p - n) + 2];
        const fg::TrigDirect d{z, z, H, 0.0, wvn};
        tab[(4 * (size_t)i + 0) * npts + p] = d.cf(i);
        tab[(4 * (size_t)i + 1) * npts + p] = d.sf(i);
        tab[(4 * (size_t)i + 2) * npts + p] = d.ctf(i);
        tab[(4 * (size_t)i + 3) * npts + p] = d.stf(i);
    }
}

__global__ void fingreen3d_zero_counts_kernel(unsigned* counts)
{
    if (threadIdx.x < 3) counts[threadIdx.x] = 0u;
}

// Pass 1 over the pairs of panels [j0, j0 + np / n): the pairs whose every
// quadrature point is in the eigenfunction region (R/h >= 0.5, the far
// field) are evaluated here, with that series only; the others are appended
// to the list of their series (wave-aggregated atomics).  Whole wavefronts
// step together so that the ballots see every lane.
__global__ void __launch_bounds__(256) fingreen3d_far_kernel(FgArgs A, int j0, int np)
{
    const int lane = threadIdx.x & 63;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    unsigned* counts = A.lists + 3 * (size_t)A.list_cap;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < np; base += stride) {
        const int p = (int)(base + lane);
        const bool valid = p < np;
        int cls = 0;
        int I = 0, J = 0;
        if (valid) {
            I = p % A.n;
            J = j0 + p / A.n;
            const double x0 = A.colloc[3 * (size_t)I], x1 = A.colloc[3 * (size_t)I + 1];
            int lo = 4, hi = -1;
            for (int q = 0; q < A.nq; ++q) {
                const double* qp = A.quad + 4 * ((size_t)J * A.nq + q);
                const double dx = x0 - qp[0], dy = x1 - qp[1];
                const int r = fg::region(sqrt(dx * dx + dy * dy), A.depth);
                lo = min(lo, r);
                hi = max(hi, r);
            }
            cls = (lo == hi && hi < 3) ? hi : 3;
        }
        if (valid && cls == 0) pair_wave<0>(A, I, J, nullptr, 0);
#pragma unroll
        for (int c = 1; c <= 3; ++c) {
            const bool want = valid && cls == c;
            const unsigned long long m = __ballot(want);
            if (m == 0ull) continue;
            const int leader = __ffsll((long long)m) - 1;
            unsigned slot = 0u;
            if (lane == leader) slot = atomicAdd(counts + (c - 1), (unsigned)__popcll(m));
            slot = __shfl(slot, leader);
            if (want) {
                const unsigned rank = (unsigned)__popcll(m & ((1ull << lane) - 1ull));
                A.lists[(size_t)(c - 1) * A.list_cap + slot + rank] = (unsigned)p;
            }
        }
    }
}

// Pass 2: the pairs of one series' list.
template <int CLASS>
__device__ __forceinline__ void list_pass(const FgArgs& A, int j0)
{
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= A.nthreads) return;
    const unsigned* list = A.lists + (size_t)(CLASS - 1) * A.list_cap;
    const int count = (int)A.lists[3 * (size_t)A.list_cap + (CLASS - 1)];
    double* ws = A.ws + tid;
    for (int e = tid; e < count; e += A.nthreads) {
        const int p = (int)list[e];
        pair_wave<CLASS>(A, p % A.n, j0 + p / A.n, ws, (size_t)A.nthreads);
    }
}

template <int CLASS>
__global__ void __launch_bounds__(256) fingreen3d_list_kernel(FgArgs A, int j0)
{
    list_pass<CLASS>(A, j0);
}

// all four series inline: one wavefront per SIMD and its whole register file
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) fingreen3d_points_kernel(int np, const double* __restrict__ rr,
                                                                const double* __restrict__ zf,
                                                                const double* __restrict__ zp, double v,
                                                                const double* __restrict__ roots, double depth,
                                                                cplx* __restrict__ grn, double* ws, int nthreads)
{
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= nthreads) return;
    for (int p = tid; p < np; p += nthreads) {
        const fg::Grn g = fg::point(rr[p], zf[p], zp[p], v, roots, depth, ws + tid, (size_t)nthreads);
        for (int c = 0; c < 3; ++c) {
            grn[3 * (size_t)p + c].re = g.g[c].re;
            grn[3 * (size_t)p + c].im = g.g[c].im;
        }
    }
}

}  // namespace

// A grid of at most 64 work-groups of 256 per 256 CUs: enough waves to hide
// the long dependent chains of the series, a workspace of 210 MB at most.
int fingreen3d_threads(int64_t items)
{
    const int64_t cap = 256 * 256;
    const int64_t t = ((items + 255) / 256) * 256;
    return (int)(t < cap ? (t > 256 ? t : 256) : cap);
}

size_t fingreen3d_table_bytes(size_t npts)
{
    return sizeof(double) * (4 * (size_t)fg::TAB_TERMS * npts + 3 * fg::TAB_TERMS);
}

size_t fingreen3d_list_bytes(int n)
{
    const size_t cap = (size_t)(n > FG_CHUNK_PAIRS ? n : FG_CHUNK_PAIRS);
    return sizeof(unsigned) * (3 * cap + 64);
}

// Over chunks of whole panel columns (at most FG_CHUNK_PAIRS pairs, or one
// column): the far pass, then the three series' lists.  The lists' counts
// stay on the device; each list kernel reads its own.
int launch_fingreen3d_pairs(const FgArgs& A, hipStream_t stream)
{
    if (A.n <= 0 || A.m <= 0) return 0;
    TimedLaunch timed(stream, TK_ASSEMBLY, (double)A.n * A.m, 32.0 * A.n * A.m);
    hipLaunchKernelGGL(fingreen3d_root_table_kernel, dim3(1), dim3(128), 0, stream, A.roots, A.depth, A.v,
                       A.root_tab);
    {
        const size_t total = A.npts * fg::TAB_TERMS;
        const size_t blocks = (total + 255) / 256;
        hipLaunchKernelGGL(fingreen3d_point_table_kernel, dim3((unsigned)(blocks < 8192 ? blocks : 8192)), dim3(256),
                           0, stream, A.colloc, A.n, A.quad, A.roots, A.depth, A.tab, A.npts);
    }
    const int cols = A.n >= FG_CHUNK_PAIRS ? 1 : FG_CHUNK_PAIRS / A.n;
    for (int j0 = 0; j0 < A.m; j0 += cols) {
        const int nc = A.m - j0 < cols ? A.m - j0 : cols;
        const int np = A.n * nc;
        hipLaunchKernelGGL(fingreen3d_zero_counts_kernel, dim3(1), dim3(64), 0, stream,
                           A.lists + 3 * (size_t)A.list_cap);
        const int far_blocks = (np + 255) / 256 < 4096 ? (np + 255) / 256 : 4096;
        hipLaunchKernelGGL(fingreen3d_far_kernel, dim3(far_blocks), dim3(256), 0, stream, A, j0, np);
        hipLaunchKernelGGL(fingreen3d_list_kernel<1>, dim3(A.nthreads / 256), dim3(256), 0, stream, A, j0);
        hipLaunchKernelGGL(fingreen3d_list_kernel<2>, dim3(A.nthreads / 256), dim3(256), 0, stream, A, j0);
        hipLaunchKernelGGL(fingreen3d_list_kernel<3>, dim3(A.nthreads / 256), dim3(256), 0, stream, A, j0);
    }
    return (int)hipGetLastError();
}

int launch_fingreen3d_points(int np, const double* rr, const double* zf, const double* zp, double v,
                             const double* roots, double depth, cplx* grn, double* ws, int nthreads,
                             hipStream_t stream)
{
    if (np <= 0) return 0;
    hipLaunchKernelGGL(fingreen3d_points_kernel, dim3(nthreads / 256), dim3(256), 0, stream, np, rr, zf, zp, v,
                       roots, depth, grn, ws, nthreads);
    return (int)hipGetLastError();
}

}  // namespace cpt
