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

__global__ void __launch_bounds__(256) fingreen3d_pairs_kernel(FgArgs A)
{
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= A.nthreads) return;
    double* ws = A.ws + tid;
    const size_t st = (size_t)A.nthreads;
    const int64_t npairs = (int64_t)A.n * A.m;
    for (int64_t p = tid; p < npairs; p += A.nthreads) {
        const int I = (int)(p % A.n), J = (int)(p / A.n);
        const double x0 = A.colloc[3 * (size_t)I], x1 = A.colloc[3 * (size_t)I + 1], x2 = A.colloc[3 * (size_t)I + 2];
        double gre = 0.0, gim = 0.0, nre[3] = {0.0, 0.0, 0.0}, nim[3] = {0.0, 0.0, 0.0};
        for (int q = 0; q < A.nq; ++q) {
            const double* qp = A.quad + 4 * ((size_t)J * A.nq + q);
            const double dx = x0 - qp[0], dy = x1 - qp[1], w = qp[3];
            const double r = sqrt(dx * dx + dy * dy);
            // the reference passes the panel's depth first for the adjoint
            // double layer and the collocation point's first otherwise
            const double zf = A.adjoint ? qp[2] : x2, zp = A.adjoint ? x2 : qp[2];
            const fg::Grn g = fg::point(r, zf, zp, A.v, A.roots, A.depth, ws, st);
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
}

__global__ void __launch_bounds__(256) fingreen3d_points_kernel(int np, const double* __restrict__ rr,
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

int launch_fingreen3d_pairs(const FgArgs& A, hipStream_t stream)
{
    if (A.n <= 0 || A.m <= 0) return 0;
    TimedLaunch timed(stream, TK_ASSEMBLY, (double)A.n * A.m, 32.0 * A.n * A.m);
    hipLaunchKernelGGL(fingreen3d_pairs_kernel, dim3(A.nthreads / 256), dim3(256), 0, stream, A);
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
