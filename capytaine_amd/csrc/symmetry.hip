// Block diagonalisation of the influence matrices of symmetric meshes on
// gfx950 (tools/block_circulant_matrices.py:131-160, engines.py:178-247).
//
// A mesh made of n copies of a reference block of w faces (reflections
// across one or two planes, or n rotations about Oz) has S and K whose
// (i, j) block of w x w depends on the group element g_i^-1 g_j only: the
// assembly computes the first block column A_0..A_{n-1} (all N = n w
// collocation points against the reference block) and the group's Fourier
// transform turns the N x N system into n independent w x w systems:
//
//   out_k = scale * sum_j chi(k, j) in_j
//
// with chi(k, j) = exp(-2 pi i j k / n) (rotations; the conjugate for the
// inverse transform) or (-1)^popcount(j & k) (one or two planes: the Walsh
// character of the block index, bit 0 the inner plane, bit 1 the outer).
// The same kernel maps right-hand sides to the transformed space and
// solutions back (scale 1/n for the inverse).
//
// One thread per position (row r, column c, batch item b) of a block: it
// reads its n inputs once into its own LDS column (n <= 60; above, from the
// caches at every output) and forms the n outputs.
// Consecutive threads take consecutive rows, so loads and stores are
// coalesced along the column-major blocks.  The n x n characters are
// formed once per work-group in LDS.  HBM-bound for small n (one read and
// one write per element), VALU-bound for large n (8 n flops per output).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "cpt_internal.hpp"

namespace cpt {

constexpr int BT_THREADS = 64;
constexpr int BT_MAX_BLOCKS = 128;
constexpr int BT_MAX_STAGED = 60;  // (n + 64 n) complex128 within 64 KB of LDS

// STAGED: the thread's n inputs go to its LDS column once; otherwise they
// are read again for every output (from the caches) -- for n > 60.
template <bool STAGED>
__global__ void __launch_bounds__(BT_THREADS)
block_transform_kernel(int xor_kind, int n, int inverse, double scale, int R, int C, int64_t total,
                       const cplx* __restrict__ in, int64_t sbi, int64_t sj, int64_t ldi, cplx* __restrict__ out,
                       int64_t sbo, int64_t sk, int64_t ldo)
{
    extern __shared__ cplx lds[];
    cplx* chi = lds;      // n: exp(-+2 pi i m / n), m = 0..n-1 (rotations only)
    cplx* col = lds + n;  // STAGED: n x BT_THREADS, input j of thread t at col[j * BT_THREADS + t]
    const int t = threadIdx.x;
    if (!xor_kind) {
        for (int m = t; m < n; m += BT_THREADS) {
            double sn, cs;
            sincospi(2.0 * (double)m / (double)n, &sn, &cs);
            chi[m] = cplx{cs, inverse ? sn : -sn};
        }
    }
    __syncthreads();
    const int64_t pos = (int64_t)blockIdx.x * BT_THREADS + t;
    if (pos >= total) return;
    const int r = (int)(pos % R);
    const int64_t rest = pos / R;
    const int c = (int)(rest % C);
    const int64_t b = rest / C;
    const cplx* src = in + b * sbi + (int64_t)c * ldi + r;
    if (STAGED)
        for (int j = 0; j < n; ++j) col[j * BT_THREADS + t] = src[(int64_t)j * sj];
    cplx* dst = out + b * sbo + (int64_t)c * ldo + r;
    for (int k = 0; k < n; ++k) {
        double re = 0.0, im = 0.0;
        int m = 0;  // (j k) mod n
        for (int j = 0; j < n; ++j) {
            const cplx x = STAGED ? col[j * BT_THREADS + t] : src[(int64_t)j * sj];
            if (xor_kind) {
                const double sg = (__popc(j & k) & 1) ? -1.0 : 1.0;
                re += sg * x.re;
                im += sg * x.im;
            } else {
                const cplx w = chi[m];
                re += w.re * x.re - w.im * x.im;
                im += w.re * x.im + w.im * x.re;
                m += k;
                if (m >= n) m -= n;
            }
        }
        dst[(int64_t)k * sk] = cplx{scale * re, scale * im};
    }
}

int launch_block_transform(int xor_kind, int n, int inverse, int R, int C, int nbatch, const cplx* in, int64_t sbi,
                           int64_t sj, int64_t ldi, cplx* out, int64_t sbo, int64_t sk, int64_t ldo, hipStream_t stream)
{
    const int64_t total = (int64_t)R * C * nbatch;
    if (total == 0) return 0;
    const int64_t grid = (total + BT_THREADS - 1) / BT_THREADS;
    const double scale = inverse ? 1.0 / n : 1.0;
    if (n <= BT_MAX_STAGED) {
        const size_t lds = sizeof(cplx) * ((size_t)n + (size_t)n * BT_THREADS);
        hipLaunchKernelGGL(block_transform_kernel<true>, dim3((unsigned)grid), dim3(BT_THREADS), lds, stream,
                           xor_kind, n, inverse, scale, R, C, total, in, sbi, sj, ldi, out, sbo, sk, ldo);
    } else {
        hipLaunchKernelGGL(block_transform_kernel<false>, dim3((unsigned)grid), dim3(BT_THREADS),
                           sizeof(cplx) * (size_t)n, stream, xor_kind, n, inverse, scale, R, C, total, in, sbi, sj,
                           ldi, out, sbo, sk, ldo);
    }
    return (int)hipGetLastError();
}

int block_transform_max_blocks() { return BT_MAX_BLOCKS; }

}  // namespace cpt
