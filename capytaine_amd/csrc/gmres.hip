// Restarted GMRES without preconditioner on the device: the
// linear_solver="gmres" of the reference's engine (src/capytaine/bem/
// engines.py:74-89, scipy.sparse.linalg.gmres(A, b, atol=1e-6) with its
// defaults rtol=1e-5, restart=20, maxiter=10 n cycles and x0=0).
//
// Arnoldi basis V (n x (restart+1)) in the matrix's precision, classical
// Gram-Schmidt applied twice per step (two projections, each a block of dot
// products and one update; the two coefficient sets are added), the
// Hessenberg column reduced with Givens rotations on the host.  The products
// A v go through the library's skinny product kernels; every reduction is
// deterministic (fixed-order partial sums in FP64), so a solve repeats bit for
// bit.  One host synchronisation per Arnoldi step reads the step's
// coefficients and the new vector's norm.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <limits>
#include <vector>

#include "cpt_internal.hpp"

namespace cpt {

int launch_zgemm(int M, int N, int K, double alpha, const cplx* A, int64_t lda, int64_t sA, const cplx* B,
                 int64_t ldb, int64_t sB, cplx* C, int64_t ldc, int64_t sC, int beta_zero, int batch,
                 hipStream_t stream);
int launch_cgemm(int M, int N, int K, float alpha, const void* A, int64_t lda, int64_t sA, const void* B,
                 int64_t ldb, int64_t sB, void* C, int64_t ldc, int64_t sC, int beta_zero, int batch,
                 hipStream_t stream);

namespace {

template <typename R>
struct gc {
    R re, im;
};

constexpr int GM_THREADS = 256;
constexpr int GM_CHUNK = 4096;  // rows per partial sum
constexpr int GM_MAX_RESTART = 128;

inline int gm_chunks(int n) { return (n + GM_CHUNK - 1) / GM_CHUNK; }

// partial[i * nchunk + c] = sum over rows of chunk c of conj(V[k, i]) w[k]
template <typename R>
__global__ void __launch_bounds__(GM_THREADS)
    gm_dots_kernel(const gc<R>* __restrict__ V, int64_t ldv, const gc<R>* __restrict__ w, int n, int nchunk,
                   double2* __restrict__ partial)
{
    const int i = blockIdx.x, c = blockIdx.y;
    const int k0 = c * GM_CHUNK, k1 = min(n, k0 + GM_CHUNK);
    const gc<R>* v = V + (size_t)i * ldv;
    double sr = 0.0, si = 0.0;
    for (int k = k0 + threadIdx.x; k < k1; k += GM_THREADS) {
        const double vr = v[k].re, vi = v[k].im, wr = w[k].re, wi = w[k].im;
        sr += vr * wr + vi * wi;
        si += vr * wi - vi * wr;
    }
    __shared__ double lr[GM_THREADS], li[GM_THREADS];
    lr[threadIdx.x] = sr;
    li[threadIdx.x] = si;
    __syncthreads();
    for (int s = GM_THREADS / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            lr[threadIdx.x] += lr[threadIdx.x + s];
            li[threadIdx.x] += li[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(size_t)i * nchunk + c] = make_double2(lr[0], li[0]);
}

// out[i] = sum over chunks (in order) of partial[i * nchunk + c]
__global__ void gm_reduce_kernel(const double2* __restrict__ partial, int ncols, int nchunk, double2* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncols) return;
    double sr = 0.0, si = 0.0;
    for (int c = 0; c < nchunk; ++c) {
        sr += partial[(size_t)i * nchunk + c].x;
        si += partial[(size_t)i * nchunk + c].y;
    }
    out[i] = make_double2(sr, si);
}

// w[k] += sign * sum_{i < ncols} V[k, i] h[i]
template <typename R>
__global__ void __launch_bounds__(GM_THREADS)
    gm_update_kernel(const gc<R>* __restrict__ V, int64_t ldv, int ncols, const double2* __restrict__ h, double sign,
                     gc<R>* __restrict__ w, int n)
{
    const int k = blockIdx.x * GM_THREADS + threadIdx.x;
    if (k >= n) return;
    double ar = 0.0, ai = 0.0;
    for (int i = 0; i < ncols; ++i) {
        const gc<R> v = V[(size_t)i * ldv + k];
        const double2 c = h[i];
        ar += (double)v.re * c.x - (double)v.im * c.y;
        ai += (double)v.re * c.y + (double)v.im * c.x;
    }
    w[k].re = (R)((double)w[k].re + sign * ar);
    w[k].im = (R)((double)w[k].im + sign * ai);
}

// dst = a * src
template <typename R>
__global__ void gm_scale_kernel(const gc<R>* __restrict__ src, double a, gc<R>* __restrict__ dst, int n)
{
    const int k = blockIdx.x * GM_THREADS + threadIdx.x;
    if (k >= n) return;
    dst[k].re = (R)(a * (double)src[k].re);
    dst[k].im = (R)(a * (double)src[k].im);
}

// r = b - r (r holds A x)
template <typename R>
__global__ void gm_residual_kernel(const gc<R>* __restrict__ b, gc<R>* __restrict__ r, int n)
{
    const int k = blockIdx.x * GM_THREADS + threadIdx.x;
    if (k >= n) return;
    r[k].re = b[k].re - r[k].re;
    r[k].im = b[k].im - r[k].im;
}

template <typename R>
__global__ void gm_fill_kernel(gc<R>* __restrict__ x, int n)
{
    const int k = blockIdx.x * GM_THREADS + threadIdx.x;
    if (k < n) x[k].re = x[k].im = (R)0;
}

template <typename R>
int matvec(const void* A, int n, int64_t lda, const gc<R>* v, gc<R>* y, hipStream_t s)
{
    if constexpr (sizeof(R) == 8)
        return launch_zgemm(n, 1, n, 1.0, (const cplx*)A, lda, 0, (const cplx*)v, n, 0, (cplx*)y, n, 0, 1, 1, s);
    else
        return launch_cgemm(n, 1, n, 1.0f, A, lda, 0, v, n, 0, y, n, 0, 1, 1, s);
}

// Device workspace layout (bytes): V (n x (restart+1)), w, r, partials,
// reduced coefficients (2 sets of restart+1) and the norm, y.
struct GmLayout {
    size_t V, w, r, part, coef, y, total;
};

template <typename R>
GmLayout layout(int n, int restart)
{
    auto a256 = [](size_t x) { return (x + 255) & ~(size_t)255; };
    GmLayout L;
    size_t p = 0;
    const size_t vec = sizeof(gc<R>) * (size_t)n;
    L.V = p;
    p += a256(vec * (restart + 1));
    L.w = p;
    p += a256(vec);
    L.r = p;
    p += a256(vec);
    L.part = p;
    p += a256(sizeof(double2) * (size_t)(restart + 1) * gm_chunks(n));
    L.coef = p;
    p += a256(sizeof(double2) * (size_t)(2 * (restart + 1) + 2));
    L.y = p;
    p += a256(sizeof(double2) * (size_t)(restart + 1));
    L.total = p;
    return L;
}

// coefficients of V[:, :ncols]^H w into coef (device), fixed-order sums
template <typename R>
int dots(const gc<R>* V, int64_t ldv, int ncols, const gc<R>* w, int n, double2* part, double2* coef, hipStream_t s)
{
    const int nch = gm_chunks(n);
    hipLaunchKernelGGL(gm_dots_kernel<R>, dim3(ncols, nch), dim3(GM_THREADS), 0, s, V, ldv, w, n, nch, part);
    hipLaunchKernelGGL(gm_reduce_kernel, dim3((ncols + 63) / 64), dim3(64), 0, s, (const double2*)part, ncols, nch,
                       coef);
    return (int)hipGetLastError();
}

template <typename R>
double vec_norm(const gc<R>* w, int n, double2* part, double2* coef, hipStream_t s, int* err)
{
    double2 v;
    if ((*err = dots<R>(w, n, 1, w, n, part, coef, s))) return 0.0;
    if ((*err = (int)hipMemcpyAsync(&v, coef, sizeof(double2), hipMemcpyDeviceToHost, s))) return 0.0;
    if ((*err = (int)hipStreamSynchronize(s))) return 0.0;
    return std::sqrt(std::max(v.x, 0.0));
}

// One right-hand side.  Returns a HIP error code (0 on success); *converged,
// *iters (inner steps) and *res (final ||b - A x|| / ||b||) are set.
template <typename R>
int gmres_one(const void* A, int n, int64_t lda, const gc<R>* b, gc<R>* x, int restart, int maxiter, double rtol,
              double atol, char* ws, hipStream_t s, int* iters, double* res, bool* converged)
{
    using C = std::complex<double>;
    const GmLayout L = layout<R>(n, restart);
    gc<R>* V = (gc<R>*)(ws + L.V);
    gc<R>* w = (gc<R>*)(ws + L.w);
    gc<R>* r = (gc<R>*)(ws + L.r);
    double2* part = (double2*)(ws + L.part);
    double2* coef = (double2*)(ws + L.coef);
    double2* ydev = (double2*)(ws + L.y);
    const dim3 g1((n + GM_THREADS - 1) / GM_THREADS), b1(GM_THREADS);
    int e = 0;
    *iters = 0;
    *converged = false;
    hipLaunchKernelGGL(gm_fill_kernel<R>, g1, b1, 0, s, x, n);
    const double bnorm = vec_norm<R>(b, n, part, coef, s, &e);
    if (e) return e;
    if (bnorm == 0.0) {  // x = 0 solves it (scipy returns at once)
        *res = 0.0;
        *converged = true;
        return (int)hipStreamSynchronize(s);
    }
    const double tol = std::max(atol, rtol * bnorm);
    // r = b (x0 = 0)
    hipLaunchKernelGGL(gm_scale_kernel<R>, g1, b1, 0, s, b, 1.0, r, n);
    double beta = bnorm;
    std::vector<C> H((size_t)(restart + 1) * restart), cs(restart), sn(restart), g(restart + 1), y(restart);
    std::vector<double2> hv(2 * (restart + 1) + 2);
    const double eps = std::numeric_limits<R>::epsilon();
    bool breakdown = false;
    for (int cycle = 0; cycle < maxiter; ++cycle) {
        if (beta <= tol) {
            *converged = true;
            break;
        }
        hipLaunchKernelGGL(gm_scale_kernel<R>, g1, b1, 0, s, (const gc<R>*)r, 1.0 / beta, V, n);
        std::fill(g.begin(), g.end(), C(0.0));
        g[0] = beta;
        int m = 0;
        for (int j = 0; j < restart; ++j) {
            gc<R>* vj = V + (size_t)j * n;
            if ((e = matvec<R>(A, n, lda, vj, w, s))) return e;
            // ||A v_j|| for the breakdown test, then two classical
            // Gram-Schmidt projections
            if ((e = dots<R>(w, n, 1, w, n, part, coef + 2 * (restart + 1) + 1, s))) return e;
            if ((e = dots<R>(V, n, j + 1, w, n, part, coef, s))) return e;
            hipLaunchKernelGGL(gm_update_kernel<R>, g1, b1, 0, s, (const gc<R>*)V, (int64_t)n, j + 1,
                               (const double2*)coef, -1.0, w, n);
            if ((e = dots<R>(V, n, j + 1, w, n, part, coef + (restart + 1), s))) return e;
            hipLaunchKernelGGL(gm_update_kernel<R>, g1, b1, 0, s, (const gc<R>*)V, (int64_t)n, j + 1,
                               (const double2*)(coef + (restart + 1)), -1.0, w, n);
            if ((e = dots<R>(w, n, 1, w, n, part, coef + 2 * (restart + 1), s))) return e;
            if ((e = (int)hipMemcpyAsync(hv.data(), coef, sizeof(double2) * hv.size(), hipMemcpyDeviceToHost, s)))
                return e;
            if ((e = (int)hipStreamSynchronize(s))) return e;
            double hn = std::sqrt(std::max(hv[2 * (restart + 1)].x, 0.0));
            const double h0 = std::sqrt(std::max(hv[2 * (restart + 1) + 1].x, 0.0));
            // exact solution in the Krylov space (scipy's test, relative to
            // ||A v_j||): no new direction
            if (hn <= eps * h0) {
                hn = 0.0;
                breakdown = true;
            }
            C* col = &H[(size_t)j * (restart + 1)];
            for (int i = 0; i <= j; ++i)
                col[i] = C(hv[i].x + hv[restart + 1 + i].x, hv[i].y + hv[restart + 1 + i].y);
            col[j + 1] = hn;
            // previous rotations, then the one that zeroes col[j + 1]
            for (int i = 0; i < j; ++i) {
                const C t = cs[i] * col[i] + sn[i] * col[i + 1];
                col[i + 1] = -std::conj(sn[i]) * col[i] + cs[i] * col[i + 1];
                col[i] = t;
            }
            const double a = std::abs(col[j]);
            const double d = std::hypot(a, hn);
            if (d == 0.0) {
                cs[j] = 1.0;
                sn[j] = 0.0;
            } else if (a == 0.0) {
                cs[j] = 0.0;
                sn[j] = 1.0;
            } else {
                const C ph = col[j] / a;
                cs[j] = a / d;
                sn[j] = ph * (hn / d);
            }
            col[j] = cs[j] * col[j] + sn[j] * col[j + 1];
            col[j + 1] = 0.0;
            g[j + 1] = -std::conj(sn[j]) * g[j];
            g[j] = cs[j] * g[j];
            m = j + 1;
            ++*iters;
            if (hn > 0.0)
                hipLaunchKernelGGL(gm_scale_kernel<R>, g1, b1, 0, s, (const gc<R>*)w, 1.0 / hn, V + (size_t)(j + 1) * n,
                                   n);
            if (std::abs(g[j + 1]) <= tol || breakdown) break;
        }
        // y = H^-1 g (upper triangular m x m), x += V y
        for (int i = m - 1; i >= 0; --i) {
            C t = g[i];
            for (int k = i + 1; k < m; ++k) t -= H[(size_t)k * (restart + 1) + i] * y[k];
            const C dgi = H[(size_t)i * (restart + 1) + i];
            y[i] = (dgi == C(0.0)) ? C(0.0) : t / dgi;
        }
        for (int i = 0; i < m; ++i) hv[i] = make_double2(y[i].real(), y[i].imag());
        if ((e = (int)hipMemcpyAsync(ydev, hv.data(), sizeof(double2) * m, hipMemcpyHostToDevice, s))) return e;
        hipLaunchKernelGGL(gm_update_kernel<R>, g1, b1, 0, s, (const gc<R>*)V, (int64_t)n, m, (const double2*)ydev,
                           1.0, x, n);
        // true residual
        if ((e = matvec<R>(A, n, lda, x, r, s))) return e;
        hipLaunchKernelGGL(gm_residual_kernel<R>, g1, b1, 0, s, b, r, n);
        beta = vec_norm<R>(r, n, part, coef, s, &e);  // synchronises: hv is free again
        if (e) return e;
        if (beta <= tol) {
            *converged = true;
            break;
        }
        // no progress possible: breakdown with the true residual still
        // above tol (scipy bails out the same way), or a non-finite residual
        if (breakdown || !std::isfinite(beta)) break;
    }
    *res = beta / bnorm;
    return (int)hipGetLastError();
}

}  // namespace

size_t gmres_workspace_bytes(int n, int restart, bool single)
{
    return single ? layout<float>(n, restart).total : layout<double>(n, restart).total;
}

int gmres_max_restart() { return GM_MAX_RESTART; }

int run_gmres(bool single, const void* A, int n, int64_t lda, const void* B, int nrhs, int64_t ldb, void* X,
              int64_t ldx, int restart, int maxiter, double rtol, double atol, void* work, hipStream_t stream,
              int* iterations, double* residual, int* not_converged)
{
    *not_converged = 0;
    for (int j = 0; j < nrhs; ++j) {
        bool ok = false;
        int e;
        if (single)
            e = gmres_one<float>(A, n, lda, (const gc<float>*)B + (size_t)j * ldb, (gc<float>*)X + (size_t)j * ldx,
                                 restart, maxiter, rtol, atol, (char*)work, stream, &iterations[j], &residual[j], &ok);
        else
            e = gmres_one<double>(A, n, lda, (const gc<double>*)B + (size_t)j * ldb,
                                  (gc<double>*)X + (size_t)j * ldx, restart, maxiter, rtol, atol, (char*)work, stream,
                                  &iterations[j], &residual[j], &ok);
        if (e) return e;
        if (!ok) ++*not_converged;
    }
    return 0;
}

}  // namespace cpt
