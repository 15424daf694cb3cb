// Construction of the Delhommeau wave-term table on the device and its
// re-layout for the assembly kernel.
//
// Restates Delhommeau_integrals.f90:43-153 (numerical_integration with the
// float32 exp_e1 of Delhommeau's thesis, p. 367-369) and :190-204
// (construct_tabulation): one thread per (r, z) cell, 1001-point Simpson
// rule in theta.  The float32 part follows the reference's operation order
// with contraction disabled, and the complex divisions use the scaled C99
// algorithm that the reference's Fortran compiler emits, so that the table
// agrees with the reference's own table to float32 rounding of the complex
// logarithm (the only operation whose device and host implementations
// differ).
#include <hip/hip_runtime.h>
#include "cpt_internal.hpp"

#pragma clang fp contract(off)

namespace cpt {
namespace {

struct cf { float re, im; };
struct cdd { double re, im; };

__device__ __forceinline__ cf mk(float a, float b) { return cf{a, b}; }
__device__ __forceinline__ cf add(cf a, cf b) { return mk(a.re + b.re, a.im + b.im); }
__device__ __forceinline__ cf addr(float s, cf a) { return mk(s + a.re, a.im); }
__device__ __forceinline__ cf mul(cf a, cf b) { return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
__device__ __forceinline__ cf mulr(float s, cf a) { return mk(s * a.re, s * a.im); }

__device__ __forceinline__ cf cdiv(cf a, cf b)
{
    int ilogbw = 0;
    float c = b.re, d = b.im;
    const float logbw = logbf(fmaxf(fabsf(c), fabsf(d)));
    if (isfinite(logbw)) {
        ilogbw = (int)logbw;
        c = scalbnf(c, -ilogbw);
        d = scalbnf(d, -ilogbw);
    }
    const float denom = c * c + d * d;
    return mk(scalbnf((a.re * c + a.im * d) / denom, -ilogbw), scalbnf((a.im * c - a.re * d) / denom, -ilogbw));
}

__device__ __forceinline__ cdd ddiv(cdd a, cdd b)
{
    int ilogbw = 0;
    double c = b.re, d = b.im;
    const double logbw = logb(fmax(fabs(c), fabs(d)));
    if (isfinite(logbw)) {
        ilogbw = (int)logbw;
        c = scalbn(c, -ilogbw);
        d = scalbn(d, -ilogbw);
    }
    const double denom = c * c + d * d;
    return cdd{scalbn((a.re * c + a.im * d) / denom, -ilogbw), scalbn((a.im * c - a.re * d) / denom, -ilogbw)};
}

// exp(z) E1(z) in float32, Delhommeau_integrals.f90:104-151.
__device__ cf exp_e1(double zr, double zi)
{
    const cf z = mk((float)zr, (float)zi);
    if (z.re < -16.0f) {
        const cf y = cdiv(mk(1.0f, 0.0f), z);
        cf t = addr(24.0f, mulr(-120.0f, y));
        t = addr(-6.0f, mul(y, t));
        t = addr(2.0f, mul(y, t));
        t = addr(-1.0f, mul(y, t));
        t = addr(1.0f, mul(y, t));
        return mul(y, t);
    }
    if (fabsf(z.im) > 10.0f) {
        return add(add(cdiv(mk(0.711093f, 0.0f), addr(0.415775f, z)), cdiv(mk(0.278518f, 0.0f), addr(2.29428f, z))),
                   cdiv(mk(0.010389f, 0.0f), addr(6.2900f, z)));
    }
    if (z.re > -0.5f) {
        const cf lz = mk(logf(hypotf(z.re, z.im)), atan2f(z.im, z.re));
        cf p1 = addr(0.763297e-03f, mulr(0.97087007e-05f, z));
        p1 = addr(0.206543e-01f, mul(z, p1));
        p1 = addr(0.23721365e+00f, mul(z, p1));
        p1 = addr(0.1e+01f, mul(z, p1));
        cf p2 = addr(0.19387339e-02f, mulr(-0.51801555e-03f, z));
        p2 = addr(-0.3000591e-01f, mul(z, p2));
        p2 = addr(0.41806426e-01f, mul(z, p2));
        p2 = addr(-0.149545886e+01f, mul(z, p2));
        p2 = addr(0.99999207e+00f, mul(z, p2));
        cf q = addr(-0.8700861e-03f, mulr(0.2989204e-03f, z));
        q = addr(0.12982719e-01f, mul(z, q));
        q = addr(-0.66786033e-01f, mul(z, q));
        q = addr(0.28388363e+00f, mul(z, q));
        q = addr(-0.76273617e+00f, mul(z, q));
        q = addr(0.1e+01f, mul(z, q));
        const cf r = cdiv(add(mul(lz, p1), mulr(0.5772156649e+00f, p2)), q);
        return mk(-r.re, -r.im);
    }
    // Rational fit in the left half plane; conjugate coefficients for Im z >= 0.
    const float s = (z.im < 0.0f) ? 1.0f : -1.0f;
    cf num = mul(mk(1.000000f, s * 1.3935496e-06f), z);
    num = mul(add(num, mk(15.82958f, s * -20.14222f)), z);
    num = mul(add(num, mk(-70.52863f, s * -227.9511f)), z);
    num = mul(add(num, mk(-985.4221f, s * -226.6272f)), z);
    num = mul(add(num, mk(-1202.318f, s * 1580.907f)), z);
    num = mul(add(num, mk(953.2441f, s * 1342.447f)), z);
    num = mul(add(num, mk(417.3716f, s * -196.6665f)), z);
    num = add(num, mk(-9.881266f, s * -24.24952f));
    cf den = mul(mk(1.000000f, s * 0.0f), z);
    den = mul(add(den, mk(16.83184f, s * -20.14481f)), z);
    den = mul(add(den, mk(-55.66969f, s * -248.1167f)), z);
    den = mul(add(den, mk(-1068.640f, s * -434.4707f)), z);
    den = mul(add(den, mk(-2082.250f, s * 1522.471f)), z);
    den = mul(add(den, mk(383.3455f, s * 2730.378f)), z);
    den = mul(add(den, mk(1216.791f, s * 351.7189f)), z);
    den = mul(add(den, mk(115.3926f, s * -161.2647f)), z);
    den = add(den, mk(-3.777369f, s * -4.510900f));
    return cdiv(num, den);
}

// Writes the table in the reference layout (nr, nz, 5), column-major.
__global__ void __launch_bounds__(256) tabulate_kernel(const double* __restrict__ r_range, int nr,
                                                       const double* __restrict__ z_range, int nz, int npts,
                                                       double* __restrict__ tab)
{
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= nr * nz) return;
    const int i = cell % nr, j = cell / nr;
    const double r = r_range[i], z = z_range[j];
    const int n = (npts % 2 == 0) ? npts + 1 : npts;
    const float nm1 = (float)n - 1.0f;
    const double d_end = PI / (double)(3.0f * nm1);
    const double d_even = 4 * PI / (double)(3.0f * nm1);
    const double d_odd = 2 * PI / (double)(3.0f * nm1);
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0;
    for (int k = 1; k <= n; ++k) {
        const double theta = -PI / 2 + PI * (double)((float)k - 1.0f) / (double)nm1;
        const double dt = (k == 1 || k == n) ? d_end : ((k % 2 == 0) ? d_even : d_odd);
        const double ct = cos(theta);
        const double zr = z, zi = r * ct;
        double er = 0.0, ei = 0.0;
        if (!(zr <= -30.0)) {
            const double e = exp(zr);
            double sn, cs;
            sincos(zi, &sn, &cs);
            er = e * cs;
            ei = e * sn;
        }
        const cf g = exp_e1(zr, zi);
        // ii*pi*exp_zeta = (0, pi) * exp_zeta
        const double pr = 0.0 * er - PI * ei, pi_ = 0.0 * ei + PI * er;
        const double jr = (double)g.re + pr, ji = (double)g.im + pi_;
        const cdd inv = ddiv(cdd{1.0, 0.0}, cdd{zr, zi});
        const double mr = jr - inv.re, mi = ji - inv.im;
        a0 += dt * jr;
        a1 += dt * mr;
        a2 += dt * er;
        a3 -= dt * ct * mi;
        a4 -= dt * ct * ei;
    }
    const size_t plane = (size_t)nr * nz;
    tab[cell] = 2 * (a0 / PI);
    tab[cell + plane] = 2 * (a1 / PI);
    tab[cell + 2 * plane] = 2 * a2;
    tab[cell + 3 * plane] = 2 * (a3 / PI);
    tab[cell + 4 * plane] = 2 * a4;
}

// (nr, nz, 5) column-major  ->  (nz, nr, 5) row-major with the five values
// of one cell contiguous.
__global__ void relayout_kernel(const double* __restrict__ src, int nr, int nz, double* __restrict__ dst)
{
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= nr * nz) return;
    const size_t plane = (size_t)nr * nz;
#pragma unroll
    for (int c = 0; c < 5; ++c) dst[(size_t)cell * 5 + c] = src[cell + c * plane];
}

}  // namespace

int launch_tabulation(const double* r_range, int nr, const double* z_range, int nz, int npts, double* tab,
                      hipStream_t stream)
{
    const int cells = nr * nz;
    hipLaunchKernelGGL(tabulate_kernel, dim3((cells + 255) / 256), dim3(256), 0, stream, r_range, nr, z_range, nz,
                       npts, tab);
    return (int)hipGetLastError();
}

int launch_table_relayout(const double* src, int nr, int nz, double* dst, hipStream_t stream)
{
    const int cells = nr * nz;
    hipLaunchKernelGGL(relayout_kernel, dim3((cells + 255) / 256), dim3(256), 0, stream, src, nr, nz, dst);
    return (int)hipGetLastError();
}

}  // namespace cpt
