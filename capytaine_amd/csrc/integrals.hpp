// exp(z) E1(z) in float32 and the Simpson integration over theta of the
// Delhommeau wave-term integrals, Delhommeau_integrals.f90:43-153, shared by
// the table construction (tabulation.hip, with the host C library's
// transcendental values: the reference's table bit for bit) and the
// untabulated evaluation of the assembly (device transcendentals, within
// float32 rounding of the reference's numerical_integration).
//
// Every function runs with floating-point contraction off, the reference's
// operations in its order.
#pragma once
#include <hip/hip_runtime.h>

namespace cpt {
namespace {

#pragma clang fp contract(off)

// constants.f90:26: PI is a float32 literal widened to double
constexpr double INT_PI = (double)3.141592653588979f;

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

// exp(z) E1(z) in float32, Delhommeau_integrals.f90:104-151.  log_z(re, im)
// gives log(z) where the logarithmic branch needs it (z.re > -0.5,
// |z.im| <= 10).
template <typename LogZ>
__device__ cf exp_e1(double zr, double zi, LogZ log_z)
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
        const float2 l = log_z(z.re, z.im);
        const cf lz = mk(l.x, l.y);
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

// numerical_integration (Delhommeau_integrals.f90:43-153) for one (r, z),
// every transcendental value from the device: cos(theta), exp(z), the sine
// and cosine of r cos(theta) and the float complex logarithm.  The
// untabulated evaluation of the assembly (the reference's branch for a
// table of at most one r or z, Green_wave.f90:226-228).
__device__ __forceinline__ void numerical_integration_device(double r, double z, int npts, double* out)
{
    const int n = (npts % 2 == 0) ? npts + 1 : npts;
    const float nm1 = (float)n - 1.0f;
    const double d_end = INT_PI / (double)(3.0f * nm1);
    const double d_even = 4 * INT_PI / (double)(3.0f * nm1);
    const double d_odd = 2 * INT_PI / (double)(3.0f * nm1);
    const bool underflow = z <= -30.0;
    const double e = underflow ? 0.0 : exp(z);
    auto log_z = [](float re, float im) { return make_float2(logf(hypotf(re, im)), atan2f(im, re)); };
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0, a4 = 0;
    for (int k = 1; k <= n; ++k) {
        const double dt = (k == 1 || k == n) ? d_end : ((k % 2 == 0) ? d_even : d_odd);
        // pi*(k-1.0)/(n-1.0), the default reals promoted left to right
        const double theta = -INT_PI / 2 + INT_PI * (double)((float)k - 1.0f) / (double)nm1;
        const double ct = cos(theta);
        const double zr = z, zi = r * ct;
        double er = 0.0, ei = 0.0;
        if (!underflow) {
            double sn, cs;
            sincos(zi, &sn, &cs);
            er = e * cs;
            ei = e * sn;
        }
        const cf g = exp_e1(zr, zi, log_z);
        const double pr = 0.0 * er - INT_PI * ei, pi_ = 0.0 * ei + INT_PI * er;
        const double jr = (double)g.re + pr, ji = (double)g.im + pi_;
        const cdd inv = ddiv(cdd{1.0, 0.0}, cdd{zr, zi});
        const double mr = jr - inv.re, mi = ji - inv.im;
        a0 += dt * jr;
        a1 += dt * mr;
        a2 += dt * er;
        a3 -= dt * ct * mi;
        a4 -= dt * ct * ei;
    }
    out[0] = 2 * (a0 / INT_PI);
    out[1] = 2 * (a1 / INT_PI);
    out[2] = 2 * a2;
    out[3] = 2 * (a3 / INT_PI);
    out[4] = 2 * a4;
}

#pragma clang fp contract(fast)

}  // namespace
}  // namespace cpt
