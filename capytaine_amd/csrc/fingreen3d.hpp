// The HAMS finite-depth Green function (FinGreen3D) of one point pair.
//
// Liu, Iwashita & Hu (2015), "A calculation method for finite depth
// free-surface Green function", Int. J. Nav. Archit. Ocean Eng. 7:375-389, and
// Liu et al. (2018), Energy Conversion and Management 174:516-536: the wave
// part of the finite-depth Green function G(R, zf, zp) from one of four series
// chosen by R/h.  The choice of series, their truncations and the constants
// follow the reference's fingreen3d_routine (FinGreen3D.f90:71-117), so that
// the values agree with it to rounding:
//   R/h >= 0.5            John's eigenfunction expansion, at most 10 terms
//                         (EIGEN, FinGreen3D.f90:297-366)
//   0.05 <= R/h < 0.5     the same with the epsilon-algorithm limit of its
//                         partial sums (EIGENE, :417-498)
//   0.0005 <= R/h < 0.05  Pidcock's form: the eigenfunction series less its
//                         logarithmic part, Newman's image sum, epsilon limit
//                         (PIDCOCK, :550-650)
//   R/h < 0.0005          Linton's Ewald-type form with error functions and
//                         an adaptive Gauss-Kronrod integral (LINTON, :702-849)
// The result is [G, dG/dR, dG/dzf] with 1/r and the free-surface image 1/r1
// taken out (the Rankine terms are integrated exactly on the panels).
//
// The special functions are the approximations the reference uses, since
// their errors are part of its values: Zhang & Jin's (1996) polynomial J0,
// J1, Y0, Y1 (errors of a few 1e-9) and series/asymptotic I0, I1, K0, K1;
// Abramowitz & Stegun 5.1.53 and 5.1.56 for E1 (errors up to 2e-7); series
// and asymptotic forms for Ei and erf.  Operations are kept in the
// reference's order and not contracted into fused multiply-adds, so that
// the same IEEE operations give the same roundings where the elementary
// functions (exp, log, sin, ...) agree.
//
// Every function here is __host__ __device__: the same code runs in the
// kernels and, on the CPU, in a test against the reference.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

// no fused multiply-adds in this file: the reference is built without them
#pragma clang fp contract(off)

namespace fg {

constexpr double PI = 3.141592653589793;
constexpr double GA = 0.5772156649015328;  // Euler's constant
// The reference writes several thresholds as default (single-precision)
// real literals; they enter its double arithmetic as these values.
constexpr double EPS_1E6 = (double)1e-6f;
constexpr double C_EIGENE_A = (double)88.89f, C_EIGENE_B = (double)54.45f;
constexpr double C_PIDCOCK_A = (double)1010.10f, C_PIDCOCK_B = 100.5;

// Up to 99 partial sums per series (Pidcock at R/h = 0.0005); a workspace of
// LIMES_LEN doubles per array, four arrays per thread.
constexpr int LIMES_LEN = 100;
constexpr int WS_ARRAYS = 4;

struct C {
    double re, im;
};
__host__ __device__ inline C cmk(double re, double im) { return C{re, im}; }
__host__ __device__ inline C operator+(C a, C b) { return C{a.re + b.re, a.im + b.im}; }
__host__ __device__ inline C operator-(C a, C b) { return C{a.re - b.re, a.im - b.im}; }
__host__ __device__ inline C operator*(C a, C b) { return C{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__host__ __device__ inline C operator*(double s, C a) { return C{s * a.re, s * a.im}; }
__host__ __device__ inline C operator/(C a, C b)
{
    const double d = b.re * b.re + b.im * b.im;
    return C{(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
__host__ __device__ inline C ccos(C z) { return C{cos(z.re) * cosh(z.im), -sin(z.re) * sinh(z.im)}; }
__host__ __device__ inline C csin(C z) { return C{sin(z.re) * cosh(z.im), cos(z.re) * sinh(z.im)}; }
__host__ __device__ inline double cabs_(C z) { return hypot(z.re, z.im); }

// ---------------------------------------------------------------------------
// Special functions

// a / b for the small integer divisors of the power series (k^2, k (k + 1),
// k <= 50), correctly rounded as the division is, in three operations:
// q0 = a y with y = RN(1/b), the exact remainder r = a - b q0 by a fused
// multiply-add, q = q0 + r y (Markstein 1990).  Checked against the division
// for 2e8 random dividends of every divisor: no difference.
struct SeriesRecip {
    double sq[51], kk1[51], k[51];
    constexpr SeriesRecip() : sq(), kk1(), k()
    {
        for (int i = 1; i <= 50; ++i) {
            sq[i] = 1.0 / (double)(i * i);
            kk1[i] = 1.0 / (double)(i * (i + 1));
            k[i] = 1.0 / (double)i;
        }
    }
};
constexpr SeriesRecip RECIP{};

__host__ __device__ inline double div_small(double a, double b, double y)
{
    const double q0 = a * y;
    const double r = fma(-q0, b, a);
    return fma(r, y, q0);
}

// I0, I1, K0, K1 as in Zhang & Jin, "Computation of Special Functions"
// (Wiley 1996), routine IK01A as the reference carries it
// (FinGreen3D.f90:3221-3310): power series of I0, I1 up to x = 18 and their
// asymptotic expansions (12, 9 or 7 terms) beyond; K0 by its series up to
// x = 9, beyond by the asymptotic expansion of the product K0 I0 (8 terms);
// K1 from the Wronskian I0 K1 + I1 K0 = 1/x.  The expansion coefficients are
// the book's 14-digit values.  Above x = 709 the reference's I0 overflows and
// its K1 is 0 * inf = NaN; here both K are 0 there (their value underflows).
__host__ __device__ inline void bessel_k01(double x, double& k0, double& k1)
{
    if (x == 0.0) {
        k0 = k1 = 1.0e300;
        return;
    }
    if (x > 709.0) {
        k0 = k1 = 0.0;
        return;
    }
    const double x2 = x * x;
    double bi0 = 1.0, bi1 = 1.0;
    if (x <= 18.0) {
        // the stopping tests |r / I| < 1e-15 written without the division
        // (the same decision but where the quotient rounds across 1e-15)
        double r = 1.0;
        for (int k = 1; k <= 50; ++k) {
            r = div_small(0.25 * r * x2, (double)(k * k), RECIP.sq[k]);
            bi0 += r;
            if (fabs(r) < 1.0e-15 * fabs(bi0)) break;
        }
        r = 1.0;
        for (int k = 1; k <= 50; ++k) {
            r = div_small(0.25 * r * x2, (double)(k * (k + 1)), RECIP.kk1[k]);
            bi1 += r;
            if (fabs(r) < 1.0e-15 * fabs(bi1)) break;
        }
        bi1 = 0.5 * x * bi1;
    } else {
        const double a[12] = {0.125, 7.03125e-2, 7.32421875e-2, 1.1215209960938e-1, 2.2710800170898e-1,
                              5.7250142097473e-1, 1.7277275025845, 6.0740420012735, 2.4380529699556e01,
                              1.1001714026925e02, 5.5133589612202e02, 3.0380905109224e03};
        const double b[12] = {-0.375, -1.171875e-1, -1.025390625e-1, -1.4419555664063e-1, -2.7757644653320e-1,
                              -6.7659258842468e-1, -1.9935317337513, -6.8839142681099, -2.7248827311269e01,
                              -1.2159789187654e02, -6.0384407670507e02, -3.3022722944809e03};
        const int nt = x >= 50.0 ? 7 : (x >= 35.0 ? 9 : 12);
        const double ca = exp(x) / sqrt(2.0 * PI * x), xr = 1.0 / x;
        double p = 1.0;
#pragma unroll
        for (int k = 1; k <= 12; ++k) {
            p *= xr;
            if (k <= nt) bi0 += a[k - 1] * p;
        }
        bi0 = ca * bi0;
        p = 1.0;
#pragma unroll
        for (int k = 1; k <= 12; ++k) {
            p *= xr;
            if (k <= nt) bi1 += b[k - 1] * p;
        }
        bi1 = ca * bi1;
    }
    if (x <= 9.0) {
        const double ct = -(log(x / 2.0) + 0.5772156649015329);
        double s = 0.0, w0 = 0.0, r = 1.0, prev = 0.0;
        for (int k = 1; k <= 50; ++k) {
            w0 += RECIP.k[k];  // 1.0 / k
            r = div_small(0.25 * r, (double)(k * k), RECIP.sq[k]) * x2;
            s += r * (w0 + ct);
            if (fabs(s - prev) < 1.0e-15 * fabs(s)) break;
            prev = s;
        }
        k0 = s + ct;
    } else {
        const double a1[8] = {0.125, 0.2109375, 1.0986328125, 1.1775970458984e01, 2.1461706161499e02,
                              5.9511522710323e03, 2.3347645606175e05, 1.2312234987631e07};
        const double xr2 = 1.0 / x2;
        double s = 1.0, p = 1.0;
#pragma unroll
        for (int k = 1; k <= 8; ++k) {
            p *= xr2;
            s += a1[k - 1] * p;
        }
        k0 = 0.5 / x * s / bi0;
    }
    k1 = (1.0 / x - bi1 * k0) / bi0;
}

// J0, J1, Y0, Y1: the polynomial approximations of Zhang & Jin's JY01B
// (FinGreen3D.f90:3346-3422), in t = x/4 up to x = 4 and in the
// modulus-phase form with t = 4/x beyond; absolute errors of a few 1e-9.
// One coefficient of Y0 is a single-precision literal in the reference and
// enters its arithmetic rounded to float.
__host__ __device__ inline void bessel_jy01(double x, double& j0_, double& j1_, double& y0_, double& y1_)
{
    if (x == 0.0) {
        j0_ = 1.0;
        j1_ = 0.0;
        y0_ = y1_ = -1.0e300;
        return;
    }
    if (x <= 4.0) {
        const double t = x / 4.0, t2 = t * t;
        j0_ = ((((((-.5014415e-3 * t2 + .76771853e-2) * t2 - .0709253492) * t2 + .4443584263) * t2 - 1.7777560599) *
                    t2 + 3.9999973021) * t2 - 3.9999998721) * t2 + 1.0;
        j1_ = t * (((((((-.1289769e-3 * t2 + .22069155e-2) * t2 - .0236616773) * t2 + .1777582922) * t2 -
                      .8888839649) * t2 + 2.6666660544) * t2 - 3.9999999710) * t2 + 1.9999999998);
        const double y0c = (double)1.0766115157f;
        y0_ = (((((((-.567433e-4 * t2 + .859977e-3) * t2 - .94855882e-2) * t2 + .0772975809) * t2 - .4261737419) *
                     t2 + 1.4216421221) * t2 - 2.3498519931) * t2 + y0c) * t2 + .3674669052;
        y0_ = 2.0 / PI * log(x / 2.0) * j0_ + y0_;
        y1_ = ((((((((.6535773e-3 * t2 - .0108175626) * t2 + .107657606) * t2 - .7268945577) * t2 + 3.1261399273) *
                      t2 - 7.3980241381) * t2 + 6.8529236342) * t2 + .3932562018) * t2 - .6366197726) / x;
        y1_ = 2.0 / PI * log(x / 2.0) * j1_ + y1_;
        return;
    }
    const double t = 4.0 / x, t2 = t * t;
    const double a0 = sqrt(2.0 / (PI * x));
    const double p0 = ((((-.9285e-5 * t2 + .43506e-4) * t2 - .122226e-3) * t2 + .434725e-3) * t2 - .4394275e-2) * t2 +
                      .999999997;
    const double q0 =
        t * (((((.8099e-5 * t2 - .35614e-4) * t2 + .85844e-4) * t2 - .218024e-3) * t2 + .1144106e-2) * t2 - .031249995);
    const double ta0 = x - .25 * PI;
    j0_ = a0 * (p0 * cos(ta0) - q0 * sin(ta0));
    y0_ = a0 * (p0 * sin(ta0) + q0 * cos(ta0));
    const double p1 = ((((.10632e-4 * t2 - .50363e-4) * t2 + .145575e-3) * t2 - .559487e-3) * t2 + .7323931e-2) * t2 +
                      1.000000004;
    const double q1 =
        t * (((((-.9173e-5 * t2 + .40658e-4) * t2 - .99941e-4) * t2 + .266891e-3) * t2 - .1601836e-2) * t2 + .093749994);
    const double ta1 = x - .75 * PI;
    j1_ = a0 * (p1 * cos(ta1) - q1 * sin(ta1));
    y1_ = a0 * (p1 * sin(ta1) + q1 * cos(ta1));
}

// E1(x), Abramowitz & Stegun 5.1.53 (x <= 1) and 5.1.56 (x > 1), the
// reference's E1XA (FinGreen3D.f90:2881-2935).
__host__ __device__ inline double e1_as(double x)
{
    if (x == 0.0) return 1.0e300;
    if (x <= 1.0)
        return -log(x) + ((((1.07857e-3 * x - 9.76004e-3) * x + 5.519968e-2) * x - 0.24991055) * x + 0.99999193) * x -
               0.57721566;
    const double p = (((x + 8.5733287401) * x + 18.059016973) * x + 8.6347608925) * x + 0.2677737343;
    const double q = (((x + 9.5733223454) * x + 25.6329561486) * x + 21.0996530827) * x + 3.9584969228;
    return exp(-x) / x * p / q;
}

// E_n(x) by the forward recurrence E_{k} = (exp(-x) - x E_{k-1})/(k-1) from
// the approximation above (ENXA, FinGreen3D.f90:2818-2877).
__host__ __device__ inline double en_rec(int n, double x)
{
    if (n == 0) return exp(-x) / x;
    double e = e1_as(x);
    for (int k = 2; k <= n; ++k) e = (exp(-x) - x * e) / (k - 1.0);
    return e;
}

// Ei(x), x > 0: gamma + ln x + sum x^k/(k k!) up to x = 40, the asymptotic
// series beyond (EIX, FinGreen3D.f90:2937-2967).
__host__ __device__ inline double ei_series(double x)
{
    if (x == 0.0) return -1.0e300;
    if (x <= 40.0) {
        double s = 1.0, r = 1.0;
        for (int k = 1; k <= 100; ++k) {
            r = r * k * x / ((k + 1.0) * (k + 1.0));
            s += r;
            if (fabs(r / s) <= 1.0e-15) break;
        }
        return GA + log(x) + x * s;
    }
    double s = 1.0, r = 1.0;
    for (int k = 1; k <= 20; ++k) {
        r = r * k / x;
        s += r;
    }
    return exp(x) / x * s;
}

// erf(x): the series 2/sqrt(pi) x exp(-x^2) sum (2x^2)^k/(1.3...(2k+1)) for
// |x| < 3.5, the 12-term asymptotic form of erfc beyond (ERROR,
// FinGreen3D.f90:2745-2777).
__host__ __device__ inline double erf_zj(double x)
{
    const double x2 = x * x;
    if (fabs(x) < 3.5) {
        double er = 1.0, r = 1.0;
        for (int k = 1; k <= 50; ++k) {
            r = r * x2 / (k + 0.5);
            er += r;
            if (fabs(r) <= fabs(er) * 1.0e-15) break;
        }
        return 2.0 / sqrt(PI) * x * exp(-x2) * er;
    }
    double er = 1.0, r = 1.0;
    for (int k = 1; k <= 12; ++k) {
        r = -r * (k - 0.5) / x2;
        er += r;
    }
    const double v = 1.0 - exp(-x2) / (fabs(x) * sqrt(PI)) * er;
    return x < 0.0 ? -v : v;
}
__host__ __device__ inline double erfc_zj(double x) { return 1.0 - erf_zj(x); }
__host__ __device__ inline double derfc(double x) { return -2.0 / sqrt(PI) * exp(-x * x); }

// ---------------------------------------------------------------------------
// The epsilon algorithm on the partial sums s[0..n] (n >= 2), in place, with
// the stopping rule of Mishonov & Penev (Int. J. Mod. Phys. B 14, 3831, 2000):
// of the table's even columns, the entry whose neighbouring odd-column entry
// is largest in magnitude is the limit (Limes, FinGreen3D.f90:3053-3186).
// s and a (n+1 entries, a scratch) are strided by st.
__host__ __device__ inline double limes(int n, double* s, double* a, size_t st)
{
    // The tables live in global memory: each pass loads a block of LB
    // entries before it stores any of them (every entry a pass reads is one
    // it has not yet overwritten), so a wavefront waits on memory once per
    // block instead of once per entry.  The arithmetic and its order are the
    // sequential algorithm's.
    constexpr int LB = 8;
    double lim = s[n * st];
    for (int i = 0; i <= n; ++i) a[i * st] = 0.0;
    double amax = 0.0;
    for (int k = 1;; ++k) {
        const int la = n - 2 * k + 1;  // last index of the odd-column pass
        if (la < 0) break;
        // odd column from the previous even one (cross rule):
        // a(i) = a(i+1) + 1 / (s(i+1) - s(i))
        double sp = s[0];
        for (int i0 = 0; i0 <= la; i0 += LB) {
            double sv[LB], av[LB];
#pragma unroll
            for (int b = 0; b < LB; ++b) {
                const int j = i0 + 1 + b;
                sv[b] = j <= n ? s[j * st] : 0.0;
                av[b] = j <= n ? a[j * st] : 0.0;
            }
#pragma unroll
            for (int b = 0; b < LB; ++b) {
                const int i = i0 + b;
                if (i <= la) {
                    const double d = sv[b] - (b == 0 ? sp : sv[b - 1]);
                    a[i * st] = (d != 0.0) ? av[b] + 1.0 / d : av[b];
                }
            }
            sp = sv[LB - 1];
        }
        const int ls = n - 2 * k;  // last index of the even-column pass
        if (ls < 0) break;
        // even column: s(i) = s(i+1) + 1 / (a(i+1) - a(i)); the entry whose
        // a is the largest so far is the limit
        double ap = a[0];
        for (int i0 = 0; i0 <= ls; i0 += LB) {
            double sv[LB], av[LB];
#pragma unroll
            for (int b = 0; b < LB; ++b) {
                const int j = i0 + 1 + b;
                sv[b] = j <= n ? s[j * st] : 0.0;
                av[b] = j <= n ? a[j * st] : 0.0;
            }
#pragma unroll
            for (int b = 0; b < LB; ++b) {
                const int i = i0 + b;
                if (i <= ls) {
                    const double ai = (b == 0 ? ap : av[b - 1]), ai1 = av[b];
                    const double next = sv[b];
                    const double v = (ai1 != ai) ? next + 1.0 / (ai1 - ai) : next;
                    s[i * st] = v;
                    if (fabs(ai) > amax) {
                        amax = fabs(ai);
                        lim = v;
                        if (next == v) return lim;
                    }
                }
            }
            ap = av[LB - 1];
        }
    }
    return lim;
}

// ---------------------------------------------------------------------------
// Pieces shared by the series

struct Grn {
    C g[3];  // G, dG/dR, dG/dzf
};

// G_wave = G - 1/r - 1/r1: subtract the direct and free-surface image
// Rankine terms and their derivatives.
__host__ __device__ inline void take_out_rankine(double R, double ZF, double ZP, Grn& o)
{
    const double sr = sqrt(R * R + (ZF - ZP) * (ZF - ZP));
    const double sr1 = sqrt(R * R + (ZF + ZP) * (ZF + ZP));
    const double dsrr = R / sr, dsrz = (ZF - ZP) / sr, dsr1r = R / sr1, dsr1z = (ZF + ZP) / sr1;
    o.g[0].re -= (1.0 / sr + 1.0 / sr1);
    o.g[1].re += (dsrr / (sr * sr) + dsr1r / (sr1 * sr1));
    o.g[2].re += (dsrz / (sr * sr) + dsr1z / (sr1 * sr1));
}

// The propagating mode pi/N0 cosh(k(zf+h)) cosh(k(zp+h)) (i J0 - Y0)(kR) and
// its R and zf derivatives.  The reference scales the zf derivative by k in
// two places of EIGEN/EIGENE and once in PIDCOCK (deriv_has_k).
__host__ __device__ inline void propagating(double R, double ZF, double ZP, double WK, double H, bool pidcock,
                                            Grn& o)
{
    const double nm0 = H / 2.0 * (1.0 + sinh(2.0 * WK * H) / (2.0 * WK * H));
    const double p0 = PI / nm0 * cosh(WK * (ZF + H)) * cosh(WK * (ZP + H));
    const double dp0 = (pidcock ? WK * PI : PI) / nm0 * sinh(WK * (ZF + H)) * cosh(WK * (ZP + H));
    double J0, J1, Y0, Y1;
    bessel_jy01(WK * R, J0, J1, Y0, Y1);
    o.g[0] = cmk(-p0 * Y0, p0 * J0);
    o.g[1] = cmk(WK * p0 * Y1, -WK * p0 * J1);
    const double s = pidcock ? dp0 : WK * dp0;
    o.g[2] = cmk(-s * Y0, s * J0);
}

// The per-term factors of the series that depend on one root or on one
// point's depth only.  TrigDirect computes them where they are used (the
// reference's expressions); TrigTable reads them from tables the kernels fill
// beforehand with the same expressions, so both give the same bits.  Term i
// (2 <= i <= 99) uses root wvn[i - 1] and umt = (i - 1) pi / h.
struct TrigDirect {
    double ZF, ZP, H, V;
    const double* wvn;
    __host__ __device__ double pm(int i) const
    {
        const double KM = wvn[i - 1];
        return 4.0 / (1.0 + sin(2.0 * KM * H) / (2.0 * KM * H)) / H;
    }
    __host__ __device__ double umt(int i) const { return (i - 1) * PI / H; }
    __host__ __device__ double ac(int i) const
    {
        const double um = wvn[i - 1];
        return (um * um + V * V) / ((um * um + V * V) * H - V);
    }
    __host__ __device__ double cf(int i) const { return cos(wvn[i - 1] * (ZF + H)); }
    __host__ __device__ double sf(int i) const { return sin(wvn[i - 1] * (ZF + H)); }
    __host__ __device__ double cp(int i) const { return cos(wvn[i - 1] * (ZP + H)); }
    __host__ __device__ double ctf(int i) const { return cos(umt(i) * (ZF + H)); }
    __host__ __device__ double stf(int i) const { return sin(umt(i) * (ZF + H)); }
    __host__ __device__ double ctp(int i) const { return cos(umt(i) * (ZP + H)); }
};

constexpr int TAB_TERMS = 100;  // term index i < 100

// root: pm, ac, umt by term index (3 x TAB_TERMS); pf, pp: the point table of
// the zf and of the zp point, entry (c, i) at [(4 i + c) npts] with c = 0..3
// cos(k (z + h)), sin(k (z + h)), cos(umt (z + h)), sin(umt (z + h)).
struct TrigTable {
    const double* root;
    const double* pf;
    const double* pp;
    size_t npts;
    __host__ __device__ double pm(int i) const { return root[i]; }
    __host__ __device__ double ac(int i) const { return root[TAB_TERMS + i]; }
    __host__ __device__ double umt(int i) const { return root[2 * TAB_TERMS + i]; }
    __host__ __device__ double cf(int i) const { return pf[(4 * (size_t)i + 0) * npts]; }
    __host__ __device__ double sf(int i) const { return pf[(4 * (size_t)i + 1) * npts]; }
    __host__ __device__ double ctf(int i) const { return pf[(4 * (size_t)i + 2) * npts]; }
    __host__ __device__ double stf(int i) const { return pf[(4 * (size_t)i + 3) * npts]; }
    __host__ __device__ double cp(int i) const { return pp[(4 * (size_t)i + 0) * npts]; }
    __host__ __device__ double ctp(int i) const { return pp[(4 * (size_t)i + 2) * npts]; }
};

// One evanescent term of the eigenfunction expansion (EIGEN/EIGENE).
template <class T>
__host__ __device__ inline void evanescent(double R, int i, double KM, const T tr, double rm[3])
{
    const double pm = tr.pm(i);
    double K0, K1;
    bessel_k01(KM * R, K0, K1);
    const double cf = tr.cf(i), cp = tr.cp(i);
    rm[0] = pm * K0 * cf * cp;
    rm[1] = -pm * KM * K1 * cf * cp;
    rm[2] = -pm * KM * K0 * tr.sf(i) * cp;
}

// R/h >= 0.5: at most 10 terms, stopped when a term is below 1e-6.
template <class T>
__host__ __device__ inline void eigen(double R, double ZF, double ZP, double WK, const double* wvn, double H,
                                     const T tr, Grn& o)
{
    propagating(R, ZF, ZP, WK, H, false, o);
    for (int i = 2; i <= 10; ++i) {
        double rm[3];
        evanescent(R, i, wvn[i - 1], tr, rm);
        o.g[0].re += rm[0];
        o.g[1].re += rm[1];
        o.g[2].re += rm[2];
        if (fabs(rm[0]) < EPS_1E6 && fabs(rm[1]) < EPS_1E6 && fabs(rm[2]) < EPS_1E6) break;
    }
    take_out_rankine(R, ZF, ZP, o);
}

// 0.05 <= R/h < 0.5: NT = int(-88.89 R/h + 54.45) terms, epsilon limit.
// ws: four arrays of LIMES_LEN doubles at stride st.
template <class T>
__host__ __device__ inline void eigene(double R, double ZF, double ZP, double WK, const double* wvn, double H,
                                      const T tr, double* ws, size_t st, Grn& o)
{
    const double rhp = fabs(R / H);
    const int nt = (int)(-C_EIGENE_A * rhp + C_EIGENE_B);
    propagating(R, ZF, ZP, WK, H, false, o);
    double* g = ws;
    double* gr = ws + LIMES_LEN * st;
    double* gz = ws + 2 * LIMES_LEN * st;
    double* a = ws + 3 * LIMES_LEN * st;
    // partial sums carried in registers, stored for the epsilon tables
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    g[0] = gr[0] = gz[0] = 0.0;
    for (int i = 2; i <= nt; ++i) {
        double rm[3];
        evanescent(R, i, wvn[i - 1], tr, rm);
        s0 = s0 + rm[0];
        s1 = s1 + rm[1];
        s2 = s2 + rm[2];
        g[(i - 1) * st] = s0;
        gr[(i - 1) * st] = s1;
        gz[(i - 1) * st] = s2;
    }
    o.g[0].re += limes(nt - 1, g, a, st);
    o.g[1].re += limes(nt - 1, gr, a, st);
    o.g[2].re += limes(nt - 1, gz, a, st);
    take_out_rankine(R, ZF, ZP, o);
}

// Newman's (1985) fit of the remainder of the bottom-image series,
// coefficients c[m][n] of (R/h)^(2m) ((zf-zp)/h)^(2n) (+ the zf+zp+2h image).
__host__ __device__ inline double newman_amn(int m, int n)
{
    // row m, column n
    const double t[5][5] = {
        {0.0, 0.05051418, 0.00230838, 0.00012934, 0.00000913},
        {-0.02525711, -0.00692380, -0.00097875, -0.00010879, -0.00001270},
        {0.00086546, 0.00073292, 0.00020597, 0.00003965, 0.00000466},
        {-0.00004063, -0.00006636, -0.00003333, -0.00000891, 0.0},
        {0.00000193, 0.00000398, 0.00000524, 0.0, 0.0},
    };
    return t[m][n];
}

// The image sum of Pidcock's form and its R and zf derivatives
// (IMGS, DGSR, DGSZ, FinGreen3D.f90:896-1084).
__host__ __device__ inline void image_sum(double R, double ZF, double ZP, double H, double out[3])
{
    const double t1 = R / H, t2 = (ZF - ZP) / H, t3 = (ZF + ZP + 2.0 * H) / H;
    double s1 = 0.0, s1r = 0.0, s1z = 0.0;
    for (int l = -1; l <= 1; ++l) {
        const double za = t2 + 2.0 * l, zb = t3 + 2.0 * l;
        const double ra = sqrt(t1 * t1 + za * za), rb = sqrt(t1 * t1 + zb * zb);
        s1 = s1 + 1.0 / ra + 1.0 / rb;
        s1r = s1r + t1 / (ra * ra * ra) + t1 / (rb * rb * rb);
        s1z = s1z + za / (ra * ra * ra) + zb / (rb * rb * rb);
    }
    double s2 = 0.0, s2r = 0.0, s2z = 0.0;
#pragma unroll
    for (int m = 0; m <= 4; ++m)
#pragma unroll
        for (int n = 0; n <= 4; ++n) {
            const double c = newman_amn(m, n);
            s2 += c * pow(t1, 2.0 * m) * (pow(t2, 2.0 * n) + pow(t3, 2.0 * n));
            if (m >= 1) s2r += 2.0 * m * c * pow(t1, 2.0 * m - 1.0) * (pow(t2, 2.0 * n) + pow(t3, 2.0 * n));
            if (n >= 1) s2z += 2.0 * n * c * pow(t1, 2.0 * m) * (pow(t2, 2.0 * n - 1.0) + pow(t3, 2.0 * n - 1.0));
        }
    out[0] = (s1 - 2.0 + s2) / H;
    out[1] = (-s1r + s2r) / (H * H);
    out[2] = (-s1z + s2z) / (H * H);
}

// 0.0005 <= R/h < 0.05: Pidcock's form, NT = int(-1010.10 R/h + 100.50).
template <class T>
__host__ __device__ inline void pidcock(double R, double ZF, double ZP, double V, double WK, const double* wvn,
                                       double H, const T tr, double* ws, size_t st, Grn& o)
{
    propagating(R, ZF, ZP, WK, H, true, o);
    o.g[0].re += 2.0 / H * (GA + log(R / 4.0 / H));
    o.g[1].re += 2.0 / (R * H);
    const double rhp = fabs(R / H);
    const int nt = (int)(-C_PIDCOCK_A * rhp + C_PIDCOCK_B);
    double* g = ws;
    double* gr = ws + LIMES_LEN * st;
    double* gz = ws + 2 * LIMES_LEN * st;
    double* a = ws + 3 * LIMES_LEN * st;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    g[0] = gr[0] = gz[0] = 0.0;
    for (int i = 2; i <= nt; ++i) {
        const double umt = tr.umt(i);  // the evanescent roots' deep-water limit m pi / h
        const double um = wvn[i - 1];
        double K0, K1, K0t, K1t;
        bessel_k01(um * R, K0, K1);
        bessel_k01(umt * R, K0t, K1t);
        const double ac = tr.ac(i);
        const double cu = tr.cp(i), ct = tr.ctp(i);
        const double am = ac * tr.cf(i) * cu;
        const double amt = tr.ctf(i) * ct / H;
        const double dam = -um * ac * tr.sf(i) * cu;
        const double damt = -umt * tr.stf(i) * ct / H;
        s0 = s0 + (am * K0 - amt * K0t);
        s1 = s1 + (-um * am * K1 + umt * amt * K1t);
        s2 = s2 + (dam * K0 - damt * K0t);
        g[(i - 1) * st] = s0;
        gr[(i - 1) * st] = s1;
        gz[(i - 1) * st] = s2;
    }
    const double l0 = limes(nt - 1, g, a, st), l1 = limes(nt - 1, gr, a, st), l2 = limes(nt - 1, gz, a, st);
    double im[3];
    image_sum(R, ZF, ZP, H, im);
    o.g[0].re = o.g[0].re + 4.0 * l0 + im[0];
    o.g[1].re = o.g[1].re + 4.0 * l1 + im[1];
    o.g[2].re = o.g[2].re + 4.0 * l2 + im[2];
    take_out_rankine(R, ZF, ZP, o);
}

// ---------------------------------------------------------------------------
// Linton's form (R/h < 0.0005)

// Gauss-Kronrod (7, 15) on [-1, 1]: node i = 1..15 in increasing order, its
// Kronrod weight and, at even i (the Gauss nodes), the Gauss weight.  The
// published values rounded to 15 digits, as the reference carries them
// (FinGreen3D.f90:2360-2395).  Switches rather than arrays keep them out of
// private memory when the index is not a constant.
__host__ __device__ inline double gk_node(int i)
{
    double v;
    switch (i < 8 ? 8 - i : i - 8) {
    case 7: v = 0.991455371120813; break;
    case 6: v = 0.949107912342759; break;
    case 5: v = 0.864864423359769; break;
    case 4: v = 0.741531185599394; break;
    case 3: v = 0.586087235467691; break;
    case 2: v = 0.405845151377397; break;
    case 1: v = 0.207784955007898; break;
    default: v = 0.0;
    }
    return i < 8 ? -v : v;
}
__host__ __device__ inline double gk_wkron(int i)
{
    switch (i < 8 ? 8 - i : i - 8) {
    case 7: return 0.022935322010529;
    case 6: return 0.063092092629979;
    case 5: return 0.104790010322250;
    case 4: return 0.140653259715525;
    case 3: return 0.169004726639267;
    case 2: return 0.190350578064785;
    case 1: return 0.204432940075298;
    default: return 0.209482141084728;
    }
}
__host__ __device__ inline double gk_wgauss(int i)  // i even
{
    switch (i < 8 ? 8 - i : i - 8) {
    case 6: return 0.129484966168870;
    case 4: return 0.279705391489277;
    case 2: return 0.381830050505119;
    default: return 0.417959183673469;
    }
}

// The integrands of Linton's form.  tag 1, 2: the coefficients' integrands
// F1, F2 of t for the evanescent (no != 1) or propagating (no == 1) mode;
// tag 3..5: G1, G2, G3 of the error-function integral (FinGreen3D.f90:1715-2174).
struct LintonF {
    double R, wvn;
    int no;
    double V, ZF, ZP, H;
    int tag;
    __host__ __device__ double operator()(double t) const
    {
        if (fabs(t) < 1.0e-8) return 0.0;
        if (tag <= 2) {
            const double e = (no == 1) ? exp(-R * R / 4.0 / t + wvn * wvn * t) : exp(-R * R / 4.0 / t - wvn * wvn * t);
            return tag == 1 ? e / 4.0 / PI / t : -R * e / 8.0 / PI / (t * t);
        }
        const double val[4] = {-ZP - ZF, 2.0 * H - ZP + ZF, 2.0 * H + ZP - ZF, 4.0 * H + ZP + ZF};
        const double dval[4] = {-1.0, 1.0, -1.0, 1.0};
        double sum = 0.0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const double arg = val[i] / 2.0 / t - V * t;
            const double ev = exp(-V * val[i]);
            if (tag == 5)
                sum = sum - V * dval[i] * ev * erfc_zj(arg) + ev * derfc(arg) * dval[i] / 2.0 / t;
            else
                sum += ev * erfc_zj(arg);
        }
        const double e = exp(V * V * t * t - R * R / 4.0 / (t * t));
        if (tag == 4) return R / 2.0 / (t * t * t) * e * sum;
        return e * sum / t;
    }
};

// Adaptive Gauss-Kronrod integral of f on [a, b]: an interval is accepted
// when (200 |K15 - G7|)^1.5 <= eps or it is no longer than eps, else halved
// (AQUAD/AQUADF, FinGreen3D.f90:2348-2702).  The recursion is walked depth
// first, left half first, without a stack: the node at depth d with path bits
// `path` is found again by halving from [a, b].
template <class F>
__host__ __device__ inline double aquad(const F& f, double a0, double b0, double eps)
{
    double total = 0.0;
    int depth = 0;
    unsigned long long path = 0;
    for (int guard = 0; guard < (1 << 20); ++guard) {
        double a = a0, b = b0;
        for (int l = depth - 1; l >= 0; --l) {
            const double mid = (a + b) / 2.0;
            if ((path >> l) & 1ull) a = mid;
            else b = mid;
        }
        double gk = 0.0, g7 = 0.0;
        for (int i = 1; i <= 15; ++i) {
            const double fi = f(1.0 / 2.0 * (b + a + (b - a) * gk_node(i)));
            if ((i & 1) == 0) g7 = g7 + gk_wgauss(i) * (b - a) / 2.0 * fi;
            gk = gk + gk_wkron(i) * (b - a) / 2.0 * fi;
        }
        const double err = pow(200.0 * fabs(gk - g7), 1.5);
        if (err > eps && fabs(a - b) > eps && depth < 60) {
            ++depth;
            path <<= 1;  // left half
            continue;
        }
        total += gk;
        // next: climb while this node is a right half, then take the sibling
        while (depth > 0 && (path & 1ull)) {
            path >>= 1;
            --depth;
        }
        if (depth == 0) break;
        path |= 1ull;
    }
    return total;
}

// The series coefficients of Linton's form for mode i (1: propagating, with
// wvn = k; else the evanescent root), and their R derivatives
// (COEF/DCOEF/SA0/DSA0/SAM/DSAM, FinGreen3D.f90:1148-1668).
__host__ __device__ inline double sa0(double R, double PA, double WK, bool deriv)
{
    double s = deriv ? 0.0 : -GA - 2.0 * log(WK * PA / 2.0);
    double sgn = -1.0, fac = 1.0, rm = 100.0;
    for (int n = 1; fabs(rm) > EPS_1E6 && n < 400; ++n) {
        double psi = 0.0;
        for (int p = 1; p <= n; ++p) psi += 1.0 / p;
        const double q = R / PA, wr = WK * R / 2.0, wp = WK * PA / 2.0;
        if (!deriv) {
            const double en = en_rec(n + 1, q * q);
            rm = sgn * pow(q, 2.0 * n) / fac / n - pow(wp, 2.0 * n) / fac * en -
                 2.0 * sgn / (fac * fac) * pow(wr, 2.0 * n) * (log(wr) + GA - psi);
        } else {
            const double en = en_rec(n, q * q);
            rm = 2.0 * sgn * pow(q, 2.0 * n - 1.0) / fac / PA + pow(wp, 2.0 * n) / fac * en * 2.0 * R / (PA * PA) -
                 2.0 * sgn * n * WK / (fac * fac) * pow(wr, 2.0 * n - 1.0) * (log(wr) + GA - psi) -
                 2.0 * sgn / (fac * fac) * pow(wr, 2.0 * n) / R;
        }
        s += rm;
        sgn = -sgn;
        fac *= (n + 1);
    }
    return s / 4.0 / PI;
}

__host__ __device__ inline double sam(double R, double PA, double WVN, bool deriv)
{
    double s = 0.0, rm = 100.0;
    const double q = R / PA, arg = (WVN * PA / 2.0) * (WVN * PA / 2.0);
    if (!deriv) {
        double sgn = 1.0, fac = 1.0;
        for (int n = 0; fabs(rm) > EPS_1E6 && n < 400; ++n) {
            rm = sgn * pow(q, 2.0 * n) / fac * en_rec(n + 1, arg);
            s += rm;
            sgn = -sgn;
            fac *= (n + 1);
        }
    } else {
        double sgn = -1.0, fac = 1.0;
        for (int n = 1; fabs(rm) > EPS_1E6 && n < 400; ++n) {
            rm = 2.0 * n * sgn * pow(q, 2.0 * n - 1.0) / fac / PA * en_rec(n + 1, arg);
            s += rm;
            sgn = -sgn;
            fac *= (n + 1);
        }
    }
    return s / 4.0 / PI;
}

__host__ __device__ __forceinline__ double linton_coef(int i, double R, double PA, double A, double wvn_i, double WK,
                                             bool deriv)
{
    const double dnt = PA * PA / 4.0;
    if (fabs(R) < EPS_1E6) {
        if (deriv) return 0.0;
        return i == 1 ? -ei_series(WK * WK * dnt) / 4.0 / PI : e1_as(wvn_i * wvn_i * dnt) / 4.0 / PI;
    }
    if (i == 1) {
        if (R / A <= 0.5) return sa0(R, PA, WK, deriv);
        LintonF f{R, WK, 1, 0.0, 0.0, 0.0, 0.0, deriv ? 2 : 1};
        const double q = aquad(f, 0.0, dnt, EPS_1E6);
        double J0, J1, Y0, Y1;
        bessel_jy01(WK * R, J0, J1, Y0, Y1);
        return deriv ? WK * Y1 / 4.0 - q : -Y0 / 4.0 - q;
    }
    if (R / A <= 1.0) return sam(R, PA, wvn_i, deriv);
    LintonF f{R, wvn_i, i, 0.0, 0.0, 0.0, 0.0, deriv ? 2 : 1};
    double p1 = dnt + 100.0;
    for (int guard = 0; fabs(f(p1)) > EPS_1E6 && guard < 100000; ++guard) p1 += 100.0;
    return aquad(f, dnt, p1, EPS_1E6);
}

__host__ __device__ inline C nm(double H, C k)
{
    const C two_kh = cmk(2.0 * H, 0.0) * k;
    return (H / 2.0) * (cmk(1.0, 0.0) + csin(two_kh) / two_kh);
}

__host__ __device__ __forceinline__ void linton(double R, double ZF, double ZP, double V, double WK, const double* wvn,
                                      double H, Grn& o)
{
    double A;
    if (V < 0.02) A = 3324.2 * V * V * V - 258.89 * V * V + 8.3958 * V + 1.3208;
    else if (V < 0.1) A = -30.303 * V * V * V - 2.8138 * V * V + 2.7740 * V + 0.05052;
    else if (V < 1.0) A = 1.5699 * V * V * V - 2.5075 * V * V + 1.8316 * V + 0.1077;
    else A = -0.000069641 * V * V * V - 0.0017082 * V * V + 0.45566 * V + 0.80551;
    A = fabs(A) / (WK * H);
    const double PA = A * H;
    const double sr = sqrt(R * R + (ZF - ZP) * (ZF - ZP));
    const double sr1 = sqrt(R * R + (2.0 * H + ZF + ZP) * (2.0 * H + ZF + ZP));
    const double sr2 = sqrt(R * R + (ZF + ZP) * (ZF + ZP));
    const double dsrr = R / sr, dsrz = (ZF - ZP) / sr;
    const double dsr1r = R / sr1, dsr1z = (2.0 * H + ZF + ZP) / sr1;
    const double dsr2r = R / sr2, dsr2z = (ZF + ZP) / sr2;

    C g[3] = {cmk(0, 0), cmk(0, 0), cmk(0, 0)};
    for (int i = 1; i <= 20; ++i) {
        const C km = (i == 1) ? cmk(0.0, -WK) : cmk(wvn[i - 1], 0.0);
        // one call site for both coefficients keeps a single inlined copy
        double cof = 0.0, dcof = 0.0;
#pragma nounroll
        for (int d = 0; d < 2; ++d) {
            const double c = linton_coef(i, R, PA, A, wvn[i - 1], WK, d != 0);
            if (d) dcof = c;
            else cof = c;
        }
        const C n_ = nm(H, km);
        const C cf = ccos(km * cmk(ZF + H, 0.0)), cp = ccos(km * cmk(ZP + H, 0.0));
        const C sf = csin(km * cmk(ZF + H, 0.0));
        const C r0 = (cmk(-cof, 0.0) / n_) * cf * cp;
        const C r1 = (cmk(-dcof, 0.0) / n_) * cf * cp;
        const C r2 = (cmk(cof, 0.0) * km / n_) * sf * cp;
        g[0] = g[0] + r0;
        g[1] = g[1] + r1;
        g[2] = g[2] + r2;
        if (i > 2 && cabs_(r0) < EPS_1E6 && cabs_(r1) < EPS_1E6 && cabs_(r2) < EPS_1E6) break;
    }
    // images of the error-function part
    const double val[4] = {-ZP - ZF, 2.0 * H - ZP + ZF, 2.0 * H + ZP - ZF, 4.0 * H + ZP + ZF};
    double e0 = 0.0, e1 = 0.0, e2 = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const double ro = sqrt(R * R + val[i] * val[i]);
        const double dror = R / ro;
        const double droz = val[i] * ((i & 1) ? 1.0 : -1.0) / ro;  // (-1)^I, I = i + 1
        const double ec = erfc_zj(ro / PA), dc = derfc(ro / PA);
        e0 = e0 + ec / (4.0 * PI * ro);
        e1 = e1 + (dc / PA * dror * ro - dror * ec) / (4.0 * PI * ro * ro);
        e2 = e2 + (dc / PA * droz * ro - droz * ec) / (4.0 * PI * ro * ro);
    }
    const double ec1 = erfc_zj(sr1 / PA), dc1 = derfc(sr1 / PA);
    if (fabs(sr) > EPS_1E6) {
        const double ec = erfc_zj(sr / PA), dc = derfc(sr / PA);
        e0 = e0 + (ec - 1.0) / (4.0 * PI * sr) + ec1 / (4.0 * PI * sr1) - 1.0 / (4.0 * PI * sr2);
        e1 = e1 - (ec - 1.0) / (4.0 * PI * sr * sr) * dsrr + dc / (4.0 * PI * sr * PA) * dsrr -
             ec1 / (4.0 * PI * sr1 * sr1) * dsr1r + dc1 / (4.0 * PI * sr1 * PA) * dsr1r +
             dsr2r / (4.0 * PI * sr2 * sr2);
        e2 = e2 - (ec - 1.0) / (4.0 * PI * sr * sr) * dsrz + dc / (4.0 * PI * sr * PA) * dsrz -
             ec1 / (4.0 * PI * sr1 * sr1) * dsr1z + dc1 / (4.0 * PI * sr1 * PA) * dsr1z +
             dsr2z / (4.0 * PI * sr2 * sr2);
    } else {
        e0 = e0 + ec1 / (4.0 * PI * sr1) - 1.0 / (4.0 * PI * sr2);
        e1 = e1 - ec1 / (4.0 * PI * sr1 * sr1) * dsr1r + dc1 / (4.0 * PI * sr1 * PA) * dsr1r +
             dsr2r / (4.0 * PI * sr2 * sr2);
        e2 = e2 - ec1 / (4.0 * PI * sr1 * sr1) * dsr1z + dc1 / (4.0 * PI * sr1 * PA) * dsr1z +
             dsr2z / (4.0 * PI * sr2 * sr2);
    }
    LintonF fq{R, 0.0, 0, V, ZF, ZP, H, 3};
    const double q0 = aquad(fq, 0.0, PA / 2.0, EPS_1E6);
    fq.tag = 4;
    const double q1 = aquad(fq, 0.0, PA / 2.0, EPS_1E6);
    fq.tag = 5;
    const double q2 = aquad(fq, 0.0, PA / 2.0, EPS_1E6);
    g[0].re = g[0].re - e0 - V / 2.0 / PI * q0;
    g[1].re = g[1].re - e1 + V / 2.0 / PI * q1;
    g[2].re = g[2].re - e2 - V / 2.0 / PI * q2;
    // the propagating mode's J0 part
    double bj0, bj1, by0, by1;
    bessel_jy01(WK * R, bj0, bj1, by0, by1);
    const C n0 = nm(H, cmk(0.0, -WK));
    const double chf = cosh(WK * (ZF + H)), chp = cosh(WK * (ZP + H)), shf = sinh(WK * (ZF + H));
    const C i4 = cmk(0.0, 1.0) / cmk(4.0, 0.0) / n0;  // i/4/N0
    const C ik4 = cmk(0.0, WK) / cmk(4.0, 0.0) / n0;  // i k/4/N0
    g[0] = g[0] - chp * (chf * (bj0 * i4));
    g[1] = g[1] + chp * (chf * (bj1 * ik4));
    g[2] = g[2] - chp * (shf * (bj0 * ik4));
    for (int c = 0; c < 3; ++c) o.g[c] = (-4.0 * PI) * g[c];
}

// The series fingreen3d_routine picks for R/h: 0 eigenfunction expansion,
// 1 the same with its epsilon limit, 2 Pidcock's form, 3 Linton's form.
__host__ __device__ inline int region(double RR, double H)
{
    const double rhp = fabs(RR / H);
    return rhp < 0.0005 ? 3 : (rhp < 0.05 ? 2 : (rhp < 0.5 ? 1 : 0));
}

// One region's series, for kernels whose pairs were sorted by region
// (REG = region(RR, H) is the caller's guarantee); the R = 0 convention of
// point() included.
template <int REG, class T>
__host__ __device__ __forceinline__ Grn point_in(double RR, double ZF, double ZP, double V, const double* wvn,
                                                 double H, const T tr, double* ws, size_t st)
{
    Grn o;
    const double WK = wvn[0];
    if (REG == 0) eigen(RR, ZF, ZP, WK, wvn, H, tr, o);
    else if (REG == 1) eigene(RR, ZF, ZP, WK, wvn, H, tr, ws, st, o);
    else if (REG == 2) pidcock(RR, ZF, ZP, V, WK, wvn, H, tr, ws, st, o);
    else linton(RR, ZF, ZP, V, WK, wvn, H, o);
    if (RR == 0.0) o.g[1] = cmk(0.0, 0.0);
    return o;
}

// fingreen3d_routine: the region choice and the R = 0 convention (dG/dR = 0).
// v = omega^2/g, wvn[0] = k the propagating root, wvn[1..] the evanescent
// ones (at least 99 of them); ws/st as for eigene.
template <class T>
__host__ __device__ __forceinline__ Grn point(double RR, double ZF, double ZP, double V, const double* wvn, double H,
                                              const T tr, double* ws, size_t st)
{
    Grn o;
    const double WK = wvn[0];
    const double rhp = fabs(RR / H);
    if (rhp < 0.0005) linton(RR, ZF, ZP, V, WK, wvn, H, o);
    else if (rhp < 0.05) pidcock(RR, ZF, ZP, V, WK, wvn, H, tr, ws, st, o);
    else if (rhp < 0.5) eigene(RR, ZF, ZP, WK, wvn, H, tr, ws, st, o);
    else eigen(RR, ZF, ZP, WK, wvn, H, tr, o);
    if (RR == 0.0) o.g[1] = cmk(0.0, 0.0);
    return o;
}

}  // namespace fg

namespace fg {
// point() with the factors computed in place
__host__ __device__ __forceinline__ Grn point(double RR, double ZF, double ZP, double V, const double* wvn, double H,
                                              double* ws, size_t st)
{
    const TrigDirect tr{ZF, ZP, H, V, wvn};
    return point(RR, ZF, ZP, V, wvn, H, tr, ws, st);
}
}  // namespace fg
