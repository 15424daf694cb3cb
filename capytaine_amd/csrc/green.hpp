// Device-side Delhommeau Green function for gfx950.
//
// Restates, for one (collocation point, panel) pair, the reference routines
//   Green_Rankine.f90:28-338   Rankine and reflected-Rankine panel integrals,
//   Green_wave.f90:50-271      infinite-depth wave term (tabulated / asymptotic),
//   Green_wave.f90:311-519     Delhommeau finite-depth images and Prony terms,
//   Delhommeau_integrals.f90:157-186, 274-402   asymptotics, nearest indices,
//                              3x3 Lagrange interpolation in the table.
// All arithmetic is FP64.  Constants follow constants.f90: PI and
// EULER_GAMMA are float32 literals widened to double.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "integrals.hpp"

namespace cpt {

constexpr double PI = (double)3.141592653588979f;
constexpr double EULER_GAMMA = (double)0.5772156649f;
constexpr double LOG_2 = 0.69314718055994530942;
constexpr double M1_4PI = -1.0 / (4.0 * PI);
constexpr double DBL_EPS = 2.220446049250313e-16;

enum { LEGACY_GRID = 0, SCALED_NEMOH3_GRID = 1 };
enum { LEGACY_FINITE_DEPTH = 0, NEWER_FINITE_DEPTH = 1 };
enum { HIGH_FREQ = 0, LOW_FREQ = 1, LOW_FREQ_WITH_RANKINE_PART = 2 };

// One panel as the assembly kernel reads it: 24 doubles (192 B) so that a
// wave-uniform panel index turns into scalar (SMEM) loads.
struct alignas(64) Face {
    double v[4][3];   // vertices (triangles repeat their last vertex)
    double c[3];      // centre
    double n[3];      // unit normal
    double area;
    double radius;
    double pad[4];
};
static_assert(sizeof(Face) == 192, "Face record must stay 192 bytes");

// Wave-term table as re-laid out on the device: cell (i, j) holds the five
// tabulated integrals contiguously at tab[(j * nr + i) * 5 + c].
struct Table {
    const double* tab;      // (nz, nr, 5)
    const double* r_range;  // nr
    const double* z_range;  // nz
    int nr, nz, grid, npts;
    double rmax;            // r_range[nr-1]
    double zmax_abs;        // |z_range[nz-1]|
    double dz;              // (log10(|z_range[nz-1]|) + 10) / nz  (scaled grid)
    int index_of_1;         // change-of-slope index of the scaled r grid
    // float copies (same layout) for the FP32 far field of the mixed form
    const float* tab_f;
    const float* r_range_f;
    const float* z_range_f;
};

struct cplx { double re, im; };

__device__ __forceinline__ double norm3(double x, double y, double z) { return sqrt(x * x + y * y + z * z); }

// NORM2 as gfortran evaluates it: a scaled sum of squares, every operation
// rounded on its own (no contraction), scale * sqrt(sum) at the end.
__device__ __noinline__ double norm2_rounded(double x, double y, double z)
{
#pragma clang fp contract(off)
    const double v[3] = {x, y, z};
    double scale = 0.0, ssq = 1.0;
    for (int i = 0; i < 3; ++i) {
        if (v[i] != 0.0) {
            const double a = fabs(v[i]);
            if (scale < a) {
                const double t = scale / a;
                ssq = 1.0 + ssq * t * t;
                scale = a;
            } else {
                const double t = a / scale;
                ssq += t * t;
            }
        }
    }
    return scale * sqrt(ssq);
}

// sqrt(a^2 + b^2) correctly rounded (a, b of moderate size): the sum of
// squares as a double-double (FMA residuals), then one Newton correction of
// its rounded square root.
__device__ __noinline__ double hypot_rounded(double a, double b)
{
#pragma clang fp contract(off)
    const double p = a * a, pe = fma(a, a, -p);
    const double q = b * b, qe = fma(b, b, -q);
    const double s = p + q, t = s - p;
    const double lo = ((p - (s - t)) + (q - t)) + pe + qe;
    const double r = sqrt(s);
    if (r == 0.0) return 0.0;
    return r + (fma(-r, r, s) + lo) / (2.0 * r);
}

// The near-field tests r0 > 7 R and r1 > 7 R (Green_Rankine.f90:48-50 and
// :224-227).  On regular meshes 7 R often equals a centre-to-centre distance
// exactly, and the last bit of the reference's distance then picks the
// branch (the exact and one-point integrals differ by ~1e-5 there).  The
// fast distance r is within a few ulps of the reference's; inside that
// window the test is taken again on the reference's expression: NORM2 of
// the difference for the direct term, HYPOT of the horizontal NORM2 and the
// vertical distance for the image (HYPOT correctly rounded; a C library's
// hypot can be one ulp off in rare cases).
__device__ __forceinline__ bool far_direct(double dx, double dy, double dz, double r, double radius)
{
    const double thr = 7 * radius;
    if (fabs(r - thr) > 16 * 2.220446049250313e-16 * thr) return r > thr;
    return norm2_rounded(dx, dy, dz) > thr;
}

__device__ __forceinline__ bool far_image(double dx, double dy, double dz, double r, double radius)
{
    const double thr = 7 * radius;
    if (fabs(r - thr) > 16 * 2.220446049250313e-16 * thr) return r > thr;
    return hypot_rounded(norm2_rounded(dx, dy, 0.0), dz) > thr;
}

// Fortran NINT (half away from zero) and INT (truncation); values outside
// the int range give INT_MIN like the reference's x86 conversion, which the
// callers then clamp.
__device__ __forceinline__ int f_nint(double x)
{
    if (!(x > -2147483648.0 && x < 2147483647.0)) return INT32_MIN;
    return (int)round(x);
}
__device__ __forceinline__ int f_int(double x)
{
    if (!(x > -2147483648.0 && x < 2147483647.0)) return INT32_MIN;
    return (int)x;
}

// ---------------------------------------------------------------------------
// Rankine part (Green_Rankine.f90)
// ---------------------------------------------------------------------------

// Integral of 1/r over a panel and its gradient.
struct RankineInt {
    double G, gx, gy, gz;
};

// exact_integral_of_Rankine, Green_Rankine.f90:68-157.  Gradient with respect
// to the first variable (M); callers flip signs for the other convention.
// Out of line (register pressure of the caller); the result is returned by
// value, in registers, rather than through pointers into private memory.
__device__ __noinline__ RankineInt exact_rankine(double mx, double my, double mz, const Face* __restrict__ Fp)
{
    double G, gx, gy, gz;
    const Face& F = *Fp;
    const double nx = F.n[0], ny = F.n[1], nz = F.n[2];
    const double GZ = (mx - F.c[0]) * nx + (my - F.c[1]) * ny + (mz - F.c[2]) * nz;
    const double aGZ = fabs(GZ);
    const double sGZ = copysign(1.0, GZ);
    double RR[4], DX[4], DY[4], DZ[4];
#pragma unroll
    for (int L = 0; L < 4; ++L) {
        const double ax = mx - F.v[L][0], ay = my - F.v[L][1], az = mz - F.v[L][2];
        RR[L] = norm3(ax, ay, az);
        const double inv = 1.0 / RR[L];
        DX[L] = ax * inv;
        DY[L] = ay * inv;
        DZ[L] = az * inv;
    }
    G = 0.0;
    gx = gy = gz = 0.0;
    const double dk_min = (double)1e-3f * F.radius;
    const double gz_min = (double)1e-4f * F.radius;
#pragma unroll
    for (int L = 0; L < 4; ++L) {
        const int Ln = (L + 1) & 3;
        const double ex = F.v[Ln][0] - F.v[L][0], ey = F.v[Ln][1] - F.v[L][1], ez = F.v[Ln][2] - F.v[L][2];
        const double DK = norm3(ex, ey, ez);
        if (DK >= dk_min) {
            const double idk = 1.0 / DK;
            const double px = ex * idk, py = ey * idk, pz = ez * idk;
            const double ax = ny * pz - nz * py;
            const double ay = nz * px - nx * pz;
            const double az = nx * py - ny * px;
            const double GY = (mx - F.v[L][0]) * ax + (my - F.v[L][1]) * ay + (mz - F.v[L][2]) * az;
            const double sR = RR[Ln] + RR[L];
            const double ANT = 2 * GY * DK;
            const double DNT = sR * sR - DK * DK + 2 * aGZ * sR;
            const double ANL = sR + DK;
            const double DNL = sR - DK;
            const double ALDEN = log(ANL / DNL);
            const double AT = (aGZ >= gz_min) ? atan(ANT / DNT) : 0.0;
            const double lx = DX[Ln] + DX[L], ly = DY[Ln] + DY[L], lz = DZ[Ln] + DZ[L];
            const double f1 = 2 * (sR + aGZ), f2 = 2 * sGZ * sR;
            const double tx = f1 * lx + f2 * nx, ty = f1 * ly + f2 * ny, tz = f1 * lz + f2 * nz;
            if (fabs(GY) < (double)1e-5f) G = G - 2 * AT * aGZ;
            else G = G + GY * ALDEN - 2 * AT * aGZ;
            const double c2 = 2 * sGZ * AT;
            const double c3 = GY * (DNL - ANL) / (ANL * DNL);
            const double c4 = 2 * aGZ / (ANT * ANT + DNT * DNT);
            const double c5 = 2 * DK;
            gx = gx + ALDEN * ax - c2 * nx + c3 * lx - c4 * (c5 * ax * DNT - tx * ANT);
            gy = gy + ALDEN * ay - c2 * ny + c3 * ly - c4 * (c5 * ay * DNT - ty * ANT);
            gz = gz + ALDEN * az - c2 * nz + c3 * lz - c4 * (c5 * az * DNT - tz * ANT);
        }
    }
    return RankineInt{G, gx, gy, gz};
}

__device__ __forceinline__ void exact_rankine(double mx, double my, double mz, const Face* __restrict__ Fp, double& G,
                                              double& gx, double& gy, double& gz)
{
    const RankineInt r = exact_rankine(mx, my, mz, Fp);
    G = r.G;
    gx = r.gx;
    gy = r.gy;
    gz = r.gz;
}

// one_point_integral_of_Rankine, Green_Rankine.f90:161-198 (first-variable gradient).
__device__ __forceinline__ void one_point_rankine(double mx, double my, double mz, double cx, double cy, double cz,
                                                  double area, double& G, double& gx, double& gy, double& gz)
{
    const double vx = cx - mx, vy = cy - my, vz = cz - mz;
    const double ro = norm3(vx, vy, vz);
    if (ro > (double)1e-7f) {
        G = area / ro;
        const double s = G / (ro * ro);
        gx = vx * s;
        gy = vy * s;
        gz = vz * s;
    } else {
        G = 0.0;
        gx = gy = gz = 0.0;
    }
}

// integral_of_Rankine, Green_Rankine.f90:28-64.  Returns the gradient w.r.t.
// the first variable when deriv_first, else w.r.t. the second.
// (cx, cy, cz, area, radius) are the panel's centre, area and radius, read
// by the caller; the vertices are only read on the exact (near-field) path.
__device__ __forceinline__ void rankine(double mx, double my, double mz, const Face* __restrict__ Fp, double cx,
                                        double cy, double cz, double area, double radius, bool deriv_first,
                                        double& G, double& gx, double& gy, double& gz)
{
    const double r0 = norm3(mx - cx, my - cy, mz - cz);
    if (far_direct(mx - cx, my - cy, mz - cz, r0, radius)) one_point_rankine(mx, my, mz, cx, cy, cz, area, G, gx, gy, gz);
    else exact_rankine(mx, my, mz, Fp, G, gx, gy, gz);
    if (!deriv_first) {
        gx = -gx;
        gy = -gy;
        gz = -gz;
    }
}

// integral_of_reflected_Rankine and its exact / one-point variants,
// Green_Rankine.f90:202-338.  The image point is (x, y, a z + b).
// If force_one_point, the one-point formula is used whatever the distance
// (matrices.f90:143-151 and :235-268 call it directly).
__device__ __forceinline__ void reflected_rankine(double mx, double my, double mz, const Face* __restrict__ Fp,
                                                  double cx, double cy, double cz, double area, double radius,
                                                  bool deriv_first, double a, double b, bool force_one_point,
                                                  double& G, double& gx, double& gy, double& gz)
{
    const double rz = a * mz + b;
    bool one_point = force_one_point;
    if (!one_point) {
        const double dx = mx - cx, dy = my - cy;
        const double r1 = sqrt(dx * dx + dy * dy + (rz - cz) * (rz - cz));
        one_point = far_image(dx, dy, rz - cz, r1, radius);
    }
    if (one_point) one_point_rankine(mx, my, rz, cx, cy, cz, area, G, gx, gy, gz);
    else exact_rankine(mx, my, rz, Fp, G, gx, gy, gz);
    gz = a * gz;
    if (!deriv_first) {
        gx = -gx;
        gy = -gy;
        if (a > 0) gz = -gz;
    }
}

// ---------------------------------------------------------------------------
// Wave part (Delhommeau_integrals.f90, Green_wave.f90)
// ---------------------------------------------------------------------------

__device__ __forceinline__ double pl2(double u1, double u2, double u3, double xu)
{
    return ((xu - u1) * (xu - u2)) / ((u3 - u1) * (u3 - u2));
}

// pick_in_default_tabulation + nearest_r_index / nearest_z_index +
// lagrange_polynomial_interpolation (Delhommeau_integrals.f90:274-402).
__device__ __forceinline__ void pick_in_table(const Table& T, double r, double z, double* out)
{
    int i, j;
    if (T.grid == LEGACY_GRID) {
        if (r < (double)1e-6f) i = 1;
        else if (r < 1.0) i = f_int(5 * (log10(r) + 6) + 1);
        else i = f_int(3 * r + 28);
        const double az = fabs(z);
        if (az > (double)1e-2f) j = f_int(8 * (log10(az) + 4.5));
        else j = f_int(5 * (log10(az) + 6));
    } else {
        if (r < (double)1e-10f) i = 1;
        else if (r < 1.0) i = f_nint((log10(r) / 10.0 + 1.0) * (T.index_of_1 - 1) + 1);
        else i = f_nint((r - 1) * (T.nr - T.index_of_1) / (T.rmax - 1) + T.index_of_1);
        j = f_nint((log10(fabs(z)) + 10) / T.dz);
    }
    i = max(2, min(T.nr - 1, i)) - 1;  // 0-based centre of the stencil
    j = max(2, min(T.nz - 1, j)) - 1;
    const double r0 = T.r_range[i - 1], r1 = T.r_range[i], r2 = T.r_range[i + 1];
    const double z0 = T.z_range[j - 1], z1 = T.z_range[j], z2 = T.z_range[j + 1];
    const double xl0 = pl2(r1, r2, r0, r), xl1 = pl2(r2, r0, r1, r), xl2 = pl2(r0, r1, r2, r);
    const double zl0 = pl2(z1, z2, z0, z), zl1 = pl2(z2, z0, z1, z), zl2 = pl2(z0, z1, z2, z);
    // out[c] = sum_b zl(b) sum_a xl(a) T(a, b, c): one stencil row (15
    // contiguous doubles) at a time keeps the live gathered values to 15.
    const double zl[3] = {zl0, zl1, zl2};
#pragma unroll
    for (int c = 0; c < 5; ++c) out[c] = 0.0;
#pragma unroll 1
    for (int b = 0; b < 3; ++b) {
        const double* t = T.tab + ((size_t)(j - 1 + b) * T.nr + (i - 1)) * 5;
        const double w = zl[b];
#pragma unroll
        for (int c = 0; c < 5; ++c) out[c] += w * (xl0 * t[c] + xl1 * t[5 + c] + xl2 * t[10 + c]);
    }
}

// asymptotic_approximations, Delhommeau_integrals.f90:157-186.
__device__ __forceinline__ void asymptotic(double r, double z, double* out)
{
    const double oor = 1.0 / fmax(r, (double)1e-10f);
    const double oor1 = 1.0 / sqrt(r * r + z * z);
    const double ez = exp(z) * sqrt(2 * PI * oor);
    double sk, ck;
    sincos(r - PI / 4, &sk, &ck);
    const double o3 = oor1 * oor1 * oor1;
    out[0] = 2 * (-ez * sk + z * o3 - oor1);
    out[1] = 2 * (-ez * sk + z * o3);
    out[2] = 2 * (ez * ck);
    out[3] = 2 * (-ez * (ck - sk * oor / 2) + r * o3);
    out[4] = 2 * (-ez * (sk + ck * oor / 2));
}

// ---------------------------------------------------------------------------
// Far field of the mixed-precision assembly: the wave term of a far panel
// pair read from an FP32 copy of the table.
//
// For a panel pair outside the reference's near-field radius (r0 > 7 R and
// r1 > 7 R, where Green_Rankine.f90:50 and :224 switch to the one-point
// formulas) the infinite-depth wave term (Green_wave.f90:162-271 with
// Delhommeau_integrals.f90:274-402) gathers its 3 x 3 x 5 stencil from the
// float table (half the bytes of the double one) and finds the stencil with
// FP32 logarithms.  The Lagrange weights and the interpolation are FP64:
// an error census of the 4352-panel barge (FP32 variants of each far-pair
// term against the FP64 pipeline, 4 frequencies) put the coefficient error
// of FP32 weights and sums at 1.1e-6 (the weights of a 0.2-wide r cell from
// an FP32 r near 13 are wrong in the 6th digit), and that of float table
// values with FP64 arithmetic at 1e-7.
//
// The stencil is the reference's: a nearest-index position whose FP32 value
// lies within 1e-3 of a rounding boundary is recomputed in FP64.

// Nearest index of Fortran NINT / INT of a position p computed in FP32,
// recomputed from the FP64 position when p is within 1e-3 of where the
// rounding changes.
__device__ __forceinline__ bool near_boundary(float p, bool nint)
{
    const float f = p - floorf(p);
    return nint ? fabsf(f - 0.5f) < 1e-3f : (f < 1e-3f || f > 1.0f - 1e-3f);
}

__device__ __forceinline__ void table_index_f(const Table& T, double r, double z, int& i, int& j)
{
    const float rf = (float)r, azf = fabsf((float)z);
    if (T.grid == LEGACY_GRID) {
        if (r < (double)1e-6f) {
            i = 1;
        } else if (r < 1.0) {
            const float p = 5 * (log10f(rf) + 6) + 1;
            i = near_boundary(p, false) ? f_int(5 * (log10(r) + 6) + 1) : (int)p;
        } else {
            i = f_int(3 * r + 28);
        }
        const double az = fabs(z);
        if (az > (double)1e-2f) {
            const float p = 8 * (log10f(azf) + 4.5f);
            j = near_boundary(p, false) ? f_int(8 * (log10(az) + 4.5)) : (int)p;
        } else {
            const float p = 5 * (log10f(azf) + 6);
            j = near_boundary(p, false) ? f_int(5 * (log10(az) + 6)) : (int)p;
        }
    } else {
        if (r < (double)1e-10f) {
            i = 1;
        } else if (r < 1.0) {
            const float p = (log10f(rf) / 10.0f + 1.0f) * (float)(T.index_of_1 - 1) + 1.0f;
            i = near_boundary(p, true) ? f_nint((log10(r) / 10.0 + 1.0) * (T.index_of_1 - 1) + 1)
                                       : (int)roundf(p);
        } else {
            i = f_nint((r - 1) * (T.nr - T.index_of_1) / (T.rmax - 1) + T.index_of_1);
        }
        const float p = (log10f(azf) + 10.0f) / (float)T.dz;
        j = near_boundary(p, true) ? f_nint((log10(fabs(z)) + 10) / T.dz) : (int)roundf(p);
    }
    i = max(2, min(T.nr - 1, i)) - 1;
    j = max(2, min(T.nz - 1, j)) - 1;
}

// pick_in_table with the stencil's values read from the float table.
__device__ __forceinline__ void pick_in_float_table(const Table& T, double r, double z, double* out)
{
    int i, j;
    table_index_f(T, r, z, i, j);
    const double r0 = T.r_range[i - 1], r1 = T.r_range[i], r2 = T.r_range[i + 1];
    const double z0 = T.z_range[j - 1], z1 = T.z_range[j], z2 = T.z_range[j + 1];
    const double xl0 = pl2(r1, r2, r0, r), xl1 = pl2(r2, r0, r1, r), xl2 = pl2(r0, r1, r2, r);
    const double zl[3] = {pl2(z1, z2, z0, z), pl2(z2, z0, z1, z), pl2(z0, z1, z2, z)};
#pragma unroll
    for (int c = 0; c < 5; ++c) out[c] = 0.0;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        const float* t = T.tab_f + ((size_t)(j - 1 + b) * T.nr + (i - 1)) * 5;
        const double w = zl[b];
#pragma unroll
        for (int c = 0; c < 5; ++c)
            out[c] += w * (xl0 * (double)t[c] + xl1 * (double)t[5 + c] + xl2 * (double)t[10 + c]);
    }
}

// WAVE_PART_INFINITE_DEPTH, Green_wave.f90:162-271, for a point x and a
// source point xi (already differenced in the horizontal plane): dx, dy are
// x - xi, zsum is x3 + xi3.  Gradient with respect to x.
// With ftab, the table's values come from its float copy (far pairs of the
// mixed-precision assembly).  UNTAB: no table, the integrals are evaluated
// for each point (Green_wave.f90:226-228), in kernels of their own so that
// the inlined integration does not weigh on the tabulated kernels' registers.
template <bool UNTAB = false>
__device__ __forceinline__ void wave_point(const Table& T, double k, int gfs, double dx, double dy, double zsum,
                                           cplx& G, cplx* nG, bool ftab = false)
{
    const double r = k * sqrt(dx * dx + dy * dy);
    const double z = k * zsum;
    const double r1 = sqrt(r * r + z * z);
    double drdx1 = 0.0, drdx2 = 0.0;
    if (fabs(r) > 16 * DBL_EPS) {
        drdx1 = k * k * dx / r;
        drdx2 = k * k * dy / r;
    }
    double I[5];
    if (UNTAB) {
        numerical_integration_device(r, z, T.npts, I);
    } else if (fabs(z) < T.zmax_abs && r < T.rmax) {
        if (ftab) pick_in_float_table(T, r, z, I);
        else pick_in_table(T, r, z, I);
    } else {
        asymptotic(r, z, I);
    }
    double gre, gim = I[2];
    double n0r, n1r, n2r, n2i;
    if (gfs == HIGH_FREQ) {
        gre = I[1];
        const double r13 = r1 * r1 * r1;
        n0r = drdx1 * I[3] - 2 * drdx1 * r / r13;
        n1r = drdx2 * I[3] - 2 * drdx2 * r / r13;
        n2r = k * gre - 2 * k * z / r13;
        n2i = k * gim;
    } else {
        gre = I[0];
        n0r = drdx1 * I[3];
        n1r = drdx2 * I[3];
        n2r = (gfs == LOW_FREQ) ? k * (gre + 2 / r1) : k * gre;
        n2i = k * gim;
    }
    G.re = k * gre;
    G.im = k * gim;
    nG[0].re = k * n0r;
    nG[0].im = k * (drdx1 * I[4]);
    nG[1].re = k * n1r;
    nG[1].im = k * (drdx2 * I[4]);
    nG[2].re = k * n2r;
    nG[2].im = k * n2i;
}

// integral_of_singularity_on_free_surface, Green_wave.f90:128-158.
__device__ __forceinline__ void free_surface_self(double area, double k, cplx& G, cplx* nG)
{
    const double re = (1 - log(k * k * area / PI)) + 2 * (EULER_GAMMA - LOG_2);
    G.re = k * area * re;
    G.im = k * area * (-2 * PI);
    nG[0].re = nG[0].im = 0.0;
    nG[1].re = nG[1].im = 0.0;
    nG[2].re = k * (G.re + 4 * sqrt(PI * area));
    nG[2].im = k * G.im;
}

}  // namespace cpt
