// FP32 far field of the mixed-precision assembly.
//
// For a panel pair outside the reference's near-field radius (r0 > 7 R,
// where Green_Rankine.f90:50 switches to the one-point formula) the Rankine
// and reflected-Rankine one-point terms (Green_Rankine.f90:161-198) and the
// infinite-depth wave term (Green_wave.f90:162-271 with
// Delhommeau_integrals.f90:274-402) are evaluated in FP32 from a float copy
// of the table.  The discrete decisions stay those of the FP64 path: the
// table / asymptotic branch is taken on FP64 r and z, and a stencil index
// whose FP32 position lies within 1e-3 of a rounding boundary is recomputed
// in FP64, so the FP32 result interpolates in the same cell as the
// reference and differs from it by FP32 rounding only (~1e-7 relative).
#pragma once
#include "green.hpp"

namespace cpt {

struct cplxf {
    float re, im;
};

__device__ __forceinline__ void one_point_rankine_f(float mx, float my, float mz, float cx, float cy, float cz,
                                                    float area, float& G, float& gx, float& gy, float& gz)
{
    const float vx = cx - mx, vy = cy - my, vz = cz - mz;
    const float ro = sqrtf(vx * vx + vy * vy + vz * vz);
    G = area / ro;
    const float s = G / (ro * ro);
    gx = vx * s;
    gy = vy * s;
    gz = vz * s;
}

__device__ __forceinline__ float pl2f(float u1, float u2, float u3, float xu)
{
    return ((xu - u1) * (xu - u2)) / ((u3 - u1) * (u3 - u2));
}

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

__device__ __forceinline__ void pick_in_table_f(const Table& T, double r, double z, float* out)
{
    int i, j;
    table_index_f(T, r, z, i, j);
    const float rf = (float)r, zf = (float)z;
    const float r0 = T.r_range_f[i - 1], r1 = T.r_range_f[i], r2 = T.r_range_f[i + 1];
    const float z0 = T.z_range_f[j - 1], z1 = T.z_range_f[j], z2 = T.z_range_f[j + 1];
    const float xl0 = pl2f(r1, r2, r0, rf), xl1 = pl2f(r2, r0, r1, rf), xl2 = pl2f(r0, r1, r2, rf);
    const float zl[3] = {pl2f(z1, z2, z0, zf), pl2f(z2, z0, z1, zf), pl2f(z0, z1, z2, zf)};
#pragma unroll
    for (int c = 0; c < 5; ++c) out[c] = 0.0f;
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        const float* t = T.tab_f + ((size_t)(j - 1 + b) * T.nr + (i - 1)) * 5;
        const float w = zl[b];
#pragma unroll
        for (int c = 0; c < 5; ++c) out[c] += w * (xl0 * t[c] + xl1 * t[5 + c] + xl2 * t[10 + c]);
    }
}

__device__ __forceinline__ void asymptotic_f(float r, float z, float* out)
{
    const float oor = 1.0f / fmaxf(r, 1e-10f);
    const float oor1 = 1.0f / sqrtf(r * r + z * z);
    const float ez = expf(z) * sqrtf(2 * (float)PI * oor);
    float sk, ck;
    sincosf(r - (float)PI / 4, &sk, &ck);
    const float o3 = oor1 * oor1 * oor1;
    out[0] = 2 * (-ez * sk + z * o3 - oor1);
    out[1] = 2 * (-ez * sk + z * o3);
    out[2] = 2 * (ez * ck);
    out[3] = 2 * (-ez * (ck - sk * oor / 2) + r * o3);
    out[4] = 2 * (-ez * (sk + ck * oor / 2));
}

// wave_point in FP32 (decisions in FP64, see the file comment).
__device__ __forceinline__ void wave_point_f(const Table& T, double k, int gfs, double dx, double dy, double zsum,
                                             cplxf& G, cplxf* nG)
{
    const double r_d = k * sqrt(dx * dx + dy * dy);
    const double z_d = k * zsum;
    const float kf = (float)k, r = (float)r_d, z = (float)z_d;
    const float r1 = sqrtf(r * r + z * z);
    float drdx1 = 0.0f, drdx2 = 0.0f;
    if (fabs(r_d) > 16 * DBL_EPS) {
        const float s = kf * kf / r;
        drdx1 = s * (float)dx;
        drdx2 = s * (float)dy;
    }
    float I[5];
    if (fabs(z_d) < T.zmax_abs && r_d < T.rmax) pick_in_table_f(T, r_d, z_d, I);
    else asymptotic_f(r, z, I);
    float gre, gim = I[2];
    float n0r, n1r, n2r, n2i;
    if (gfs == HIGH_FREQ) {
        gre = I[1];
        const float r13 = r1 * r1 * r1;
        n0r = drdx1 * I[3] - 2 * drdx1 * r / r13;
        n1r = drdx2 * I[3] - 2 * drdx2 * r / r13;
        n2r = kf * gre - 2 * kf * z / r13;
        n2i = kf * gim;
    } else {
        gre = I[0];
        n0r = drdx1 * I[3];
        n1r = drdx2 * I[3];
        n2r = (gfs == LOW_FREQ) ? kf * (gre + 2 / r1) : kf * gre;
        n2i = kf * gim;
    }
    G.re = kf * gre;
    G.im = kf * gim;
    nG[0].re = kf * n0r;
    nG[0].im = kf * (drdx1 * I[4]);
    nG[1].re = kf * n1r;
    nG[1].im = kf * (drdx2 * I[4]);
    nG[2].re = kf * n2r;
    nG[2].im = kf * n2i;
}

}  // namespace cpt
