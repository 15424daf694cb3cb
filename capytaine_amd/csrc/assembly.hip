// Panel-pair assembly of the single-layer matrix S and the (adjoint)
// double-layer matrix K for one or several wavenumbers at once.
//
// Restates matrices.f90:28-325 (build_matrices) and :329-364
// (add_diagonal_term, fused here) on gfx950:
//   * one thread per collocation point I (64 consecutive I per wavefront),
//     so the column-major stores S(I,J), K(I,J) are 1 KiB coalesced runs;
//   * the panel index J is uniform across the work-group: the 192-byte panel
//     record and its quadrature points are read with scalar loads and held
//     in SGPRs;
//   * the frequency-independent Rankine and reflected-Rankine integrals are
//     computed once per pair and reused for every wavenumber of the batch,
//     with the same order of additions as the reference for each entry.
#include "green.hpp"
#include "cpt_internal.hpp"

namespace cpt {

struct WaveOut {
    cplx G;
    cplx nG[3];
};

// integral_of_wave_part_infinite_depth, Green_wave.f90:50-124.
// (x, xi) = collocation point and panel; qp points to nq records (x,y,z,w);
// zsign/zshift map the panel's z coordinates for the sea-bottom image
// (z -> zsign*z + zshift) used by the finite-depth images.
template <bool UNTAB = false>
__device__ __forceinline__ void wave_panel(const Table& T, double k, int gfs, bool deriv_first, double x0, double x1,
                                           double x2, double cx, double cy, double cz, double area, const double* qp,
                                           int nq, double zsign, double zshift, WaveOut& o, bool ftab = false)
{
    const double czi = zsign * cz + zshift;
    const double dxc = x0 - cx, dyc = x1 - cy;
    const double r = k * sqrt(dxc * dxc + dyc * dyc);
    const double z = k * (x2 + czi);
    if (fabs(r) < (double)1e-10f && fabs(z) < (double)1e-10f) {
        free_surface_self(area, k, o.G, o.nG);
    } else {
        o.G.re = o.G.im = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) o.nG[c].re = o.nG[c].im = 0.0;
        for (int q = 0; q < nq; ++q) {
            const double qx = qp[4 * q], qy = qp[4 * q + 1], qz = zsign * qp[4 * q + 2] + zshift, w = qp[4 * q + 3];
            cplx g, ng[3];
            wave_point<UNTAB>(T, k, gfs, x0 - qx, x1 - qy, x2 + qz, g, ng, ftab);
            o.G.re += g.re * w;
            o.G.im += g.im * w;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                o.nG[c].re += ng[c].re * w;
                o.nG[c].im += ng[c].im * w;
            }
        }
    }
    if (!deriv_first) {
        o.nG[0].re = -o.nG[0].re;
        o.nG[0].im = -o.nG[0].im;
        o.nG[1].re = -o.nG[1].re;
        o.nG[1].im = -o.nG[1].im;
    }
}

__device__ __forceinline__ void acc(WaveOut& a, const WaveOut& b, bool flip_z)
{
    a.G.re += b.G.re;
    a.G.im += b.G.im;
    a.nG[0].re += b.nG[0].re;
    a.nG[0].im += b.nG[0].im;
    a.nG[1].re += b.nG[1].re;
    a.nG[1].im += b.nG[1].im;
    if (flip_z) {
        a.nG[2].re += -b.nG[2].re;
        a.nG[2].im += -b.nG[2].im;
    } else {
        a.nG[2].re += b.nG[2].re;
        a.nG[2].im += b.nG[2].im;
    }
}

// integral_of_wave_parts_finite_depth, Green_wave.f90:311-436.
template <bool UNTAB = false>
__device__ __forceinline__ void wave_panel_finite_depth(const Table& T, double k, double depth, int gfs,
                                                        bool deriv_first, double x0, double x1, double x2, double cx,
                                                        double cy, double cz, double area, const double* qp, int nq,
                                                        WaveOut& o)
{
    WaveOut t;
    o.G.re = o.G.im = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) o.nG[c].re = o.nG[c].im = 0.0;
    const double xs2 = -x2 - 2 * depth;
    wave_panel<UNTAB>(T, k, gfs, deriv_first, x0, x1, x2, cx, cy, cz, area, qp, nq, 1.0, 0.0, t);
    acc(o, t, false);
    wave_panel<UNTAB>(T, k, gfs, deriv_first, x0, x1, xs2, cx, cy, cz, area, qp, nq, 1.0, 0.0, t);
    acc(o, t, deriv_first);
    wave_panel<UNTAB>(T, k, gfs, deriv_first, x0, x1, x2, cx, cy, cz, area, qp, nq, -1.0, -2 * depth, t);
    acc(o, t, !deriv_first);
    wave_panel<UNTAB>(T, k, gfs, deriv_first, x0, x1, xs2, cx, cy, cz, area, qp, nq, -1.0, -2 * depth, t);
    acc(o, t, true);
    const double amh = k * depth;
    const double akh = amh * tanh(amh);
    const double a = (amh + akh) * (amh + akh) / (4 * amh * (amh * amh - akh * akh + akh));
    o.G.re *= a;
    o.G.im *= a;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        o.nG[c].re *= a;
        o.nG[c].im *= a;
    }
}

// One tile: collocation points iblk * 256 + threadIdx.x against panels
// jc * j_chunk .. + j_chunk, for every wavenumber of A.
template <bool FINITE_DEPTH, bool FAR32 = false, bool UNTAB = false>
__device__ __forceinline__ void assemble_tile(const AssemblyArgs& A, int iblk, int jc);

// One entry of S or K: complex128, or complex64 for the single-precision
// Green function (the FP64 value rounded once, at the store).
__device__ __forceinline__ void store_entry(cplx* base, size_t idx, double re, double im, bool c64)
{
    if (c64) {
        reinterpret_cast<float2*>(base)[idx] = make_float2((float)re, (float)im);
    } else {
        base[idx].re = re;
        base[idx].im = im;
    }
}

// Infinite depth: two wavefronts per SIMD (about 220 VGPRs, no spills of the
// main path); finite depth carries more live state and runs at one.
// UNTAB: the wave term integrated for each point instead of read from a table.
template <bool UNTAB>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
assemble_infinite_depth_kernel(AssemblyArgs A)
{
    assemble_tile<false, false, UNTAB>(A, blockIdx.x, blockIdx.y);
}

template <bool UNTAB>
__global__ void __launch_bounds__(256) assemble_finite_depth_kernel(AssemblyArgs A)
{
    assemble_tile<true, false, UNTAB>(A, blockIdx.x, blockIdx.y);
}

// Mixed-precision form: far pairs read the float table.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
assemble_infinite_depth_mixed_kernel(AssemblyArgs A)
{
    assemble_tile<false, true>(A, blockIdx.x, blockIdx.y);
}

// Persistent form over many bodies: each work-group walks the work-list of
// (body, row block, panel chunk) tiles with a grid stride.  The tiles are
// ordered by decreasing cost on the host so the stride balances the load.
template <bool FAR32, bool UNTAB>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
assemble_worklist_kernel(const AssemblyArgs* __restrict__ bodies, const AsmTile* __restrict__ tiles, int ntiles)
{
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const AsmTile tl = tiles[t];
        assemble_tile<false, FAR32, UNTAB>(bodies[tl.body], tl.iblk, tl.jc);
    }
}

template <bool FINITE_DEPTH, bool FAR32, bool UNTAB>
__device__ __forceinline__ void assemble_tile(const AssemblyArgs& A, int iblk, int jc)
{
    const int I = iblk * blockDim.x + threadIdx.x;
    const bool active = I < A.n;
    const int Ic = active ? I : 0;
    const double x0 = A.colloc[3 * (size_t)Ic], x1 = A.colloc[3 * (size_t)Ic + 1], x2 = A.colloc[3 * (size_t)Ic + 2];
    double nx = 0.0, ny = 0.0, nz = 0.0;
    if (A.adjoint) {
        nx = A.dpn[3 * (size_t)Ic];
        ny = A.dpn[3 * (size_t)Ic + 1];
        nz = A.dpn[3 * (size_t)Ic + 2];
    }
    const bool deriv_first = A.adjoint != 0;
    const int j_begin = jc * A.j_chunk;
    const int j_end = min(A.m, j_begin + A.j_chunk);

    for (int J = j_begin; J < j_end; ++J) {
        const Face* Fp = A.faces + J;
        const double cx = Fp->c[0], cy = Fp->c[1], cz = Fp->c[2], area = Fp->area, radius = Fp->radius;
        const double* qp = A.quad + (size_t)4 * A.nq * J;
        if (!A.adjoint) {
            nx = A.dpn[3 * (size_t)J];
            ny = A.dpn[3 * (size_t)J + 1];
            nz = A.dpn[3 * (size_t)J + 2];
        }
        // mixed form: a far pair (outside the reference's near-field radius
        // for both the direct and the reflected Rankine terms,
        // Green_Rankine.f90:50 and :224) reads the float table
        bool ftab = false;
        if (FAR32) {
            const double r0 = norm3(x0 - cx, x1 - cy, x2 - cz);
            const double r1 = sqrt((x0 - cx) * (x0 - cx) + (x1 - cy) * (x1 - cy) + (-x2 - cz) * (-x2 - cz));
            ftab = far_direct(x0 - cx, x1 - cy, x2 - cz, r0, radius) &&
                   far_image(x0 - cx, x1 - cy, -x2 - cz, r1, radius);
        }
        // Rankine term (matrices.f90:122-134)
        double gd, gdx, gdy, gdz;
        rankine(x0, x1, x2, Fp, cx, cy, cz, area, radius, deriv_first, gd, gdx, gdy, gdz);
        // Reflected Rankine term through the free surface (matrices.f90:141-165)
        double gr = 0.0, grx = 0.0, gry = 0.0, grz = 0.0;
        if (!A.rankine_only)
            reflected_rankine(x0, x1, x2, Fp, cx, cy, cz, area, radius, deriv_first, -1.0, 0.0,
                              FINITE_DEPTH && A.fd_method == LEGACY_FINITE_DEPTH, gr, grx, gry, grz);
        // Sea-bottom images (matrices.f90:218-268), frequency independent
        double gb = 0.0, gbx = 0.0, gby = 0.0, gbz = 0.0;
        double gs[3] = {0.0, 0.0, 0.0}, gsx[3] = {0.0, 0.0, 0.0}, gsy[3] = {0.0, 0.0, 0.0}, gsz[3] = {0.0, 0.0, 0.0};
        if (FINITE_DEPTH) {
            const double h = A.depth;
            reflected_rankine(x0, x1, x2, Fp, cx, cy, cz, area, radius, deriv_first, -1.0, -2 * h, false, gb, gbx, gby, gbz);
            reflected_rankine(x0, x1, x2, Fp, cx, cy, cz, area, radius, deriv_first, 1.0, -2 * h, true, gs[0], gsx[0], gsy[0], gsz[0]);
            reflected_rankine(x0, x1, x2, Fp, cx, cy, cz, area, radius, deriv_first, 1.0, 2 * h, true, gs[1], gsx[1], gsy[1], gsz[1]);
            reflected_rankine(x0, x1, x2, Fp, cx, cy, cz, area, radius, deriv_first, -1.0, -4 * h, true, gs[2], gsx[2], gsy[2], gsz[2]);
        }
        const bool on_diag = A.diag && I == J && I < A.n_diag;
        double diag_coef = 0.0;
        if (on_diag) diag_coef = (fabs(cz - A.free_surface) < (double)1e-8f) ? 1.0 : 0.5;

        for (int b = 0; b < A.nb; ++b) {
            const FreqArgs& P = A.freq[b];
            const double sgn = (P.gfs == HIGH_FREQ) ? -1.0 : 1.0;
            WaveOut o;
            o.G.re = 0.0 + gd;
            o.G.im = 0.0;
            o.nG[0].re = 0.0 + gdx;
            o.nG[1].re = 0.0 + gdy;
            o.nG[2].re = 0.0 + gdz;
            o.nG[0].im = o.nG[1].im = o.nG[2].im = 0.0;
            if (!A.rankine_only) {
                o.G.re += sgn * gr;
                o.nG[0].re += sgn * grx;
                o.nG[1].re += sgn * gry;
                o.nG[2].re += sgn * grz;
                if (P.gfs == LOW_FREQ_WITH_RANKINE_PART) o.nG[2].re += 2 * P.k * gr;
            }
            if (!FINITE_DEPTH) {
                if (P.wave) {
                    WaveOut w;
                    wave_panel<UNTAB>(A.T, P.k, P.gfs, deriv_first, x0, x1, x2, cx, cy, cz, area, qp, A.nq,
                                      1.0, 0.0, w, ftab);
                    acc(o, w, false);
                }
            } else {
                o.G.re += gb;
                o.nG[0].re += gbx;
                o.nG[1].re += gby;
                o.nG[2].re += gbz;
#pragma unroll
                for (int s = 0; s < 3; ++s) {
                    o.G.re += sgn * gs[s];
                    o.nG[0].re += sgn * gsx[s];
                    o.nG[1].re += sgn * gsy[s];
                    o.nG[2].re += sgn * gsz[s];
                }
                if (P.wave) {
                    WaveOut w;
                    wave_panel_finite_depth<UNTAB>(A.T, P.k, A.depth, P.gfs, deriv_first, x0, x1, x2, cx,
                                                   cy, cz, area, qp, A.nq, w);
                    acc(o, w, false);
                }
                // integral_of_prony_decomp_finite_depth, Green_wave.f90:438-519
                double pg = 0.0, px = 0.0, py = 0.0, pz = 0.0;
                const double h = A.depth;
                for (int e = 0; e < P.nexp; ++e) {
                    const double lk = P.prony_lambda[e], ak2 = P.prony_a[e] / 2;
                    const double cb[4] = {h * lk - 2 * h, -h * lk, h * lk - 4 * h, -h * lk + 2 * h};
                    const double ca[4] = {1.0, -1.0, -1.0, 1.0};
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        double g, gx, gy, gz;
                        reflected_rankine(x0, x1, x2, Fp, cx, cy, cz, area, radius, deriv_first, ca[t], cb[t], false, g, gx, gy, gz);
                        pg += ak2 * g;
                        px += ak2 * gx;
                        py += ak2 * gy;
                        pz += ak2 * gz;
                    }
                }
                o.G.re += pg;
                o.nG[0].re += px;
                o.nG[1].re += py;
                o.nG[2].re += pz;
            }
            if (active) {
                const size_t ij = (size_t)I + (size_t)A.ldS * J;
                const bool c64 = A.c64;
                store_entry(A.S, (size_t)b * A.strideS + ij, M1_4PI * o.G.re, M1_4PI * o.G.im, c64);
                const size_t k0 = (size_t)b * A.strideK;
                if (A.nk == 1) {
                    const size_t kij = (size_t)I + (size_t)A.ldK * J;
                    double kre = M1_4PI * (nx * o.nG[0].re + ny * o.nG[1].re + nz * o.nG[2].re);
                    const double kim = M1_4PI * (nx * o.nG[0].im + ny * o.nG[1].im + nz * o.nG[2].im);
                    if (on_diag) kre += diag_coef;
                    store_entry(A.K, k0 + kij, kre, kim, c64);
                } else {
                    // add_diagonal_term with early_dot_product=False adds the
                    // coefficient times the dot-product normal of row I (== J).
                    const size_t kij = 3 * ((size_t)I + (size_t)A.ldK * J);
                    const double dn[3] = {nx, ny, nz};
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        double kre = M1_4PI * o.nG[c].re;
                        if (on_diag) kre += diag_coef * dn[c];
                        store_entry(A.K, k0 + kij + c, kre, M1_4PI * o.nG[c].im, c64);
                    }
                }
            }
        }
    }
}

int launch_assembly(const AssemblyArgs& A, hipStream_t stream)
{
    if (A.n <= 0 || A.m <= 0 || A.nb <= 0) return 0;
    dim3 block(256);
    const int gx = (A.n + 255) / 256;
    const int gy = (A.m + A.j_chunk - 1) / A.j_chunk;
    dim3 grid(gx, gy);
    // stores of S and K: 32 bytes per pair and wavenumber (16 in complex64)
    TimedLaunch timed(stream, TK_ASSEMBLY, (double)A.n * A.m * A.nb, (A.c64 ? 16.0 : 32.0) * A.n * A.m * A.nb);
    // without a table there is no float copy either: the FP64 integration
    // serves both forms
    const bool untab = A.T.nr <= 1 || A.T.nz <= 1;
    if (A.finite_depth) {
        if (untab) hipLaunchKernelGGL(assemble_finite_depth_kernel<true>, grid, block, 0, stream, A);
        else hipLaunchKernelGGL(assemble_finite_depth_kernel<false>, grid, block, 0, stream, A);
    } else if (untab) {
        hipLaunchKernelGGL(assemble_infinite_depth_kernel<true>, grid, block, 0, stream, A);
    } else if (A.far_fp32) {
        hipLaunchKernelGGL(assemble_infinite_depth_mixed_kernel, grid, block, 0, stream, A);
    } else {
        hipLaunchKernelGGL(assemble_infinite_depth_kernel<false>, grid, block, 0, stream, A);
    }
    return (int)hipGetLastError();
}

int launch_assembly_worklist(const AssemblyArgs* bodies, const AsmTile* tiles, int ntiles, int grid, double pairs,
                             int far_fp32, bool untabulated, hipStream_t stream)
{
    if (ntiles <= 0) return 0;
    TimedLaunch timed(stream, TK_ASSEMBLY, pairs, 32.0 * pairs);
    if (untabulated)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(assemble_worklist_kernel<false, true>), dim3(grid), dim3(256), 0, stream,
                           bodies, tiles, ntiles);
    else if (far_fp32)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(assemble_worklist_kernel<true, false>), dim3(grid), dim3(256), 0, stream,
                           bodies, tiles, ntiles);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(assemble_worklist_kernel<false, false>), dim3(grid), dim3(256), 0, stream,
                           bodies, tiles, ntiles);
    return (int)hipGetLastError();
}

}  // namespace cpt
