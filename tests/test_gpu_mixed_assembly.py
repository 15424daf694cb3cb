"""Far field of the mixed-precision assembly (far pairs read the float copy
of the table, FP64 arithmetic) against the FP64 assembly: near pairs
(r0 <= 7 R) bit-equal, far pairs within the float table's rounding,
coefficients within the 1e-6 relative target.

Tolerances (written here): max |dS| / max |S| and max |dK| / max |K| below
1e-7, normwise below 1e-7 (a stencil of float values carries 6e-8 relative
rounding); forces of every problem within 1e-6 of its largest FP64 force
(the mixed solve refines to a 1e-12 residual against the mixed-assembled
matrix), on a 640-panel barge and on the 4352-panel benchmark barge at
three frequencies (0.2, 1.4 and 2.5 rad/s, the last the hardest)."""
import numpy as np
import pytest

from gpu_helpers import device, ref_table_gf

pytestmark = pytest.mark.gpu


def _assemble(gf, mesh, ks, fp32):
    import torch
    from capytaine_amd import _lib
    ctx = gf._context()
    _lib.check(_lib.load().cpt_set_assembly_precision(ctx.handle, int(fp32)), "precision")
    try:
        S, K = gf.evaluate_on_device(mesh, mesh, wavenumber=ks)
        torch.cuda.synchronize()
    finally:
        _lib.check(_lib.load().cpt_set_assembly_precision(ctx.handle, 0), "precision")
    return S.cpu().numpy(), K.cpu().numpy()


@pytest.mark.parametrize("case", ["barge", "sphere", "sphere_gl2"])
def test_far_field_fp32_matrices(default_table, case):
    import capytaine_amd as cpa
    print(f"[gpu] FP32 far field, {case}", flush=True)
    gf = ref_table_gf(default_table)
    if case == "barge":
        mesh = cpa.mesh_parallelepiped(size=(10, 4, 2), center=(0, 0, -1), resolution=(20, 8, 4),
                                       missing_sides={"top"})
    else:
        mesh = cpa.mesh_sphere(radius=1.0, resolution=(20, 20)).immersed_part()
        if case == "sphere_gl2":
            mesh = cpa.Mesh(mesh.vertices, mesh.faces, quadrature_method="Gauss-Legendre 2")
    ks = np.array([0.004, 0.3, 2.0, 10.0])
    S64, K64 = _assemble(gf, mesh, ks, False)
    S32, K32 = _assemble(gf, mesh, ks, True)
    # stored as (nb, m, n): [b, J, I]
    c = mesh.faces_centers
    R = mesh.faces_radiuses
    d0 = np.linalg.norm(c[None, :, :] - c[:, None, :], axis=2)  # [J, I] = |x_I - c_J|
    near = d0 <= 7 * R[:, None]
    for b in range(len(ks)):
        for name, A64, A32 in (("S", S64[b], S32[b]), ("K", K64[b], K32[b])):
            assert np.array_equal(A64[near], A32[near]), (name, b)
            scale = np.abs(A64).max()
            err = np.abs(A32 - A64).max() / scale
            nerr = np.linalg.norm(A32 - A64) / np.linalg.norm(A64)
            print(f"[gpu]   k={ks[b]} {name}: max {err:.2e} normwise {nerr:.2e}", flush=True)
            assert err < 1e-7 and nerr < 1e-7, (name, b, err, nerr)


@pytest.mark.parametrize("case", ["barge640", "barge4352"])
def test_mixed_sweep_coefficients(default_table, case):
    import capytaine_amd as cpa
    from capytaine_amd import workloads
    from capytaine_amd.sweep import FrequencySweep, boundary_conditions
    print(f"[gpu] mixed sweep (float-table far field) coefficients vs FP64, {case}", flush=True)
    gf = ref_table_gf(default_table)
    if case == "barge640":
        mesh = cpa.mesh_parallelepiped(size=(10, 4, 2), center=(0, 0, -1), resolution=(20, 8, 4),
                                       missing_sides={"top"})
        body = cpa.FloatingBody(mesh, name="b").add_all_rigid_body_dofs(rotation_center=(0, 0, 0))
        om = np.array([0.3, 0.8, 1.5, 2.4])
        headings = np.arange(4) * np.pi / 2
    else:
        body, _, headings = workloads.barge()
        mesh = body.mesh
        om = np.array([0.2, 1.4, 2.5])
    pbs = workloads.problems_by_frequency(body, om, headings, rho=1025.0)
    r = len(pbs[0])
    bc = boundary_conditions(pbs, r, mesh.nb_faces, device())
    out = {}
    for prec in ("fp64", "mixed"):
        sw = FrequencySweep(gf, body, om ** 2 / 9.81, r, precision=prec, far_field_fp32=True)
        sw.run(bc)
        out[prec] = sw.forces(om, 1025.0)
        if prec == "mixed":
            assert sw.far_field_fp32 and sw.residual <= 1e-12
    scale = np.abs(out["fp64"]).max(axis=2, keepdims=True)
    err = np.max(np.abs(out["mixed"] - out["fp64"]) / scale)
    print(f"[gpu]   max coefficient error {err:.2e}", flush=True)
    assert err < 1e-6
