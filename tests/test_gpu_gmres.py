"""linear_solver="gmres" on the device (cpt_zgmres / cpt_cgmres), the
reference's solve_gmres (engines.py:74-89, 340-341): scipy's
sparse.linalg.gmres(A, b, atol=1e-6) with rtol 1e-5, restart 20.

Checks: the stopping rule itself (||b - A x|| <= max(1e-6, 1e-5 ||b||),
recomputed in numpy from the returned x), agreement with the LU solution to
what that rule allows (relative 1e-4 on these well-conditioned BEM systems),
scipy's own gmres on the same matrix as a second opinion (same rule, Arnoldi
steps within a few), the reference's test_all_linear_solvers_work
(test_bem_engines_basic_matrix_engine.py:97-125), complex64 matrices,
symmetric meshes (block by block in the transformed space) and the
non-convergence error."""
import numpy as np
import pytest

from gpu_helpers import device

pytestmark = pytest.mark.gpu


def _bem_K(gf, mesh, k):
    S, K = gf.evaluate(mesh, mesh, wavenumber=k)
    return S, K


def _gmres(K, b, **kw):
    import torch
    from capytaine_amd.engines import gmres_device
    dt = torch.complex64 if K.dtype == np.complex64 else torch.complex128
    A = torch.as_tensor(np.ascontiguousarray(K.T), dtype=dt).to(device())[None]
    bd = torch.as_tensor(np.ascontiguousarray(np.atleast_2d(b.T)), dtype=dt).to(device())[None]
    x, its, res = gmres_device(A, bd, **kw)
    return x[0].cpu().numpy().T, its[0], res[0]


def test_gmres_on_bem_matrix_vs_lu_and_scipy():
    import scipy.sparse.linalg as ssl
    from scipy.linalg import lu_factor, lu_solve
    import capytaine_amd as cpa
    print("[gpu] gmres on the sphere's K", flush=True)
    gf = cpa.Delhommeau(tabulation_cache_dir=None, device=device())
    mesh = cpa.mesh_sphere(radius=1.0, resolution=(30, 30)).immersed_part()
    rng = np.random.default_rng(3)
    for k in (0.3, 1.0, 3.0):
        _, K = _bem_K(gf, mesh, k)
        n = K.shape[0]
        B = rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3))
        B[:, 2] *= 1e-7  # below atol / rtol ||b|| crossover: the absolute tolerance rules
        X, its, res = _gmres(K, B)
        Xlu = lu_solve(lu_factor(K), B)
        for j in range(3):
            b, x = B[:, j], X[:, j]
            tol = max(1e-6, 1e-5 * np.linalg.norm(b))
            r = np.linalg.norm(b - K @ x)
            xs, info = ssl.gmres(K, b, atol=1e-6)
            d_lu = np.linalg.norm(x - Xlu[:, j]) / np.linalg.norm(Xlu[:, j])
            print(f"[gpu]   k {k} rhs {j}: {its[j]} steps, ||r|| {r:.2e} (tol {tol:.1e}), vs LU {d_lu:.1e}; "
                  f"scipy info {info}, vs LU {np.linalg.norm(xs - Xlu[:, j]) / np.linalg.norm(Xlu[:, j]):.1e}",
                  flush=True)
            assert r <= tol * (1 + 1e-9)
            assert abs(res[j] - r / np.linalg.norm(b)) <= 1e-6 * res[j] + 1e-15
            assert d_lu <= 1e-4 or np.linalg.norm(x - Xlu[:, j]) <= 10 * tol
            assert info == 0


def test_all_linear_solvers_work():
    """test_bem_engines_basic_matrix_engine.py:97-125 (with "gmres" actually
    used for the third solver)."""
    import capytaine_amd as cpa
    print("[gpu] all linear solvers", flush=True)
    sphere = cpa.FloatingBody(mesh=cpa.mesh_sphere(radius=1.0, resolution=(4, 3)).immersed_part())
    sphere.add_translation_dof(direction=(1, 0, 0), name="Surge")
    problem = cpa.RadiationProblem(body=sphere, omega=1.0, water_depth=np.inf)
    gf = cpa.Delhommeau(tabulation_cache_dir=None, device=device())
    out = {}
    for ls in ("lu_decomposition", "lu_decomposition_with_overwrite", "gmres"):
        res = cpa.BEMSolver(engine=cpa.DefaultMatrixEngine(green_function=gf, linear_solver=ls)).solve(problem)
        out[ls] = res.added_masses["Surge"]
    print(f"[gpu]   {out}", flush=True)
    assert np.isclose(out["lu_decomposition"], out["lu_decomposition_with_overwrite"])
    assert np.isclose(out["lu_decomposition"], out["gmres"])


def test_gmres_solver_larger_body():
    """A 1280-panel barge: every rigid-body radiation problem and a
    diffraction problem by the gmres engine against the LU engine (forces
    within 1e-5 of the largest)."""
    import capytaine_amd as cpa
    print("[gpu] gmres engine, barge", flush=True)
    mesh = cpa.mesh_parallelepiped(size=(20.0, 8.0, 2.0), center=(0, 0, -1.0), resolution=(40, 16, 4),
                                   missing_sides={"top"})
    body = cpa.FloatingBody(mesh).add_all_rigid_body_dofs(rotation_center=(0, 0, 0))
    gf = cpa.Delhommeau(tabulation_cache_dir=None, device=device())
    pbs = [cpa.RadiationProblem(body=body, radiating_dof=d, omega=0.9) for d in body.dofs]
    pbs.append(cpa.DiffractionProblem(body=body, wave_direction=0.4, omega=0.9))
    lu = cpa.BEMSolver(engine=cpa.DefaultMatrixEngine(green_function=gf))
    gm = cpa.BEMSolver(engine=cpa.DefaultMatrixEngine(green_function=gf, linear_solver="gmres"))
    for pb in pbs:
        f_lu = np.array([lu.solve(pb).forces[d] for d in body.dofs])
        f_gm = np.array([gm.solve(pb).forces[d] for d in body.dofs])
        err = np.abs(f_gm - f_lu).max() / np.abs(f_lu).max()
        print(f"[gpu]   {pb}: {err:.1e}", flush=True)
        assert err <= 1e-5


def test_gmres_complex64():
    """cpt_cgmres on the float32 Green function's K, and the float32 engine:
    the stopping rule holds on the complex64 system."""
    import capytaine_amd as cpa
    print("[gpu] gmres complex64", flush=True)
    gf = cpa.Delhommeau(floating_point_precision="float32", tabulation_cache_dir=None, device=device())
    mesh = cpa.mesh_sphere(radius=1.0, resolution=(20, 20)).immersed_part()
    _, K = _bem_K(gf, mesh, 1.0)
    assert K.dtype == np.complex64
    b = np.ones(K.shape[0], dtype=np.complex64)
    x, its, res = _gmres(K, b)
    x = x[:, 0]
    assert x.dtype == np.complex64
    r = np.linalg.norm(b - K.astype(np.complex128) @ x.astype(np.complex128))
    print(f"[gpu]   {its[0]} steps, ||r|| {r:.2e}", flush=True)
    assert r <= max(1e-6, 1e-5 * np.linalg.norm(b)) * 1.01  # the FP64 residual of the FP32 iterate
    body = cpa.FloatingBody(mesh).add_translation_dof(name="Heave")
    pb = cpa.RadiationProblem(body=body, radiating_dof="Heave", omega=1.0)
    res32 = cpa.BEMSolver(engine=cpa.DefaultMatrixEngine(green_function=gf, linear_solver="gmres")).solve(pb)
    ref = cpa.BEMSolver(engine=cpa.DefaultMatrixEngine(green_function=gf)).solve(pb)
    assert res32.pressure.dtype == np.complex64
    assert abs(res32.forces["Heave"] - ref.forces["Heave"]) <= 1e-4 * abs(ref.forces["Heave"])


@pytest.mark.parametrize("kind", ["two_planes", "rotation"])
def test_gmres_symmetric_mesh(kind):
    """Meshes with two symmetry planes and with a rotation symmetry: gmres
    block by block in the transformed space against the LU of the same
    blocks, forces within 1e-5 of the largest."""
    import capytaine_amd as cpa
    print(f"[gpu] gmres on a symmetric mesh ({kind})", flush=True)
    if kind == "two_planes":
        mesh = cpa.mesh_parallelepiped(size=(10.0, 4.0, 2.0), center=(0, 0, -1.0), resolution=(20, 8, 4),
                                       missing_sides={"top"}, reflection_symmetry=True)
    else:
        mesh = cpa.mesh_sphere(radius=1.0, resolution=(12, 16), axial_symmetry=True).immersed_part()
    body = cpa.FloatingBody(mesh).add_all_rigid_body_dofs(rotation_center=(0, 0, 0))
    gf = cpa.Delhommeau(tabulation_cache_dir=None, device=device())
    lu = cpa.BEMSolver(engine=cpa.DefaultMatrixEngine(green_function=gf))
    gm = cpa.BEMSolver(engine=cpa.DefaultMatrixEngine(green_function=gf, linear_solver="gmres"))
    for pb in (cpa.RadiationProblem(body=body, radiating_dof="Pitch", omega=1.1),
               cpa.DiffractionProblem(body=body, wave_direction=0.7, omega=1.1)):
        f_lu = np.array([lu.solve(pb).forces[d] for d in body.dofs])
        f_gm = np.array([gm.solve(pb).forces[d] for d in body.dofs])
        err = np.abs(f_gm - f_lu).max() / np.abs(f_lu).max()
        print(f"[gpu]   {pb}: {err:.1e}", flush=True)
        assert err <= 1e-5


def test_gmres_zero_rhs_and_no_convergence():
    rng = np.random.default_rng(1)
    n = 200
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    x, its, res = _gmres(A, np.zeros(n, dtype=complex))
    assert np.all(x == 0) and its[0] == 0 and res[0] == 0
    with pytest.raises(RuntimeError, match="No convergence of the GMRES"):
        _gmres(A, np.ones(n, dtype=complex), restart=2, maxiter=1)
