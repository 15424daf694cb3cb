"""Bodies with a lid on the device (irregular-frequency removal): the lid's
faces join the hull's in S and K (free-surface panels: diagonal
coefficient 1, matrices.f90:348-352), the boundary condition is zero on
them and the forces integrate the hull only.

Parity: against the oracle pipeline on mesh_including_lid (oracle S, K +
scipy LU) at 1e-10; the hull block of S and K with a lid equals the
matrices without it; away from irregular frequencies the lid changes
forces and fields by less than 5 % (pytest/test_bem_irregular_
frequencies_removal.py:66-196); near the first irregular frequency the
coefficients without lid jump and those with lid stay smooth."""
import numpy as np
import pytest
from conftest import rel_err

from gpu_helpers import ref_table_gf

pytestmark = pytest.mark.gpu


def _bodies():
    import capytaine_amd as cpa
    hull = cpa.mesh_parallelepiped(size=(1.0, 1.0, 1.0), center=(0.0, 0.0, -0.5), resolution=(6, 6, 6),
                                   missing_sides={"top"})
    lid = cpa.mesh_rectangle(size=(1.0, 1.0), resolution=(6, 6), center=(0.0, 0.0, 0.0), normal=(0.0, 0.0, -1.0))
    with_lid = cpa.FloatingBody(hull, cpa.rigid_body_dofs(), name="box with lid", lid_mesh=lid)
    without = cpa.FloatingBody(hull, cpa.rigid_body_dofs(), name="box")
    return with_lid, without


@pytest.fixture(scope="module")
def solver(default_table):
    import capytaine_amd as cpa
    print("[gpu] lids: solver with the reference table", flush=True)
    gf = ref_table_gf(default_table)
    assert gf.gf_singularities == "low_freq"  # the only choice with free-surface panels
    return cpa.BEMSolver(engine=cpa.DefaultMatrixEngine(green_function=gf))


def test_regular_mesh_ties_vs_oracle(solver, orc, default_table):
    """On a regular mesh 7 R equals some centre-to-centre distances exactly;
    the branch (exact or one-point Rankine integral) is then the one the
    oracle's NORM2 picks, entry by entry."""
    from oracle import mesh_arrays
    body, _ = _bodies()
    mesh = body.mesh_including_lid
    k = 0.4
    S, K = solver.engine.green_function.evaluate(mesh, mesh, wavenumber=k)
    So, Ko = orc.build_matrices(**mesh_arrays(mesh), wavenumber=k, depth=np.inf, table=default_table)
    for a, b in ((S, So), (K, Ko)):
        assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()


def test_hull_block_unchanged_by_lid(solver):
    with_lid, without = _bodies()
    n = without.mesh.nb_faces
    gp = dict(free_surface=0.0, water_depth=np.inf, wavenumber=1.0)
    for adjoint in (True, False):
        Sw, Kw = (np.asarray(a) for a in solver.engine.build_matrices(
            with_lid.mesh_including_lid, with_lid.mesh_including_lid, adjoint_double_layer=adjoint, **gp))
        So, Ko = (np.asarray(a) for a in solver.engine.build_matrices(
            without.mesh, without.mesh, adjoint_double_layer=adjoint, **gp))
        assert rel_err(Sw[:n, :n], So) < 1e-14 and rel_err(Kw[:n, :n], Ko) < 1e-14


def test_lid_solve_vs_oracle(solver, orc, default_table):
    """Radiation and diffraction of the box with its lid against the oracle's
    S, K on the joined mesh, scipy LU, pressure integrated on the hull."""
    import capytaine_amd as cpa
    from oracle import mesh_arrays
    from scipy.linalg import lu_factor, lu_solve
    print("[gpu] lids: solve vs oracle", flush=True)
    body, _ = _bodies()
    mesh = body.mesh_including_lid
    pbs = [cpa.RadiationProblem(body=body, radiating_dof=d, omega=2.0) for d in ("Heave", "Pitch")]
    pbs.append(cpa.DiffractionProblem(body=body, omega=2.0, wave_direction=0.4))
    S, K = orc.build_matrices(**mesh_arrays(mesh), wavenumber=pbs[0].wavenumber, depth=np.inf,
                              table=default_table, gf_singularities=1)
    lu = lu_factor(K)
    for pb, res in zip(pbs, solver.solve_all(pbs, keep_details=True)):
        pressure = 1j * pb.omega * pb.rho * (S @ lu_solve(lu, pb.boundary_condition))
        ref = body.integrate_pressure(pressure[body.hull_mask])
        scale = max(abs(v) for v in ref.values())
        for dof in ref:
            assert abs(res.forces[dof] - ref[dof]) <= 1e-10 * scale, (pb, dof)
        single = solver.solve(pb)
        for dof in ref:
            assert abs(single.forces[dof] - ref[dof]) <= 1e-10 * scale, (pb, dof)


@pytest.mark.parametrize("water_depth", [np.inf, 10.0])
def test_effect_of_lid_at_a_regular_frequency(solver, water_depth):
    import capytaine_amd as cpa
    with_lid, without = _bodies()
    k = 2 * np.pi / 5.0  # wavelength 5 m: omega about half the box's first irregular frequency
    out = {}
    for name, body in (("with", with_lid), ("without", without)):
        pb = cpa.DiffractionProblem(body=body, wavenumber=k, water_depth=water_depth)
        res = solver.solve(pb, keep_details=True)
        out[name] = (res.forces["Heave"], solver.compute_free_surface_elevation(np.array([[4.0, 4.0]]), res),
                     solver.compute_velocity(np.array([[4.0, 4.0, -2.0]]), res))
    for a, b in zip(out["with"], out["without"]):
        assert np.all(np.abs(np.asarray(a) - np.asarray(b)) <= 5e-2 * np.abs(np.asarray(b)))


def test_irregular_frequency_removed(solver):
    """Heave added mass over a band around the box's first irregular
    frequency: the curve without lid is rougher (largest second difference
    7x that of the curve with lid on MI355X), the lid removes the jump."""
    import capytaine_amd as cpa
    print("[gpu] lids: irregular-frequency sweep", flush=True)
    with_lid, without = _bodies()
    w_irr = without.first_irregular_frequency_estimate()
    omegas = np.linspace(0.8 * w_irr, 1.2 * w_irr, 41)
    curves = {}
    for name, body in (("with", with_lid), ("without", without)):
        pbs = [cpa.RadiationProblem(body=body, radiating_dof="Heave", omega=w) for w in omegas]
        curves[name] = np.array([r.added_mass["Heave"] for r in solver.solve_all(pbs)])
    rough = {k: float(np.abs(np.diff(c, 2)).max() / np.abs(c).max()) for k, c in curves.items()}
    print(f"[gpu] lids: w_irr estimate {w_irr:.3f}; max |second difference| / max |A33|: {rough}", flush=True)
    assert rough["without"] > 4 * rough["with"]
