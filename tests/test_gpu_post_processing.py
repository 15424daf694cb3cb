"""Post-processing of a solved problem on the device: potential, velocity,
pressure and free-surface elevation at points (BEMSolver.compute_*),
against the oracle's S and full gradient at the same points times the same
sources, and against what the boundary-integral equation fixes (the normal
velocity on the hull is the boundary condition, the integrated pressure is
the force, the gradient is the derivative of the potential).

Tolerance: 1e-10 relative (FP64 path), normwise and per value."""
import numpy as np
import pytest
from conftest import elem_ok, rel_err

from gpu_helpers import ref_table_gf

pytestmark = pytest.mark.gpu

RTOL = 1e-10


@pytest.fixture(scope="module")
def solved(default_table):
    """Diffraction and heave radiation around a floating sphere (256 panels)."""
    import capytaine_amd as cpa
    print("[gpu] post-processing: solving the sphere problems", flush=True)
    mesh = cpa.mesh_sphere(radius=1.0, resolution=(16, 32)).immersed_part()
    body = cpa.FloatingBody(mesh, cpa.rigid_body_dofs(only=["Heave", "Surge"]))
    solver = cpa.BEMSolver(engine=cpa.DefaultMatrixEngine(green_function=ref_table_gf(default_table)))
    diff = solver.solve(cpa.DiffractionProblem(body=body, omega=1.0, wave_direction=np.pi / 4), keep_details=True)
    rad = solver.solve(cpa.RadiationProblem(body=body, radiating_dof="Heave", omega=1.0), keep_details=True)
    return solver, diff, rad


def _fluid_points(n, seed):
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n:
        p = rng.uniform((-3.0, -3.0, -3.0), (3.0, 3.0, -0.05))
        if np.linalg.norm(p) > 1.1:
            pts.append(p)
    return np.array(pts)


def _oracle(orc, table, points, dpn, mesh, k, *, full):
    from oracle import mesh_arrays
    arr = dict(mesh_arrays(mesh), colloc=points, dpn=dpn)
    return orc.build_matrices(**arr, wavenumber=k, depth=np.inf, table=table, early_dot_product=not full)


def test_potential_at_points_vs_oracle(solved, orc, default_table):
    solver, diff, _ = solved
    pts = _fluid_points(50, 1)
    phi = solver.compute_potential(pts, diff)
    S, _ = _oracle(orc, default_table, pts, np.zeros_like(pts), diff.body.mesh, diff.wavenumber, full=False)
    ref = S @ diff.sources
    assert phi.shape == (50,)
    assert rel_err(phi, ref) < RTOL and elem_ok(phi, ref, RTOL, 1e-6)


def test_velocity_at_points_vs_oracle(solved, orc, default_table):
    solver, _, rad = solved
    pts = _fluid_points(40, 2)
    u = solver.compute_velocity(pts, rad)
    _, K = _oracle(orc, default_table, pts, np.zeros_like(pts), rad.body.mesh, rad.wavenumber, full=True)
    ref = np.stack([K[c] @ rad.sources for c in range(3)], axis=-1)
    assert u.shape == (40, 3)
    assert rel_err(u, ref) < RTOL and elem_ok(u, ref, RTOL, 1e-6)


def test_velocity_on_another_mesh_vs_oracle(solved, orc, default_table):
    """A mesh as points keeps its normals: the diagonal term of
    build_fullK_matrix lands on its first min(n, m) faces."""
    import capytaine_amd as cpa
    solver, diff, _ = solved
    cyl = cpa.mesh_vertical_cylinder(radius=2.0, length=5.0, resolution=(2, 8, 10)).immersed_part()
    u = solver.compute_velocity(cyl, diff)
    _, K = _oracle(orc, default_table, cyl.faces_centers, cyl.faces_normals, diff.body.mesh, diff.wavenumber,
                   full=True)
    ref = np.stack([K[c] @ diff.sources for c in range(3)], axis=-1)
    assert u.shape == (cyl.nb_faces, 3)
    assert rel_err(u, ref) < RTOL and elem_ok(u, ref, RTOL, 1e-6)


@pytest.mark.parametrize("which", ["diffraction", "radiation"])
def test_reconstruction_of_given_boundary_condition(solved, which):
    """Velocity on the hull dotted with the normals is the Neumann condition
    the sources were solved for (pytest/test_bem_potential_velocity_and_free_
    surface_elevation.py:214-217)."""
    solver, diff, rad = solved
    res = diff if which == "diffraction" else rad
    u = solver.compute_velocity(res.body.mesh, res)
    un = np.einsum("...k,...k->...", u, res.body.mesh.faces_normals)
    bc = res.problem.boundary_condition
    assert np.abs(un - bc).max() < 1e-9 * np.abs(bc).max()


def test_pressure_integration_and_potential_on_hull(solved):
    """The pressure on the hull integrates to the solver's forces, and the
    potential on the hull is the one of the solve."""
    solver, diff, rad = solved
    for res in (diff, rad):
        pot = solver.compute_potential(res.body.mesh, res)
        assert rel_err(pot, res.potential) < 1e-12
        f = res.body.integrate_pressure(solver.compute_pressure(res.body.mesh, res))
        for dof in res.forces:
            assert abs(f[dof] - res.forces[dof]) <= 1e-10 * max(abs(v) for v in res.forces.values())


def test_free_surface_elevation(solved):
    """i omega / g times the potential at z = 0; x, y points equal x, y, 0."""
    solver, diff, _ = solved
    pts2 = np.array([[0.0, 3.0], [1.5, -2.0], [-2.2, -1.0]])
    pts3 = np.concatenate([pts2, np.zeros((3, 1))], axis=1)
    fse = solver.compute_free_surface_elevation(pts2, diff)
    assert fse.shape == (3,)
    np.testing.assert_array_equal(fse, solver.compute_free_surface_elevation(pts3, diff))
    phi = solver.compute_potential(pts3, diff)
    assert rel_err(fse, 1j * diff.omega / diff.g * phi) < 1e-15
    grid = np.meshgrid(np.linspace(2.0, 3.0, 2), np.linspace(3.0, 4.0, 3))
    assert solver.compute_free_surface_elevation(grid, diff).shape == grid[0].shape
    assert solver.compute_free_surface_elevation((0.0, 3.0), diff).shape == (1,)


def test_output_shapes(solved):
    """pytest/test_bem_potential_velocity_and_free_surface_elevation.py:45-204."""
    import capytaine_amd as cpa
    solver, diff, _ = solved
    assert solver.compute_potential((0.0, 0.0, -3.0), diff).shape == (1,)
    assert solver.compute_velocity((0.0, 0.0, -3.0), diff).shape == (1, 3)
    pts = [(0.0, 0.0, -3.0), (0.0, 1.0, -2.0), (1.0, 1.0, -1.5)]
    assert solver.compute_potential(pts, diff).shape == (3,)
    assert solver.compute_velocity(pts, diff).shape == (3, 3)
    grid = np.meshgrid(np.linspace(2.0, 3.0, 2), np.linspace(-1.0, 1.0, 3), np.linspace(-2.0, -1.0, 4))
    assert solver.compute_potential(grid, diff).shape == grid[0].shape
    assert solver.compute_velocity(grid, diff).shape == (*grid[0].shape, 3)
    fs = cpa.mesh_rectangle(center=(0, 0, 0), normal=(0, 0, -1), resolution=(3, 3), size=(2.0, 2.0))
    assert solver.compute_free_surface_elevation(fs, diff).shape == (fs.nb_faces,)


def test_chunks_and_several_sources(solved):
    """Chunks of points (crossing the diagonal term of a mesh) and several
    source vectors at once give what one chunk and one vector give."""
    solver, diff, rad = solved
    eng = solver.engine
    mesh = diff.body.mesh
    gp = dict(free_surface=0.0, water_depth=np.inf, wavenumber=diff.wavenumber)
    src = np.stack([diff.sources, rad.sources], axis=1)
    small = 16 * mesh.nb_faces * 4 * 7  # 7 points per chunk of the gradient
    for gradient in (False, True):
        one = eng.field_at_points(mesh, mesh, src, gradient=gradient, **gp)
        many = eng.field_at_points(mesh, mesh, src, gradient=gradient, max_chunk_bytes=small, **gp)
        assert rel_err(many, one) < 1e-14
        for q in range(2):
            single = eng.field_at_points(mesh, mesh, src[:, q], gradient=gradient, **gp)
            assert rel_err(one[:, q], single) < 1e-14


def test_velocity_is_gradient_of_potential(default_table):
    """Without free surface (Rankine part only, exact closed forms) the
    velocity is the central difference of the potential."""
    import capytaine_amd as cpa
    print("[gpu] post-processing: gradient vs finite differences", flush=True)
    mesh = cpa.mesh_sphere(radius=1.0, resolution=(10, 20), center=(0, 0, -3.0))
    body = cpa.FloatingBody(mesh, cpa.rigid_body_dofs(only=["Heave"]))
    solver = cpa.BEMSolver(engine=cpa.DefaultMatrixEngine(green_function=ref_table_gf(default_table)))
    res = solver.solve(cpa.RadiationProblem(body=body, omega=1.0, free_surface=np.inf), keep_details=True)
    rng = np.random.default_rng(4)
    pts = rng.normal(size=(20, 3))
    pts = pts / np.linalg.norm(pts, axis=1)[:, None] * rng.uniform(1.2, 2.5, (20, 1)) + (0, 0, -3.0)
    u = solver.compute_velocity(pts, res)
    h = 1e-4
    fd = np.empty_like(u)
    for c in range(3):
        e = np.zeros(3)
        e[c] = h
        fd[:, c] = (solver.compute_potential(pts + e, res) - solver.compute_potential(pts - e, res)) / (2 * h)
    assert rel_err(fd, u) < 1e-6
    # a gradient with a wrong sign or swapped components fails
    assert rel_err(fd, -u) > 1.0 and rel_err(fd, u[:, [1, 0, 2]]) > 1e-2


def test_direct_method_has_no_sources(solved):
    solver, diff, _ = solved
    res = solver.solve(diff.problem, keep_details=True, method="direct")
    with pytest.raises(Exception, match="direct method"):
        solver.compute_potential([(0.0, 0.0, -3.0)], res)
