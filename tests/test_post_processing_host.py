"""Host side of the post-processing: the point lists it accepts and their
output shapes (tools/lists_of_points.py), the undisturbed Airy wave at
points (bem/airy_waves.py), and the error raised for a result without
sources.  No GPU."""
import numpy as np
import pytest

from capytaine_amd import (FloatingBody, RadiationProblem, DiffractionProblem, mesh_rectangle, mesh_sphere,
                           rigid_body_dofs, airy_waves_potential, airy_waves_velocity,
                           airy_waves_free_surface_elevation, airy_waves_pressure)
from capytaine_amd.lists_of_points import _normalize_points, _normalize_free_surface_points


@pytest.fixture(scope="module")
def body():
    return FloatingBody(mesh_sphere(resolution=(4, 4)).immersed_part(), rigid_body_dofs(only=["Heave"]))


def test_normalize_points_shapes(body):
    p, s = _normalize_points((0.0, 0.0, -3.0))
    assert p.shape == (1, 3) and s == (1,)
    p, s = _normalize_points([(0.0, 0.0, -3.0), (0.0, 1.0, -2.0)])
    assert p.shape == (2, 3) and s == (2,)
    grid = np.meshgrid(np.linspace(0, 1, 2), np.linspace(-1, 1, 3), np.linspace(-2, 0, 4))
    p, s = _normalize_points(grid)
    assert p.shape == (24, 3) and s == grid[0].shape
    np.testing.assert_array_equal(p[5], [g.reshape(-1)[5] for g in grid])
    mesh = body.mesh
    p, s = _normalize_points(mesh)
    assert p is mesh.faces_centers and s == (mesh.nb_faces,)
    p, s = _normalize_points(body, keep_mesh=True)
    assert p is mesh and s == (mesh.nb_faces,)


def test_normalize_free_surface_points():
    p, s = _normalize_free_surface_points((0.0, 3.0))
    np.testing.assert_array_equal(p, [[0.0, 3.0, 0.0]])
    assert s == (1,)
    grid = np.meshgrid(np.linspace(2, 3, 2), np.linspace(3, 4, 3))
    p, s = _normalize_free_surface_points(grid)
    assert p.shape == (6, 3) and s == grid[0].shape and np.all(p[:, 2] == 0.0)
    fs = mesh_rectangle(center=(0, 0, 0), normal=(0, 0, -1), resolution=(3, 3), size=(2.0, 2.0))
    p, s = _normalize_free_surface_points(fs, keep_mesh=True)
    assert p is fs and s == (9,)


def test_airy_wave_free_surface_elevation_values(body):
    """One wavelength of 1 m in deep water: elevation 1, 0, -1 at x = 0,
    1/4, 1/2 (pytest/test_bem_potential_velocity_and_free_surface_elevation.
    py:219-225)."""
    pb = DiffractionProblem(body=body, wavenumber=2 * np.pi, wave_direction=0.0)
    assert np.isclose(np.real(airy_waves_free_surface_elevation([0, 0], pb)), 1.0)
    assert np.isclose(np.real(airy_waves_free_surface_elevation([0.25, 0], pb)), 0.0, atol=1e-5)
    assert np.isclose(np.real(airy_waves_free_surface_elevation([0.5, 0], pb)), -1.0)
    pts = [(0.0, 3.0), (1.0, -2.0), (1.0, -1.0)]
    np.testing.assert_array_equal(airy_waves_free_surface_elevation(pts, pb),
                                  airy_waves_free_surface_elevation([p + (0.0,) for p in pts], pb))


def test_airy_wave_velocity_is_gradient_of_potential(body):
    """Central differences of the potential, finite and deep water."""
    rng = np.random.default_rng(0)
    pts = rng.uniform((-3, -3, -2.5), (3, 3, -0.1), (10, 3))
    for depth in (np.inf, 3.0):
        pb = DiffractionProblem(body=body, omega=1.3, wave_direction=0.4, water_depth=depth)
        u = airy_waves_velocity(pts, pb)
        h = 1e-6
        fd = np.stack([(airy_waves_potential(pts + h * e, pb) - airy_waves_potential(pts - h * e, pb)) / (2 * h)
                       for e in np.eye(3)], axis=-1)
        assert np.abs(fd - u).max() < 1e-6 * np.abs(u).max()
        np.testing.assert_allclose(airy_waves_pressure(pts, pb), 1j * pb.omega * pb.rho * airy_waves_potential(pts, pb))


def test_airy_wave_shapes_and_limits(body):
    pb = DiffractionProblem(body=body, omega=1.0)
    assert airy_waves_potential((0.0, 0.0, -3.0), pb).shape == (1,)
    assert airy_waves_velocity([(0.0, 0.0, -3.0)] * 3, pb).shape == (3, 3)
    grid = np.meshgrid(np.linspace(0, 1, 2), np.linspace(-1, 1, 3), np.linspace(-2, 0, 4))
    assert airy_waves_velocity(grid, pb).shape == (*grid[0].shape, 3)
    pb0 = RadiationProblem(body=body, omega=0.0)
    assert np.all(np.isnan(airy_waves_potential([(0.0, 0.0, -1.0)], pb0)))


class _Engine:
    exportable_settings = {}


def test_result_without_sources_raises(body):
    from capytaine_amd import BEMSolver
    solver = BEMSolver(engine=_Engine())
    res = RadiationProblem(body=body, omega=1.0).make_results_container({"Heave": 0j})
    with pytest.raises(Exception, match="direct method"):
        solver.compute_potential([(0.0, 0.0, -3.0)], res)
    with pytest.raises(Exception, match="keep_details"):
        solver.compute_velocity([(0.0, 0.0, -3.0)], res)

