"""Lids (bodies.py:85-111, problems_and_results.py:373-447): the joined
mesh, the lid turned over, boundary conditions zero on the lid, the
Froude-Krylov force unchanged, the irregular-frequency estimate and the
serving request's batched terms.  No GPU."""
import numpy as np
import pytest

from capytaine_amd import FloatingBody, RadiationProblem, DiffractionProblem, mesh_rectangle, rigid_body_dofs

def _box():
    from capytaine_amd import mesh_parallelepiped
    return mesh_parallelepiped(center=(0.0, 0.0, -0.5)).immersed_part()  # vertices on z = 0


def test_lid_with_upward_normals_is_turned_over():
    lid = mesh_rectangle(size=(1, 1), resolution=(4, 4), center=(0, 0, -0.1), normal=(0, 0, 1))
    body = FloatingBody(_box(), rigid_body_dofs(), lid_mesh=lid)
    np.testing.assert_allclose(body.lid_mesh.faces_normals[:, 2], -1.0)
    assert body.lid_mesh is not lid
    np.testing.assert_allclose(lid.faces_normals[:, 2], 1.0)  # the given lid is unchanged
    np.testing.assert_allclose(body.lid_mesh.faces_centers, lid.faces_centers, atol=1e-15)
    np.testing.assert_allclose(body.lid_mesh.faces_areas, lid.faces_areas)


def test_lid_faces_and_boundary_conditions():
    hull = _box()
    lid = mesh_rectangle(size=(1, 1), resolution=(4, 4), center=(0, 0, 0), normal=(0, 0, -1))
    body = FloatingBody(hull, rigid_body_dofs(), lid_mesh=lid)
    plain = FloatingBody(hull, rigid_body_dofs())
    n, nl = hull.nb_faces, lid.nb_faces
    assert body.mesh_including_lid.nb_faces == n + nl and body.hull_mask.sum() == n
    assert body.lid_mesh is lid  # already facing down
    np.testing.assert_array_equal(body.mesh_including_lid.faces_centers[n:], lid.faces_centers)
    for pb_cls, kw in ((RadiationProblem, dict(radiating_dof="Heave")), (DiffractionProblem, dict(wave_direction=0.3))):
        with_lid = pb_cls(body=body, omega=1.2, **kw).boundary_condition
        without = pb_cls(body=plain, omega=1.2, **kw).boundary_condition
        assert with_lid.shape == (n + nl,)
        np.testing.assert_array_equal(with_lid[:n], without)
        assert np.all(with_lid[n:] == 0)
    dn = body.dof_normal_matrix(including_lid=True)
    assert dn.shape == (6, n + nl) and np.all(dn[:, n:] == 0)
    np.testing.assert_array_equal(dn[:, :n], plain.dof_normal_matrix())


def test_effect_of_lid_on_froude_krylov_force():
    from capytaine_amd import froude_krylov_force
    lid = mesh_rectangle(size=(1, 1), resolution=(4, 4), center=(0, 0, 0), normal=(0, 0, -1))
    f_with = froude_krylov_force(DiffractionProblem(body=FloatingBody(_box(), rigid_body_dofs(), lid_mesh=lid)))
    f_without = froude_krylov_force(DiffractionProblem(body=FloatingBody(_box(), rigid_body_dofs())))
    assert f_with["Heave"] == pytest.approx(f_without["Heave"])


def test_first_irregular_frequency_estimate():
    """pytest/test_bem_irregular_frequencies_removal.py:9-27."""
    from capytaine_amd import mesh_parallelepiped
    mesh = mesh_parallelepiped(size=(1, 1, 1)).immersed_part()
    assert 6.0 < FloatingBody(mesh, rigid_body_dofs()).first_irregular_frequency_estimate() < 7.0
    sub = mesh_rectangle(size=(1, 1), center=(0, 0, -0.1))
    assert 10.0 < FloatingBody(mesh, rigid_body_dofs(), lid_mesh=sub).first_irregular_frequency_estimate() < 11.0
    top = mesh_rectangle(size=(1, 1), center=(0, 0, 0))
    assert FloatingBody(mesh, rigid_body_dofs(), lid_mesh=top).first_irregular_frequency_estimate() == np.inf


def test_serving_request_with_lid():
    """A request's batched boundary conditions carry the lid's zeros like the
    problem objects."""
    from capytaine_amd.serving import _Request
    lid = mesh_rectangle(size=(1, 1), resolution=(4, 4), center=(0, 0, 0), normal=(0, 0, -1))
    body = FloatingBody(_box(), rigid_body_dofs(), lid_mesh=lid)
    om, hd = np.array([0.8, 1.5]), np.array([0.0, 0.9])
    req = _Request(0, body, om, hd, 1025.0, 9.81, "mixed")
    bc = req.boundary_conditions(np.arange(2))
    assert bc.shape == (2, 6 + 2, body.mesh_including_lid.nb_faces)
    for i, w in enumerate(om):
        for r, d in enumerate(body.dofs):
            ref = RadiationProblem(body=body, radiating_dof=d, omega=w, rho=1025.0).boundary_condition
            np.testing.assert_allclose(bc[i, r], ref, rtol=0, atol=1e-15 * max(1.0, np.abs(ref).max()))
        for h, beta in enumerate(hd):
            ref = DiffractionProblem(body=body, omega=w, wave_direction=beta, rho=1025.0).boundary_condition
            assert np.abs(bc[i, 6 + h] - ref).max() <= 1e-14 * np.abs(ref).max()
