"""Forward speed on the host (problems_and_results.py:96-105, :451-464,
dofs.py:60-96): encounter frequency and direction, the m-terms of the
radiation conditions, and the errors for what is not implemented.  Mirrors
pytest/test_forward_speed.py:38-95.  No GPU."""
import numpy as np
import pytest
from pytest import approx

import capytaine_amd as cpa


@pytest.fixture(scope="module")
def body():
    mesh = cpa.mesh_vertical_cylinder(radius=1.0, length=1.0, center=(0.0, 0.0, -0.5), resolution=(2, 20, 10))
    return cpa.FloatingBody(mesh, cpa.rigid_body_dofs(only=["Surge", "Heave", "Pitch"], rotation_center=(0, 0, -0.5)),
                            name="body")


def test_encounter_frequency_along_waves(body):
    pb = cpa.DiffractionProblem(body=body, omega=2.0, forward_speed=1.0, wave_direction=0.0)
    assert pb.forward_speed == 1.0 and pb.encounter_omega < pb.omega
    assert pb.encounter_wave_direction == pb.wave_direction
    # omega_e = omega - k U cos(beta), k_e from the dispersion relation at omega_e
    assert pb.encounter_omega == approx(2.0 - 4.0 / 9.81 * 1.0)
    assert pb.encounter_wavenumber == approx(pb.encounter_omega ** 2 / 9.81)


def test_encounter_frequency_against_waves(body):
    pb = cpa.DiffractionProblem(body=body, omega=2.0, forward_speed=-1.0, wave_direction=0.0)
    assert pb.encounter_omega > pb.omega and pb.encounter_wave_direction == pb.wave_direction


def test_encounter_frequency_faster_than_waves(body):
    pb = cpa.DiffractionProblem(body=body, omega=2.0, forward_speed=10.0, wave_direction=0.0)
    assert pb.encounter_omega >= 0.0
    assert pb.encounter_wave_direction == approx(pb.wave_direction + np.pi)


def test_encounter_frequency_orthogonal_waves(body):
    pb = cpa.DiffractionProblem(body=body, omega=2.0, forward_speed=1.0, wave_direction=np.pi / 2)
    assert pb.encounter_omega == approx(pb.omega)


def test_encounter_frequency_finite_depth(body):
    pb = cpa.RadiationProblem(body=body, omega=2.0, forward_speed=1.0, water_depth=1.5, radiating_dof="Surge")
    k_e = pb.encounter_wavenumber
    assert pb.encounter_omega < pb.omega
    assert k_e * np.tanh(k_e * 1.5) == approx(pb.encounter_omega ** 2 / 9.81)


def test_m_terms():
    """pytest/test_forward_speed.py:66-82."""
    mesh = cpa.mesh_sphere().immersed_part()
    body = cpa.FloatingBody(mesh, cpa.rigid_body_dofs(rotation_center=(0, 0, 0)))
    x, y, z = body.mesh.faces_centers.T
    nx, ny, nz = body.mesh.faces_normals.T
    pb = cpa.RadiationProblem(body=body, omega=1.0, forward_speed=1.0, wave_direction=0.0, radiating_dof="Surge")
    assert np.allclose(pb.boundary_condition.real, 0.0)
    assert np.allclose(pb.boundary_condition.imag, -pb.encounter_omega * nx)
    pb = cpa.RadiationProblem(body=body, omega=1.0, forward_speed=1.0, wave_direction=0.0, radiating_dof="Roll")
    assert np.allclose(pb.boundary_condition.real, 0.0)
    assert np.allclose(pb.boundary_condition.imag, -pb.encounter_omega * (-z * ny + y * nz))
    pb = cpa.RadiationProblem(body=body, omega=1.0, forward_speed=1.0, wave_direction=0.0, radiating_dof="Pitch")
    assert np.allclose(pb.boundary_condition.real, pb.forward_speed * nz)
    assert np.allclose(pb.boundary_condition.imag, -pb.encounter_omega * (z * nx - x * nz))
    pb = cpa.RadiationProblem(body=body, omega=1.0, forward_speed=1.0, wave_direction=0.0, radiating_dof="Yaw")
    assert np.allclose(pb.boundary_condition.real, -pb.forward_speed * ny)
    assert np.allclose(pb.boundary_condition.imag, -pb.encounter_omega * (-y * nx + x * ny))


def test_non_rigid_body(body):
    b = cpa.FloatingBody(body.mesh, dict(body.dofs), name="body")
    b.dofs["Shear"] = np.array([[z, 0, 0] for (x, y, z) in body.mesh.faces_centers])
    with pytest.raises(NotImplementedError):
        cpa.RadiationProblem(body=b, omega=2.0, forward_speed=1.0, radiating_dof="Shear")


def test_zero_frequency_with_forward_speed(body):
    with pytest.raises(NotImplementedError):
        cpa.RadiationProblem(body=body, omega=0.0, forward_speed=1.0, radiating_dof="Surge")


def test_direct_method_with_forward_speed(body):
    class _Engine:
        exportable_settings = {}
    solver = cpa.BEMSolver(engine=_Engine())
    pb = cpa.DiffractionProblem(body=body, omega=2.0, forward_speed=1.0)
    with pytest.raises(NotImplementedError, match="forward speed"):
        solver._solve(pb, method="direct")
