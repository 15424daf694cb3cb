"""CPU tests of the host side: the C-ABI library loads and exports what
include/capytaine_amd.h declares, the grid and geometry restatements, the
problem definitions, and the loud failure without a GPU."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "capytaine_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(cpt_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_header_symbols():
    from capytaine_amd import _lib
    lib = _lib.load()
    syms = header_symbols()
    assert len(syms) >= 11
    for s in syms:
        assert hasattr(lib, s), s
        assert s in _lib.SIGNATURES, s
    assert lib.cpt_version() == 1


def test_grid_spacing_bit_exact(core_golden):
    from capytaine_amd.green_functions import default_r_spacing, default_z_spacing
    for name, (nr, rmax, nz, zmin, method) in {"scaled": (676, 100.0, 372, -251.0, 1),
                                               "legacy": (328, 100.0, 46, -16.0, 0),
                                               "scaled_small": (100, 50.0, 100, -50.0, 1)}.items():
        assert np.array_equal(default_r_spacing(nr, rmax, method), core_golden[f"grid_{name}_r"]), name
        assert np.array_equal(default_z_spacing(nz, zmin, method), core_golden[f"grid_{name}_z"]), name


def test_geometry_of_simple_panels():
    from capytaine_amd.meshes import Mesh
    v = np.array([[0, 0, 0], [2, 0, 0], [2, 1, 0], [0, 1, 0], [0, 0, 1.0]])
    m = Mesh(v, [[0, 1, 2, 3], [0, 1, 4]])
    assert m.faces.tolist()[1] == [0, 1, 4, 4]
    np.testing.assert_allclose(m.faces_areas, [2.0, 1.0])
    np.testing.assert_allclose(m.faces_normals, [[0, 0, 1], [0, -1, 0]], atol=1e-15)
    np.testing.assert_allclose(m.faces_centers, [[1, 0.5, 0], [2 / 3, 0, 1 / 3]])
    np.testing.assert_allclose(m.faces_radiuses, [np.hypot(1, 0.5), np.linalg.norm([2 / 3, 0, 1 / 3])])
    qp, qw = m.quadrature_points
    assert qp.shape == (2, 1, 3) and np.allclose(qw[:, 0], m.faces_areas)
    gl = m.with_quadrature("Gauss-Legendre 2").quadrature_points
    np.testing.assert_allclose(gl[1][0].sum(), 2.0)


def test_config_meshes():
    import capytaine_amd as cpa
    s = cpa.mesh_sphere(radius=1.0, resolution=(30, 30)).immersed_part()
    assert s.nb_faces == 450
    assert np.all(np.sum(s.faces_normals * s.faces_centers, axis=1) > 0)
    b = cpa.mesh_parallelepiped(size=(20, 8, 2), center=(0, 0, -1), resolution=(80, 32, 8), missing_sides={"top"})
    assert b.nb_faces == 4352 and np.isclose(b.faces_areas.sum(), 272.0)
    c = cpa.mesh_vertical_cylinder(radius=2, length=4, center=(0, 0, -2), resolution=(8, 40, 12))
    assert c.nb_faces == 800


def test_radiation_and_diffraction_conditions():
    import capytaine_amd as cpa
    mesh = cpa.mesh_sphere(radius=1.0, resolution=(6, 12)).immersed_part()
    body = cpa.FloatingBody(mesh).add_all_rigid_body_dofs()
    pb = cpa.RadiationProblem(body=body, radiating_dof="Heave", omega=2.0)
    np.testing.assert_allclose(pb.boundary_condition, -2j * mesh.faces_normals[:, 2])
    assert np.isclose(pb.wavenumber, 4 / 9.81)
    d = cpa.DiffractionProblem(body=body, omega=1.0, wave_direction=0.3)
    k = 1 / 9.81
    x, y, z = mesh.faces_centers.T
    u = 9.81 * k / 1.0 * np.exp(1j * k * (x * np.cos(0.3) + y * np.sin(0.3))) * np.exp(k * z)
    vel = np.stack([np.cos(0.3) * u, np.sin(0.3) * u, -1j * u], axis=1)
    np.testing.assert_allclose(d.boundary_condition, -(vel * mesh.faces_normals).sum(axis=1))
    fin = cpa.RadiationProblem(body=body, omega=1.0, water_depth=10.0)
    assert np.isclose(fin.wavenumber * np.tanh(fin.wavenumber * 10.0), 1 / 9.81)


def test_integrate_pressure_closed_form():
    """Uniform pressure on a closed body gives no net force."""
    import capytaine_amd as cpa
    mesh = cpa.mesh_sphere(radius=1.0, resolution=(20, 20))
    body = cpa.FloatingBody(mesh).add_all_rigid_body_dofs()
    f = body.integrate_pressure(np.ones(mesh.nb_faces))
    assert all(abs(v) < 1e-12 for v in f.values())


def test_prony_decomposition_reproducible():
    from capytaine_amd.green_functions import find_best_exponential_decomposition
    f = lambda x: np.exp(-0.5 * x) + 0.3 * np.exp(-3 * x)
    a1, l1 = find_best_exponential_decomposition(f, -0.1, 20.0, range(4, 31, 2))
    a2, l2 = find_best_exponential_decomposition(f, -0.1, 20.0, range(4, 31, 2))
    assert np.array_equal(a1, a2) and np.array_equal(l1, l2)
    X = np.linspace(-0.1, 20, 300)
    approx = np.sum(np.asarray(a1)[:, None] * np.exp(np.asarray(l1)[:, None] * X), axis=0)
    assert np.mean((approx - f(X)) ** 2) < 1e-4


def test_no_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import capytaine_amd as cpa
    with pytest.raises(cpa.CptError):
        cpa.Delhommeau()


def test_context_creation_does_not_deadlock():
    """Context.get takes the module lock and Context.__init__ loads the
    library under it: the lock must be re-entrant.  Without a GPU the
    creation raises instead of blocking."""
    import threading
    from capytaine_amd import _lib
    outcome = []

    def target():
        try:
            _lib.Context.get("cuda:0")
            outcome.append("created")
        except Exception as e:
            outcome.append(type(e).__name__)
    t = threading.Thread(target=target, daemon=True)
    t.start()
    t.join(30)
    assert outcome, "Context.get blocked"


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors of cpt_body and cpt_lu_group have the C layout of
    include/capytaine_amd.h (offsets and sizes from gcc on the header)."""
    import ctypes
    import shutil
    import subprocess
    from capytaine_amd._lib import CptBody, CptLuGroup
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "capytaine_amd.h"', 'int main(void) {']
    for cname, cls in (("cpt_body", CptBody), ("cpt_lu_group", CptLuGroup)):
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(root, "include"), "-I", "/opt/rocm/include",
                    "-D__HIP_PLATFORM_AMD__", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            c, f, v = line.split()
            got[(c, f)] = int(v)
    for cname, cls in (("cpt_body", CptBody), ("cpt_lu_group", CptLuGroup)):
        assert got[(cname, "size")] == ctypes.sizeof(cls)
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)


def test_delhommeau_only_paths_refuse_other_green_functions():
    """The sweeps, the per-frequency graphs and the server assemble with the
    Delhommeau kernels; with FinGreen3D they say so before touching the
    device."""
    import pytest
    import capytaine_amd as cpa
    from capytaine_amd.sweep import FrequencySweep
    from capytaine_amd.batch import BatchedSweep
    from capytaine_amd.graphs import SweepGraph
    from capytaine_amd.serving import BEMServer
    gf = cpa.FinGreen3D(device="cpu")
    with pytest.raises(NotImplementedError, match="FrequencySweep runs the Delhommeau"):
        FrequencySweep(gf, None, [1.0], 1)
    with pytest.raises(NotImplementedError, match="BatchedSweep runs the Delhommeau"):
        BatchedSweep(gf, [], [], 1)
    with pytest.raises(NotImplementedError, match="SweepGraph runs the Delhommeau"):
        SweepGraph(gf, None, 1, 1)
    with pytest.raises(NotImplementedError, match="BEMServer runs the Delhommeau"):
        BEMServer(gf)
