"""Symmetric meshes on the host (meshes/symmetric_meshes.py,
predefined/rectangles.py:157-175, spheres.py:26-66): the merged faces equal
those of the plain generators, mirrored and rotated copies keep outward
normals, the block order the device path assumes, the hull mask of a
symmetric lid join, and the group characters.  No GPU."""
import numpy as np
import pytest

import capytaine_amd as cpa
from capytaine_amd.meshes import join_with_mask
from capytaine_amd.symmetry import Symmetry, symmetry_of
from mesh_helpers import dihedral_sphere


def _by_center(mesh):
    return {tuple(np.round(c, 9)): (n, a) for c, n, a in zip(mesh.faces_centers, mesh.faces_normals,
                                                             mesh.faces_areas)}


@pytest.mark.parametrize("sym_mesh,plain", [
    (lambda: cpa.mesh_parallelepiped(size=(2, 1, 1), center=(0, 0, -0.5), resolution=(4, 2, 2),
                                     missing_sides={"top"}, reflection_symmetry=True),
     lambda: cpa.mesh_parallelepiped(size=(2, 1, 1), center=(0, 0, -0.5), resolution=(4, 2, 2),
                                     missing_sides={"top"})),
    (lambda: cpa.mesh_sphere(resolution=(6, 8), axial_symmetry=True), lambda: cpa.mesh_sphere(resolution=(6, 8))),
    (lambda: cpa.mesh_vertical_cylinder(resolution=(2, 12, 3), axial_symmetry=True),
     lambda: cpa.mesh_vertical_cylinder(resolution=(2, 12, 3))),
])
def test_symmetric_generators_equal_plain(sym_mesh, plain):
    a, b = _by_center(sym_mesh()), _by_center(plain())
    assert set(a) == set(b)
    for k in a:
        np.testing.assert_allclose(a[k][0], b[k][0], atol=1e-14)
        np.testing.assert_allclose(a[k][1], b[k][1], rtol=1e-13)


def test_block_order_and_symmetry_detection():
    q = cpa.mesh_rectangle(size=(1, 1), resolution=(2, 2), center=(0.5, 0.5, -1))
    two = cpa.ReflectionSymmetricMesh(cpa.ReflectionSymmetricMesh(q, "yOz"), "xOz")
    c = two.faces_centers.reshape(4, 4, 3)
    np.testing.assert_allclose(c[1], c[0] * (-1, 1, 1))   # bit 0: inner plane yOz
    np.testing.assert_allclose(c[2], c[0] * (1, -1, 1))   # bit 1: outer plane xOz
    np.testing.assert_allclose(c[3], c[0] * (-1, -1, 1))
    s = symmetry_of(two)
    assert (s.kind, s.n, s.w) == ("xor", 4, 4) and s.block_mesh is q
    assert symmetry_of(cpa.ReflectionSymmetricMesh(q, "xOz")).n == 2
    rot = cpa.RotationSymmetricMesh(q, 5)
    c = rot.faces_centers.reshape(5, 4, 3)
    ang = 2 * np.pi / 5
    R = np.array([[np.cos(ang), -np.sin(ang), 0], [np.sin(ang), np.cos(ang), 0], [0, 0, 1]])
    np.testing.assert_allclose(c[1], c[0] @ R.T, atol=1e-15)
    assert symmetry_of(rot).kind == "cyclic" and symmetry_of(rot).n == 5
    assert symmetry_of(q) is None
    # mirrored faces keep outward normals: reflected, not reversed
    np.testing.assert_allclose(two.faces_normals.reshape(4, 4, 3)[1], two.faces_normals[:4] * (-1, 1, 1), atol=1e-15)


@pytest.mark.parametrize("kind,n", [("xor", 2), ("xor", 4), ("cyclic", 5)])
def test_group_characters_diagonalise_the_block_matrix(kind, n):
    """The forward characters turn the matrix with blocks A_(index(i, j))
    into the block diagonal of the transformed blocks, and the inverse
    characters undo the forward ones."""
    rng = np.random.default_rng(n)
    w = 3

    class _Block:
        nb_faces = w
    sym = Symmetry(kind, n, _Block())
    A = rng.normal(size=(n, w, w)) + 1j * rng.normal(size=(n, w, w))
    M = np.block([[A[sym.block_index(i, j)] for j in range(n)] for i in range(n)])
    F = np.kron(sym.characters(), np.eye(w))
    Finv = np.kron(sym.characters(inverse=True), np.eye(w))
    np.testing.assert_allclose(F @ Finv, np.eye(n * w), atol=1e-14)
    D = F @ M @ Finv
    Ahat = np.einsum("kj,jab->kab", sym.characters(), A)
    for k in range(n):
        np.testing.assert_allclose(D[k * w:(k + 1) * w, k * w:(k + 1) * w], Ahat[k], atol=1e-12)
        off = np.delete(D[k * w:(k + 1) * w], np.s_[k * w:(k + 1) * w], axis=1)
        assert np.abs(off).max() < 1e-12


def test_symmetric_lid_join_mask():
    hull = cpa.mesh_parallelepiped(size=(2, 1, 1), center=(0, 0, -0.5), resolution=(4, 2, 2),
                                   missing_sides={"top"}, reflection_symmetry=True)
    lid_q = cpa.mesh_rectangle(size=(1.0, 0.5), resolution=(2, 1), center=(-0.5, -0.25, 0.0), normal=(0, 0, -1))
    lid = cpa.ReflectionSymmetricMesh(cpa.ReflectionSymmetricMesh(lid_q, "yOz"), "xOz")
    joined, mask = join_with_mask(hull, lid)
    assert symmetry_of(joined).n == 4 and mask.sum() == hull.nb_faces
    np.testing.assert_array_equal(joined.faces_centers[mask], hull.faces_centers)
    np.testing.assert_array_equal(joined.faces_centers[~mask], lid.faces_centers)
    body = cpa.FloatingBody(hull, cpa.rigid_body_dofs(), lid_mesh=lid)
    bc = cpa.RadiationProblem(body=body, radiating_dof="Heave", omega=1.0).boundary_condition
    plain = cpa.RadiationProblem(body=cpa.FloatingBody(hull, cpa.rigid_body_dofs()), radiating_dof="Heave",
                                 omega=1.0).boundary_condition
    assert np.all(bc[~body.hull_mask] == 0)
    np.testing.assert_array_equal(bc[body.hull_mask], plain)


def test_dihedral_mesh_and_detection():
    mesh = dihedral_sphere(6, 8)
    s = symmetry_of(mesh)
    assert s.kind == "cyclic" and s.n == 6 and s.half is mesh.wedge.half and s.w == 2 * mesh.wedge.half.nb_faces
    # the merged faces are those of a 12-meridian sphere, normals outward
    c, nrm = mesh.faces_centers, mesh.faces_normals
    assert mesh.nb_faces == 12 * 8
    assert np.all(np.einsum("ij,ij->i", c, nrm) > 0)
    ang = np.sort(np.round(np.arctan2(c[:, 1], c[:, 0]) % (2 * np.pi), 9))
    np.testing.assert_allclose(np.unique(ang), (np.arange(12) + 0.5) * np.pi / 6, atol=1e-9)


def test_dihedral_macro_column_is_the_group_rule():
    """The macro column equals the first two block columns of the matrix
    whose block (p, q) is the evaluated block of T_q^-1 T_p (p = 2 i + s:
    R^i M^s), built here directly from that rule."""
    import torch
    from capytaine_amd.symmetry import Symmetry, macro_column
    n, h, nb = 5, 3, 2
    N = 2 * n * h
    rng = np.random.default_rng(1)
    cols = rng.normal(size=(nb, h, N)) + 1j * rng.normal(size=(nb, h, N))

    class _Half:
        nb_faces = h
    sym = Symmetry("cyclic", n, None, half=_Half())
    mac = macro_column(sym, torch.as_tensor(cols)).numpy()   # (nb, 2h, N)

    def rel(p, q):  # index of T_q^-1 T_p in the first block column
        (i, s), (j, t) = divmod(p, 2), divmod(q, 2)
        return 2 * ((i - j) % n) + s if t == 0 else 2 * ((j - i) % n) + 1 - s
    blocks = cols.transpose(0, 2, 1).reshape(nb, 2 * n, h, h)  # [b, p, row, col]
    for q in (0, 1):
        for p in range(2 * n):
            want = blocks[:, rel(p, q)]
            got = mac[:, q * h:(q + 1) * h, p * h:(p + 1) * h].transpose(0, 2, 1)
            np.testing.assert_array_equal(got, want)
