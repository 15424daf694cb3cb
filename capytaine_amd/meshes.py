"""Panel meshes: the geometry the assembly kernel consumes.

Geometric definitions follow capytaine's ``meshes/geometry.py:35-118`` and
``meshes/meshes.py:409-488``:

* quadrilaterals are stored as 4 vertex indices, triangles repeat their last
  vertex (``faces`` property, meshes.py:453-469);
* normal of a quadrilateral: normalised cross product of its diagonals;
  of a triangle: of two edges (geometry.py:121-168);
* area and centre of a quadrilateral from its split into the triangles
  (a, b, c) and (a, c, d) (geometry.py:46-76);
* radius: distance from the first vertex to the centre (geometry.py:79-85);
* default quadrature: one point per face, the centre, with the area as
  weight (meshes.py:481-482); "Gauss-Legendre 2" as geometry.py:88-118.

The generators rebuild the reference's predefined meshes
(``meshes/predefined/spheres.py``, ``rectangles.py``, ``cylinders.py``) with
the same panel layout, plus an ellipsoid hull generator for batched and
serving workloads (the reference has no ellipsoid mesher).
"""
import threading

import logging

import numpy as np

__all__ = ["Mesh", "mesh_sphere", "mesh_parallelepiped", "mesh_rectangle", "mesh_ellipsoid",
           "mesh_vertical_cylinder", "mesh_horizontal_rectangle_lid"]


_device_cache_lock = threading.Lock()


LOG = logging.getLogger(__name__)


def _cross(a, b):
    return np.cross(a, b)


def _norm(a):
    return np.sqrt(np.sum(a * a, axis=-1))


class Mesh:
    """A mesh of planar quadrilateral and triangular panels.

    Parameters
    ----------
    vertices: (nv, 3) float array
    faces: (m, 3 or 4) int array, or list of 3- or 4-lists of vertex indices
    quadrature_method: None (one point per face) or "Gauss-Legendre 2"
    name: str, optional
    """

    _counter = 0

    def __init__(self, vertices, faces, *, quadrature_method=None, name=None):
        self.vertices = np.ascontiguousarray(np.asarray(vertices, dtype=np.float64).reshape(-1, 3))
        faces_list = [list(f) for f in faces] if not isinstance(faces, np.ndarray) else None
        if faces_list is not None:
            padded = [f if len(f) == 4 else list(f) + [f[-1]] for f in faces_list]
            faces = np.array(padded, dtype=np.int64).reshape(-1, 4)
        else:
            faces = np.asarray(faces, dtype=np.int64)
            if faces.ndim == 2 and faces.shape[1] == 3:
                faces = np.concatenate([faces, faces[:, 2:3]], axis=1)
        self.faces = np.ascontiguousarray(faces.reshape(-1, 4))
        if self.faces.size and (self.faces.min() < 0 or self.faces.max() >= len(self.vertices)):
            raise ValueError("Mesh faces refer to vertices that do not exist.")
        self.quadrature_method = quadrature_method
        if name is None:
            Mesh._counter += 1
            name = f"mesh_{Mesh._counter}"
        self.name = name
        self._cache = {}
        self._device_cache = {}

    # ------------------------------------------------------------------
    # sizes
    @property
    def nb_faces(self):
        return self.faces.shape[0]

    @property
    def nb_vertices(self):
        return self.vertices.shape[0]

    def __repr__(self):
        return f"Mesh(name={self.name!r}, nb_faces={self.nb_faces}, nb_vertices={self.nb_vertices})"

    # Meshes are compared by identity, like the reference's cache keys.
    def __eq__(self, other):
        return self is other

    def __hash__(self):
        return id(self)

    # ------------------------------------------------------------------
    # geometry (geometry.py:35-118)
    def _geometry(self):
        if "geom" in self._cache:
            return self._cache["geom"]
        v = self.vertices
        f = self.faces
        a, b, c, d = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]], v[f[:, 3]]
        is_quad = f[:, 2] != f[:, 3]
        # normals
        n_quad = _cross(c - a, d - b)
        n_tri = _cross(b - a, c - a)
        n = np.where(is_quad[:, None], n_quad, n_tri)
        normals = n / _norm(n)[:, None]
        # areas and centres
        area1 = 0.5 * _norm(_cross(b - a, c - a))
        area2 = 0.5 * _norm(_cross(c - a, d - a))
        c1 = (a + b + c) / 3
        c2 = (a + c + d) / 3
        with np.errstate(invalid="ignore", divide="ignore"):
            center_quad = (c1 * area1[:, None] + c2 * area2[:, None]) / (area1 + area2)[:, None]
        centers = np.where(is_quad[:, None], center_quad, c1)
        areas = np.where(is_quad, area1 + area2, area1)
        radii = _norm(centers - a)
        geom = dict(normals=np.ascontiguousarray(normals), centers=np.ascontiguousarray(centers),
                    areas=np.ascontiguousarray(areas), radii=np.ascontiguousarray(radii), is_quad=is_quad)
        self._cache["geom"] = geom
        return geom

    @property
    def faces_normals(self):
        return self._geometry()["normals"]

    @property
    def faces_centers(self):
        return self._geometry()["centers"]

    @property
    def faces_areas(self):
        return self._geometry()["areas"]

    @property
    def faces_radiuses(self):
        return self._geometry()["radii"]

    @property
    def quadrature_points(self):
        """(points (m, nq, 3), weights (m, nq)), meshes.py:471-486."""
        if self.quadrature_method is None:
            return (self.faces_centers.reshape((-1, 1, 3)), self.faces_areas.reshape(-1, 1))
        if self.quadrature_method == "Gauss-Legendre 2":
            if "gl2" not in self._cache:
                self._cache["gl2"] = _gauss_legendre_2(self.vertices, self.faces)
            return self._cache["gl2"]
        raise ValueError(f"Unknown quadrature_method: {self.quadrature_method}")

    def with_quadrature(self, quadrature_method):
        return Mesh(self.vertices, self.faces, quadrature_method=quadrature_method, name=self.name)

    # ------------------------------------------------------------------
    # transformations
    def translated(self, shift, name=None):
        return Mesh(self.vertices + np.asarray(shift, dtype=float), self.faces,
                    quadrature_method=self.quadrature_method, name=name or self.name)

    def merged(self):
        return self

    def with_normal_vector_going_down(self):
        """meshes.py:767-798: a horizontal mesh whose normals all point up is
        returned with its faces reversed (normals down); any other mesh is
        returned as is.  Triangles keep their repeated vertex last."""
        if not np.allclose(self.faces_normals[:, 2], np.ones(self.nb_faces)):
            return self
        LOG.warning("Inverting the direction of the normal vectors of %s to be downward.", self)
        quad = self.faces[:, 2] != self.faces[:, 3]
        faces = np.where(quad[:, None], self.faces[:, ::-1], self.faces[:, [2, 1, 0, 0]])
        return Mesh(self.vertices, faces, quadrature_method=self.quadrature_method, name=self.name)

    def join(self, other, name=None):
        """Union of two meshes (faces of self first)."""
        faces = np.concatenate([self.faces, other.faces + self.nb_vertices])
        same = self.quadrature_method == other.quadrature_method  # else dropped, as meshes.py:620-624
        return Mesh(np.concatenate([self.vertices, other.vertices]), faces,
                    quadrature_method=self.quadrature_method if same else None,
                    name=name or f"{self.name}+{other.name}")

    def __add__(self, other):
        return self.join(other)

    def immersed_part(self, free_surface=0.0, tol=1e-8):
        """Faces lying entirely below the free surface (within tol).

        The generators of this module put a row of vertices exactly on the
        free surface, so no face needs to be split; a face that crosses the
        free surface raises.
        """
        z = self.vertices[self.faces, 2]
        below = np.all(z <= free_surface + tol, axis=1)
        above = np.all(z >= free_surface - tol, axis=1)
        crossing = ~below & ~above
        if np.any(crossing):
            raise NotImplementedError("Faces crossing the free surface would need splitting.")
        return Mesh(self.vertices, self.faces[below], quadrature_method=self.quadrature_method,
                    name=self.name)

    # ------------------------------------------------------------------
    # device copies for the assembly kernel (cached per device)
    def device_arrays(self, device):
        import torch
        key = str(device)
        with _device_cache_lock:
            return self._device_arrays_locked(device, key)

    def _device_arrays_locked(self, device, key):
        import torch
        if key not in self._device_cache:
            qp, qw = self.quadrature_points
            as_t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt).to(device)
            self._device_cache[key] = dict(
                vertices=as_t(self.vertices),
                faces=as_t(self.faces.astype(np.int32), torch.int32),
                centers=as_t(self.faces_centers),
                normals=as_t(self.faces_normals),
                areas=as_t(self.faces_areas),
                radii=as_t(self.faces_radiuses),
                quad_points=as_t(qp),
                quad_weights=as_t(qw),
                nq=int(qw.shape[1]),
            )
        return self._device_cache[key]


def _gauss_legendre_2(vertices, faces):
    """geometry.py:88-118, vectorised."""
    s = 1 / np.sqrt(3)
    local = np.array([(+s, +s), (+s, -s), (-s, +s), (-s, -s)])
    fv = vertices[faces]  # (m, 4, 3)
    pts = np.empty((faces.shape[0], 4, 3))
    wts = np.empty((faces.shape[0], 4))
    for q, (xk, yk) in enumerate(local):
        pts[:, q, :] = ((1 + xk) * (1 + yk) * fv[:, 0] + (1 + xk) * (1 - yk) * fv[:, 1]
                        + (1 - xk) * (1 - yk) * fv[:, 2] + (1 - xk) * (1 + yk) * fv[:, 3]) / 4
        dxidx = ((1 + yk) * fv[:, 0] + (1 - yk) * fv[:, 1] - (1 - yk) * fv[:, 2] - (1 + yk) * fv[:, 3]) / 4
        dxidy = ((1 + xk) * fv[:, 0] - (1 + xk) * fv[:, 1] - (1 - xk) * fv[:, 2] + (1 - xk) * fv[:, 3]) / 4
        wts[:, q] = 0.25 * 4 * _norm(np.cross(dxidx, dxidy))
    return pts, wts


# ----------------------------------------------------------------------
# generators
# ----------------------------------------------------------------------
def _revolution(profile, n, name, scale=(1.0, 1.0, 1.0)):
    """Surface of revolution of `profile` (points (x, 0, z) sorted by
    increasing z) repeated n times around Oz, as
    RotationSymmetricMesh.from_profile_points(...).merged()
    (symmetric_meshes.py:386-423): panel (k, r) joins profile points k, k+1
    on the meridians r, r+1; faces ordered meridian by meridian; poles
    (points on the axis) are single vertices and their panels triangles."""
    profile = np.asarray(profile, dtype=float)
    npts = len(profile)
    on_axis = np.abs(profile[:, 0]) < 1e-12
    vid = -np.ones((npts, n), dtype=np.int64)
    verts = []
    for k in range(npts):
        if on_axis[k]:
            vid[k, :] = len(verts)
            verts.append([0.0, 0.0, profile[k, 2]])
        else:
            for r in range(n):
                ang = 2 * np.pi * r / n
                vid[k, r] = len(verts)
                verts.append([profile[k, 0] * np.cos(ang), profile[k, 0] * np.sin(ang), profile[k, 2]])
    verts = np.array(verts) * np.asarray(scale)[None, :]
    faces = []
    for r in range(n):
        r1 = (r + 1) % n
        for k in range(npts - 1):
            quad = [vid[k, r], vid[k, r1], vid[k + 1, r1], vid[k + 1, r]]
            dedup = []
            for v in quad:
                if not dedup or dedup[-1] != v:
                    dedup.append(v)
            if len(dedup) > 1 and dedup[-1] == dedup[0]:
                dedup.pop()
            if len(dedup) >= 3:
                faces.append(dedup)
    return Mesh(verts, faces, name=name)


def mesh_sphere(*, radius=1.0, center=(0.0, 0.0, 0.0), resolution=(10, 10), name=None):
    """Sphere, as predefined/spheres.py:26-66 (theta = linspace(0, pi, ntheta+1)
    on a meridian, nphi rotations)."""
    ntheta, nphi = resolution
    theta = np.linspace(0.0, np.pi, ntheta + 1)
    profile = radius * np.stack([np.sin(theta), np.zeros_like(theta), -np.cos(theta)], axis=1)
    profile[0, 0] = profile[-1, 0] = 0.0
    return _revolution(profile, nphi, name or "sphere").translated(center)


def mesh_ellipsoid(*, semi_axes=(2.0, 1.0, 1.0), center=(0.0, 0.0, 0.0), resolution=(10, 10), name=None):
    """Ellipsoid x^2/a^2 + y^2/b^2 + z^2/c^2 = 1: a sphere mesh of the given
    resolution scaled by the semi-axes (panels stay planar quadrilaterals)."""
    ntheta, nphi = resolution
    theta = np.linspace(0.0, np.pi, ntheta + 1)
    profile = np.stack([np.sin(theta), np.zeros_like(theta), -np.cos(theta)], axis=1)
    profile[0, 0] = profile[-1, 0] = 0.0
    return _revolution(profile, nphi, name or "ellipsoid", scale=semi_axes).translated(center)


def mesh_rectangle(*, size=(5.0, 5.0), center=(0.0, 0.0, 0.0), resolution=(5, 5), normal=(0.0, 0.0, 1.0),
                   name=None):
    """One-sided rectangle orthogonal to a coordinate axis (predefined/
    rectangles.py:26-100).  `size` spans the two other axes in increasing
    axis order; panels are ordered along the first of them, then the second."""
    normal = np.asarray(normal, dtype=float)
    axis = int(np.argmax(np.abs(normal)))
    if not np.isclose(abs(normal[axis]), 1.0):
        raise NotImplementedError("Only axis-aligned rectangles are generated.")
    u_ax, v_ax = [a for a in range(3) if a != axis]
    nu, nv = resolution
    u = np.linspace(-size[0] / 2, size[0] / 2, nu + 1)
    v = np.linspace(-size[1] / 2, size[1] / 2, nv + 1)
    verts = np.zeros(((nu + 1) * (nv + 1), 3))
    uu, vv = np.meshgrid(u, v, indexing="ij")
    verts[:, u_ax] = uu.ravel()
    verts[:, v_ax] = vv.ravel()
    idx = lambda i, j: i * (nv + 1) + j
    faces = []
    for i in range(nu):
        for j in range(nv):
            faces.append([idx(i, j), idx(i, j + 1), idx(i + 1, j + 1), idx(i + 1, j)])
    faces = np.array(faces, dtype=np.int64)
    mesh = Mesh(verts, faces, name=name)
    if np.dot(mesh.faces_normals[0], normal) < 0:
        faces = faces[:, ::-1].copy()
    return Mesh(verts + np.asarray(center, dtype=float), faces, name=name or "rectangle")


def mesh_parallelepiped(size=(1.0, 1.0, 1.0), center=(0.0, 0.0, 0.0), resolution=(4, 4, 4), missing_sides=(),
                        name=None):
    """Six rectangles (predefined/rectangles.py:103-220), outward normals,
    in the order left, right, front, back, top, bottom."""
    width, thickness, height = size
    nw, nth, nh = resolution
    missing = set(missing_sides)
    sides = []
    if "left" not in missing:
        sides.append(mesh_rectangle(size=(thickness, height), resolution=(nth, nh), center=(-width / 2, 0, 0),
                                    normal=(-1, 0, 0)))
    if "right" not in missing:
        sides.append(mesh_rectangle(size=(thickness, height), resolution=(nth, nh), center=(width / 2, 0, 0),
                                    normal=(1, 0, 0)))
    if "front" not in missing:
        sides.append(mesh_rectangle(size=(width, height), resolution=(nw, nh), center=(0, -thickness / 2, 0),
                                    normal=(0, -1, 0)))
    if "back" not in missing:
        sides.append(mesh_rectangle(size=(width, height), resolution=(nw, nh), center=(0, thickness / 2, 0),
                                    normal=(0, 1, 0)))
    if "top" not in missing:
        sides.append(mesh_rectangle(size=(width, thickness), resolution=(nw, nth), center=(0, 0, height / 2),
                                    normal=(0, 0, 1)))
    if "bottom" not in missing:
        sides.append(mesh_rectangle(size=(width, thickness), resolution=(nw, nth), center=(0, 0, -height / 2),
                                    normal=(0, 0, -1)))
    mesh = sides[0]
    for s in sides[1:]:
        mesh = mesh.join(s)
    mesh = mesh.translated(center, name=name or "parallelepiped")
    return mesh


def mesh_vertical_cylinder(*, radius=1.0, length=1.0, center=(0.0, 0.0, 0.0), resolution=(2, 8, 10), name=None):
    """Vertical cylinder open at the top (its top sits at center_z + length/2),
    closed by a disk at the bottom: resolution = (nr rings on the bottom disk,
    ntheta panels around, nz panels along the side), as predefined/
    cylinders.py:108-180 with the top disk removed."""
    nr, ntheta, nz = resolution
    z_top, z_bot = length / 2, -length / 2
    side_profile = [(radius, 0.0, z) for z in np.linspace(z_bot, z_top, nz + 1)]
    disk_profile = [(r, 0.0, z_bot) for r in np.linspace(0.0, radius, nr + 1)]
    # bottom disk from the axis outwards, then the side upwards
    profile = np.array(disk_profile + side_profile[1:])
    mesh = _revolution(profile, ntheta, name or "cylinder")
    return mesh.translated(center)


def mesh_horizontal_rectangle_lid(*, size, center, resolution, name=None):
    """Horizontal rectangle with upward normals (an interior lid at z = center_z)."""
    return mesh_rectangle(size=size, center=center, resolution=resolution, normal=(0, 0, 1), name=name)
