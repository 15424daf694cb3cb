"""Points at which post-processing evaluates a solved problem.

Same inputs and output shapes as capytaine's tools/lists_of_points.py
(:19-64): a single point (3,), a list of points (N, 3), the arrays of a
``np.meshgrid`` stacked along a first axis (3, nx, ny, ...), a mesh (its
face centres), or a body (its mesh).  For the free surface, points may be
given by x and y only; z is then 0.
"""
import numpy as np

from .meshes import Mesh, SymmetricMesh
from .problems import FloatingBody


def _normalize_points(points, keep_mesh=False):
    """Returns (points, output_shape): points an (N, 3) array, or the mesh
    itself when keep_mesh (the gradient needs the face normals of a mesh)."""
    if isinstance(points, FloatingBody):
        points = points.mesh
    if isinstance(points, (Mesh, SymmetricMesh)):
        return (points if keep_mesh else points.faces_centers), (points.nb_faces,)
    points = np.asarray(points)
    if points.ndim == 1:  # one point
        return points.reshape((1, points.shape[0])), (1,)
    if points.ndim == 2:
        return points, (points.shape[0],)
    if points.ndim > 2:  # meshgrid: (d, nx, ny, ...)
        return points.reshape(points.shape[0], -1).T, points.shape[1:]
    raise ValueError(f"Expected a list of points or a mesh, but got instead: {points}")


def _normalize_free_surface_points(points, keep_mesh=False):
    """As _normalize_points; (N, 2) points get z = 0."""
    if keep_mesh and isinstance(points, (FloatingBody, Mesh, SymmetricMesh)):
        return _normalize_points(points, keep_mesh=True)
    points, output_shape = _normalize_points(points, keep_mesh)
    if points.ndim == 2 and points.shape[1] == 2:
        points = np.concatenate([points, np.zeros((points.shape[0], 1))], axis=1)
    return points, output_shape
