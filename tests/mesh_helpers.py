"""Meshes the tests build from the mesh classes directly."""
import numpy as np

import capytaine_amd as cpa


def dihedral_sphere(n=6, ntheta=8):
    """A sphere as n rotations of a wedge that is symmetric across xOz: the
    half-wedge spans the angles 0 to pi / n."""
    theta = np.linspace(0.0, np.pi, ntheta + 1)
    pts = np.stack([np.sin(theta), np.zeros_like(theta), -np.cos(theta)], axis=1)
    pts[0, 0] = pts[-1, 0] = 0.0
    a = np.pi / n
    R = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    m = len(pts)
    half = cpa.Mesh(np.concatenate([pts, pts @ R.T]), [(i, i + m, i + m + 1, i + 1) for i in range(m - 1)])
    return cpa.RotationSymmetricMesh(cpa.ReflectionSymmetricMesh(half, "xOz"), n)
