"""Synthetic workloads: the bodies, frequencies and problems the benchmark
and the tests run, generated from seeds (no data files).

  barge()            rectangular barge 20 x 8 x 2 m, 0.25 m panels (4352),
                     40 frequencies in 0.2-2.5 rad/s, 6 dof + 8 headings
  hull(a, b, c, n)   half-immersed ellipsoid hull with about n panels
  hull_fleet(seed)   64 hulls, a~U[2,10], b~U[1,4], c~U[0.5,3] m, panel
                     counts log-uniform in 300-6000; 16 frequencies
                     log-uniform in 0.3-3.0 rad/s; 6 dof + 4 headings
  serving_request()  one request of the serving stream: a fleet-generator
                     hull with 500-12000 panels, 1-32 frequencies in
                     0.2-3.0 rad/s, a non-empty subset of the 6 dof, 0-8
                     headings
  cylinder_array()   5 x 5 truncated vertical cylinders, radius 2 m, draft
                     4 m, 12 m apart, 800 panels each (20000), depth 40 m,
                     20 frequencies in 0.3-1.8 rad/s, 150 dof + 8 headings
"""
import numpy as np

from .meshes import Mesh, _revolution, mesh_parallelepiped, mesh_vertical_cylinder
from .problems import DiffractionProblem, FloatingBody, RadiationProblem, rigid_body_dofs


def barge():
    mesh = mesh_parallelepiped(size=(20, 8, 2), center=(0, 0, -1), resolution=(80, 32, 8), missing_sides={"top"})
    body = FloatingBody(mesh, name="barge").add_all_rigid_body_dofs(rotation_center=(0, 0, 0))
    return body, np.linspace(0.2, 2.5, 40), np.arange(8) * np.pi / 4


def hull_mesh(a, b, c, n_target, name="hull"):
    """Half-immersed ellipsoid x^2/a^2 + y^2/b^2 + z^2/c^2 = 1, z <= 0: a
    revolution mesh of the lower half-sphere scaled by (a, b, c), with
    nt rows of latitude and nphi meridians, nt * nphi ~ n_target (nphi = 2
    nt gives near-square panels on the sphere)."""
    nt = max(2, int(round(np.sqrt(n_target / 2.0))))
    nphi = max(3, int(round(n_target / nt)))
    theta = np.linspace(0.0, np.pi / 2, nt + 1)
    profile = np.stack([np.sin(theta), np.zeros_like(theta), -np.cos(theta)], axis=1)
    profile[0, 0] = 0.0
    profile[-1, 2] = 0.0
    return _revolution(profile, nphi, name, scale=(a, b, c))


def hull(a, b, c, n_target, name="hull"):
    mesh = hull_mesh(a, b, c, n_target, name)
    return FloatingBody(mesh, name=name).add_all_rigid_body_dofs(rotation_center=(0, 0, 0))


def hull_fleet(seed=0, count=64, n_range=(300, 6000)):
    """(bodies, omegas, headings) of the fleet; seed recorded by the caller."""
    rng = np.random.default_rng(seed)
    bodies = []
    for i in range(count):
        a, b, c = rng.uniform(2, 10), rng.uniform(1, 4), rng.uniform(0.5, 3)
        n = float(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]))))
        bodies.append(hull(a, b, c, n, name=f"hull{i}"))
    omegas = np.exp(np.linspace(np.log(0.3), np.log(3.0), 16))
    headings = np.arange(4) * np.pi / 2
    return bodies, omegas, headings


def serving_request(rng, n_range=(500, 12000)):
    """(body, omegas, dofs, headings) of one request of the serving stream."""
    a, b, c = rng.uniform(2, 10), rng.uniform(1, 4), rng.uniform(0.5, 3)
    n = float(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]))))
    body = FloatingBody(hull_mesh(a, b, c, n), name=f"req{rng.integers(1 << 30)}")
    all_dofs = rigid_body_dofs(rotation_center=(0, 0, 0))
    names = list(all_dofs)
    mask = rng.random(6) < 0.5
    if not mask.any():
        mask[rng.integers(6)] = True
    for name, keep in zip(names, mask):
        if keep:
            body.dofs[name] = all_dofs[name]
    omegas = np.sort(rng.uniform(0.2, 3.0, int(rng.integers(1, 33))))
    headings = np.arange(int(rng.integers(0, 9))) * np.pi / 4
    return body, omegas, headings


def cylinder_array(rows=5, cols=5, spacing=12.0, radius=2.0, draft=4.0, resolution=(8, 40, 12)):
    """One FloatingBody with a rigid-body dof set per cylinder (names
    "c<i>_<dof>"), the 5 x 5 array by default."""
    meshes, dofs = [], {}
    k = 0
    for i in range(rows):
        for j in range(cols):
            x, y = (i - (rows - 1) / 2) * spacing, (j - (cols - 1) / 2) * spacing
            m = mesh_vertical_cylinder(radius=radius, length=draft, center=(x, y, -draft / 2),
                                       resolution=resolution, name=f"c{k}")
            meshes.append(m)
            k += 1
    vertices, faces, offset = [], [], 0
    face_owner = []
    for idx, m in enumerate(meshes):
        vertices.append(m.vertices)
        faces.append(m.faces + offset)
        face_owner.append(np.full(m.nb_faces, idx))
        offset += m.nb_vertices
    mesh = Mesh(np.concatenate(vertices), np.concatenate(faces), name="cylinder_array")
    owner = np.concatenate(face_owner)
    body = FloatingBody(mesh, name="cylinder_array")
    for idx, m in enumerate(meshes):
        c = m.faces_centers.mean(axis=0)
        mask = (owner == idx)[:, None]
        for name, dof in rigid_body_dofs(rotation_center=(c[0], c[1], 0.0)).items():
            # a dof given as its motion on every face, zero off this cylinder
            dofs[f"c{idx}_{name}"] = dof.evaluate_motion(mesh) * mask
    body.dofs.update(dofs)
    return body, np.linspace(0.3, 1.8, 20), np.arange(8) * np.pi / 4, 40.0


def problems_by_frequency(body, omegas, headings, rho=1000.0, g=9.81, water_depth=np.inf):
    """Per frequency: the radiation problems of every dof, then the
    diffraction problems of every heading."""
    out = []
    for om in omegas:
        pbs = [RadiationProblem(body=body, radiating_dof=d, omega=om, rho=rho, g=g, water_depth=water_depth)
               for d in body.dofs]
        pbs += [DiffractionProblem(body=body, omega=om, wave_direction=h, rho=rho, g=g, water_depth=water_depth)
                for h in headings]
        out.append(pbs)
    return out
