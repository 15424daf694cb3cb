"""Bodies, problems and results: the host-side data around the device path.

Restates the parts of capytaine the BEM solve needs:
* rigid-body degrees of freedom (bodies/dofs.py:60-172) and
  ``FloatingBody.integrate_pressure`` (bodies/bodies.py:238-250);
* frequency/wavenumber relations (bem/problems_and_results.py:115-177);
* radiation and diffraction Neumann conditions (:371-381, :434-449) with the
  Airy incident wave (bem/airy_waves.py:19-102);
* results: added mass, radiation damping (:583-592), diffraction and
  Froude-Krylov forces (airy_waves.py:101-102).
"""
import logging

import numpy as np
from scipy.optimize import newton

LOG = logging.getLogger(__name__)

_default_parameters = dict(rho=1000.0, g=9.81, omega=1.0, free_surface=0.0, water_depth=np.inf,
                           wave_direction=0.0, forward_speed=0.0)


class TranslationDof:
    def __init__(self, direction):
        self.direction = np.asarray(direction, dtype=float)

    def evaluate_motion(self, mesh):
        return np.tile(self.direction, (mesh.nb_faces, 1))

    def evaluate_gradient_of_motion(self, mesh):
        """(faces, 3, 3) Jacobian of the motion, zero (dofs.py:71-72)."""
        return np.zeros((mesh.nb_faces, 3, 3))


class RotationDof:
    def __init__(self, rotation_center, direction):
        self.direction = np.asarray(direction, dtype=float)
        self.rotation_center = np.zeros(3) if rotation_center is None else np.asarray(rotation_center, float)

    def evaluate_motion(self, mesh):
        return np.cross(self.direction, mesh.faces_centers - self.rotation_center)

    def evaluate_gradient_of_motion(self, mesh):
        """(faces, 3, 3) Jacobian d(motion_i)/d(x_j) = (direction x e_j)_i
        (dofs.py:92-96)."""
        grad = np.cross(self.direction, np.eye(3)).T
        return np.tile(grad, (mesh.nb_faces, 1, 1))


class DofOnSubmesh:
    """`dof` on the faces of `mask`, zero elsewhere (dofs.py:99-128): the
    dofs of a body inside a joined multibody mesh."""

    def __init__(self, dof, mask):
        self.dof = dof
        self.mask = np.asarray(mask, dtype=bool)

    def evaluate_motion(self, mesh):
        out = np.zeros((mesh.nb_faces, 3))
        out[self.mask] = self.dof.evaluate_motion(mesh)[self.mask]
        return out

    def evaluate_gradient_of_motion(self, mesh):
        out = np.zeros((mesh.nb_faces, 3, 3))
        out[self.mask] = self.dof.evaluate_gradient_of_motion(mesh)[self.mask]
        return out

    def translated(self, shift):
        return DofOnSubmesh(_translated_dof(self.dof, shift), self.mask)


def _translated_dof(dof, shift):
    if isinstance(dof, RotationDof):
        return RotationDof(dof.rotation_center + np.asarray(shift, float), dof.direction)
    if hasattr(dof, "translated"):
        return dof.translated(shift)
    return dof


def normalize_name(name):
    """dofs.py:175-176."""
    return name[0].upper() + name[1:].lower()


def rigid_body_dofs(only=None, rotation_center=None):
    """dofs.py:131-172."""
    rc = np.zeros(3) if rotation_center is None else np.asarray(rotation_center, float)
    dofs = {
        "Surge": TranslationDof((1, 0, 0)),
        "Sway": TranslationDof((0, 1, 0)),
        "Heave": TranslationDof((0, 0, 1)),
        "Roll": RotationDof(rc, (1, 0, 0)),
        "Pitch": RotationDof(rc, (0, 1, 0)),
        "Yaw": RotationDof(rc, (0, 0, 1)),
    }
    if only is not None:
        dofs = {k: v for k, v in dofs.items() if k in only}
    return dofs


class FloatingBody:
    """A mesh with degrees of freedom and an optional lid (bodies/bodies.py:
    85-250).

    lid_mesh: panels of the interior free surface (or just below it) that
    remove the irregular frequencies; a lid whose normals all point up is
    turned over (bodies.py:91-99).  The BEM problem is set on
    ``mesh_including_lid`` (hull faces first); ``hull_mask`` selects the
    hull's faces, on which the boundary conditions and the forces live."""

    def __init__(self, mesh, dofs=None, name=None, *, lid_mesh=None, center_of_mass=None, mass=None):
        self.mesh = mesh
        self.dofs = {} if dofs is None else dict(dofs)
        self.name = name or mesh.name
        self.center_of_mass = None if center_of_mass is None else np.asarray(center_of_mass, dtype=float)
        self.mass = mass
        if lid_mesh is not None and lid_mesh.nb_faces == 0:
            LOG.warning("Lid mesh %s provided for body initialization is empty. The lid mesh is ignored.", lid_mesh)
            lid_mesh = None
        self.lid_mesh = None if lid_mesh is None else lid_mesh.with_normal_vector_going_down()
        if self.lid_mesh is None:
            self.mesh_including_lid = mesh
            self.hull_mask = np.ones(mesh.nb_faces, dtype=bool)
        else:
            from .meshes import join_with_mask
            self.mesh_including_lid, self.hull_mask = join_with_mask(mesh, self.lid_mesh,
                                                                     name=f"{mesh.name}+{self.lid_mesh.name}")

    def first_irregular_frequency_estimate(self, *, g=9.81):
        """bodies.py:566-601: lowest irregular frequency of a box of the
        body's horizontal spans and draft (the lid's depth when there is a
        lid; infinite for a lid on the free surface), for a body of one
        connected waterplane."""
        if self.lid_mesh is None:
            draft = abs(self.mesh.vertices[:, 2].min())
        else:
            draft = abs(self.lid_mesh.vertices[:, 2].min())
            if draft < 1e-6:
                return np.inf
        v = self.mesh.vertices
        p = np.hypot(1 / (v[:, 0].max() - v[:, 0].min()), 1 / (v[:, 1].max() - v[:, 1].min()))
        return float(np.sqrt(np.pi * g * p / np.tanh(np.pi * draft * p)))

    @property
    def nb_dofs(self):
        return len(self.dofs)

    @property
    def minimal_computable_wavelength(self):
        """bodies.py:558-563: eight times the largest panel radius."""
        r = self.mesh.faces_radiuses.max()
        if self.lid_mesh is not None:
            r = max(r, self.lid_mesh.faces_radiuses.max())
        return 8 * r

    def add_translation_dof(self, direction=None, name=None):
        """bodies.py:176-194: a translation; the direction of Surge, Sway or
        Heave is inferred from the name when not given."""
        if name is None:
            name = f"dof_{self.nb_dofs}_translation"
        if direction is None and normalize_name(name) in {"Surge", "Sway", "Heave"}:
            self.dofs[name] = rigid_body_dofs()[normalize_name(name)]
        else:
            self.dofs[name] = TranslationDof(direction)
        return self

    def add_rotation_dof(self, rotation_center=None, direction=None, name=None):
        """bodies.py:196-227: a rotation about an axis through
        `rotation_center` (default: the body's centre of mass if given, else
        the origin); Roll, Pitch or Yaw infer the direction from the name."""
        if name is None:
            name = f"dof_{self.nb_dofs}_rotation"
        if rotation_center is None and self.center_of_mass is not None:
            rotation_center = self.center_of_mass
        if direction is None and normalize_name(name) in {"Roll", "Pitch", "Yaw"}:
            self.dofs[name] = rigid_body_dofs(rotation_center=rotation_center)[normalize_name(name)]
        else:
            self.dofs[name] = RotationDof(rotation_center, direction)
        return self

    def add_all_rigid_body_dofs(self, rotation_center=None):
        self.dofs.update(rigid_body_dofs(rotation_center=rotation_center))
        return self

    def copy(self, name=None):
        return FloatingBody(self.mesh, dict(self.dofs), name=name or self.name, lid_mesh=self.lid_mesh,
                            center_of_mass=self.center_of_mass, mass=self.mass)

    def translated(self, shift, name=None):
        """bodies.py:389-400: mesh, lid, rotation centres and centre of mass
        moved by `shift`."""
        shift = np.asarray(shift, dtype=float)
        dofs = {k: _translated_dof(v, shift) for k, v in self.dofs.items()}
        com = None if self.center_of_mass is None else np.asarray(self.center_of_mass) + shift
        return FloatingBody(self.mesh.translated(shift), dofs, name=name or self.name,
                            lid_mesh=None if self.lid_mesh is None else self.lid_mesh.translated(shift),
                            center_of_mass=com, mass=self.mass)

    def translated_x(self, dx, name=None):
        return self.translated((dx, 0.0, 0.0), name=name)

    def translated_y(self, dy, name=None):
        return self.translated((0.0, dy, 0.0), name=name)

    def translated_z(self, dz, name=None):
        return self.translated((0.0, 0.0, dz), name=name)

    def immersed_part(self, free_surface=0.0, name=None):
        """The body with the faces of its mesh below the free surface; rigid
        dofs are kept (their motion is evaluated on the new mesh)."""
        dofs = {k: v for k, v in self.dofs.items() if hasattr(v, "evaluate_motion")}
        return FloatingBody(self.mesh.immersed_part(free_surface), dofs, name=name or self.name,
                            lid_mesh=self.lid_mesh, center_of_mass=self.center_of_mass, mass=self.mass)

    def join_bodies(self, *others, name=None):
        """The bodies as one (multibodies.py:34-131): meshes joined, each
        body's dofs restricted to its faces and renamed body__dof, lids
        joined.  Names must be distinct."""
        from .meshes import SymmetricMesh
        bodies = [self, *others]
        if len({b.name for b in bodies}) < len(bodies):
            raise ValueError(f"In Multibody, all component bodies must have a distinct name.\n"
                             f"Got: {[b.name for b in bodies]}")
        meshes = [b.mesh.merged() if isinstance(b.mesh, SymmetricMesh) else b.mesh for b in bodies]
        mesh = meshes[0]
        for m in meshes[1:]:
            mesh = mesh.join(m)
        sizes = np.cumsum([0] + [m.nb_faces for m in meshes])
        dofs = {}
        for b, lo, hi in zip(bodies, sizes[:-1], sizes[1:]):
            mask = np.zeros(mesh.nb_faces, dtype=bool)
            mask[lo:hi] = True
            for k, dof in b.dofs.items():
                new_name = k if "__" in k else f"{b.name}__{k}"
                if hasattr(dof, "evaluate_motion"):
                    dofs[new_name] = DofOnSubmesh(dof, mask)
                else:
                    arr = np.zeros((mesh.nb_faces, 3))
                    arr[mask] = np.asarray(dof)
                    dofs[new_name] = arr
        lids = [b.lid_mesh for b in bodies if b.lid_mesh is not None]
        lid = None
        if lids:
            lid = lids[0]
            for m in lids[1:]:
                lid = lid.join(m)
        mesh.name = name or "+".join(b.name for b in bodies)
        return FloatingBody(mesh, dofs, name=name or "+".join(b.name for b in bodies), lid_mesh=lid)

    def __add__(self, other):
        return self.join_bodies(other)

    def dof_motion(self, dof_name):
        dof = self.dofs[dof_name]
        return dof.evaluate_motion(self.mesh) if hasattr(dof, "evaluate_motion") else np.asarray(dof)

    def dof_normal_matrix(self, including_lid=False):
        """(ndof, faces) matrix of dof . n on each face of the hull, dofs in
        insertion order; with including_lid, over mesh_including_lid with
        zero columns on the lid."""
        n = self.mesh.faces_normals
        dn = np.array([np.sum(self.dof_motion(d) * n, axis=1) for d in self.dofs]).reshape(len(self.dofs), -1)
        if not including_lid or self.lid_mesh is None:
            return dn
        out = np.zeros((len(self.dofs), self.mesh_including_lid.nb_faces))
        out[:, self.hull_mask] = dn
        return out

    def integrate_pressure(self, pressure):
        """bodies.py:238-250: force on each dof, F = -sum p (dof . n) A."""
        forces = {}
        for dof_name in self.dofs:
            dof = self.dof_motion(dof_name)
            normal_dof_amplitude_on_face = - np.sum(dof * self.mesh.faces_normals, axis=1)
            forces[dof_name] = np.sum(pressure * normal_dof_amplitude_on_face * self.mesh.faces_areas)
        return forces

    def __repr__(self):
        return f"FloatingBody(name={self.name!r}, nb_faces={self.mesh.nb_faces}, dofs={list(self.dofs)})"


def wavenumber_from_omega(omega, g, water_depth):
    """problems_and_results.py:159-162."""
    if water_depth == np.inf:
        return omega ** 2 / g
    return newton(lambda k: k * np.tanh(k * water_depth) - omega ** 2 / g, x0=1.0)


class LinearPotentialFlowProblem:
    def __init__(self, *, body=None, free_surface=0.0, water_depth=np.inf, omega=None, wavenumber=None,
                 freq=None, period=None, wavelength=None, rho=1000.0, g=9.81, wave_direction=0.0,
                 forward_speed=0.0):
        self.body = body
        self.free_surface = float(free_surface)
        self.water_depth = float(water_depth)
        self.rho = float(rho)
        self.g = float(g)
        self.wave_direction = float(wave_direction)
        self.forward_speed = float(forward_speed)
        given = {k: v for k, v in dict(omega=omega, freq=freq, period=period, wavenumber=wavenumber,
                                           wavelength=wavelength).items() if v is not None}
        if len(given) > 1:
            raise ValueError("Settings a problem requires at most one of the following: omega (angular frequency) "
                             "OR freq (in Hz) OR period OR wavenumber OR wavelength.\n"
                             f"Received {len(given)} of them: {given}")
        # problems_and_results.py:115-177: the other forms of the frequency
        if freq is not None:
            omega = 2 * np.pi * float(freq)
        elif period is not None:
            omega = 2 * np.pi / float(period) if float(period) not in (0.0, np.inf) else \
                (np.inf if float(period) == 0.0 else 0.0)
        elif wavelength is not None:
            wavenumber = 2 * np.pi / float(wavelength) if float(wavelength) not in (0.0, np.inf) else \
                (np.inf if float(wavelength) == 0.0 else 0.0)
        if wavenumber is not None:
            self.wavenumber = float(wavenumber)
            self.omega = float(np.sqrt(self.g * self.wavenumber * np.tanh(self.wavenumber * self.water_depth))) \
                if self.wavenumber not in (0.0, np.inf) else self.wavenumber
        else:
            self.omega = float(_default_parameters["omega"] if omega is None else omega)
            if self.omega in (0.0, np.inf):
                self.wavenumber = self.omega
            else:
                self.wavenumber = float(wavenumber_from_omega(self.omega, self.g, self.water_depth))
        if self.omega in (0.0, np.inf) and self.forward_speed != 0.0:
            raise NotImplementedError(f"omega={self.omega} is only implemented without forward speed "
                                      f"(provided forward speed: {self.forward_speed}).")
        if self.forward_speed != 0.0:
            # problems_and_results.py:96-105: Doppler-shifted frequency; a
            # body faster than the waves meets them from the other side
            dopplered = self.omega - self.wavenumber * self.forward_speed * np.cos(self.wave_direction)
            self.encounter_omega = float(abs(dopplered))
            self.encounter_wavenumber = (self.encounter_omega if self.encounter_omega == 0.0 else
                                         float(wavenumber_from_omega(self.encounter_omega, self.g, self.water_depth)))
            self.encounter_wave_direction = self.wave_direction + (0.0 if dopplered >= 0.0 else np.pi)
        else:
            self.encounter_omega = self.omega
            self.encounter_wavenumber = self.wavenumber
            self.encounter_wave_direction = self.wave_direction

    @property
    def influenced_dofs(self):
        return self.body.dofs

    freq = property(lambda self: self.omega / (2 * np.pi))
    period = property(lambda self: 2 * np.pi / self.omega if self.omega != 0.0 else np.inf)
    wavelength = property(lambda self: 2 * np.pi / self.wavenumber if self.wavenumber != 0.0 else np.inf)

    def _group_key(self):
        """problems_and_results.py:266-275 grouping key."""
        return (id(self.body), self.water_depth, self.omega, self.rho, self.g, self.free_surface, self.forward_speed)


def _airy_depth_terms(k, h, z):
    """cosh(k(z+h))/cosh(kh) and sinh(k(z+h))/cosh(kh), or exp(kz) for both
    from kh >= 20 on (airy_waves.py:40-46)."""
    if 0 <= k * h < 20:
        return np.cosh(k * (z + h)) / np.cosh(k * h), np.sinh(k * (z + h)) / np.cosh(k * h)
    return np.exp(k * z), np.exp(k * z)


def airy_waves_velocity(points, pb):
    """airy_waves.py:56-94: velocity of the undisturbed incident wave,
    shape (*output_shape, 3); NaN at zero or infinite wavenumber."""
    from .lists_of_points import _normalize_points
    points, output_shape = _normalize_points(points)
    if float(pb.wavenumber) in (0.0, np.inf):
        return np.nan * np.ones((*output_shape, 3))
    x, y, z = points.T
    k, h, beta = pb.wavenumber, pb.water_depth, pb.encounter_wave_direction
    wbar = x * np.cos(beta) + y * np.sin(beta)
    cih, sih = _airy_depth_terms(k, h, z)
    v = pb.g * k / pb.omega * np.exp(1j * k * wbar) * \
        np.array([np.cos(pb.wave_direction) * cih, np.sin(pb.wave_direction) * cih, -1j * sih])
    return v.T.reshape((*output_shape, 3))


def airy_waves_potential(points, pb):
    """airy_waves.py:19-53: potential of the undisturbed incident wave,
    shape output_shape; NaN at zero or infinite wavenumber."""
    from .lists_of_points import _normalize_points
    points, output_shape = _normalize_points(points)
    if float(pb.wavenumber) in (0.0, np.inf):
        return np.nan * np.ones(output_shape)
    x, y, z = points.T
    k, h, beta = pb.wavenumber, pb.water_depth, pb.encounter_wave_direction
    wbar = x * np.cos(beta) + y * np.sin(beta)
    cih, _ = _airy_depth_terms(k, h, z)
    return (-1j * pb.g / pb.omega * cih * np.exp(1j * k * wbar)).reshape(output_shape)


def airy_waves_pressure(points, pb):
    """airy_waves.py:97-98."""
    return 1j * float(pb.omega) * pb.rho * airy_waves_potential(points, pb)


def airy_waves_free_surface_elevation(points, pb):
    """airy_waves.py:105-122: elevation of the undisturbed incident wave;
    points of two coordinates are taken at z = 0."""
    from .lists_of_points import _normalize_free_surface_points
    points, output_shape = _normalize_free_surface_points(points)
    return 1j * pb.omega / pb.g * airy_waves_potential(points, pb).reshape(output_shape)


def froude_krylov_force(pb):
    """airy_waves.py:97-102."""
    return pb.body.integrate_pressure(airy_waves_pressure(pb.body.mesh.faces_centers, pb))


class RadiationProblem(LinearPotentialFlowProblem):
    def __init__(self, *, body=None, radiating_dof=None, **kwargs):
        super().__init__(body=body, **kwargs)
        if body is None or len(body.dofs) == 0:
            raise ValueError("A radiation problem needs a body with degrees of freedom")
        self.radiating_dof = next(iter(body.dofs)) if radiating_dof is None else radiating_dof
        if self.radiating_dof not in body.dofs:
            raise ValueError(f"Unknown radiating dof {self.radiating_dof!r}")
        displacement = np.sum(body.dof_motion(self.radiating_dof) * body.mesh.faces_normals, axis=1)
        # problems_and_results.py:436-447: zero on the lid
        self.boundary_condition = np.zeros(body.mesh_including_lid.nb_faces, dtype=np.complex128)
        self.boundary_condition[body.hull_mask] = -1j * self.encounter_omega * displacement
        if self.forward_speed != 0.0:
            # the m-terms (problems_and_results.py:451-464): -U (d motion / dx) . n
            dof = body.dofs[self.radiating_dof]
            if not hasattr(dof, "evaluate_gradient_of_motion"):
                raise NotImplementedError(
                    "Radiation problem with forward speed is currently only implemented for rigid bodies.\n"
                    "Only radiating dofs instantiating a TranslationDof or a RotationDof are supported.\n"
                    f"Got instead `radiating_dof={self.radiating_dof}`")
            ddofdx = dof.evaluate_gradient_of_motion(body.mesh)[:, :, 0]
            self.boundary_condition[body.hull_mask] += -self.forward_speed * np.sum(ddofdx * body.mesh.faces_normals,
                                                                                     axis=1)

    def make_results_container(self, forces=None, sources=None, potential=None, pressure=None):
        return RadiationResult(self, forces, sources, potential, pressure)

    def make_failed_results_container(self, exception):
        r = RadiationResult(self, {d: np.nan for d in self.body.dofs})
        r.exception = exception
        return r

    def __repr__(self):
        return f"RadiationProblem(body={self.body.name}, omega={self.omega:.3f}, radiating_dof={self.radiating_dof})"


class DiffractionProblem(LinearPotentialFlowProblem):
    def __init__(self, *, body=None, **kwargs):
        super().__init__(body=body, **kwargs)
        # problems_and_results.py:373-381: zero on the lid
        self.boundary_condition = np.zeros(body.mesh_including_lid.nb_faces, dtype=np.complex128)
        self.boundary_condition[body.hull_mask] = -(airy_waves_velocity(body.mesh.faces_centers, self)
                                                    * body.mesh.faces_normals).sum(axis=1)

    def make_results_container(self, forces=None, sources=None, potential=None, pressure=None):
        return DiffractionResult(self, forces, sources, potential, pressure)

    def make_failed_results_container(self, exception):
        r = DiffractionResult(self, {d: np.nan for d in self.body.dofs})
        r.exception = exception
        return r

    def __repr__(self):
        return f"DiffractionProblem(body={self.body.name}, omega={self.omega:.3f}, wave_direction={self.wave_direction:.3f})"


class LinearPotentialFlowResult:
    def __init__(self, problem, forces=None, sources=None, potential=None, pressure=None):
        self.problem = problem
        self.forces = forces if forces is not None else {}
        self.force = self.forces  # alias of the reference (problems_and_results.py:529-531)
        self.sources = sources
        self.potential = potential
        self.pressure = pressure
        self.omega = problem.omega
        self.encounter_omega = problem.encounter_omega
        self.encounter_wave_direction = problem.encounter_wave_direction
        self.wavenumber = problem.wavenumber
        self.encounter_wavenumber = problem.encounter_wavenumber
        self.body = problem.body
        self.free_surface = problem.free_surface
        self.water_depth = problem.water_depth
        self.forward_speed = problem.forward_speed
        self.rho = problem.rho
        self.g = problem.g


class RadiationResult(LinearPotentialFlowResult):
    @property
    def radiating_dof(self):
        return self.problem.radiating_dof

    added_masses = property(lambda self: self.added_mass)
    radiation_dampings = property(lambda self: self.radiation_damping)

    @property
    def added_mass(self):
        w = self.encounter_omega
        return {dof: float(np.real(f) / (w * w)) for dof, f in self.forces.items()}

    @property
    def radiation_damping(self):
        w = self.encounter_omega
        if float(w) in {0.0, np.inf} and self.forward_speed == 0.0:
            return {dof: 0.0 for dof in self.forces}
        return {dof: float(np.imag(f) / w) for dof, f in self.forces.items()}


class DiffractionResult(LinearPotentialFlowResult):
    @property
    def diffraction_force(self):
        return dict(self.forces)

    @property
    def froude_krylov_force(self):
        return froude_krylov_force(self.problem)

    @property
    def excitation_force(self):
        fk = self.froude_krylov_force
        return {dof: self.forces[dof] + fk[dof] for dof in self.forces}
