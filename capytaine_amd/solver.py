"""BEM solver on the GPU engine.

Drop-in for ``capytaine.bem.solver.BEMSolver`` (src/capytaine/bem/solver.py:
59-423) for the problems the hot path covers: ``solve`` follows ``_solve``
(:145-215) step by step (build S and K through the engine, LU solve of the
boundary condition, potential = S @ sources, pressure, forces), and
``solve_all`` (:268-372) solves the problems of a group -- same body, depth,
frequency, rho and g (problems_and_results.py:266-275) -- with one assembly,
one factorisation and one multi-right-hand-side solve.  ``solve_batched``
goes further: every frequency of a body in one assembly launch, one batched
LU and one batched solve.  ``compute_potential``, ``compute_velocity``,
``compute_pressure`` and ``compute_free_surface_elevation`` (:426-561)
evaluate a solved problem at points with the same assembly kernel, points
as collocation, and take the sum over the panels on the device.
"""
import logging
import time
from collections import OrderedDict, defaultdict

import numpy as np

from . import _lib
from ._lib import check, ptr, stream_handle
from .engines import (DefaultMatrixEngine, DeviceMatrix, LUDecomposedDeviceMatrix, MixedPrecisionSolver,
                      check_if_nan_in_matrix)
from .lists_of_points import _normalize_points, _normalize_free_surface_points
from .symmetry import LUDecomposedSymmetricMatrix, MixedSymmetricSolver, evaluate_symmetric, symmetry_of
from .problems import DiffractionProblem, RadiationProblem
from .problems_checks import check_problems

LOG = logging.getLogger(__name__)


def _complex_dtype(A):
    """Complex type of the right-hand side for the matrix A (solver.py:52-57,
    184, 195): complex64 for the single-precision Green function's
    matrices, complex128 otherwise."""
    return np.complex64 if np.dtype(getattr(A, "dtype", np.complex128)) in (np.float32, np.complex64) \
        else np.complex128


def _single_precision(engine):
    gf = getattr(engine, "green_function", None)
    return getattr(gf, "floating_point_precision", "float64") == "float32"


class Timer:
    """Accumulated wall time per step, with the reference's step names
    (tools/timer.py, solver.py:176-198)."""

    def __init__(self):
        self.totals = defaultdict(float)
        self.counts = defaultdict(int)

    def __call__(self, step):
        timer = self

        class _Ctx:
            def __enter__(self_):
                self_.t = time.perf_counter()

            def __exit__(self_, *exc):
                timer.totals[step] += time.perf_counter() - self_.t
                timer.counts[step] += 1
        return _Ctx()

    def summary(self):
        return dict(self.totals)


class BEMSolver:
    def __init__(self, *, green_function=None, engine=None, method="indirect"):
        if engine is None:
            self.engine = DefaultMatrixEngine(green_function=green_function)
        else:
            if green_function is not None:
                raise ValueError("If you are not using the default engine, set the Green function in the engine.")
            self.engine = engine
        if method.lower() not in {"direct", "indirect"}:
            raise ValueError(f"Unrecognized method when initializing solver: {method!r}.")
        self.method = method.lower()
        self.timer = Timer()
        self.exportable_settings = {**self.engine.exportable_settings, "method": self.method}

    def __str__(self):
        return f"BEMSolver(engine={self.engine}, method={self.method})"

    def reset_timer(self):
        """solver.py:113-114."""
        self.timer = Timer()

    def timer_summary(self):
        """Total wall time and count per step (solver.py:116-126)."""
        return {step: {"total_s": t, "count": self.timer.counts[step]} for step, t in self.timer.totals.items()}

    __repr__ = __str__

    # ------------------------------------------------------------------
    def _solve(self, problem, method=None, keep_details=True):
        if isinstance(problem, DiffractionProblem) and float(problem.encounter_omega) in {0.0, np.inf}:
            raise ValueError("Diffraction problems at zero or infinite frequency are not defined")
        # solver.py:163-167: with forward speed, the encounter frequency
        omega, wavenumber = problem.encounter_omega, problem.encounter_wavenumber
        gf_params = dict(free_surface=problem.free_surface, water_depth=problem.water_depth, wavenumber=wavenumber)
        mesh = problem.body.mesh_including_lid
        method = method if method is not None else self.method
        if method == "direct" and problem.forward_speed != 0.0:
            raise NotImplementedError("Direct solver is not able to solve problems with forward speed.")
        if method == "direct":
            with self.timer("Green function"):
                S, D = self.engine.build_matrices(mesh, mesh, **gf_params, adjoint_double_layer=False,
                                                  diagonal_term_in_double_layer=True)
            with self.timer("Matrix-vector product"):
                rhs = S @ problem.boundary_condition
            with self.timer("Linear solver"):
                potential = self.engine.linear_solver(D, rhs.astype(_complex_dtype(D)))
            sources = None
        else:
            with self.timer("Green function"):
                S, K = self.engine.build_matrices(mesh, mesh, **gf_params, adjoint_double_layer=True,
                                                  diagonal_term_in_double_layer=True)
            with self.timer("Linear solver"):
                sources = self.engine.linear_solver(K, problem.boundary_condition.astype(_complex_dtype(K)))
            with self.timer("Matrix-vector product"):
                potential = S @ sources
        # in the potential's precision (complex64 for the single-precision
        # Green function), whatever the scalars' types
        pressure = (1j * omega * problem.rho * potential).astype(potential.dtype, copy=False)
        if problem.forward_speed != 0.0:
            # solver.py:200-204: + rho U d(phi)/dx, the gradient on the panels
            # of the mesh itself (with the diagonal term of build_fullK_matrix)
            nabla_phi = self._compute_potential_gradient(mesh, problem.make_results_container(sources=sources))
            pressure = pressure + problem.rho * problem.forward_speed * nabla_phi[:, 0]
        forces = problem.body.integrate_pressure(pressure[problem.body.hull_mask])
        if not keep_details:
            return problem.make_results_container(forces)
        return problem.make_results_container(forces, sources, potential, pressure)

    def solve(self, problem, method=None, keep_details=True, n_threads=None, _check_wavelength=True):
        """solver.py:217-254 (n_threads accepted and unused)."""
        if _check_wavelength:
            check_problems([problem])
        with self.timer("Total solve function"):
            return self._solve(problem, method=method, keep_details=keep_details)

    def _solve_and_catch_errors(self, problem, **kwargs):
        try:
            return self.solve(problem, **kwargs)
        except Exception as e:  # solver.py:256-265
            LOG.warning("Problem %s failed: %s", problem, e)
            return problem.make_failed_results_container(e)

    # ------------------------------------------------------------------
    def solve_all(self, problems, *, method=None, keep_details=False, batch_frequencies=True,
                  max_batch_bytes=None, n_jobs=1, n_threads=None, progress_bar=None):
        """Solve several problems.  Problems sharing a body, depth, rho, g and
        free surface are solved together: with batch_frequencies, all their
        frequencies at once (`solve_batched`); otherwise one frequency at a
        time with all its right-hand sides at once.  Results come back in
        the order of `problems`.  The engine's linear_solver
        ("mixed_precision" or not) chooses the batched factorisation.
        n_jobs, n_threads and progress_bar are accepted for the reference's
        signature (solver.py:268) and unused: the device batches instead."""
        method = method if method is not None else self.method
        check_problems(problems)
        # the batched paths hold complex128 matrices: the single-precision
        # Green function solves problem by problem, as the reference does
        if method != "indirect" or _single_precision(self.engine):
            return [self._solve_and_catch_errors(pb, method=method, keep_details=keep_details,
                                                 _check_wavelength=False) for pb in problems]
        results = [None] * len(problems)
        groups = OrderedDict()
        for i, pb in enumerate(problems):
            groups.setdefault((id(pb.body), pb.water_depth, pb.rho, pb.g, pb.free_surface, pb.forward_speed),
                              []).append(i)
        for key, idx in groups.items():
            sub = [problems[i] for i in idx]
            if batch_frequencies:
                res = self.solve_batched(sub, keep_details=keep_details, max_batch_bytes=max_batch_bytes)
            else:
                res = []
                by_omega = OrderedDict()
                for j, pb in enumerate(sub):
                    by_omega.setdefault(pb.encounter_omega, []).append(j)
                res = [None] * len(sub)
                for om, js in by_omega.items():
                    out = self.solve_batched([sub[j] for j in js], keep_details=keep_details,
                                             max_batch_bytes=max_batch_bytes)
                    for j, r in zip(js, out):
                        res[j] = r
            for i, r in zip(idx, res):
                results[i] = r
        return results

    def solve_batched(self, problems, *, keep_details=False, max_batch_bytes=None, return_device=False):
        """All problems share body, depth, rho, g, free surface and forward
        speed; frequencies may differ.  Per chunk of encounter frequencies: one
        assembly launch for all of them, one batched LU of the K matrices
        (overwritten), one batched solve for all right-hand sides, one batched
        S @ sources, and the forces as a (dof x panels) product on the device.
        With forward speed, one more assembly launch gives the full gradient
        on the panels for the pressure term rho U d(phi)/dx
        (solver.py:200-204)."""
        import torch
        if _single_precision(self.engine):
            raise NotImplementedError("solve_batched holds complex128 matrices; with the single-precision Green "
                                      "function use solve or solve_all")
        body = problems[0].body
        mesh = body.mesh_including_lid
        gf = self.engine.green_function
        dev = gf.device
        n = mesh.nb_faces
        speed = problems[0].forward_speed
        omegas = OrderedDict()
        for j, pb in enumerate(problems):
            if isinstance(pb, DiffractionProblem) and float(pb.encounter_omega) in {0.0, np.inf}:
                raise ValueError("Diffraction problems at zero or infinite frequency are not defined")
            if pb.forward_speed != speed:
                raise ValueError("solve_batched takes problems of one forward speed")
            omegas.setdefault(pb.encounter_omega, []).append(j)
        omega_list = list(omegas)
        if max_batch_bytes is None:
            free, _ = torch.cuda.mem_get_info(dev)
            max_batch_bytes = int(0.6 * free)
        sym = symmetry_of(mesh)
        per_freq = (3 * n * n * 16 if speed != 0.0 else 0) + 2 * n * n * 16 // (sym.n if sym else 1)
        chunk = max(1, min(len(omega_list), max_batch_bytes // per_freq))
        # W[i, d] = -(dof_d . n_i) A_i  (integrate_pressure, bodies.py:238-250)
        dofs = list(body.dofs)
        W = -(body.dof_normal_matrix(including_lid=True) * mesh.faces_areas[None, :]).T  # (n, ndof)
        W_d = torch.as_tensor(np.ascontiguousarray(W), dtype=torch.complex128).to(dev)
        ndof = len(dofs)
        results = [None] * len(problems)
        ctx = _lib.Context.get(dev)
        lib = _lib.load()
        for c0 in range(0, len(omega_list), chunk):
            oms = omega_list[c0:c0 + chunk]
            ks = [problems[omegas[om][0]].encounter_wavenumber for om in oms]
            pb0 = problems[omegas[oms[0]][0]]
            with self.timer("Green function"):
                gp = dict(free_surface=pb0.free_surface, water_depth=pb0.water_depth, wavenumber=ks,
                          adjoint_double_layer=True, diagonal_term_in_double_layer=True)
                if sym:  # the first block column only (engines.py:178-247)
                    S, K = evaluate_symmetric(gf, mesh, sym, **gp)
                else:
                    S, K = (DeviceMatrix(t) for t in gf.evaluate_on_device(mesh, mesh, **gp))
                if self.engine.check_nan:
                    check_if_nan_in_matrix([S, K])
            nb = len(oms)
            R = max(len(omegas[om]) for om in oms)
            bc = np.zeros((nb, R, n), dtype=np.complex128)
            for b, om in enumerate(oms):
                for r, j in enumerate(omegas[om]):
                    bc[b, r] = problems[j].boundary_condition
            with self.timer("Linear solver"):
                X = torch.as_tensor(bc).to(dev)
                if getattr(self.engine, "_linear_solver", None) == "mixed_precision":
                    # complex64 LU with FP64 refinement, all frequencies at once
                    lu = MixedSymmetricSolver(K) if sym else MixedPrecisionSolver(K)
                    X = lu.solve_device(X)
                else:
                    lu = LUDecomposedSymmetricMatrix(K) if sym else LUDecomposedDeviceMatrix(K, overwrite_a=True)
                    lu.solve_device(X)
            with self.timer("Matrix-vector product"):
                Phi = S.matmul_device(X)
                # G (ndof x R) = W^T (ndof x n) Phi (n x R), per frequency
                G = torch.empty((nb, R, ndof), dtype=torch.complex128, device=dev)
                check(lib.cpt_zgemm(ctx.handle, ndof, R, n, 1.0, ptr(W_d), ndof, 0, ptr(Phi), n, n * R, ptr(G),
                                    ndof, ndof * R, 1, nb, stream_handle(dev)), "cpt_zgemm")
            Dx_h = None
            if speed != 0.0:
                with self.timer("Post-processing velocity"):
                    # d(phi)/dx on the panels: the x row of the full gradient
                    # (build_fullK_matrix(mesh, mesh), diagonal term included);
                    # a symmetric mesh's on its merged faces, in the same order
                    S = K = None
                    gm = mesh.merged() if sym else mesh
                    _, K3 = gf.evaluate_on_device(gm, gm, **dict(gp, early_dot_product=False))
                    D3 = DeviceMatrix(K3.view(nb, n, 3 * n)).matmul_device(X)  # (nb, R, 3n)
                    del K3
                    Dx = D3.view(nb, R, n, 3)[..., 0].contiguous()
                    G2 = torch.empty_like(G)
                    check(lib.cpt_zgemm(ctx.handle, ndof, R, n, 1.0, ptr(W_d), ndof, 0, ptr(Dx), n, n * R, ptr(G2),
                                        ndof, ndof * R, 1, nb, stream_handle(dev)), "cpt_zgemm")
                    Dx_h, G2_h = Dx.cpu().numpy(), G2.cpu().numpy()
            if hasattr(lu, "check_info"):
                lu.check_info()
            G_h = G.cpu().numpy()
            if keep_details:
                X_h, Phi_h = X.cpu().numpy(), Phi.cpu().numpy()
            for b, om in enumerate(oms):
                for r, j in enumerate(omegas[om]):
                    pb = problems[j]
                    w = pb.encounter_omega
                    f = 1j * w * pb.rho * G_h[b, r]
                    if Dx_h is not None:
                        f = f + pb.rho * speed * G2_h[b, r]
                    forces = {d: f[q] for q, d in enumerate(dofs)}
                    if keep_details:
                        pot = Phi_h[b, r]
                        pressure = 1j * w * pb.rho * pot
                        if Dx_h is not None:
                            pressure = pressure + pb.rho * speed * Dx_h[b, r]
                        results[j] = pb.make_results_container(forces, X_h[b, r], pot, pressure)
                    else:
                        results[j] = pb.make_results_container(forces)
            S = K = lu = X = Phi = None  # free this chunk's matrices before the next one's
        return results


    # ------------------------------------------------------------------
    # Post-processing of a solved problem (solver.py:426-561)
    @staticmethod
    def _check_sources(result):
        if result.sources is None:
            raise Exception(f"The values of the sources of {result} cannot been found.\n"
                            "They probably have not been stored by the solver because the option keep_details=True "
                            "have not been set or the direct method has been used.\n"
                            "Please re-run the resolution with the indirect method and keep_details=True.")
        if float(result.encounter_omega) in (0.0, np.inf):
            raise NotImplementedError("Post-processing at zero or infinite frequency is not supported.")

    def _field(self, points, result, gradient):
        """Sum over the panels of the sources times S (gradient False) or
        times the gradient of the Green function (gradient True), at the
        points: (N,) or (N, 3)."""
        gf_params = dict(free_surface=result.free_surface, water_depth=result.water_depth,
                         wavenumber=result.encounter_wavenumber)
        mesh = result.body.mesh_including_lid
        if hasattr(self.engine, "field_at_points"):
            return self.engine.field_at_points(points, mesh, result.sources, gradient=gradient, **gf_params)
        if gradient:  # an engine of the reference's interface
            gradG = self.engine.build_fullK_matrix(points, mesh, **gf_params)
            return np.stack([gradG[c] @ result.sources for c in range(3)], axis=-1)
        return self.engine.build_S_matrix(points, mesh, **gf_params) @ result.sources

    def compute_potential(self, points, result):
        """Potential at points (a point, a list of points, a meshgrid or a
        mesh) of a problem solved with keep_details (solver.py:426-456)."""
        points, output_shape = _normalize_points(points)
        self._check_sources(result)
        with self.timer("Post-processing potential"):
            potential = self._field(points, result, gradient=False)
        return potential.reshape(output_shape)

    def _compute_potential_gradient(self, points, result):
        """solver.py:458-479; a mesh keeps its normals for the diagonal term."""
        points, output_shape = _normalize_points(points, keep_mesh=True)
        self._check_sources(result)
        with self.timer("Post-processing velocity"):
            velocities = self._field(points, result, gradient=True)
        return velocities.reshape((*output_shape, 3))

    def compute_velocity(self, points, result):
        """Velocity at points, shape (*output_shape, 3) (solver.py:481-503)."""
        nabla_phi = self._compute_potential_gradient(points, result)
        if result.forward_speed != 0.0:
            nabla_phi[..., 0] -= result.forward_speed
        return nabla_phi

    def compute_pressure(self, points, result):
        """Pressure i omega rho phi at points (solver.py:505-530)."""
        if result.forward_speed != 0:
            pressure = 1j * result.encounter_omega * result.rho * self.compute_potential(points, result)
            nabla_phi = self._compute_potential_gradient(points, result)
            pressure += result.rho * result.forward_speed * nabla_phi[..., 0]
        else:
            pressure = 1j * result.omega * result.rho * self.compute_potential(points, result)
        return pressure

    def compute_free_surface_elevation(self, points, result):
        """Free surface elevation i omega / g phi at points given by x, y (z = 0)
        or x, y, z (solver.py:533-561)."""
        points, output_shape = _normalize_free_surface_points(points)
        if result.forward_speed != 0:
            fs_elevation = -1 / result.g * (-1j * result.encounter_omega) * self.compute_potential(points, result)
            nabla_phi = self._compute_potential_gradient(points, result)
            fs_elevation += -1 / result.g * result.forward_speed * nabla_phi[..., 0]
        else:
            fs_elevation = -1 / result.g * (-1j * result.omega) * self.compute_potential(points, result)
        return fs_elevation.reshape(output_shape)


def problems_for_body(body, omegas, *, headings=(), rho=1000.0, g=9.81, water_depth=np.inf, dofs=None):
    """Radiation problems for every dof plus diffraction problems for every
    heading (radians), for every frequency: the problem set of a
    fill_dataset call over (omega, radiating_dof, wave_direction)."""
    out = []
    dofs = list(body.dofs) if dofs is None else list(dofs)
    for om in omegas:
        for d in dofs:
            out.append(RadiationProblem(body=body, radiating_dof=d, omega=om, rho=rho, g=g, water_depth=water_depth))
        for beta in headings:
            out.append(DiffractionProblem(body=body, omega=om, wave_direction=beta, rho=rho, g=g,
                                          water_depth=water_depth))
    return out
