"""Device-resident frequency sweep: the whole per-body BEM job on the GPU.

For one body and a list of wavenumbers, with the boundary conditions of all
problems already on the device, ``FrequencySweep.run`` does what capytaine's
``BEMSolver.solve_all`` does for the same problems (solver.py:268-372 ->
_solve :145-215 per problem), batched:

  1. one assembly launch for every wavenumber (cpt_build_matrices),
  2. one batched LU of the K matrices, in place (cpt_lu_factor), or their
     complex64 factors (cpt_mixed_factor) in the mixed form,
  3. one batched solve for all right-hand sides (cpt_lu_solve), or the FP64
     refinement with those factors (cpt_mixed_refine),
  4. potentials Phi = S @ sources for all of them (cpt_zgemm),
  5. generalised forces W^T Phi with W = -(dof . n) A (cpt_zgemm), to be
     scaled by i omega rho per frequency (bodies.py:238-250).

Buffers are allocated once and reused across calls.
"""
import numpy as np

from . import _lib
from ._lib import check, ptr, stream_handle


class FrequencySweep:
    def __init__(self, green_function, body, wavenumbers, nrhs, *, free_surface=0.0, water_depth=np.inf,
                 precision="fp64", max_iter=10, tol=1e-12, far_field_fp32=None):
        from .green_functions import require_delhommeau
        require_delhommeau(green_function, "FrequencySweep")
        import torch
        self.gf = green_function
        self.body = body
        self.mesh = body.mesh_including_lid
        self.dev = green_function.device
        self.ks = np.asarray(wavenumbers, dtype=np.float64)
        self.nb = len(self.ks)
        self.n = self.mesh.nb_faces
        self.nrhs = nrhs
        self.free_surface = free_surface
        self.water_depth = water_depth
        if precision not in ("fp64", "mixed"):
            raise ValueError(f"precision must be 'fp64' or 'mixed', got {precision!r}")
        self.precision = precision
        # mixed form: far pairs read the float copy of the table (FP64
        # arithmetic; coefficients within ~1e-7 of the FP64 assembly) unless
        # far_field_fp32 is False
        self.far_field_fp32 = precision == "mixed" and far_field_fp32 is not False
        self.max_iter = max_iter
        self.tol = tol
        self.iterations = 0
        self.residual = 0.0
        n, nb = self.n, self.nb
        cplx = torch.complex128
        self.S = torch.empty((nb, n, n), dtype=cplx, device=self.dev)
        self.K = torch.empty((nb, n, n), dtype=cplx, device=self.dev)
        self.X = torch.empty((nb, nrhs, n), dtype=cplx, device=self.dev)
        self.Phi = torch.empty((nb, nrhs, n), dtype=cplx, device=self.dev)
        self.ipiv = torch.empty((nb, n), dtype=torch.int32, device=self.dev)
        self.info = torch.empty((nb,), dtype=torch.int32, device=self.dev)
        dofs = list(body.dofs)
        self.dofs = dofs
        W = -(body.dof_normal_matrix(including_lid=True) * self.mesh.faces_areas[None, :]).T  # (n, ndof)
        self.W = torch.as_tensor(np.ascontiguousarray(W), dtype=cplx).to(self.dev)
        self.G = torch.empty((nb, nrhs, len(dofs)), dtype=cplx, device=self.dev)
        self.ctx = _lib.Context.get(self.dev)
        self.stream = stream_handle(self.dev)
        # warm the mesh and table uploads
        self.mesh.device_arrays(self.dev)
        self.gf._context()

    # individual phases, so that callers can time them
    def assemble(self):
        lib = _lib.load()
        check(lib.cpt_set_assembly_precision(self.ctx.handle, int(self.far_field_fp32)), "assembly precision")
        try:
            self.gf.evaluate_on_device(self.mesh, self.mesh, free_surface=self.free_surface,
                                       water_depth=self.water_depth, wavenumber=self.ks, adjoint_double_layer=True,
                                       diagonal_term_in_double_layer=True, out=(self.S, self.K))
        finally:
            check(lib.cpt_set_assembly_precision(self.ctx.handle, 0), "assembly precision")

    def factor(self):
        """FP64: LU of K in place.  Mixed: complex64 factor of K
        (cpt_mixed_factor) into a buffer of its own; K is kept for the FP64
        residuals of the refinement."""
        if self.precision == "mixed":
            from .engines import MixedFactor
            self.mixed_factor = MixedFactor(self.K, stream=self.stream)
            return
        n = self.n
        check(_lib.load().cpt_lu_factor(self.ctx.handle, ptr(self.K), n, n, n * n, ptr(self.ipiv), n,
                                        ptr(self.info), self.nb, self.stream), "cpt_lu_factor")

    def solve(self, bc):
        """bc: device tensor (nb, nrhs, n) of boundary conditions (not modified)."""
        n = self.n
        lib = _lib.load()
        if self.precision == "mixed":
            try:
                self.mixed_factor.refine(bc, self.X, self.max_iter, self.tol, stream=self.stream)
            finally:
                self.iterations, self.residual = self.mixed_factor.iterations, self.mixed_factor.residual
            return
        self.X.copy_(bc)
        _lib.call_with_memory_retry(lambda: lib.cpt_lu_solve(
            self.ctx.handle, ptr(self.K), n, n, n * n, ptr(self.ipiv), n, ptr(self.X), self.nrhs, n, n * self.nrhs,
            self.nb, self.stream), "cpt_lu_solve")

    def post(self):
        n, r, nd = self.n, self.nrhs, len(self.dofs)
        lib = _lib.load()
        check(lib.cpt_zgemm(self.ctx.handle, n, r, n, 1.0, ptr(self.S), n, n * n, ptr(self.X), n, n * r,
                            ptr(self.Phi), n, n * r, 1, self.nb, self.stream), "cpt_zgemm S@sources")
        check(lib.cpt_zgemm(self.ctx.handle, nd, r, n, 1.0, ptr(self.W), nd, 0, ptr(self.Phi), n, n * r,
                            ptr(self.G), nd, nd * r, 1, self.nb, self.stream), "cpt_zgemm forces")

    def run(self, bc):
        self.assemble()
        self.factor()
        self.solve(bc)
        self.post()
        return self.G

    def forces(self, omegas, rho):
        """Host copy of the forces (nb, nrhs, ndof), scaled by i omega rho."""
        G = self.G.cpu().numpy()
        return 1j * np.asarray(omegas)[:, None, None] * rho * G


def boundary_conditions(problems_by_freq, nrhs, n, device):
    """(nb, nrhs, n) device tensor from lists of problems per frequency."""
    import torch
    bc = np.zeros((len(problems_by_freq), nrhs, n), dtype=np.complex128)
    for b, pbs in enumerate(problems_by_freq):
        for r, pb in enumerate(pbs):
            bc[b, r] = pb.boundary_condition
    return torch.as_tensor(bc).to(device)
