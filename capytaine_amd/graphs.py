"""Per-frequency-batch HIP graphs of the FP64 sweep step.

``SweepGraph`` captures once, in a HIP graph, the whole device job of a
batch of frequencies of one body -- the assembly of S and K for every
wavenumber (cpt_build_matrices_dev, which reads the wavenumbers and the
Prony coefficients from device arrays), the batched LU of K in place
(cpt_lu_factor), the batched solve for all right-hand sides (cpt_lu_solve),
the potentials S @ sources and the generalised forces (cpt_zgemm) -- and
replays it for new wavenumbers and boundary conditions: the same results as
``FrequencySweep.run`` (solver.py:268-372 batched), with one graph launch
in place of the several hundred kernel launches of the blocked LU and its
solves.  This is what pays on small bodies, whose kernels are shorter than
their launches.

The wavenumbers of a replay must be finite and positive (the singularity
variant of the capture then holds for all of them); in finite depth each
replay brings its Prony decomposition, padded with zero terms to
``NEXP_MAX``.  Buffers, workspaces and the graph stay allocated for the
object's lifetime.
"""
import numpy as np

from . import _lib
from ._lib import check, ptr, stream_handle

# the Prony fit returns at most 30 exponentials (delhommeau.py:313,
# n_exp_range=range(4, 31, 2))
NEXP_MAX = 30


class SweepGraph:
    def __init__(self, green_function, body, nb, nrhs, *, free_surface=0.0, water_depth=np.inf):
        from .green_functions import require_delhommeau
        require_delhommeau(green_function, "SweepGraph")
        import torch
        from .green_functions import _SING
        if green_function.floating_point_precision != "float64":
            raise NotImplementedError("SweepGraph runs the FP64 path")
        self.gf = green_function
        self.body = body
        self.mesh = body.mesh_including_lid
        self.dev = green_function.device
        self.nb, self.nrhs = int(nb), int(nrhs)
        self.n = n = self.mesh.nb_faces
        self.free_surface = float(free_surface)
        self.water_depth = float(water_depth)
        self.finite = np.isfinite(self.water_depth)
        c = dict(dtype=torch.complex128, device=self.dev)
        self.S = torch.empty((nb, n, n), **c)
        self.K = torch.empty((nb, n, n), **c)
        self.bc = torch.zeros((nb, nrhs, n), **c)
        self.X = torch.empty((nb, nrhs, n), **c)
        self.Phi = torch.empty((nb, nrhs, n), **c)
        self.dofs = list(body.dofs)
        W = -(body.dof_normal_matrix(including_lid=True) * self.mesh.faces_areas[None, :]).T  # (n, ndof)
        self.W = torch.as_tensor(np.ascontiguousarray(W), **c)
        self.G = torch.empty((nb, nrhs, len(self.dofs)), **c)
        self.ipiv = torch.empty((nb, n), dtype=torch.int32, device=self.dev)
        self.info = torch.empty((nb,), dtype=torch.int32, device=self.dev)
        self.ks_d = torch.ones((nb,), dtype=torch.float64, device=self.dev)
        gfs = _SING[self.gf._gf_singularities_for(self.water_depth, 1.0, self.mesh)]
        self.gfs_d = torch.full((nb,), gfs, dtype=torch.int32, device=self.dev)
        self.nexp = NEXP_MAX if self.finite else 0
        self.prony_d = torch.zeros((nb, 2, max(self.nexp, 1)), dtype=torch.float64, device=self.dev)
        self.md = self.mesh.device_arrays(self.dev)
        self.ctx = self.gf._context()  # table uploaded before any capture
        if self.finite:
            self._set_prony(np.ones(nb))
        # one eager step on a side stream reserves every workspace, then the
        # capture (torch's capture stream is current inside the block)
        side = torch.cuda.Stream(self.dev)
        side.wait_stream(torch.cuda.current_stream(self.dev))
        with torch.cuda.stream(side):
            self._step()
        torch.cuda.current_stream(self.dev).wait_stream(side)
        torch.cuda.synchronize(self.dev)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._step()
        self.replays = 0

    def _set_prony(self, ks):
        import torch
        pr = np.zeros((self.nb, 2, self.nexp))
        for b, k in enumerate(ks):
            d = self.gf.find_best_exponential_decomposition(float(k) * self.water_depth)
            if d.shape[1] > self.nexp:
                raise ValueError(f"Prony decomposition of {d.shape[1]} terms > {self.nexp}")
            pr[b, :, :d.shape[1]] = d  # missing terms have a = 0
        self.prony_d.copy_(torch.as_tensor(pr))

    def _step(self):
        """The captured work, on the current stream."""
        lib = _lib.load()
        st = stream_handle(self.dev)
        n, r, nb, nd, md = self.n, self.nrhs, self.nb, len(self.dofs), self.md
        h = self.ctx.handle
        check(lib.cpt_build_matrices_dev(
            h, n, ptr(md["centers"]), ptr(md["normals"]), self.mesh.nb_vertices, ptr(md["vertices"]), n,
            ptr(md["faces"]), ptr(md["centers"]), ptr(md["normals"]), ptr(md["areas"]), ptr(md["radii"]), md["nq"],
            ptr(md["quad_points"]), ptr(md["quad_weights"]), nb, ptr(self.ks_d), ptr(self.gfs_d),
            self.water_depth, self.gf.finite_depth_method_index, self.nexp, ptr(self.prony_d), 1, 1, 1,
            self.free_surface, ptr(self.S), n, n * n, ptr(self.K), n, n * n, st), "cpt_build_matrices_dev")
        self.X.copy_(self.bc)
        check(lib.cpt_lu_factor(h, ptr(self.K), n, n, n * n, ptr(self.ipiv), n, ptr(self.info), nb, st),
              "cpt_lu_factor")
        check(lib.cpt_lu_solve(h, ptr(self.K), n, n, n * n, ptr(self.ipiv), n, ptr(self.X), r, n, n * r, nb, st),
              "cpt_lu_solve")
        check(lib.cpt_zgemm(h, n, r, n, 1.0, ptr(self.S), n, n * n, ptr(self.X), n, n * r, ptr(self.Phi), n,
                            n * r, 1, nb, st), "cpt_zgemm S@sources")
        check(lib.cpt_zgemm(h, nd, r, n, 1.0, ptr(self.W), nd, 0, ptr(self.Phi), n, n * r, ptr(self.G), nd,
                            nd * r, 1, nb, st), "cpt_zgemm forces")

    def run(self, wavenumbers, bc):
        """Replay for `nb` wavenumbers (finite, positive) and boundary
        conditions bc ((nb, nrhs, n), numpy or device tensor); returns the
        device tensor of generalised forces (nb, nrhs, ndof), W^T Phi, to be
        scaled by i omega rho (``forces``)."""
        import torch
        ks = np.asarray(wavenumbers, dtype=np.float64)
        if ks.shape != (self.nb,) or not np.all(np.isfinite(ks) & (ks > 0)):
            raise ValueError(f"SweepGraph replays {self.nb} finite positive wavenumbers, got {ks}")
        self.ks_d.copy_(torch.as_tensor(ks))
        if self.finite:
            self._set_prony(ks)
        self.bc.copy_(torch.as_tensor(bc).to(self.dev, dtype=torch.complex128))
        self.graph.replay()
        self.replays += 1
        return self.G

    def check_info(self):
        """LAPACK's info of the last replay's factorisations (waits)."""
        return self.info.cpu().numpy()

    def forces(self, omegas, rho):
        G = self.G.cpu().numpy()
        return 1j * np.asarray(omegas)[:, None, None] * rho * G
