"""Batched BEM over many bodies of uneven panel counts and their frequencies.

`BatchedSweep` does, for a list of bodies each with its own wavenumbers and
right-hand sides, what capytaine's ``BEMSolver.solve_all`` does for the same
problems (solver.py:268-372 -> _solve :145-215 per problem), on the device:

  1. one persistent assembly launch over every (body, wavenumber) pair
     (cpt_build_matrices_batched: a work-list of (body, row block, panel
     chunk) tiles, most expensive first), far pairs of the mixed form from the
     float table;
  2. a variable-size batched LU: one group per body (its matrices share the
     body's order), the groups running concurrently on the library's lane
     streams -- complex128 (cpt_lu_factor_grouped + cpt_lu_solve_grouped) or
     complex64 with FP64 refinement (cpt_lu_solve_mixed_grouped);
  3. potentials S @ sources and generalised forces per body (cpt_zgemm).

Each factorisation is computed once and serves all right-hand sides of its
(body, wavenumber).  Infinite depth, indirect method, adjoint double layer.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import CptBody, CptLuGroup, call_with_memory_retry, check, ptr, stream_handle
from .green_functions import _SING


def footprint_bytes(n, nb_wavenumbers, nrhs, precision):
    """Device bytes of one body's matrices in a BatchedSweep."""
    per = 2 * 16 * n * n  # S and K, complex128
    if precision == "mixed":
        per += 8 * n * n + 16 * n * nrhs  # complex64 copy of K, FP64 residuals
    return nb_wavenumbers * (per + 3 * 16 * n * nrhs)


def chunks_by_memory(sizes, budget):
    """Split indices 0..len(sizes)-1 into consecutive chunks whose summed
    `sizes` stay under `budget` (a single oversized item forms its own
    chunk)."""
    out, cur, tot = [], [], 0
    for i, s in enumerate(sizes):
        if cur and tot + s > budget:
            out.append(cur)
            cur, tot = [], 0
        cur.append(i)
        tot += s
    if cur:
        out.append(cur)
    return out


def _arena(shapes, dtype, dev):
    """Views of the given shapes into one device allocation, each starting
    on a 256-byte boundary."""
    import torch
    item = torch.empty((), dtype=dtype).element_size()
    align = max(1, 256 // item)
    sizes = [int(np.prod(sh)) for sh in shapes]
    offs, total = [], 0
    for n in sizes:
        offs.append(total)
        total += -(-n // align) * align
    buf = torch.empty(max(total, 1), dtype=dtype, device=dev)
    return [buf[o:o + n].view(sh) for o, n, sh in zip(offs, sizes, shapes)]


class BatchedSweep:
    """Device-resident solve of many bodies, each at its own wavenumbers.

    bodies: FloatingBody list; wavenumbers: one array per body (or one array
    for all); nrhs: right-hand sides per (body, wavenumber), one int per body
    or one for all.
    """

    def __init__(self, green_function, bodies, wavenumbers, nrhs, *, precision="mixed", max_iter=10, tol=1e-12,
                 free_surface=0.0, context=None, far_field_fp32=None):
        from .green_functions import require_delhommeau
        require_delhommeau(green_function, "BatchedSweep")
        import torch
        if precision not in ("fp64", "mixed"):
            raise ValueError(f"precision must be 'fp64' or 'mixed', got {precision!r}")
        self.gf = green_function
        self.dev = green_function.device
        self.bodies = list(bodies)
        nbody = len(self.bodies)
        if len(wavenumbers) and np.ndim(wavenumbers[0]) == 0:
            wavenumbers = [wavenumbers] * nbody
        self.ks = [np.asarray(k, dtype=np.float64) for k in wavenumbers]
        self.nrhs = [int(nrhs)] * nbody if np.ndim(nrhs) == 0 else [int(r) for r in nrhs]
        if len(self.ks) != nbody or len(self.nrhs) != nbody:
            raise ValueError("one wavenumber list and one nrhs per body")
        self.precision = precision
        # mixed form: far pairs read the float copy of the table (FP64
        # arithmetic; coefficients within ~1e-7 of the FP64 assembly) unless
        # far_field_fp32 is False
        self.far_field_fp32 = precision == "mixed" and far_field_fp32 is not False
        self.max_iter, self.tol = max_iter, tol
        self.free_surface = free_surface
        self.meshes = [b.mesh_including_lid for b in self.bodies]
        self.n = [m.nb_faces for m in self.meshes]
        cplx = torch.complex128
        dev = self.dev
        # each kind of array in one allocation, the bodies' arrays as views
        # into it: one contiguous range per kind, whatever the allocator's
        # cache holds (separate allocations per body measured up to 2.5x
        # slower solves when they came back scattered from the cache)
        self.S = _arena([(len(k), n, n) for k, n in zip(self.ks, self.n)], cplx, dev)
        self.K = _arena([(len(k), n, n) for k, n in zip(self.ks, self.n)], cplx, dev)
        self.X = _arena([(len(k), r, n) for k, r, n in zip(self.ks, self.nrhs, self.n)], cplx, dev)
        self.Phi = _arena([tuple(x.shape) for x in self.X], cplx, dev)
        if precision == "fp64":
            self.ipiv = _arena([(len(k), n) for k, n in zip(self.ks, self.n)], torch.int32, dev)
            self.info = _arena([(len(k),) for k in self.ks], torch.int32, dev)
        self.W, self.G = [], []
        for b, mesh, k, r in zip(self.bodies, self.meshes, self.ks, self.nrhs):
            W = -(b.dof_normal_matrix(including_lid=True) * mesh.faces_areas[None, :]).T
            self.W.append(torch.as_tensor(np.ascontiguousarray(W), dtype=cplx).to(dev))
            self.G.append(torch.empty((len(k), r, W.shape[1]), dtype=cplx, device=dev))
        # a private context (own workspace) lets several sweeps run at once
        # on different streams; the work goes to the current stream
        self.ctx = self.gf._context(context)  # with the table uploaded
        self.stream = stream_handle(dev)
        self.iterations = np.zeros(nbody, dtype=np.int32)
        self.residual = np.zeros(nbody)
        self._gfs = [np.array([_SING[self.gf._gf_singularities_for(np.inf, float(kk), m)] for kk in k],
                              dtype=np.int32) for k, m in zip(self.ks, self.meshes)]
        self._bodies_c = (CptBody * nbody)()
        self._mesh_arrays = [mesh.device_arrays(dev) for mesh in self.meshes]  # kept alive with the sweep
        for i, (mesh, k) in enumerate(zip(self.meshes, self.ks)):
            md = self._mesh_arrays[i]
            n = self.n[i]
            self._bodies_c[i] = CptBody(
                nb_faces=n, nb_vertices=mesh.nb_vertices, nb_quad_points=md["nq"], nb_wavenumbers=len(k),
                vertices=md["vertices"].data_ptr(), faces=md["faces"].data_ptr(),
                centers=md["centers"].data_ptr(), normals=md["normals"].data_ptr(),
                areas=md["areas"].data_ptr(), radiuses=md["radii"].data_ptr(),
                quad_points=md["quad_points"].data_ptr(), quad_weights=md["quad_weights"].data_ptr(),
                wavenumbers=k.ctypes.data, gf_singularities=self._gfs[i].ctypes.data,
                S=self.S[i].data_ptr(), ldS=n, strideS=n * n, K=self.K[i].data_ptr(), ldK=n, strideK=n * n)

    @property
    def pairs(self):
        """Panel pairs of one assembly: sum over (body, wavenumber) of n^2."""
        return int(sum(len(k) * n * n for k, n in zip(self.ks, self.n)))

    @property
    def problems(self):
        return int(sum(len(k) * r for k, r in zip(self.ks, self.nrhs)))

    def assemble(self):
        lib = _lib.load()
        check(lib.cpt_set_assembly_precision(self.ctx.handle, int(self.far_field_fp32)), "assembly precision")
        try:
            call_with_memory_retry(lambda: lib.cpt_build_matrices_batched(
                self.ctx.handle, len(self.bodies), self._bodies_c, 1, 1, float(self.free_surface), self.stream),
                "cpt_build_matrices_batched")
        finally:
            check(lib.cpt_set_assembly_precision(self.ctx.handle, 0), "assembly precision")

    def _groups(self, bcs):
        g = (CptLuGroup * len(self.bodies))()
        for i, (n, k, r) in enumerate(zip(self.n, self.ks, self.nrhs)):
            nb = len(k)
            fields = dict(n=n, batch=nb, A=self.K[i].data_ptr(), lda=n, strideA=n * n, nrhs=r, ldb=n,
                          strideB=n * r, ldx=n, strideX=n * r)
            if self.precision == "fp64":
                fields.update(ipiv=self.ipiv[i].data_ptr(), strideP=n, info=self.info[i].data_ptr(),
                              B=self.X[i].data_ptr())
            else:
                fields.update(B=bcs[i].data_ptr(), X=self.X[i].data_ptr())
            g[i] = CptLuGroup(**fields)
        return g

    def solve(self, bcs):
        """bcs: one device tensor (nb, nrhs, n) of boundary conditions per body
        (not modified).  FP64 factors K in place."""
        lib = _lib.load()
        nbody = len(self.bodies)
        if self.precision == "fp64":
            for x, bc in zip(self.X, bcs):
                x.copy_(bc)
            g = self._groups(bcs)
            call_with_memory_retry(lambda: lib.cpt_lu_factor_grouped(self.ctx.handle, nbody, g, self.stream),
                                   "cpt_lu_factor_grouped")
            call_with_memory_retry(lambda: lib.cpt_lu_solve_grouped(self.ctx.handle, nbody, g, self.stream),
                                   "cpt_lu_solve_grouped")
            return
        import torch
        g = self._groups(bcs)
        its = np.zeros(nbody, dtype=np.int32)
        res = np.zeros(nbody)
        # the complex64 factors live in a torch buffer on this stream
        nbytes = ctypes.c_size_t()
        check(lib.cpt_mixed_workspace_size(self.ctx.handle, nbody, g, ctypes.byref(nbytes)), "workspace size")
        work = torch.empty(nbytes.value, dtype=torch.uint8, device=self.dev)
        check(lib.cpt_set_workspace(self.ctx.handle, work.data_ptr(), nbytes.value), "cpt_set_workspace")
        try:
            call_with_memory_retry(lambda: lib.cpt_lu_solve_mixed_grouped(
                self.ctx.handle, nbody, g, self.max_iter, self.tol, ptr(its), ptr(res), self.stream),
                "cpt_lu_solve_mixed_grouped")
        finally:
            check(lib.cpt_set_workspace(self.ctx.handle, None, 0), "cpt_set_workspace")
            self.iterations, self.residual = its, res

    def post(self):
        lib = _lib.load()
        for i, (n, k, r) in enumerate(zip(self.n, self.ks, self.nrhs)):
            nb, nd = len(k), self.W[i].shape[1]
            check(lib.cpt_zgemm(self.ctx.handle, n, r, n, 1.0, ptr(self.S[i]), n, n * n, ptr(self.X[i]), n, n * r,
                                ptr(self.Phi[i]), n, n * r, 1, nb, self.stream), "cpt_zgemm S@sources")
            check(lib.cpt_zgemm(self.ctx.handle, nd, r, n, 1.0, ptr(self.W[i]), nd, 0, ptr(self.Phi[i]), n, n * r,
                                ptr(self.G[i]), nd, nd * r, 1, nb, self.stream), "cpt_zgemm forces")

    def run(self, bcs):
        self.assemble()
        self.solve(bcs)
        self.post()
        return self.G

    def forces(self, omegas, rho):
        """Host forces per body, (nb, nrhs, ndof) each, scaled by i omega rho."""
        return [1j * np.asarray(om)[:, None, None] * rho * G.cpu().numpy() for om, G in zip(omegas, self.G)]
