"""Matrix engine keeping the influence matrices and their LU factors on the GPU.

Drop-in for ``capytaine.bem.engines.DefaultMatrixEngine``
(src/capytaine/bem/engines.py:95-398): ``build_matrices`` with the
last-computed cache keyed by (mesh1, mesh2, gf_params) (:258-298),
``linear_solver(A, b)`` with the LU factorisation cached in place of K
(:300-346), ``build_S_matrix`` / ``build_fullK_matrix`` (:142-176),
``compute_ram_estimation`` (:348-398) and the NaN check (:400-405), plus
``field_at_points``: the post-processing products S @ sources and
grad G @ sources at points, evaluated chunk by chunk on the device.  The
matrices it returns are :class:`DeviceMatrix` objects: they expose
``shape``, ``dtype``, ``@`` and ``__array__`` (a Fortran-ordered numpy copy),
which is all the solver needs from them.
"""
import logging

import numpy as np

from . import _lib
from ._lib import check, ptr, stream_handle
from .green_functions import Delhommeau, GreenFunctionEvaluationError

LOG = logging.getLogger(__name__)


class DeviceMatrix:
    """Dense complex128 (or complex64, for the single-precision Green
    function) matrix or batch of matrices on the GPU.

    `data` is a torch complex tensor of shape (nb, cols, rows) (or
    (cols, rows)): each matrix is stored column-major, leading dimension
    `rows`, like LAPACK and the reference's Fortran arrays.
    """

    def __init__(self, data, device=None):
        self.data = data if data.dim() == 3 else data.unsqueeze(0)
        self.device = self.data.device if device is None else device

    @property
    def batch(self):
        return self.data.shape[0]

    @property
    def shape(self):
        if self.batch == 1:
            return (self.data.shape[2], self.data.shape[1])
        return (self.batch, self.data.shape[2], self.data.shape[1])

    @property
    def single(self):
        import torch
        return self.data.dtype == torch.complex64

    @property
    def dtype(self):
        return np.dtype(np.complex64 if self.single else np.complex128)

    @property
    def torch_dtype(self):
        return self.data.dtype

    ndim = property(lambda self: len(self.shape))

    def __array__(self, dtype=None, copy=None):
        a = self.data.cpu().numpy().transpose(0, 2, 1)
        if self.batch == 1:
            a = a[0]
        return a if dtype is None else a.astype(dtype)

    def numpy(self):
        return self.__array__()

    def matmul_device(self, x):
        """self @ x for a device tensor x of shape (nb, nrhs, cols)
        (column-major (cols, nrhs) blocks); returns (nb, nrhs, rows)."""
        import torch
        nb, cols, rows = self.data.shape
        nrhs = x.shape[1]
        if x.dtype != self.data.dtype:
            raise ValueError(f"matmul_device: {x.dtype} operand for a {self.data.dtype} matrix")
        y = torch.empty((nb, nrhs, rows), dtype=self.data.dtype, device=self.data.device)
        ctx = _lib.Context.get(self.data.device)
        name = "cpt_cgemm" if self.single else "cpt_zgemm"
        check(getattr(_lib.load(), name)(ctx.handle, rows, nrhs, cols, 1.0, ptr(self.data), rows, rows * cols, ptr(x),
                                         cols, cols * nrhs, ptr(y), rows, rows * nrhs, 1, nb,
                                         stream_handle(self.data.device)), name)
        return y

    def __matmul__(self, x):
        import torch
        x = np.asarray(x)
        vec = x.ndim == 1
        xb = x.reshape(x.shape[0], -1) if self.batch == 1 else x
        if self.batch == 1:
            xd = torch.as_tensor(np.ascontiguousarray(xb.T), dtype=self.data.dtype).to(self.data.device)[None]
        else:
            xd = torch.as_tensor(np.ascontiguousarray(np.swapaxes(xb, -1, -2)),
                                 dtype=self.data.dtype).to(self.data.device)
        y = self.matmul_device(xd).cpu().numpy()
        y = np.swapaxes(y, -1, -2)
        if self.batch == 1:
            y = y[0]
            return y[:, 0] if vec else y
        return y


class LUDecomposedDeviceMatrix:
    """LU factors P A = L U of a batch of DeviceMatrix, computed by
    cpt_lu_factor (cpt_clu_factor for complex64 matrices, as LAPACK's cgetrf
    serves scipy for them); `solve` replaces LUDecomposedMatrix.solve
    (tools/block_circulant_matrices.py:374-385)."""

    def __init__(self, A: DeviceMatrix, overwrite_a=False):
        import torch
        data = A.data if overwrite_a else A.data.clone()
        nb, cols, rows = data.shape
        if rows != cols:
            raise ValueError("LU decomposition of a non-square matrix")
        self.n = rows
        self.lu = DeviceMatrix(data)
        self.ipiv = torch.empty((nb, rows), dtype=torch.int32, device=data.device)
        self.info = torch.empty((nb,), dtype=torch.int32, device=data.device)
        ctx = _lib.Context.get(data.device)
        name = "cpt_clu_factor" if self.lu.single else "cpt_lu_factor"
        check(getattr(_lib.load(), name)(ctx.handle, ptr(data), rows, rows, rows * rows, ptr(self.ipiv), rows,
                                         ptr(self.info), nb, stream_handle(data.device)), name)
        self._checked = False

    shape = property(lambda self: self.lu.shape)
    dtype = property(lambda self: self.lu.dtype)

    def check_info(self):
        if not self._checked:
            info = self.info.cpu().numpy()
            if np.any(info != 0):
                LOG.warning("LU factorisation: exactly singular matrix (info=%s)", info.tolist())
            self._checked = True

    def solve_device(self, b):
        """Solve in place for a device tensor b of shape (nb, nrhs, n)."""
        nb, nrhs, n = b.shape
        if b.dtype != self.lu.data.dtype:
            raise ValueError(f"solve_device: {b.dtype} right-hand sides for a {self.lu.data.dtype} factor")
        ctx = _lib.Context.get(b.device)
        name = "cpt_clu_solve" if self.lu.single else "cpt_lu_solve"
        _lib.call_with_memory_retry(lambda: getattr(_lib.load(), name)(
            ctx.handle, ptr(self.lu.data), n, n, n * n, ptr(self.ipiv), n, ptr(b), nrhs, n, n * nrhs, nb,
            stream_handle(b.device)), name)
        return b

    def solve(self, b):
        import torch
        b = np.asarray(b)
        vec = b.ndim == 1
        bb = b.reshape(b.shape[0], -1) if self.lu.batch == 1 else b
        if self.lu.batch == 1:
            bd = torch.as_tensor(np.ascontiguousarray(bb.T), dtype=self.lu.torch_dtype).to(self.lu.device)[None]
        else:
            bd = torch.as_tensor(np.ascontiguousarray(np.swapaxes(bb, -1, -2)),
                                 dtype=self.lu.torch_dtype).to(self.lu.device)
        x = self.solve_device(bd).cpu().numpy()
        self.check_info()
        x = np.swapaxes(x, -1, -2)
        if self.lu.batch == 1:
            x = x[0]
            return x[:, 0] if vec else x
        return x


class MixedFactor:
    """Complex64 LU factor of a batch of complex128 matrices K (torch tensor
    (nb, n, n), column-major blocks), in a caller-owned device buffer
    (cpt_mixed_factor), refined against K by cpt_mixed_refine.  K must stay
    unchanged while the factor is used."""

    def __init__(self, K, stream=None):
        import ctypes
        import torch
        self.K = K
        nb, n, _ = K.shape
        self.n, self.nb = n, nb
        size = ctypes.c_size_t()
        check(_lib.load().cpt_mixed_factor_size(n, nb, ctypes.byref(size)), "cpt_mixed_factor_size")
        # torch's allocator returns 512-byte aligned blocks; the library asks 256
        self.buf = torch.empty(size.value, dtype=torch.uint8, device=K.device)
        self.nbytes = size.value
        self.ctx = _lib.Context.get(K.device)
        stream = stream_handle(K.device) if stream is None else stream
        _lib.call_with_memory_retry(lambda: _lib.load().cpt_mixed_factor(
            self.ctx.handle, ptr(K), n, n, n * n, nb, ptr(self.buf), self.nbytes, stream), "cpt_mixed_factor")

    def refine(self, b, x, max_iter, tol, stream=None):
        """x = K^-1 b for device tensors b, x of shape (nb, nrhs, n); returns
        (sweeps, final relative residual).  Raises ConvergenceError when the
        residual stays above tol."""
        import ctypes
        nb, nrhs, n = b.shape
        it, res = ctypes.c_int(), ctypes.c_double()
        stream = stream_handle(b.device) if stream is None else stream
        try:
            _lib.call_with_memory_retry(lambda: _lib.load().cpt_mixed_refine(
                self.ctx.handle, ptr(self.K), n, n, n * n, nb, ptr(self.buf), self.nbytes, ptr(b), nrhs, n,
                n * nrhs, ptr(x), n, n * nrhs, max_iter, tol, ctypes.byref(it), ctypes.byref(res), stream),
                "cpt_mixed_refine")
        finally:
            self.iterations, self.residual = it.value, res.value
        return it.value, res.value

    def factors(self):
        """Host copies (lu (nb, n, n) complex64 row-major views of the
        column-major factors transposed back, ipiv (nb, n), info (nb)) of the
        buffer's parts, for inspection."""
        import numpy as np
        n, nb = self.n, self.nb
        raw = self.buf.cpu().numpy()
        a256 = lambda x: (x + 255) & ~255
        off = 256
        lu = raw[off:off + 8 * nb * n * n].view(np.complex64).reshape(nb, n, n).transpose(0, 2, 1)
        off += a256(8 * nb * n * n)
        ints = raw[off:off + 4 * (nb * n + nb)].view(np.int32)
        return lu, ints[:nb * n].reshape(nb, n), ints[nb * n:]


class MixedPrecisionSolver:
    """Complex64 LU on FP32 MFMA with FP64 iterative refinement against the
    FP64 matrix.  The factor is computed once, at the first solve, and kept
    for every later right-hand side, as the reference keeps the LU of K
    (engines.py:326-338).  Raises ConvergenceError when the refinement
    stays above tol."""

    def __init__(self, A: DeviceMatrix, max_iter=10, tol=1e-12):
        self.A = A
        self.max_iter = max_iter
        self.tol = tol
        self.iterations = None
        self.residual = None
        self.factor = None
        self.factorisations = 0

    shape = property(lambda self: self.A.shape)
    dtype = property(lambda self: self.A.dtype)

    def solve_device(self, b):
        import torch
        if self.factor is None:
            self.factor = MixedFactor(self.A.data)
            self.factorisations += 1
        x = torch.empty_like(b)
        try:
            self.factor.refine(b, x, self.max_iter, self.tol)
        finally:
            self.iterations, self.residual = self.factor.iterations, self.factor.residual
        return x

    def solve(self, b):
        import torch
        b = np.asarray(b)
        vec = b.ndim == 1
        bb = b.reshape(b.shape[0], -1)
        bd = torch.as_tensor(np.ascontiguousarray(bb.T), dtype=torch.complex128).to(self.A.device)[None]
        x = self.solve_device(bd).cpu().numpy()[0].T
        return x[:, 0] if vec else x


def gmres_device(mats, b, *, rtol=1e-5, atol=1e-6, restart=None, maxiter=None):
    """Device GMRES (cpt_zgmres / cpt_cgmres) for the column-major matrices
    `mats` (torch (nb, n, n)) and right-hand sides b (torch (nb, r, n)):
    returns x (nb, r, n), the Arnoldi steps and final relative residuals
    ((nb, r) numpy).  Raises RuntimeError with the reference's message when
    a right-hand side does not converge (engines.py:84-86)."""
    import torch
    nb, n, _ = mats.shape
    r = b.shape[1]
    if b.dtype != mats.dtype:
        raise ValueError(f"gmres: {b.dtype} right-hand sides for {mats.dtype} matrices")
    single = mats.dtype == torch.complex64
    name = "cpt_cgmres" if single else "cpt_zgmres"
    fn = getattr(_lib.load(), name)
    x = torch.empty_like(b)
    its = np.zeros((nb, r), dtype=np.int32)
    res = np.zeros((nb, r))
    ctx = _lib.Context.get(mats.device)
    atol = np.broadcast_to(np.asarray(atol, dtype=float), (nb,))
    for i in range(nb):
        status = fn(ctx.handle, ptr(mats[i]), n, n, ptr(b[i]), r, n, ptr(x[i]), n, int(restart or 0),
                    int(maxiter or 0), float(rtol), float(atol[i]), its[i].ctypes.data, res[i].ctypes.data,
                    stream_handle(mats.device))
        if status == _lib.CPT_ERR_NOT_CONVERGED:
            raise RuntimeError(f"No convergence of the GMRES after {int(its[i].max())} iterations.\n"
                               "This can be due to overlapping panels or irregular frequencies.")
        check(status, name)
    LOG.info("End of GMRES after %s iterations.", its.tolist())
    return x, its, res


def solve_gmres(A, b):
    """engines.py:74-89 on the device: scipy.sparse.linalg.gmres(A, b,
    atol=1e-6) for each column of b, with its defaults (rtol 1e-5, restart
    20, 10 n restart cycles).  A symmetric mesh's matrix is solved block by
    block in the transformed space, each block to the absolute tolerance
    max(1e-6, 1e-5 ||b||) of the whole system: the transform multiplies
    norms by sqrt(n) and each block's residual is at most that tolerance, so
    the residual of the whole system is too."""
    import torch
    from .symmetry import BlockSymmetricDeviceMatrix, transform_vectors
    b = np.asarray(b)
    vec = b.ndim == 1
    bb = b.reshape(b.shape[0], -1)
    if isinstance(A, BlockSymmetricDeviceMatrix):
        if A.batch != 1:
            raise ValueError("gmres: one matrix at a time")
        bd = torch.as_tensor(np.ascontiguousarray(bb.T), dtype=torch.complex128).to(A.device)[None]
        tau = max(1e-6, 1e-5 * float(np.linalg.norm(bb, axis=0).max()))
        bh = transform_vectors(A.sym, bd)  # (n, r, w): one system per block
        xh, _, _ = gmres_device(A.diagonal_blocks(), bh.contiguous(), rtol=0.0, atol=tau)
        x = transform_vectors(A.sym, xh, inverse=True)[0].cpu().numpy().T
    elif isinstance(A, DeviceMatrix):
        if A.batch != 1:
            raise ValueError("gmres: one matrix at a time")
        bd = torch.as_tensor(np.ascontiguousarray(bb.T), dtype=A.torch_dtype).to(A.device)[None]
        xd, _, _ = gmres_device(A.data, bd)
        x = xd[0].cpu().numpy().T
    else:
        raise TypeError(f"gmres: unsupported matrix {type(A).__name__}")
    return x[:, 0] if vec else x


def has_been_lu_decomposed(A):
    from .symmetry import LUDecomposedSymmetricMatrix
    return isinstance(A, (LUDecomposedDeviceMatrix, LUDecomposedSymmetricMatrix))


def lu_decompose(A, overwrite_a=False):
    from .symmetry import BlockSymmetricDeviceMatrix, LUDecomposedSymmetricMatrix
    if isinstance(A, BlockSymmetricDeviceMatrix):
        return LUDecomposedSymmetricMatrix(A, overwrite_a=overwrite_a)
    if not isinstance(A, DeviceMatrix):
        raise TypeError("The device engine factorises DeviceMatrix objects")
    return LUDecomposedDeviceMatrix(A, overwrite_a=overwrite_a)


def check_if_nan_in_matrix(matrices):
    import torch
    for m in matrices:
        data = m.data if isinstance(m, DeviceMatrix) else getattr(m, "cols", m)
        bad = bool(torch.isnan(torch.view_as_real(data)).any().item()) if hasattr(data, "device") else \
            bool(np.any(np.isnan(data)))
        if bad:
            raise GreenFunctionEvaluationError("Green function returned a NaN in the interaction matrix.\n"
                                               "It could be due to overlapping panels.")


class DefaultMatrixEngine:
    """GPU matrix engine.

    Parameters
    ----------
    green_function: Delhommeau (GPU), default: Delhommeau()
    linear_solver: "lu_decomposition" (default), "lu_decomposition_with_overwrite",
        "mixed_precision" (complex64 LU + FP64 refinement), or a callable
        (A, b) -> x that receives numpy arrays.
    check_nan: scan the matrices for NaN after assembly (engines.py:255).
    """

    def __init__(self, *, green_function=None, linear_solver='lu_decomposition', check_nan=True):
        self.green_function = Delhommeau() if green_function is None else green_function
        self._linear_solver = linear_solver
        self.check_nan = check_nan
        self.last_computed_inputs = None
        self.last_computed_matrices = None
        self.exportable_settings = {
            'engine': 'DefaultMatrixEngine',
            'linear_solver': str(linear_solver),
            **self.green_function.exportable_settings,
        }

    def __str__(self):
        return f"DefaultMatrixEngine(green_function={self.green_function}, linear_solver={self._linear_solver!r})"

    __repr__ = __str__

    @property
    def device(self):
        return self.green_function.device

    def build_S_matrix(self, points, mesh, **gf_params):
        S, _ = self.green_function.evaluate(points, mesh.merged(), **gf_params)
        check_if_nan_in_matrix([S])
        return S

    def build_fullK_matrix(self, mesh1, mesh2, **gf_params):
        gf_params.setdefault("diagonal_term_in_double_layer", True)
        gf_params.setdefault("adjoint_double_layer", True)
        gf_params.setdefault("early_dot_product", False)
        _, fullK = self.green_function.evaluate(mesh1, mesh2.merged(), **gf_params)
        check_if_nan_in_matrix([fullK])
        return fullK

    def field_at_points(self, points, mesh, sources, *, gradient=False, max_chunk_bytes=None, **gf_params):
        """The products build_S_matrix(points, mesh) @ sources (gradient
        False) or build_fullK_matrix(points, mesh)[c] @ sources for c = 0, 1, 2
        (gradient True) of the post-processing (solver.py:446-479), formed on
        the device: the points are evaluated in chunks of at most
        max_chunk_bytes of S and K (default: a quarter of the free device
        memory, at most 8 GiB), each chunk's product taken at once by
        cpt_zgemm, so the matrices never reach host memory.

        points: (N, 3) array, or a Mesh whose face centres are the points;
            with gradient, a Mesh also brings the diagonal term of
            build_fullK_matrix (matrices.f90:329-361): half (one on the free
            surface) its face normal for the faces i < min(N, faces) of
            `mesh`, times the source of face i.
        sources: (faces,) or (faces, r) complex.
        Returns (N,) or (N, r) for the potential, (N, 3) or (N, r, 3) for
        the gradient, as numpy arrays.
        """
        import torch
        gf = self.green_function
        mesh = mesh.merged()
        normals = None
        if hasattr(points, "faces_centers"):
            normals = points.faces_normals if gradient else None
            points = points.faces_centers
        points = np.ascontiguousarray(points, dtype=np.float64)
        # complex64 matrices and products for the single-precision Green
        # function, as its matrices are (abstract_green_function.py:60-70)
        f32 = gf.floating_point_precision == "float32"
        cdt, ndt = (torch.complex64, np.complex64) if f32 else (torch.complex128, np.complex128)
        src = np.asarray(sources, dtype=ndt)
        single = src.ndim == 1
        src = src.reshape(src.shape[0], -1)
        n, m, r = points.shape[0], mesh.nb_faces, src.shape[1]
        if src.shape[0] != m:
            raise ValueError(f"{src.shape[0]} sources for a mesh of {m} faces")
        dev = gf.device
        nk = 3 if gradient else 1
        per_point = (8 if f32 else 16) * m * (1 + nk)
        if max_chunk_bytes is None:
            free, _ = torch.cuda.mem_get_info(dev)
            max_chunk_bytes = min(8 << 30, free // 4)
        chunk = int(max(1, min(n, max_chunk_bytes // per_point)))
        chunk = -(-n // -(-n // chunk))  # chunks of equal size: the product's grid is one row slab per chunk
        S_buf = torch.empty(m * chunk, dtype=cdt, device=dev)
        K_buf = torch.empty(m * chunk * nk, dtype=cdt, device=dev)
        X = torch.as_tensor(np.ascontiguousarray(src.T), dtype=cdt).to(dev)[None]  # (1, r, m)
        P = torch.as_tensor(points).to(dev)  # one upload: a pageable copy per chunk would wait for the device
        out = torch.empty((r, n, nk), dtype=cdt, device=dev)
        for p0 in range(0, n, chunk):
            nc = min(chunk, n - p0)
            S = S_buf[:m * nc].view(1, m, nc)
            K = K_buf[:m * nc * nk].view((1, m, nc) if nk == 1 else (1, m, nc, 3))
            gf.evaluate_on_device(P[p0:p0 + nc], mesh, adjoint_double_layer=True, early_dot_product=not gradient,
                                  diagonal_term_in_double_layer=False, out=(S, K), **gf_params)
            A = K.view(1, m, nc * 3) if gradient else S
            out[:, p0:p0 + nc, :] = DeviceMatrix(A).matmul_device(X)[0].view(r, nc, nk)
        if self.check_nan:
            # a NaN entry of the matrix reaches every product it enters (NaN
            # times any source is NaN): one scan of the (r, N) result instead
            # of the (faces, N) matrix
            check_if_nan_in_matrix([out])
        if normals is not None:
            nd = min(n, m)
            fs = gf_params.get("free_surface", 0.0)
            on_fs = np.abs(mesh.faces_centers[:nd, 2] - fs) < float(np.float32(1e-8))
            coef = np.where(on_fs, 1.0, 0.5)[:, None] * normals[:nd]  # (nd, 3)
            coef = coef.astype(np.float32 if f32 else np.float64)
            out[:, :nd, :] += (torch.as_tensor(coef, device=dev)[None] * X[0, :, :nd, None])
        host = torch.empty(out.shape, dtype=out.dtype, pin_memory=True)
        host.copy_(out)
        res = host.numpy()
        res = res[..., 0] if not gradient else res
        res = np.moveaxis(res, 0, 1)  # (N, r) or (N, r, 3)
        return res[:, 0] if single else res

    def _build(self, mesh1, mesh2, *, diagonal_term_in_double_layer=True, **gf_params):
        gf_params.setdefault("early_dot_product", True)
        if not gf_params["early_dot_product"]:
            S, K = self.green_function.evaluate(mesh1, mesh2, diagonal_term_in_double_layer=diagonal_term_in_double_layer,
                                                **gf_params)
            check_if_nan_in_matrix([S, K])
            return S, K
        from .symmetry import evaluate_symmetric, symmetry_of
        sym = symmetry_of(mesh1) if mesh1 is mesh2 else None
        if sym is not None and self.green_function.floating_point_precision == "float32":
            raise NotImplementedError("The single-precision Green function is not available for symmetric "
                                      "meshes on the device; use mesh.merged() or floating_point_precision="
                                      "'float64'.")
        if sym is not None:
            # engines.py:178-247: the first block column, all collocation
            # points against the reference block
            S, K = evaluate_symmetric(self.green_function, mesh1, sym,
                                      diagonal_term_in_double_layer=diagonal_term_in_double_layer, **gf_params)
            if self.check_nan:
                check_if_nan_in_matrix([S, K])
            return S, K
        S, K = self.green_function.evaluate_on_device(mesh1, mesh2,
                                                      diagonal_term_in_double_layer=diagonal_term_in_double_layer,
                                                      **gf_params)
        S, K = DeviceMatrix(S), DeviceMatrix(K)
        if self.check_nan:
            check_if_nan_in_matrix([S, K])
        return S, K

    def build_matrices(self, mesh1, mesh2, **gf_params):
        """engines.py:258-298: returns the cached pair for identical inputs."""
        if (mesh1, mesh2, gf_params) == self.last_computed_inputs:
            LOG.debug("%s: reading cache.", self.__class__.__name__)
            return self.last_computed_matrices
        self.last_computed_matrices = None
        S, K = self._build(mesh1, mesh2, **dict(gf_params))
        self.last_computed_inputs = (mesh1, mesh2, dict(gf_params))
        self.last_computed_matrices = (S, K)
        return self.last_computed_matrices

    def _lu_solve(self, A, b, overwrite_a):
        """engines.py:326-338: factor once, cached in place of K."""
        if not has_been_lu_decomposed(A):
            luA = lu_decompose(A, overwrite_a=overwrite_a)
            if self.last_computed_matrices is not None and A is self.last_computed_matrices[1]:
                self.last_computed_matrices = (self.last_computed_matrices[0], luA)
        else:
            luA = A
        return luA.solve(b)

    def linear_solver(self, A, b):
        """engines.py:300-346."""
        if not isinstance(self._linear_solver, str):
            x = self._linear_solver(np.asarray(A), b)
            if not x.shape == b.shape:
                raise ValueError(f"Error in linear solver of {self}: the shape of the output ({x.shape}) "
                                 f"does not match the expected shape ({b.shape})")
            return x
        if self._linear_solver in ("lu_decomposition", "lu_decomposition_with_overwrite"):
            return self._lu_solve(A, b, overwrite_a=(self._linear_solver == "lu_decomposition_with_overwrite"))
        if self._linear_solver == "mixed_precision":
            from .symmetry import BlockSymmetricDeviceMatrix, MixedSymmetricSolver
            if isinstance(A, (DeviceMatrix, LUDecomposedDeviceMatrix)) and A.dtype == np.complex64:
                # complex64 matrices (single-precision Green function): the
                # complex64 LU is the whole solve, there is no FP64 matrix to
                # refine against
                return self._lu_solve(A, b, overwrite_a=False)
            if not isinstance(A, (MixedPrecisionSolver, MixedSymmetricSolver)):
                solver = MixedSymmetricSolver(A) if isinstance(A, BlockSymmetricDeviceMatrix) else \
                    MixedPrecisionSolver(A)
                if self.last_computed_matrices is not None and A is self.last_computed_matrices[1]:
                    self.last_computed_matrices = (self.last_computed_matrices[0], solver)
            else:
                solver = A
            return solver.solve(b)
        if self._linear_solver == "gmres":
            return solve_gmres(A, b)
        raise NotImplementedError(f"Unknown `linear_solver` in DefaultMatrixEngine: {self._linear_solver}")

    def compute_ram_estimation(self, problem):
        """Device memory estimate in GB (engines.py:348-398, plain meshes)."""
        nb_faces = problem.body.mesh.nb_faces
        nb_matrices = 3 if self._linear_solver == "lu_decomposition" else 2
        nb_bytes = 8 if self.green_function.floating_point_precision == "float32" else 16
        return nb_faces ** 2 * nb_matrices * nb_bytes / 1e9


BasicMatrixEngine = DefaultMatrixEngine
