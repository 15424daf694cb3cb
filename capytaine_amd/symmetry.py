"""Symmetric meshes on the device: block-diagonalised S and K.

Restates capytaine's ``_build_matrices_with_symmetries`` (bem/engines.py:
178-247) and its block matrices (tools/block_circulant_matrices.py:42-385)
for the symmetries whose group is abelian:

* one reflection plane (``ReflectionSymmetricMesh`` of a plain mesh): two
  blocks, block (i, j) = A_(i xor j);
* two reflection planes (a ``ReflectionSymmetricMesh`` of one across the
  other plane): four blocks, A_(i xor j), bit 0 the inner plane;
* n rotations about Oz (``RotationSymmetricMesh`` of a plain wedge), n <= 128:
  block (i, j) = A_((i - j) mod n), block circulant;
* n rotations of a wedge symmetric across a vertical plane (dihedral): the
  half-wedge's column is evaluated and the wedge's rebuilt from it, then the
  n wedges are block circulant (NestedBlockCirculantMatrix).

The assembly evaluates the first block column only (all N = n w collocation
points against the w faces of the reference block: a factor n fewer pairs).
The group's Fourier transform (``cpt_block_transform``) turns it into n
independent w x w matrices, factorised by one batched LU (n^2 fewer flops than
the N x N factorisation).  Right-hand sides are transformed, solved and
transformed back.  Other nestings are evaluated on the merged mesh.
"""
import numpy as np

from . import _lib
from ._lib import check, ptr, stream_handle
from .meshes import ReflectionSymmetricMesh, RotationSymmetricMesh, SymmetricMesh


class Symmetry:
    """kind "xor" (reflections) or "cyclic" (rotations), n blocks, the
    reference block's mesh.  For a dihedral mesh (n rotations of a wedge
    symmetric across a vertical plane), `half` is the wedge's half: the
    assembly evaluates against it and `macro_column` rebuilds the wedge's
    column, after which the n wedges are cyclic blocks."""

    def __init__(self, kind, n, block_mesh, half=None):
        self.kind, self.n, self.block_mesh, self.half = kind, int(n), block_mesh, half

    @property
    def evaluated_mesh(self):
        """The faces the assembly evaluates all collocation points against."""
        return self.half if self.half is not None else self.block_mesh

    @property
    def w(self):
        return self.block_mesh.nb_faces

    def block_index(self, i, j):
        """Index in the first block column of block (i, j)."""
        return (i ^ j) if self.kind == "xor" else (i - j) % self.n

    def characters(self, inverse=False):
        """(n, n) matrix chi[k, j] of the forward (or inverse, scaled) transform."""
        k, j = np.meshgrid(np.arange(self.n), np.arange(self.n), indexing="ij")
        if self.kind == "xor":
            bits = np.vectorize(lambda a: bin(a).count("1"))(k & j)
            chi = (-1.0) ** bits + 0j
        else:
            chi = np.exp((2j if inverse else -2j) * np.pi * k * j / self.n)
        return chi / self.n if inverse else chi

    def __repr__(self):
        return f"Symmetry({self.kind}, n={self.n}, w={self.w})"


def symmetry_of(mesh):
    """The Symmetry the device path uses for `mesh`, or None (plain mesh, or a
    nesting evaluated on the merged mesh)."""
    if isinstance(mesh, ReflectionSymmetricMesh):
        h = mesh.half
        if not isinstance(h, SymmetricMesh):
            return Symmetry("xor", 2, h)
        if isinstance(h, ReflectionSymmetricMesh) and h.plane != mesh.plane and not isinstance(h.half, SymmetricMesh):
            return Symmetry("xor", 4, h.half)
        return None
    if isinstance(mesh, RotationSymmetricMesh) and mesh.n <= 128:
        wd = mesh.wedge
        if not isinstance(wd, SymmetricMesh):
            return Symmetry("cyclic", mesh.n, wd)
        if isinstance(wd, ReflectionSymmetricMesh) and not isinstance(wd.half, SymmetricMesh):
            return Symmetry("cyclic", mesh.n, wd.merged(), half=wd.half)
    return None


def macro_column(sym, cols):
    """Dihedral mesh: from cols (nb, h, N), all collocation points against
    the wedge's half (N = 2 n h), the column of the n wedge blocks (nb, 2h, N).

    Block p = 2 i + s of the merged faces is R^i M^s applied to the half
    (R the rotation by 2 pi / n, M the wedge's mirror), and G is invariant
    under both, so block (p, q) of the full matrix is the evaluated block of
    T_q^-1 T_p.  For the wedge's second half (q = 1, T = M) that is
    M R^i M^s = R^-i M^(1-s): block 2 ((n - i) mod n) + 1 - s.  This is
    NestedBlockCirculantMatrix.to_BlockCirculantMatrix
    (block_circulant_matrices.py:222-249)."""
    import torch
    n = sym.n
    nb, h, N = cols.shape
    p = np.arange(2 * n)
    perm = torch.as_tensor(2 * ((n - p // 2) % n) + 1 - p % 2, device=cols.device)
    out = torch.empty((nb, 2 * h, N), dtype=cols.dtype, device=cols.device)
    out[:, :h, :] = cols
    out[:, h:, :] = cols.view(nb, h, 2 * n, h)[:, :, perm, :].reshape(nb, h, N)
    return out


def evaluate_symmetric(gf, mesh, sym, **gf_params):
    """S and K of a symmetric mesh as BlockSymmetricDeviceMatrix: the first
    block column (engines.py:178-247), for a dihedral mesh rebuilt from the
    half-wedge's."""
    S, K = gf.evaluate_on_device(mesh.merged(), sym.evaluated_mesh, **gf_params)
    if sym.half is not None:
        S, K = macro_column(sym, S), macro_column(sym, K)
    return BlockSymmetricDeviceMatrix(S, sym), BlockSymmetricDeviceMatrix(K, sym)


def _transform(sym, inp, out, R, C, NB, sbi, sj, ldi, sbo, sk, ldo, inverse):
    dev = inp.device
    ctx = _lib.Context.get(dev)
    kind = _lib.CPT_BLOCKS_XOR if sym.kind == "xor" else _lib.CPT_BLOCKS_CYCLIC
    check(_lib.load().cpt_block_transform(ctx.handle, kind, sym.n, int(inverse), R, C, NB, ptr(inp), sbi, sj, ldi,
                                          ptr(out), sbo, sk, ldo, stream_handle(dev)), "cpt_block_transform")
    return out


def transform_vectors(sym, x, inverse=False):
    """Forward: x (nb, r, N) -> (nb n, r, w), the transformed blocks of each
    batch item in turn.  Inverse: (nb n, r, w) -> (nb, r, N)."""
    import torch
    n, w = sym.n, sym.w
    if not inverse:
        nb, r, N = x.shape
        out = torch.empty((nb * n, r, w), dtype=torch.complex128, device=x.device)
        return _transform(sym, x, out, w, r, nb, r * N, w, N, n * r * w, r * w, w, False)
    nbn, r, _ = x.shape
    nb = nbn // n
    N = n * w
    out = torch.empty((nb, r, N), dtype=torch.complex128, device=x.device)
    return _transform(sym, x, out, w, r, nb, n * r * w, r * w, w, r * N, w, N, True)


class BlockSymmetricDeviceMatrix:
    """N x N complex128 matrix (or batch of them) of a symmetric mesh, held
    as its first block column: `cols` (nb, w, N) is the column-major N x w
    matrix of all collocation points against the reference block."""

    def __init__(self, cols, sym):
        self.cols = cols if cols.dim() == 3 else cols.unsqueeze(0)
        self.sym = sym
        self.device = self.cols.device
        self._diag = None

    batch = property(lambda self: self.cols.shape[0])
    dtype = property(lambda self: np.dtype(np.complex128))

    @property
    def shape(self):
        N = self.cols.shape[2]
        return (N, N) if self.batch == 1 else (self.batch, N, N)

    ndim = property(lambda self: len(self.shape))

    def diagonal_blocks(self):
        """(nb n, w, w): the transformed blocks, column-major, batch item b's
        block k at index b n + k."""
        import torch
        if self._diag is None:
            n, w = self.sym.n, self.sym.w
            nb, N = self.batch, self.cols.shape[2]
            out = torch.empty((nb * n, w, w), dtype=torch.complex128, device=self.device)
            self._diag = _transform(self.sym, self.cols, out, w, w, nb, w * N, w, N, n * w * w, w * w, w, False)
        return self._diag

    def __array__(self, dtype=None, copy=None):
        cols = self.cols.cpu().numpy().transpose(0, 2, 1)  # (nb, N, w)
        n, w = self.sym.n, self.sym.w
        nb, N = cols.shape[0], cols.shape[1]
        full = np.empty((nb, N, N), dtype=np.complex128)
        for i in range(n):
            for j in range(n):
                a = self.sym.block_index(i, j)
                full[:, i * w:(i + 1) * w, j * w:(j + 1) * w] = cols[:, a * w:(a + 1) * w, :]
        full = full[0] if nb == 1 else full
        return full if dtype is None else full.astype(dtype)

    def numpy(self):
        return self.__array__()

    def matmul_device(self, x):
        """self @ x for x (nb, r, N) on the device; returns (nb, r, N)."""
        from .engines import DeviceMatrix
        xh = transform_vectors(self.sym, x)
        yh = DeviceMatrix(self.diagonal_blocks()).matmul_device(xh)
        return transform_vectors(self.sym, yh, inverse=True)

    def __matmul__(self, x):
        import torch
        x = np.asarray(x)
        vec = x.ndim == 1
        xb = x.reshape(x.shape[0], -1) if self.batch == 1 else x
        if self.batch == 1:
            xd = torch.as_tensor(np.ascontiguousarray(xb.T), dtype=torch.complex128).to(self.device)[None]
        else:
            xd = torch.as_tensor(np.ascontiguousarray(np.swapaxes(xb, -1, -2)),
                                 dtype=torch.complex128).to(self.device)
        y = np.swapaxes(self.matmul_device(xd).cpu().numpy(), -1, -2)
        if self.batch == 1:
            y = y[0]
            return y[:, 0] if vec else y
        return y


class LUDecomposedSymmetricMatrix:
    """LU factors of the n transformed blocks of a BlockSymmetricDeviceMatrix
    (one batched factorisation); `solve` transforms, solves and transforms
    back (block_circulant_matrices.py:150-160, :361-385)."""

    def __init__(self, A, overwrite_a=False):
        from .engines import DeviceMatrix, LUDecomposedDeviceMatrix
        self.sym = A.sym
        self.shape, self.dtype, self.device, self.batch = A.shape, A.dtype, A.device, A.batch
        D = A.diagonal_blocks()
        A._diag = None  # the factorisation overwrites the transformed blocks
        self.lu = LUDecomposedDeviceMatrix(DeviceMatrix(D), overwrite_a=True)

    def check_info(self):
        self.lu.check_info()

    def solve_device(self, b):
        """Solve in place for b (nb, r, N) on the device."""
        bh = transform_vectors(self.sym, b)
        self.lu.solve_device(bh)
        b.copy_(transform_vectors(self.sym, bh, inverse=True))
        return b

    def solve(self, b):
        import torch
        b = np.asarray(b)
        vec = b.ndim == 1
        bb = b.reshape(b.shape[0], -1) if self.batch == 1 else b
        if self.batch == 1:
            bd = torch.as_tensor(np.ascontiguousarray(bb.T), dtype=torch.complex128).to(self.device)[None]
        else:
            bd = torch.as_tensor(np.ascontiguousarray(np.swapaxes(bb, -1, -2)),
                                 dtype=torch.complex128).to(self.device)
        x = np.swapaxes(self.solve_device(bd).cpu().numpy(), -1, -2)
        self.check_info()
        if self.batch == 1:
            x = x[0]
            return x[:, 0] if vec else x
        return x


class MixedSymmetricSolver:
    """The mixed-precision solve of a BlockSymmetricDeviceMatrix: right-hand
    sides transformed, the n transformed blocks solved by the complex64 LU
    with FP64 refinement against the complex128 blocks (cpt_mixed_factor once,
    cpt_mixed_refine per call), solutions transformed back."""

    def __init__(self, A, max_iter=10, tol=1e-12):
        from .engines import DeviceMatrix, MixedPrecisionSolver
        self.sym = A.sym
        self.shape, self.dtype, self.device, self.batch = A.shape, A.dtype, A.device, A.batch
        self.inner = MixedPrecisionSolver(DeviceMatrix(A.diagonal_blocks()), max_iter=max_iter, tol=tol)

    iterations = property(lambda self: self.inner.iterations)
    residual = property(lambda self: self.inner.residual)

    def solve_device(self, b):
        """x for b (nb, r, N) on the device (b is left unchanged)."""
        xh = self.inner.solve_device(transform_vectors(self.sym, b))
        return transform_vectors(self.sym, xh, inverse=True)

    def solve(self, b):
        import torch
        b = np.asarray(b)
        vec = b.ndim == 1
        bb = b.reshape(b.shape[0], -1)
        bd = torch.as_tensor(np.ascontiguousarray(bb.T), dtype=torch.complex128).to(self.device)[None]
        x = self.solve_device(bd).cpu().numpy()[0].T
        return x[:, 0] if vec else x

