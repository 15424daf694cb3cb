"""cpt_lu_solve_noalloc (the solve that allocates no device memory, its
scratch a static device array of the library) against cpt_lu_solve, through
the C-ABI: two calls in a row, the second smaller, both bit-identical to
the allocating entry point (same kernels, same operands); and the limit of
its scratch is an argument error."""
import numpy as np
import pytest

from gpu_helpers import device

pytestmark = pytest.mark.gpu


def _case(rng, n, nb, r):
    A = rng.normal(size=(nb, n, n)) + 1j * rng.normal(size=(nb, n, n))
    A = A / np.sqrt(n) + 2.0 * np.eye(n)
    B = rng.normal(size=(nb, n, r)) + 1j * rng.normal(size=(nb, n, r))
    return A, B


def test_lu_solve_noalloc_matches_lu_solve():
    import torch
    from scipy.linalg import lu_factor, lu_solve
    from capytaine_amd import _lib
    from capytaine_amd._lib import check, ptr, stream_handle
    lib = _lib.load()
    dev = device()
    ctx = _lib.Context.get(dev)
    st = stream_handle(dev)
    rng = np.random.default_rng(7)
    for n, nb, r in [(1100, 3, 14), (517, 2, 5)]:  # the second call smaller
        print(f"[gpu] noalloc solve n={n} batch={nb} nrhs={r}", flush=True)
        A, B = _case(rng, n, nb, r)
        Ad = torch.as_tensor(np.ascontiguousarray(np.swapaxes(A, 1, 2))).to(dev)
        ipiv = torch.empty((nb, n), dtype=torch.int32, device=dev)
        info = torch.empty(nb, dtype=torch.int32, device=dev)
        check(lib.cpt_lu_factor(ctx.handle, ptr(Ad), n, n, n * n, ptr(ipiv), n, ptr(info), nb, st), "factor")
        B1 = torch.as_tensor(np.ascontiguousarray(np.swapaxes(B, 1, 2))).to(dev)
        B2 = B1.clone()
        check(lib.cpt_lu_solve(ctx.handle, ptr(Ad), n, n, n * n, ptr(ipiv), n, ptr(B1), r, n, n * r, nb, st),
              "cpt_lu_solve")
        check(lib.cpt_lu_solve_noalloc(ctx.handle, ptr(Ad), n, n, n * n, ptr(ipiv), n, ptr(B2), r, n, n * r, nb,
                                       st), "cpt_lu_solve_noalloc")
        torch.cuda.synchronize()
        assert not info.cpu().numpy().any()
        X1, X2 = B1.cpu().numpy(), B2.cpu().numpy()
        assert np.array_equal(X1, X2)
        X = np.swapaxes(X2, 1, 2)
        for b in range(nb):
            Xr = lu_solve(lu_factor(A[b]), B[b])
            assert np.linalg.norm(X[b] - Xr) / np.linalg.norm(Xr) < 1e-10


def test_lu_solve_noalloc_rejects_beyond_scratch():
    import torch
    from capytaine_amd import _lib
    from capytaine_amd._lib import ptr, stream_handle
    dev = device()
    ctx = _lib.Context.get(dev)
    dummy = torch.zeros(16, dtype=torch.complex128, device=dev)
    piv = torch.zeros(16, dtype=torch.int32, device=dev)
    n = 4096
    batch = 32768 // n + 1  # one matrix beyond the scratch
    status = _lib.load().cpt_lu_solve_noalloc(ctx.handle, ptr(dummy), n, n, n * n, ptr(piv), n, ptr(dummy), 1, n, n,
                                              batch, stream_handle(dev))
    assert status == -1
    assert b"exceeds the static scratch" in _lib.load().cpt_last_error()
