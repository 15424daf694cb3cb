"""Batching over bodies of uneven panel counts (persistent multi-body
assembly + variable-size grouped LU): equal to the one-body path.

Tolerances: S and K bit-equal to cpt_build_matrices per body (the same
per-pair arithmetic); forces within 1e-12 relative of the one-body sweep of
the same form; mixed-precision forces (FP32 far field, complex64 factors,
refinement to a 1e-12 residual) within 1e-6 of FP64; grouped LU solutions
within 1e-10 of scipy."""
import numpy as np
import pytest

from gpu_helpers import device, ref_table_gf

pytestmark = pytest.mark.gpu


def _fleet():
    from capytaine_amd import workloads
    bodies = [workloads.hull(3.0, 1.5, 1.0, 130, "h0"), workloads.hull(6.0, 2.5, 2.0, 420, "h1"),
              workloads.hull(2.0, 1.0, 0.6, 60, "h2"), workloads.hull(8.0, 3.0, 1.5, 260, "h3")]
    omegas = [np.array([0.5, 1.1]), np.array([0.3, 0.9, 2.0]), np.array([1.4]), np.array([0.8, 2.6])]
    headings = [np.array([0.0, np.pi / 2]), np.array([0.0]), np.array([]), np.array([np.pi, 0.3])]
    return bodies, omegas, headings


def _bcs(bodies, omegas, headings, rho=1025.0):
    import torch
    from capytaine_amd import workloads
    from capytaine_amd.sweep import boundary_conditions
    out, nrhs = [], []
    for b, om, hd in zip(bodies, omegas, headings):
        pbs = workloads.problems_by_frequency(b, om, hd, rho=rho)
        r = len(pbs[0])
        nrhs.append(r)
        out.append(boundary_conditions(pbs, r, b.mesh.nb_faces, device()))
    return out, nrhs


@pytest.mark.parametrize("far_field_fp32", [False, True])
def test_batched_assembly_equals_single(default_table, far_field_fp32):
    import torch
    from capytaine_amd import _lib
    from capytaine_amd.batch import BatchedSweep
    print(f"[gpu] batched assembly vs one body at a time (FP32 far field: {far_field_fp32})", flush=True)
    gf = ref_table_gf(default_table)
    bodies, omegas, headings = _fleet()
    ks = [om ** 2 / 9.81 for om in omegas]
    bs = BatchedSweep(gf, bodies, ks, 1, precision="mixed" if far_field_fp32 else "fp64",
                      far_field_fp32=far_field_fp32)
    bs.assemble()
    torch.cuda.synchronize()
    ctx = gf._context()
    for i, b in enumerate(bodies):
        _lib.check(_lib.load().cpt_set_assembly_precision(ctx.handle, int(far_field_fp32)), "precision")
        try:
            S, K = gf.evaluate_on_device(b.mesh, b.mesh, wavenumber=ks[i])
            torch.cuda.synchronize()
        finally:
            _lib.check(_lib.load().cpt_set_assembly_precision(ctx.handle, 0), "precision")
        assert torch.equal(S, bs.S[i]) and torch.equal(K, bs.K[i]), b.name


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_forces_match_sweep(default_table, precision):
    import torch
    from capytaine_amd.batch import BatchedSweep
    from capytaine_amd.sweep import FrequencySweep
    print(f"[gpu] batched {precision} forces vs per-body sweeps", flush=True)
    gf = ref_table_gf(default_table)
    bodies, omegas, headings = _fleet()
    ks = [om ** 2 / 9.81 for om in omegas]
    bcs, nrhs = _bcs(bodies, omegas, headings)
    bs = BatchedSweep(gf, bodies, ks, nrhs, precision=precision)
    bs.run(bcs)
    got = bs.forces(omegas, 1025.0)
    if precision == "mixed":
        assert np.all(bs.residual <= 1e-12) and np.all(bs.iterations >= 1)
    for i, b in enumerate(bodies):
        # the same form one body at a time: same arithmetic
        sw = FrequencySweep(gf, b, ks[i], nrhs[i], precision=precision)
        sw.run(bcs[i])
        ref = sw.forces(omegas[i], 1025.0)
        scale = np.abs(ref).max()
        assert np.abs(got[i] - ref).max() <= 1e-12 * scale, (b.name, np.abs(got[i] - ref).max() / scale)
        if precision == "mixed":
            # FP32 far field + complex64 factors against the FP64 form
            sw = FrequencySweep(gf, b, ks[i], nrhs[i], precision="fp64")
            sw.run(bcs[i])
            ref = sw.forces(omegas[i], 1025.0)
            assert np.abs(got[i] - ref).max() <= 1e-6 * np.abs(ref).max(), b.name
    torch.cuda.synchronize()


def test_grouped_lu_random():
    """cpt_lu_factor_grouped / cpt_lu_solve_grouped / cpt_lu_solve_mixed_grouped
    on groups of orders 1, 65, 300 and 700 against scipy."""
    import ctypes
    import torch
    from scipy.linalg import lu_factor, lu_solve
    from capytaine_amd import _lib
    from capytaine_amd._lib import CptLuGroup, check, ptr, stream_handle
    print("[gpu] grouped LU", flush=True)
    rng = np.random.default_rng(7)
    dev = device()
    sizes, batches, r = [1, 65, 300, 700], [3, 2, 2, 1], 5
    As, Bs = [], []
    for n, nb in zip(sizes, batches):
        A = rng.normal(size=(nb, n, n)) + 1j * rng.normal(size=(nb, n, n))
        A = A / np.sqrt(n) + 3 * np.eye(n)
        As.append(A)
        Bs.append(rng.normal(size=(nb, n, r)) + 1j * rng.normal(size=(nb, n, r)))
    t = lambda a: torch.as_tensor(np.ascontiguousarray(np.swapaxes(a, 1, 2))).to(dev)
    K = [t(A) for A in As]
    K0 = [k.clone() for k in K]
    B = [t(b) for b in Bs]
    Bm = [b.clone() for b in B]
    X = [torch.empty_like(b) for b in B]
    ipiv = [torch.empty((nb, n), dtype=torch.int32, device=dev) for n, nb in zip(sizes, batches)]
    info = [torch.empty((nb,), dtype=torch.int32, device=dev) for nb in batches]
    g = (CptLuGroup * len(sizes))()
    gm = (CptLuGroup * len(sizes))()
    for i, (n, nb) in enumerate(zip(sizes, batches)):
        g[i] = CptLuGroup(n=n, batch=nb, A=K[i].data_ptr(), lda=n, strideA=n * n, ipiv=ipiv[i].data_ptr(),
                          strideP=n, info=info[i].data_ptr(), B=B[i].data_ptr(), nrhs=r, ldb=n, strideB=n * r)
        gm[i] = CptLuGroup(n=n, batch=nb, A=K0[i].data_ptr(), lda=n, strideA=n * n, B=Bm[i].data_ptr(), nrhs=r,
                           ldb=n, strideB=n * r, X=X[i].data_ptr(), ldx=n, strideX=n * r)
    lib = _lib.load()
    ctx = _lib.Context.get(dev)
    st = stream_handle(dev)
    check(lib.cpt_lu_factor_grouped(ctx.handle, len(sizes), g, st), "factor")
    check(lib.cpt_lu_solve_grouped(ctx.handle, len(sizes), g, st), "solve")
    its = np.zeros(len(sizes), dtype=np.int32)
    res = np.zeros(len(sizes))
    check(lib.cpt_lu_solve_mixed_grouped(ctx.handle, len(sizes), gm, 10, 1e-12, ptr(its), ptr(res), st), "mixed")
    torch.cuda.synchronize()
    assert np.all(res <= 1e-12)
    for i in range(len(sizes)):
        assert not info[i].cpu().numpy().any()
        Xg = np.swapaxes(B[i].cpu().numpy(), 1, 2)
        Xm = np.swapaxes(X[i].cpu().numpy(), 1, 2)
        for b in range(batches[i]):
            Xr = lu_solve(lu_factor(As[i][b]), Bs[i][b])
            assert np.linalg.norm(Xg[b] - Xr) / np.linalg.norm(Xr) < 1e-10
            assert np.linalg.norm(Xm[b] - Xr) / np.linalg.norm(Xr) < 1e-10


def test_grouped_mixed_fp16_products_fall_back_to_fp32():
    """cpt_lu_solve_mixed_grouped with FP16-input products and one sweep
    allowed: the groups left short of tol (orders 300 and 700, whose
    factors use the tile products) are solved again with FP32-input
    products; every group meets tol and matches scipy."""
    import torch
    from scipy.linalg import lu_factor, lu_solve
    from capytaine_amd import _lib
    from capytaine_amd._lib import CptLuGroup, check, ptr, stream_handle
    print("[gpu] grouped mixed: FP16 products, one sweep, FP32 retry", flush=True)
    rng = np.random.default_rng(8)
    dev = device()
    sizes, batches, r = [65, 300, 700], [2, 2, 1], 4
    As, Bs = [], []
    for n, nb in zip(sizes, batches):
        A = rng.normal(size=(nb, n, n)) + 1j * rng.normal(size=(nb, n, n))
        As.append(A / np.sqrt(n) + 3 * np.eye(n))
        Bs.append(rng.normal(size=(nb, n, r)) + 1j * rng.normal(size=(nb, n, r)))
    t = lambda a: torch.as_tensor(np.ascontiguousarray(np.swapaxes(a, 1, 2))).to(dev)
    K, B = [t(A) for A in As], [t(b) for b in Bs]
    X = [torch.empty_like(b) for b in B]
    gm = (CptLuGroup * len(sizes))()
    for i, (n, nb) in enumerate(zip(sizes, batches)):
        gm[i] = CptLuGroup(n=n, batch=nb, A=K[i].data_ptr(), lda=n, strideA=n * n, B=B[i].data_ptr(), nrhs=r,
                           ldb=n, strideB=n * r, X=X[i].data_ptr(), ldx=n, strideX=n * r)
    lib = _lib.load()
    ctx = _lib.Context.get(dev)
    saved = _lib.get_tuning(ctx)
    its = np.zeros(len(sizes), dtype=np.int32)
    res = np.zeros(len(sizes))
    try:
        _lib.set_tuning(ctx, gemm_tiles=4)
        check(lib.cpt_lu_solve_mixed_grouped(ctx.handle, len(sizes), gm, 1, 1e-12, ptr(its), ptr(res),
                                             stream_handle(dev)), "mixed")
        torch.cuda.synchronize()
    finally:
        _lib.set_tuning(ctx, **saved)
    assert np.all(res <= 1e-12)
    assert its[1] == 2 and its[2] == 2
    for i in range(len(sizes)):
        Xm = np.swapaxes(X[i].cpu().numpy(), 1, 2)
        for b in range(batches[i]):
            Xr = lu_solve(lu_factor(As[i][b]), Bs[i][b])
            assert np.linalg.norm(Xm[b] - Xr) / np.linalg.norm(Xr) < 1e-10
