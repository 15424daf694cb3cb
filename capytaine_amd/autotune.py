"""Autotuner of the LU block width, panel leaf width, look-ahead and assembly
work-group chunk.

The reference has no counterpart (its LAPACK and OpenMP loops take no
tuning).  Candidates are timed on the device with HIP events around the
library calls, on synthetic inputs of the requested shape; the fastest
setting per (device, precision, size class) is kept in a JSON file in the
cache directory and applied to a context with `apply`.  Every candidate
computes the same factorisation up to rounding (LAPACK's partial pivoting
rule for all of them), so tuning changes speed only.

    from capytaine_amd import autotune
    best = autotune.tune_lu(4352, 40, precision="mixed")   # times, stores
    autotune.apply(ctx, n=4352, batch=40, precision="mixed")
"""
import itertools
import json
import os
import threading

import numpy as np

from . import _lib
from ._lib import check, ptr, stream_handle

LU_BLOCKS = (128, 256)      # 64 loses at every measured size above 1000 rows
# wider blocks, for speculative panels (their panels and triangular solves
# recurse into deep, compute-bound products)
SPEC_LU_BLOCKS = (512, 1024)
PANEL_LEAVES = (8, 16)
# look-ahead modes: 0 (in order) and 1 (next panel on a side stream)
LOOKAHEAD = (0, 1)
# complex64 product tiles: 128 x 64, 128 x 128, the three-product (Gauss)
# forms 128 x 128 (one wavefront per SIMD) and 128 x 64, and the Gauss
# 128 x 64 form with FP16 inputs (more refinement sweeps).  A candidate whose
# solve was factored again with FP32-input products (the context's
# mixed_retries count grew) or missed the tolerance is discarded.
GEMM_TILES = (0, 1, 2, 3, 4)
# threads of a panel leaf's work-group timed with look-ahead 1 (the only
# mode where a leaf runs beside the trailing update).  256 (one wavefront
# per SIMD) fits on a CU beside one work-group of the tile product, but
# measured slower on MI355X at 4352 rows x 40 (FP64 228 vs 204 ms; mixed
# 127.7 vs 127.0 ms with the Gauss tiles), so only 512 is timed by default;
# pass candidates with a sixth element 256 to time it
PANEL_THREADS = (512,)
ASSEMBLY_CHUNKS = (0, 4, 8, 16, 32, 64)

_lock = threading.Lock()


def size_class(n):
    """Size classes of the tuning table: powers of two of the matrix order."""
    return int(2 ** round(np.log2(max(int(n), 1))))


def batch_class(batch):
    return int(2 ** round(np.log2(max(int(batch), 1))))


def _device_name(device):
    import torch
    return torch.cuda.get_device_name(device).replace(" ", "_")


def table_path():
    from .green_functions import cache_directory
    return os.path.join(cache_directory(), "tuning.json")


def load_table():
    path = table_path()
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def _store(key, value):
    with _lock:
        table = load_table()
        table[key] = value
        tmp = table_path() + f".{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        os.replace(tmp, table_path())


def _key(device, kind, precision, n, batch):
    return f"{_device_name(device)}/{kind}/{precision}/n{size_class(n)}/b{batch_class(batch)}"


def _test_matrices(n, batch, device, seed=0):
    """Well-conditioned complex matrices shaped like BEM K (1/2 I + smooth
    part), column-major batches (batch, n, n)."""
    import torch
    g = torch.Generator(device=device).manual_seed(seed)
    A = torch.randn((batch, n, n), dtype=torch.complex128, generator=g, device=device)
    A /= 4 * np.sqrt(n)
    A.diagonal(dim1=1, dim2=2).add_(0.5)
    return A


def _time(fn, device, reps):
    import torch
    stream = torch.cuda.current_stream(device)
    fn()  # warm: workspace allocation, first-launch costs
    torch.cuda.synchronize(device)
    best = float("inf")
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        torch.cuda.synchronize(device)
        best = min(best, e0.elapsed_time(e1))
    return best


def tune_lu(n, batch, precision="fp64", device=None, reps=2, nrhs=14, candidates=None, store=True, log=None):
    """Time every (lu_block, panel_leaf, lookahead, gemm_tiles, panel_threads)
    candidate on `batch` synthetic n x n matrices and return
    the fastest as a dict, with the timings of all candidates under
    "timings_ms"."""
    import torch
    device = torch.device("cuda", 0) if device is None else torch.device(device)
    if precision not in ("fp64", "mixed"):
        raise ValueError(f"precision must be 'fp64' or 'mixed', got {precision!r}")
    ctx = _lib.Context.get(device)
    lib = _lib.load()
    saved = _lib.get_tuning(ctx)
    A0 = _test_matrices(n, batch, device)
    A = torch.empty_like(A0)
    ipiv = torch.empty((batch, n), dtype=torch.int32, device=device)
    info = torch.empty((batch,), dtype=torch.int32, device=device)
    B = torch.ones((batch, nrhs, n), dtype=torch.complex128, device=device)
    X = torch.empty_like(B)
    st = stream_handle(device)
    import ctypes
    it, res = ctypes.c_int(), ctypes.c_double()

    failed = []

    def run():
        if precision == "fp64":
            A.copy_(A0)
            check(lib.cpt_lu_factor(ctx.handle, ptr(A), n, n, n * n, ptr(ipiv), n, ptr(info), batch, st),
                  "cpt_lu_factor")
        else:
            try:
                check(lib.cpt_lu_solve_mixed(ctx.handle, ptr(A0), n, n, n * n, batch, ptr(B), nrhs, n, n * nrhs,
                                             ptr(X), n, n * nrhs, 10, 1e-12, ctypes.byref(it), ctypes.byref(res),
                                             st), "cpt_lu_solve_mixed")
            except _lib.ConvergenceError:
                failed.append(True)

    timings = {}
    if candidates is None:
        tiles = GEMM_TILES if precision == "mixed" else (0,)
        # speculative panels (the context's default) have no pivot-search
        # leaves: their leaf width and threads are fixed
        spec = saved.get("pivot_speculation", 0)
        leaves, threads = ((16,), (512,)) if spec else (PANEL_LEAVES, PANEL_THREADS)
        blocks = LU_BLOCKS + SPEC_LU_BLOCKS if spec else LU_BLOCKS
        candidates = [(blk, leaf, la, tl, nt)
                      for blk, leaf, la, tl in itertools.product(blocks, leaves, LOOKAHEAD, tiles)
                      for nt in (threads if la == 1 else threads[:1])]
    try:
        for cand in candidates:
            blk, leaf, la, tl = cand[:4]
            nt = cand[4] if len(cand) > 4 else 512
            if leaf > blk // 2:
                continue
            _lib.set_tuning(ctx, lu_block=blk, panel_leaf=leaf, lookahead=la, gemm_tiles=tl, panel_threads=nt)
            retries0 = _lib.mixed_retries(ctx)
            del failed[:]
            t = _time(run, device, reps)
            if precision == "mixed" and (failed or _lib.mixed_retries(ctx) != retries0):
                if log:
                    log(f"[tune] n={n} batch={batch} mixed gemm_tiles={tl}: the refinement needed the FP32 "
                        f"retry or missed 1e-12 (residual {res.value:.2e}, {it.value} sweeps), discarded")
                continue
            timings[f"{blk}/{leaf}/{la}/{tl}/{nt}"] = t
            if log:
                log(f"[tune] n={n} batch={batch} {precision} block={blk} leaf={leaf} lookahead={la}"
                    f" gemm_tiles={tl} panel_threads={nt}: {t:.2f} ms")
    finally:
        _lib.set_tuning(ctx, **saved)
    best_key = min(timings, key=timings.get)
    blk, leaf, la, tl, nt = (int(x) for x in best_key.split("/"))
    best = {"lu_block": blk, "panel_leaf": leaf, "lookahead": la, "gemm_tiles": tl,
            "panel_threads": nt, "ms": timings[best_key], "timings_ms": timings}
    if store:
        _store(_key(device, "lu", precision, n, batch), best)
    return best


def tune_assembly(green_function, mesh, wavenumbers, reps=2, chunks=ASSEMBLY_CHUNKS, store=True, log=None):
    """Time the S/K assembly of `mesh` for `wavenumbers` with each work-group
    chunk and return the fastest."""
    import torch
    device = green_function.device
    ctx = _lib.Context.get(device)
    saved = _lib.get_tuning(ctx)
    n, nb = mesh.nb_faces, len(wavenumbers)
    S = torch.empty((nb, n, n), dtype=torch.complex128, device=device)
    K = torch.empty_like(S)

    def run():
        green_function.evaluate_on_device(mesh, mesh, wavenumber=list(wavenumbers), out=(S, K))

    timings = {}
    try:
        for c in chunks:
            _lib.set_tuning(ctx, assembly_chunk=c)
            timings[str(c)] = _time(run, device, reps)
            if log:
                log(f"[tune] assembly n={n} nb={nb} chunk={c}: {timings[str(c)]:.2f} ms")
    finally:
        _lib.set_tuning(ctx, **saved)
    best_key = min(timings, key=timings.get)
    best = {"assembly_chunk": int(best_key), "ms": timings[best_key], "timings_ms": timings}
    if store:
        _store(_key(device, "assembly", "fp64", n, nb), best)
    return best


def lookup(device, kind, precision, n, batch):
    return load_table().get(_key(device, kind, precision, n, batch))


def apply(ctx, n, batch, precision="fp64", nb_wavenumbers=None):
    """Set the stored best tuning of this size class on a context; returns
    the parameters applied (empty if the class was never tuned)."""
    import torch
    device = torch.device("cuda", ctx.device_index)
    applied = {}
    best = lookup(device, "lu", precision, n, batch)
    if best:
        applied.update(lu_block=best["lu_block"], panel_leaf=best["panel_leaf"],
                       lookahead=min(int(best["lookahead"]), 1), gemm_tiles=best.get("gemm_tiles", 1),
                       panel_threads=best.get("panel_threads", 512))
    best = lookup(device, "assembly", "fp64", n, nb_wavenumbers or batch)
    if best:
        applied["assembly_chunk"] = best["assembly_chunk"]
    if applied:
        _lib.set_tuning(ctx, **applied)
    return applied
