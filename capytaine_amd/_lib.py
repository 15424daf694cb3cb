"""ctypes binding of libcapytaine_amd.so (declared in include/capytaine_amd.h).

The library is built in place by ``capytaine_amd.build_lib``.  There is no
CPU fallback: every entry point that computes requires the HIP library and a
GPU, and raises otherwise.
"""
import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcapytaine_amd.so")

CPT_LEGACY_GRID, CPT_SCALED_NEMOH3_GRID = 0, 1
CPT_FINGREEN3D_METHOD, CPT_LEGACY_FINITE_DEPTH, CPT_NEWER_FINITE_DEPTH = -1, 0, 1
CPT_HIGH_FREQ, CPT_LOW_FREQ, CPT_LOW_FREQ_WITH_RANKINE_PART = 0, 1, 2
CPT_BLOCKS_CYCLIC, CPT_BLOCKS_XOR = 0, 1
CPT_ASSEMBLY_FAR_FIELD_FP32, CPT_ASSEMBLY_COMPLEX64 = 1, 2

# Every symbol of include/capytaine_amd.h with its argument types.
_P, _I, _L, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
SIGNATURES = {
    "cpt_last_error": (ctypes.c_char_p, []),
    "cpt_version": (_I, []),
    "cpt_create": (_I, [_I, ctypes.POINTER(_P)]),
    "cpt_destroy": (_I, [_P]),
    "cpt_tabulate": (_I, [_P, _P, _I, _P, _I, _I, _P, _P]),
    "cpt_set_tabulation": (_I, [_P, _P, _I, _P, _I, _P, _I, _I, _P]),
    "cpt_build_matrices": (_I, [_P, _I, _P, _P, _I, _P, _I, _P, _P, _P, _P, _P, _I, _P, _P, _I, _P, _P, _D, _I, _I,
                                _P, _I, _I, _I, _D, _P, _L, _L, _P, _L, _L, _P]),
    "cpt_build_matrices_dev": (_I, [_P, _I, _P, _P, _I, _P, _I, _P, _P, _P, _P, _P, _I, _P, _P, _I, _P, _P, _D, _I, _I,
                                    _P, _I, _I, _I, _D, _P, _L, _L, _P, _L, _L, _P]),
    "cpt_wave_part_infinite_depth": (_I, [_P, _I, _P, _P, _D, _I, _P, _P, _P]),
    "cpt_build_matrices_fingreen3d": (_I, [_P, _I, _P, _P, _I, _P, _I, _P, _P, _P, _P, _P, _I, _P, _P, _D, _D, _I,
                                           _P, _I, _I, _I, _I, _D, _P, _L, _P, _L, _P]),
    "cpt_fingreen3d_points": (_I, [_P, _I, _P, _P, _P, _D, _I, _P, _D, _P, _P]),
    "cpt_lu_factor": (_I, [_P, _P, _I, _L, _L, _P, _L, _P, _I, _P]),
    "cpt_lu_solve": (_I, [_P, _P, _I, _L, _L, _P, _L, _P, _I, _L, _L, _I, _P]),
    "cpt_zgemm": (_I, [_P, _I, _I, _I, _D, _P, _L, _L, _P, _L, _L, _P, _L, _L, _I, _I, _P]),
    "cpt_clu_factor": (_I, [_P, _P, _I, _L, _L, _P, _L, _P, _I, _P]),
    "cpt_clu_solve": (_I, [_P, _P, _I, _L, _L, _P, _L, _P, _I, _L, _L, _I, _P]),
    "cpt_cgemm": (_I, [_P, _I, _I, _I, ctypes.c_float, _P, _L, _L, _P, _L, _L, _P, _L, _L, _I, _I, _P]),
    "cpt_zgmres": (_I, [_P, _P, _I, _L, _P, _I, _L, _P, _L, _I, _I, _D, _D, _P, _P, _P]),
    "cpt_cgmres": (_I, [_P, _P, _I, _L, _P, _I, _L, _P, _L, _I, _I, _D, _D, _P, _P, _P]),
    "cpt_lu_solve_mixed": (_I, [_P, _P, _I, _L, _L, _I, _P, _I, _L, _L, _P, _L, _L, _I, _D,
                                ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double), _P]),
    "cpt_profile_begin": (_I, [_P]),
    "cpt_profile_end": (_I, [_P, _P, _P, _P, _P]),
    "cpt_build_matrices_batched": (_I, [_P, _I, _P, _I, _I, _D, _P]),
    "cpt_lu_factor_grouped": (_I, [_P, _I, _P, _P]),
    "cpt_lu_solve_grouped": (_I, [_P, _I, _P, _P]),
    "cpt_lu_solve_mixed_grouped": (_I, [_P, _I, _P, _I, _D, _P, _P, _P]),
    "cpt_release_workspace": (_I, [_P]),
    "cpt_set_assembly_precision": (_I, [_P, _I]),
    "cpt_set_workspace": (_I, [_P, _P, ctypes.c_size_t]),
    "cpt_mixed_workspace_size": (_I, [_P, _I, _P, ctypes.POINTER(ctypes.c_size_t)]),
    "cpt_set_tuning": (_I, [_P, _I, _I]),
    "cpt_get_tuning": (_I, [_P, _I, ctypes.POINTER(ctypes.c_int)]),
    "cpt_block_transform": (_I, [_P, _I, _I, _I, _I, _I, _I, _P, _L, _L, _L, _P, _L, _L, _L, _P]),
    "cpt_lu_solve_noalloc": (_I, [_P, _P, _I, _L, _L, _P, _L, _P, _I, _L, _L, _I, _P]),
    "cpt_mixed_factor_size": (_I, [_I, _I, ctypes.POINTER(ctypes.c_size_t)]),
    "cpt_mixed_factor": (_I, [_P, _P, _I, _L, _L, _I, _P, ctypes.c_size_t, _P]),
    "cpt_mixed_refine": (_I, [_P, _P, _I, _L, _L, _I, _P, ctypes.c_size_t, _P, _I, _L, _L, _P, _L, _L, _I, _D,
                              ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double), _P]),
}

# Tuning parameters of cpt_set_tuning, in the order of the CPT_TUNE_* values
# (None: a retired value).
TUNING_PARAMS = ("lu_block", "panel_leaf", "lookahead", "assembly_chunk", None, "gemm_tiles",
                 "lookahead_active", "panel_threads", "mixed_retries", "panel_wgs", "panel_spin",
                 "pivot_speculation", "spec_repairs")
READ_ONLY_TUNING = ("lookahead_active", "mixed_retries", "spec_repairs")

# Kernel classes of cpt_profile_end, in the order of the CPT_TIMED_* values.
TIMED_KERNELS = ("zgemm", "cgemm", "panel_z", "panel_c", "assembly", "skinny_z", "skinny_c", "cgemm_packed")



class CptBody(ctypes.Structure):
    """cpt_body of include/capytaine_amd.h."""
    _fields_ = [("nb_faces", ctypes.c_int), ("nb_vertices", ctypes.c_int), ("nb_quad_points", ctypes.c_int),
                ("nb_wavenumbers", ctypes.c_int), ("vertices", _P), ("faces", _P), ("centers", _P),
                ("normals", _P), ("areas", _P), ("radiuses", _P), ("quad_points", _P), ("quad_weights", _P),
                ("wavenumbers", _P), ("gf_singularities", _P), ("S", _P), ("ldS", _L), ("strideS", _L),
                ("K", _P), ("ldK", _L), ("strideK", _L)]


class CptLuGroup(ctypes.Structure):
    """cpt_lu_group of include/capytaine_amd.h."""
    _fields_ = [("n", ctypes.c_int), ("batch", ctypes.c_int), ("A", _P), ("lda", _L), ("strideA", _L),
                ("ipiv", _P), ("strideP", _L), ("info", _P), ("B", _P), ("nrhs", ctypes.c_int), ("ldb", _L),
                ("strideB", _L), ("X", _P), ("ldx", _L), ("strideX", _L)]


_lib = None
_lock = threading.RLock()  # Context.get -> Context.__init__ -> load() re-enters it


CPT_ERR_ALLOC = -3
CPT_ERR_UNSUPPORTED = -5
CPT_ERR_NOT_CONVERGED = -6


class CptError(RuntimeError):
    """Error reported by libcapytaine_amd."""

    def __init__(self, msg, status=None):
        super().__init__(msg)
        self.status = status


class ConvergenceError(CptError):
    """An iterative solve (mixed-precision refinement, GMRES) ended above its
    tolerance (CPT_ERR_NOT_CONVERGED).  This can be due to overlapping panels
    or irregular frequencies, as the reference says of its GMRES
    (engines.py:84-86)."""


def call_with_memory_retry(fn, what):
    """fn() -> status; on a device allocation failure, return torch's cached
    device memory to the driver and try once more."""
    status = fn()
    if status == CPT_ERR_ALLOC:
        import torch
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        status = fn()
    check(status, what)


def load():
    """Load the shared library (no GPU needed) and declare its signatures."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise CptError(f"{LIB_PATH} is missing: build it with `python -m capytaine_amd.build_lib`.")
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def check(status, what):
    if status != 0:
        msg = load().cpt_last_error().decode(errors="replace")
        cls = ConvergenceError if status == CPT_ERR_NOT_CONVERGED else CptError
        raise cls(f"{what} failed with status {status}: {msg}", status)


def ptr(t):
    """Device (or host) address of a torch tensor or numpy array."""
    if isinstance(t, np.ndarray):
        return ctypes.c_void_p(t.ctypes.data)
    return ctypes.c_void_p(t.data_ptr())


def stream_handle(device):
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


class Context:
    """One library context (cpt_handle) per device, created lazily."""

    _contexts = {}

    def __init__(self, device_index):
        self.device_index = device_index
        h = ctypes.c_void_p()
        check(load().cpt_create(device_index, ctypes.byref(h)), "cpt_create")
        self.handle = h
        self.table_key = None  # the device table tensor last uploaded

    def close(self):
        """Destroy a context made with Context(device_index) (not one of the
        shared per-device contexts of Context.get)."""
        if self.handle is not None and self._contexts.get(self.device_index) is not self:
            check(load().cpt_destroy(self.handle), "cpt_destroy")
            self.handle = None

    @classmethod
    def get(cls, device):
        import torch
        dev = torch.device(device)
        if dev.type != "cuda":
            raise CptError(f"capytaine_amd computes on a GPU; got device {dev}")
        idx = dev.index if dev.index is not None else torch.cuda.current_device()
        with _lock:
            ctx = cls._contexts.get(idx)
            if ctx is None:
                ctx = cls(idx)
                cls._contexts[idx] = ctx
        return ctx


def require_gpu():
    import torch
    if not torch.cuda.is_available():
        raise CptError("capytaine_amd needs a ROCm GPU (torch.cuda.is_available() is False); "
                       "there is no CPU fallback.")
    load()


def default_device():
    import torch
    require_gpu()
    return torch.device("cuda", torch.cuda.current_device())


def profile_begin(ctx):
    """Start the live HIP-event timing of the hot kernels launched through ctx."""
    check(load().cpt_profile_begin(ctx.handle), "cpt_profile_begin")


def profile_end(ctx):
    """Stop the timing; {class: {"ms", "work", "launches", "bytes"}} summed since begin."""
    n = len(TIMED_KERNELS)
    ms, work, launches, nbytes = np.zeros(n), np.zeros(n), np.zeros(n, dtype=np.int64), np.zeros(n)
    check(load().cpt_profile_end(ctx.handle, ptr(ms), ptr(work), ptr(launches), ptr(nbytes)), "cpt_profile_end")
    return {k: {"ms": float(ms[i]), "work": float(work[i]), "launches": int(launches[i]), "bytes": float(nbytes[i])}
            for i, k in enumerate(TIMED_KERNELS)}


def set_tuning(ctx, **params):
    """Set tuning parameters of a context (names of TUNING_PARAMS)."""
    lib = load()
    for name, value in params.items():
        check(lib.cpt_set_tuning(ctx.handle, TUNING_PARAMS.index(name), int(value)), f"cpt_set_tuning({name})")


def _get_one(ctx, name):
    v = ctypes.c_int()
    check(load().cpt_get_tuning(ctx.handle, TUNING_PARAMS.index(name), ctypes.byref(v)), "cpt_get_tuning")
    return v.value


def lookahead_active(ctx):
    """Look-ahead mode the context's factorisations run."""
    return _get_one(ctx, "lookahead_active")


def mixed_retries(ctx):
    """Mixed solves (or groups) of the context factored again with
    FP32-input products after FP16-input products fell short."""
    return _get_one(ctx, "mixed_retries")


def spec_repairs(ctx):
    """Panels of the context's factorisations factored again with partial
    pivoting after the speculative panel's check failed (waits for the
    device)."""
    return _get_one(ctx, "spec_repairs")


def get_tuning(ctx):
    lib = load()
    out = {}
    for i, name in enumerate(TUNING_PARAMS):
        if name is None or name in READ_ONLY_TUNING:
            continue
        v = ctypes.c_int()
        check(lib.cpt_get_tuning(ctx.handle, i, ctypes.byref(v)), "cpt_get_tuning")
        out[name] = v.value
    return out
