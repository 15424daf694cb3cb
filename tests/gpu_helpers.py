"""Helpers of the GPU tests: direct calls through the C-ABI."""
import numpy as np

from capytaine_amd import _lib
from capytaine_amd._lib import check, ptr, stream_handle


def device():
    import torch
    return torch.device("cuda", 0)


def ref_table_gf(default_table):
    """A Delhommeau whose device table is the reference table (bit-exact
    oracle copy) instead of the one computed on the GPU."""
    from capytaine_amd import Delhommeau
    gf = Delhommeau(tabulation_cache_dir=None, device=device())
    r, z, tab = default_table
    gf._set_tabulation(r, z, np.asfortranarray(tab))
    return gf


def build_raw(gf, arr, wavenumbers, *, depth=np.inf, prony=None, gfs=1, finite_depth_method=1, adjoint=True,
              early_dot_product=True, diagonal_term=True, free_surface=0.0):
    """cpt_build_matrices on explicit arrays (faces 0-based); returns
    Fortran-ordered numpy S (nb, n, m) and K (nb, n, m) / (nb, 3, n, m)."""
    import torch
    dev = device()
    t = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a), dtype=dt).to(dev)
    ks = np.atleast_1d(np.asarray(wavenumbers, float))
    nb = len(ks)
    colloc, dpn = t(arr["colloc"]), t(arr["dpn"])
    n = arr["colloc"].shape[0]
    m = arr["faces"].shape[0]
    qw = np.asarray(arr["quad_weights"])
    gfs_arr = np.full(nb, gfs, dtype=np.int32) if np.ndim(gfs) == 0 else np.asarray(gfs, np.int32)
    if prony is None:
        nexp, pr = 0, np.zeros(1)
    else:
        prony = np.asarray(prony, float)
        nexp = prony.shape[1]
        pr = np.ascontiguousarray(np.broadcast_to(prony, (nb, 2, nexp)))
    nk = 1 if early_dot_product else 3
    single = gf.floating_point_precision == "float32"  # complex64 stores
    cdt = torch.complex64 if single else torch.complex128
    S = torch.empty((nb, m, n), dtype=cdt, device=dev)
    K = torch.empty((nb, m, n) if nk == 1 else (nb, m, n, 3), dtype=cdt, device=dev)
    ctx = gf._context()
    check(_lib.load().cpt_set_assembly_precision(ctx.handle, _lib.CPT_ASSEMBLY_COMPLEX64 if single else 0),
          "cpt_set_assembly_precision")
    keep = [t(arr["vertices"]), t(arr["faces"], torch.int32), t(arr["centers"]), t(arr["normals"]),
            t(arr["areas"]), t(arr["radii"]), t(arr["quad_points"]), t(qw)]
    check(_lib.load().cpt_build_matrices(
        ctx.handle, n, ptr(colloc), ptr(dpn), arr["vertices"].shape[0], ptr(keep[0]), m, ptr(keep[1]),
        ptr(keep[2]), ptr(keep[3]), ptr(keep[4]), ptr(keep[5]), qw.shape[1], ptr(keep[6]), ptr(keep[7]),
        nb, ptr(ks), ptr(gfs_arr), float(depth), finite_depth_method, nexp, ptr(pr), int(adjoint),
        int(early_dot_product), int(diagonal_term), float(free_surface), ptr(S), n, n * m, ptr(K), n, n * m * nk,
        stream_handle(dev)), "cpt_build_matrices")
    check(_lib.load().cpt_set_assembly_precision(ctx.handle, 0), "cpt_set_assembly_precision")
    torch.cuda.synchronize()
    S = S.cpu().numpy().transpose(0, 2, 1)
    K = K.cpu().numpy().transpose(0, 2, 1) if nk == 1 else K.cpu().numpy().transpose(0, 3, 2, 1)
    return S, K
