"""Delhommeau Green function evaluated by the gfx950 kernels.

Drop-in for ``capytaine.green_functions.delhommeau.Delhommeau``
(src/capytaine/green_functions/delhommeau.py:44-520): same constructor
keywords, same ``evaluate`` signature and semantics, same tabulation cache
file names and contents, same ``exportable_settings``.  The matrices are
assembled on the GPU by ``cpt_build_matrices``; ``evaluate`` returns them as
Fortran-ordered numpy arrays like the reference, ``evaluate_on_device``
returns the device tensors for the engine.
"""
import logging
import os
from functools import lru_cache
from pathlib import Path

import numpy as np

from . import _lib
from ._lib import CptError, check, ptr, stream_handle

LOG = logging.getLogger(__name__)

__version__ = "0.1.0"


class GreenFunctionEvaluationError(Exception):
    """Same role as capytaine's GreenFunctionEvaluationError
    (abstract_green_function.py:21-22)."""


_default_parameters = dict(
    tabulation_nr=676,
    tabulation_rmax=100.0,
    tabulation_nz=372,
    tabulation_zmin=-251.0,
    tabulation_nb_integration_points=1001,
    tabulation_grid_shape="scaled_nemoh3",
    finite_depth_method="newer",
    finite_depth_prony_decomposition_method="python",
    floating_point_precision="float64",
    gf_singularities="low_freq",
)

_GRID = {"legacy": _lib.CPT_LEGACY_GRID, "scaled_nemoh3": _lib.CPT_SCALED_NEMOH3_GRID}
_SING = {"high_freq": _lib.CPT_HIGH_FREQ, "low_freq": _lib.CPT_LOW_FREQ,
         "low_freq_with_rankine_part": _lib.CPT_LOW_FREQ_WITH_RANKINE_PART}
_FD = {"legacy": _lib.CPT_LEGACY_FINITE_DEPTH, "newer": _lib.CPT_NEWER_FINITE_DEPTH}


def cache_directory():
    """Tabulation cache directory (tools/cache_on_disk.py:25-40):
    $CAPYTAINE_CACHE_DIR/<version> or the XDG cache."""
    if "CAPYTAINE_CACHE_DIR" in os.environ:
        path = os.path.join(os.environ["CAPYTAINE_CACHE_DIR"], "capytaine_amd-" + __version__)
    else:
        base = os.environ.get("XDG_CACHE_HOME", "").strip() or os.path.expanduser("~/.cache")
        path = os.path.join(base, "capytaine_amd", __version__)
    Path(path).mkdir(parents=True, exist_ok=True)
    return path


# ----------------------------------------------------------------------
# Tabulation grid (Delhommeau_integrals.f90:208-270).  The reference
# evaluates the exponential part of the grids in default (float32) real
# arithmetic with the C library's powf; the same libm routines are called
# here, so the grids are bit-identical.
# ----------------------------------------------------------------------
import ctypes as _ct

_libm = _ct.CDLL("libm.so.6")
_libm.powf.restype, _libm.powf.argtypes = _ct.c_float, [_ct.c_float, _ct.c_float]
_libm.pow.restype, _libm.pow.argtypes = _ct.c_double, [_ct.c_double, _ct.c_double]
_libm.log10.restype, _libm.log10.argtypes = _ct.c_double, [_ct.c_double]


def _powf(a, b):
    return np.float32(_libm.powf(float(np.float32(a)), float(np.float32(b))))


def default_r_spacing(nr, rmax, method):
    f32 = np.float32
    out = np.zeros(nr)
    if method == _lib.CPT_LEGACY_GRID:
        for i in range(2, nr + 1):
            a = _powf(10, (f32(i) - f32(1)) / f32(5) - f32(6))
            b = f32(abs(i - 32)) / f32(3) + f32(4) / f32(3)
            out[i - 1] = float(min(a, b))
        return out
    index_of_1 = int(np.round(f32(nr) * f32(1.0) / f32(676) * f32(81)))
    for i in range(1, nr + 1):
        if i < index_of_1:
            e = f32(-10) * (f32(1) - f32(i - 1) / f32(index_of_1 - 1))
            out[i - 1] = float(_powf(10, e))
        else:
            out[i - 1] = (rmax - 1.0) / (nr - index_of_1) * (i - index_of_1) + 1.0
    return out


def default_z_spacing(nz, zmin, method):
    f32 = np.float32
    out = np.zeros(nz)
    if method == _lib.CPT_LEGACY_GRID:
        for j in range(1, nz + 1):
            a = _powf(10, f32(j) / f32(5) - f32(6))
            b = _powf(10, f32(j) / f32(8) - f32(4.5))
            out[j - 1] = -float(min(a, b))
        if nz == 46:
            out[45] = -16.0
        return out
    if nz == 0:
        return out
    dz = (_libm.log10(abs(zmin)) + 10.0) / nz
    for j in range(1, nz + 1):
        out[j - 1] = -min(_libm.pow(10.0, dz * j - 10.0), abs(zmin))
    return out


# ----------------------------------------------------------------------
# Prony decomposition for Delhommeau's finite-depth Green function
# (tools/prony_decomposition.py:26-161, delhommeau.py:269-334).  The
# reference jitters the sampling interval with an unseeded generator; here
# the generator is seeded so that the decomposition is reproducible.
# ----------------------------------------------------------------------
class PronyDecompositionFailure(Exception):
    pass


def _exponential_decomposition(X, F, m):
    from numpy.polynomial import polynomial
    from scipy.linalg import toeplitz
    from scipy.optimize import curve_fit
    A = toeplitz(c=F[m - 1:-1], r=F[:m][::-1])
    P, *_ = np.linalg.lstsq(A, F[m:], rcond=None)
    coeffs = np.ones(m + 1)
    coeffs[:m] = -P[::-1]
    roots = polynomial.polyroots(coeffs)
    roots = roots[np.logical_or(np.imag(roots) != 0.0, np.real(roots) >= 0.0)]
    lamda = np.real(np.log(roots) / (X[1] - X[0]))
    lamda = np.unique(lamda)
    lamda = lamda[np.logical_and(-20.0 < lamda, lamda < 0.0)]

    def f(x, *ar):
        return np.sum(np.asarray(ar)[:, None] * np.exp(lamda[:, None] * x), axis=0)
    a, *_ = curve_fit(f, X, F, p0=np.zeros(lamda.shape))
    return a, lamda


def find_best_exponential_decomposition(f, x_min, x_max, n_exp_range, *, tol=1e-4,
                                        noise_on_domain_points_std=0.01, rng=None):
    rng = np.random.default_rng(0) if rng is None else rng
    for n_exp in n_exp_range:
        x_max_iter = (1 + noise_on_domain_points_std * rng.uniform()) * x_max
        try:
            X = np.linspace(x_min, x_max_iter, 4 * n_exp + 1)
            a, lamda = _exponential_decomposition(X, f(X), n_exp)
            X = np.linspace(x_min, x_max_iter, 8 * n_exp + 1)
            err = np.square(np.sum(np.asarray(a)[:, None] * np.exp(np.asarray(lamda)[:, None] * X), axis=0)
                            - f(X)).mean()
            if err < tol:
                return a, lamda
        except Exception:
            continue
    raise PronyDecompositionFailure(f"No suitable Prony decomposition found in [{x_min}, {x_max}] for tol={tol}")


class AbstractGreenFunction:
    """Collocation points / matrix allocation helpers
    (abstract_green_function.py:25-86)."""

    floating_point_precision = "float64"

    def _get_colocation_points_and_normals(self, mesh1, mesh2, adjoint_double_layer):
        if hasattr(mesh1, "faces_centers"):
            pts = mesh1.faces_centers
            normals = mesh2.faces_normals if not adjoint_double_layer else mesh1.faces_normals
            return pts, normals
        if isinstance(mesh1, np.ndarray) and mesh1.ndim == 2 and mesh1.shape[1] == 3:
            if not adjoint_double_layer:
                normals = mesh2.faces_normals
            else:
                normals = np.zeros((mesh1.shape[0], 3))
            return mesh1, normals
        raise ValueError(f"Unrecognized first input for {self.__class__.__name__}.evaluate:\n{mesh1}")


class Delhommeau(AbstractGreenFunction):
    """The Delhommeau Green function, assembled on the GPU.

    Parameters are those of the reference (delhommeau.py:116-128), plus
    `device` (a torch device, default: the current GPU).
    """

    def __init__(self, *,
                 tabulation_nr=_default_parameters["tabulation_nr"],
                 tabulation_rmax=_default_parameters["tabulation_rmax"],
                 tabulation_nz=_default_parameters["tabulation_nz"],
                 tabulation_zmin=_default_parameters["tabulation_zmin"],
                 tabulation_nb_integration_points=_default_parameters["tabulation_nb_integration_points"],
                 tabulation_grid_shape=_default_parameters["tabulation_grid_shape"],
                 tabulation_cache_dir="default",
                 finite_depth_method=_default_parameters["finite_depth_method"],
                 finite_depth_prony_decomposition_method=_default_parameters[
                     "finite_depth_prony_decomposition_method"],
                 floating_point_precision=_default_parameters["floating_point_precision"],
                 gf_singularities=_default_parameters["gf_singularities"],
                 device=None):
        if floating_point_precision not in ("float32", "float64"):
            raise NotImplementedError(f"Unsupported floating point precision: {floating_point_precision}")
        if tabulation_grid_shape not in _GRID:
            raise KeyError(tabulation_grid_shape)
        if gf_singularities not in _SING:
            raise KeyError(gf_singularities)
        if finite_depth_method not in _FD:
            raise KeyError(finite_depth_method)
        if finite_depth_prony_decomposition_method.lower() != "python":
            raise NotImplementedError("Only the 'python' Prony decomposition is available.")
        self.device = _lib.default_device() if device is None else device
        self.tabulation_grid_shape = tabulation_grid_shape
        self.tabulation_grid_shape_index = _GRID[tabulation_grid_shape]
        self.gf_singularities = gf_singularities
        self.gf_singularities_fortran_enum = dict(_SING)
        self.finite_depth_method = finite_depth_method
        self.finite_depth_method_index = _FD[finite_depth_method]
        self.floating_point_precision = floating_point_precision
        self.tabulation_nb_integration_points = tabulation_nb_integration_points
        self.finite_depth_prony_decomposition_method = finite_depth_prony_decomposition_method
        if tabulation_cache_dir == "default":
            tabulation_cache_dir = cache_directory()
        self.tabulation_cache_dir = tabulation_cache_dir
        self._create_or_load_tabulation(tabulation_nr, float(tabulation_rmax), tabulation_nz,
                                        float(tabulation_zmin), tabulation_nb_integration_points,
                                        tabulation_cache_dir)
        self.exportable_settings = {
            'green_function': self.__class__.__name__,
            'tabulation_nr': tabulation_nr,
            'tabulation_rmax': tabulation_rmax,
            'tabulation_nz': tabulation_nz,
            'tabulation_zmin': tabulation_zmin,
            'tabulation_nb_integration_points': tabulation_nb_integration_points,
            'tabulation_grid_shape': tabulation_grid_shape,
            'finite_depth_method': finite_depth_method,
            'finite_depth_prony_decomposition_method': finite_depth_prony_decomposition_method,
            'floating_point_precision': floating_point_precision,
            'gf_singularities': gf_singularities,
        }
        self._hash = hash(tuple(self.exportable_settings.values()))

    def __hash__(self):
        return self._hash

    def __str__(self):
        shown = [f"{k}={v!r}" for k, v in self.exportable_settings.items()
                 if k in _default_parameters and v != _default_parameters[k]]
        return f"{self.__class__.__name__}({', '.join(shown)})"

    def __repr__(self):
        shown = [f"{k}={v!r}" for k, v in self.exportable_settings.items() if k in _default_parameters]
        return f"{self.__class__.__name__}({', '.join(shown)})"

    # ------------------------------------------------------------------
    # tabulation (delhommeau.py:208-261)
    def _create_or_load_tabulation(self, nr, rmax, nz, zmin, npts, cache_dir):
        filename = "tabulation_{}_{}_{}_{}_{}_{}_{}.npz".format(
            self.floating_point_precision, self.tabulation_grid_shape, nr, rmax, nz, zmin, npts)
        self.tabulation_source = "computed"
        if cache_dir is not None:
            filepath = os.path.join(cache_dir, filename)
            if os.path.exists(filepath):
                try:
                    loaded = np.load(filepath)  # allow_pickle=False
                    r, z, v = loaded["r_range"], loaded["z_range"], loaded["values"]
                    if v.shape == (nr, nz, 5):
                        self._set_tabulation(np.asarray(r, self._real), np.asarray(z, self._real),
                                             np.asarray(v, self._real))
                        self.tabulation_source = filepath
                        return filename
                except (EOFError, FileNotFoundError, KeyError, ValueError, OSError):
                    LOG.warning("Error loading tabulation from %s", filepath)
        self._create_tabulation(nr, rmax, nz, zmin, npts)
        if cache_dir is not None:
            try:
                np.savez_compressed(os.path.join(cache_dir, filename), r_range=self.tabulated_r_range,
                                    z_range=self.tabulated_z_range, values=self.tabulated_integrals)
            except OSError:
                LOG.warning("Could not save the tabulation in %s", cache_dir)
        return filename

    @property
    def _real(self):
        return np.float32 if self.floating_point_precision == "float32" else np.float64

    def _create_tabulation(self, nr, rmax, nz, zmin, npts):
        """The table on the device (bit-identical to the reference's in
        float64).  In float32, the grid is rounded to float32 first and the
        integrals are tabulated in FP64 at those points, then rounded: a table
        consistent with its grid, within float rounding of the exact values
        (the reference's float32 core sums them in float32 instead)."""
        import torch
        r = default_r_spacing(nr, rmax, self.tabulation_grid_shape_index).astype(self._real)
        z = default_z_spacing(nz, zmin, self.tabulation_grid_shape_index).astype(self._real)
        if nr <= 1 or nz <= 1:
            # no usable table: the device integrates the wave term for each
            # point (Green_wave.f90:226-228); the table keeps the reference's
            # shape, its values unused
            self._set_tabulation(r, z, np.zeros((nr, nz, 5), dtype=self._real, order="F"))
            return
        ctx = _lib.Context.get(self.device)
        r_d = torch.as_tensor(r, dtype=torch.float64).to(self.device)
        z_d = torch.as_tensor(z, dtype=torch.float64).to(self.device)
        # (nr, nz, 5) column-major == torch (5, nz, nr) row-major
        tab_d = torch.empty((5, nz, nr), dtype=torch.float64, device=self.device)
        check(_lib.load().cpt_tabulate(ctx.handle, ptr(r_d), nr, ptr(z_d), nz, npts, ptr(tab_d),
                                       stream_handle(self.device)), "cpt_tabulate")
        values = tab_d.cpu().numpy().transpose(2, 1, 0)  # (nr, nz, 5), Fortran-ordered view
        if self._real is np.float32:
            self._set_tabulation(r, z, np.asfortranarray(values, dtype=np.float32))
        else:
            self._set_tabulation(r, z, np.asfortranarray(values), device_copies=(r_d, z_d, tab_d))

    def _set_tabulation(self, r, z, values, device_copies=None):
        """Host copies in the Green function's precision; the device copies
        are FP64 (float32 values widened exactly)."""
        import torch
        self.tabulated_r_range = r
        self.tabulated_z_range = z
        self.tabulated_integrals = values
        if device_copies is None:
            r_d = torch.as_tensor(r, dtype=torch.float64).to(self.device)
            z_d = torch.as_tensor(z, dtype=torch.float64).to(self.device)
            tab_d = torch.as_tensor(np.ascontiguousarray(values.transpose(2, 1, 0)),
                                    dtype=torch.float64).to(self.device)
            device_copies = (r_d, z_d, tab_d)
        self._table_device = device_copies

    @property
    def all_tabulation_parameters(self):
        return (self.tabulation_nb_integration_points, self.tabulation_grid_shape_index,
                self.tabulated_r_range, self.tabulated_z_range, self.tabulated_integrals)

    def _context(self, ctx=None):
        """Library context (the device's shared one by default) with this
        Green function's table loaded."""
        ctx = _lib.Context.get(self.device) if ctx is None else ctx
        # the context holds the uploaded table tensor itself: an identity
        # test on a live object cannot be fooled by a reused id()
        key = self._table_device[2]
        if ctx.table_key is not key:
            r_d, z_d, tab_d = self._table_device
            check(_lib.load().cpt_set_tabulation(ctx.handle, ptr(r_d), r_d.numel(), ptr(z_d), z_d.numel(),
                                                 ptr(tab_d), self.tabulation_grid_shape_index,
                                                 self.tabulation_nb_integration_points,
                                                 stream_handle(self.device)), "cpt_set_tabulation")
            ctx.table_key = key
        return ctx

    # ------------------------------------------------------------------
    @lru_cache(maxsize=128)
    def find_best_exponential_decomposition(self, dimensionless_wavenumber):
        """(2, nexp) array [lambda; a] (delhommeau.py:269-324)."""
        kh = dimensionless_wavenumber
        if kh <= 0.1:
            raise NotImplementedError(f"{self} cannot evaluate finite depth Green function for kh<0.1 (kh={kh})")
        elif kh < 1e5:
            sing_coef = (1 + np.tanh(kh)) ** 2 / (1 - np.tanh(kh) ** 2 + np.tanh(kh) / kh)

            def ref_function(x):
                return ((x + kh * np.tanh(kh)) * np.exp(x)) / (x * np.sinh(x) - kh * np.tanh(kh) * np.cosh(x)) \
                    - sing_coef / (x - kh) - 2
        else:
            def ref_function(x):
                return -2 / (1 + np.exp(-2 * x)) + 2
        try:
            a, lamda = find_best_exponential_decomposition(ref_function, x_min=-0.1, x_max=20.0,
                                                           n_exp_range=range(4, 31, 2), tol=1e-4)
            return np.stack([lamda, a])
        except PronyDecompositionFailure as e:
            raise GreenFunctionEvaluationError(f"{self} cannot evaluate finite depth Green function "
                                               f"for kh={dimensionless_wavenumber}") from e

    def _gf_singularities_for(self, water_depth, wavenumber, mesh2):
        """Overrides of delhommeau.py:466-490."""
        if water_depth < np.inf and self.finite_depth_method == 'legacy' and self.gf_singularities != 'low_freq':
            return "low_freq"
        if wavenumber == 0.0 and self.gf_singularities != 'low_freq':
            return "low_freq"
        if wavenumber == np.inf and self.gf_singularities != 'high_freq':
            return "high_freq"
        if np.any(abs(mesh2.faces_centers[:, 2]) < 1e-6) and self.gf_singularities != 'low_freq':
            LOG.warning(f"Overriding gf_singularities='{self.gf_singularities}' because of free surface panels, "
                        "which are currently only supported by gf_singularities='low_freq'")
            return "low_freq"
        return self.gf_singularities

    # ------------------------------------------------------------------
    def evaluate_on_device(self, mesh1, mesh2, *, free_surface=0.0, water_depth=np.inf, wavenumber=1.0,
                           adjoint_double_layer=True, early_dot_product=True,
                           diagonal_term_in_double_layer=True, out=None):
        """Assemble S and K on the device for one wavenumber or a sequence of
        wavenumbers.

        Returns (S, K) as torch complex128 tensors (complex64 for
        floating_point_precision="float32") holding column-major
        matrices: S has shape (nb, m, n) (that is S[b].T is the (n, m)
        matrix), K has shape (nb, m, n) if early_dot_product else
        (nb, m, n, 3).  A scalar wavenumber gives nb = 1.
        """
        import torch
        lib = _lib.load()
        ks = np.atleast_1d(np.asarray(wavenumber, dtype=np.float64))
        nb = len(ks)
        if free_surface == np.inf and water_depth != np.inf:
            raise ValueError("When setting free_surface=inf, the water depth should also be infinite "
                             f"(got water_depth={water_depth})")
        dev = self.device
        md = mesh2.device_arrays(dev)
        m = mesh2.nb_faces
        if isinstance(mesh1, torch.Tensor):
            # points already in device memory, (n, 3) float64: collocation
            # points of the post-processing, with zero normals (only S and the
            # full gradient are meaningful then, abstract_green_function.py:46-53)
            if not (mesh1.dtype == torch.float64 and mesh1.dim() == 2 and mesh1.shape[1] == 3
                    and mesh1.is_contiguous() and mesh1.device == torch.device(dev)):
                raise ValueError("device points must be a contiguous (n, 3) float64 tensor on the Green "
                                 "function's device")
            if not adjoint_double_layer:
                raise ValueError("device points need adjoint_double_layer=True")
            n = mesh1.shape[0]
            pts_d = mesh1
            dpn_d = torch.zeros_like(mesh1)
        else:
            pts, dpn = self._get_colocation_points_and_normals(mesh1, mesh2, adjoint_double_layer)
            n = pts.shape[0]
            if mesh1 is mesh2:
                pts_d = md["centers"]
                dpn_d = md["normals"]
            else:
                pts_d = torch.as_tensor(np.ascontiguousarray(pts), dtype=torch.float64).to(dev)
                dpn_d = torch.as_tensor(np.ascontiguousarray(dpn), dtype=torch.float64).to(dev)
        gfs = np.array([_SING[self._gf_singularities_for(water_depth, float(k), mesh2)] for k in ks],
                       dtype=np.int32)
        if water_depth == np.inf or free_surface == np.inf:
            nexp, prony = 0, np.zeros(1)
        else:
            decs = [self.find_best_exponential_decomposition(float(k) * water_depth) for k in ks]
            nexp = max(d.shape[1] for d in decs)
            prony = np.zeros((nb, 2, nexp))
            for b, d in enumerate(decs):
                prony[b, :, :d.shape[1]] = d  # missing terms have a = 0
            prony = np.ascontiguousarray(prony)
        nk = 1 if early_dot_product else 3
        single = self.floating_point_precision == "float32"
        cdtype = torch.complex64 if single else torch.complex128
        if out is None:
            S = torch.empty((nb, m, n), dtype=cdtype, device=dev)
            K = torch.empty((nb, m, n) if nk == 1 else (nb, m, n, 3), dtype=cdtype, device=dev)
        else:
            S, K = out
            if S.dtype != cdtype or K.dtype != cdtype:
                raise ValueError(f"out must hold {cdtype} matrices for {self.floating_point_precision}")
        ctx = self._context()
        # complex64 stores for the single-precision form, reset after the
        # call (the context is shared); the FP64 form leaves the context's
        # setting alone (the mixed sweeps set its far field before calling)
        if single:
            check(lib.cpt_set_assembly_precision(ctx.handle, _lib.CPT_ASSEMBLY_COMPLEX64),
                  "cpt_set_assembly_precision")
        status = lib.cpt_build_matrices(
            ctx.handle, n, ptr(pts_d), ptr(dpn_d), mesh2.nb_vertices, ptr(md["vertices"]), m, ptr(md["faces"]),
            ptr(md["centers"]), ptr(md["normals"]), ptr(md["areas"]), ptr(md["radii"]), md["nq"],
            ptr(md["quad_points"]), ptr(md["quad_weights"]), nb, ptr(ks), ptr(gfs), float(water_depth),
            self.finite_depth_method_index, nexp, ptr(prony), int(bool(adjoint_double_layer)),
            int(bool(early_dot_product)), int(bool(diagonal_term_in_double_layer)), float(free_surface),
            ptr(S), n, n * m, ptr(K), n, n * m * nk, stream_handle(dev))
        if single:
            check(lib.cpt_set_assembly_precision(ctx.handle, 0), "cpt_set_assembly_precision")
        check(status, "cpt_build_matrices")
        return S, K

    def evaluate(self, mesh1, mesh2, *, free_surface=0.0, water_depth=np.inf, wavenumber=1.0,
                 adjoint_double_layer=True, early_dot_product=True, diagonal_term_in_double_layer=True):
        """Same contract as the reference's Delhommeau.evaluate
        (delhommeau.py:405-520): returns Fortran-ordered numpy arrays S (n, m)
        and K (n, m) (or (3, n, m) if not early_dot_product); real arrays when
        free_surface is infinite (evaluate_rankine_only, :336-402)."""
        if free_surface == np.inf:
            # evaluate_rankine_only is called without the diagonal flag (delhommeau.py:450-454)
            diagonal_term_in_double_layer = True
        wavenumber = float(wavenumber)
        S, K = self.evaluate_on_device(mesh1, mesh2, free_surface=free_surface, water_depth=water_depth,
                                       wavenumber=wavenumber, adjoint_double_layer=adjoint_double_layer,
                                       early_dot_product=early_dot_product,
                                       diagonal_term_in_double_layer=diagonal_term_in_double_layer)
        S = S[0].cpu().numpy().T
        if early_dot_product:
            K = K[0].cpu().numpy().T
        else:
            K = K[0].cpu().numpy().transpose(2, 1, 0)
        if free_surface == np.inf:
            S, K = np.real(S), np.real(K)
            if np.any(np.isnan(S)) or np.any(np.isnan(K)):
                raise GreenFunctionEvaluationError("Green function returned a NaN in the interaction matrix.\n"
                                                   "It could be due to overlapping panels.")
        return S, K


class FinGreen3D(AbstractGreenFunction):
    """The HAMS finite-depth Green function (hams.py:109-209), assembled on
    the GPU: the Rankine term and its free-surface image integrated on the
    panels, the FinGreen3D wave part (Liu et al. 2018) at the panels'
    quadrature points.  Finite depth with a free surface, finite non-zero
    frequencies only (as the reference); float64.

    nb_dispersion_roots: roots of the dispersion relation passed to the
    series (the reference's default 200; the series read the first 99).
    `device`: a torch device, default the current GPU.
    """

    floating_point_precision = "float64"
    finite_depth_method = "fingreen3d"
    finite_depth_method_index = _lib.CPT_FINGREEN3D_METHOD
    gf_singularities = "low_freq"
    gf_singularities_index = _SING["low_freq"]

    def __init__(self, *, nb_dispersion_roots=200, device=None):
        if nb_dispersion_roots < 100:
            raise ValueError("FinGreen3D's series read the first 99 evanescent roots: "
                             f"nb_dispersion_roots={nb_dispersion_roots} < 100")
        self.nb_dispersion_roots = int(nb_dispersion_roots)
        self.device = _lib.default_device() if device is None else device
        self.exportable_settings = {
            'green_function': "FinGreen3D",
            'nb_dispersion_roots': self.nb_dispersion_roots,
        }

    def __str__(self):
        return f"FinGreen3D(nb_dispersion_roots={self.nb_dispersion_roots})"

    __repr__ = __str__

    def __hash__(self):
        return hash(("FinGreen3D", self.nb_dispersion_roots))

    @staticmethod
    @lru_cache(maxsize=256)
    def _roots(nk, wavenumber, depth):
        from scipy.optimize import brentq
        omega2_h_over_g = wavenumber * np.tanh(wavenumber * depth) * depth

        def root(i_root):
            return brentq(lambda y: omega2_h_over_g + y * np.tan(y), (2 * i_root + 1) * np.pi / 2 + 1e-10,
                          (2 * i_root + 2) * np.pi / 2 - 1e-10) / depth
        out = np.array([wavenumber] + [root(i) for i in range(nk - 1)])
        out.flags.writeable = False
        return out

    def compute_dispersion_relation_roots(self, nk, wavenumber, depth):
        """[k, then the nk - 1 evanescent roots m_i of omega^2/g = -m tan(m h)],
        by Brent's method in each branch of the tangent (hams.py:152-156)."""
        return np.array(self._roots(int(nk), float(wavenumber), float(depth)))

    def evaluate_on_device(self, mesh1, mesh2, *, free_surface=0.0, water_depth=np.inf, wavenumber=1.0,
                           adjoint_double_layer=True, early_dot_product=True,
                           diagonal_term_in_double_layer=True, out=None):
        """As Delhommeau.evaluate_on_device: S (nb, m, n) and K (nb, m, n) or
        (nb, m, n, 3) complex128 device tensors holding column-major
        matrices, one per wavenumber."""
        import torch
        if free_surface == np.inf or water_depth == np.inf:
            raise NotImplementedError("FinGreen3D is only implemented for finite depth with a free surface.")
        ks = np.atleast_1d(np.asarray(wavenumber, dtype=np.float64))
        if np.any(ks == 0.0) or np.any(ks == np.inf):
            raise NotImplementedError("FinGreen3D is only implemented for non-zero and non-infinite frequencies")
        lib = _lib.load()
        nb = len(ks)
        dev = self.device
        md = mesh2.device_arrays(dev)
        m = mesh2.nb_faces
        if isinstance(mesh1, torch.Tensor):
            if not (mesh1.dtype == torch.float64 and mesh1.dim() == 2 and mesh1.shape[1] == 3
                    and mesh1.is_contiguous() and mesh1.device == torch.device(dev)):
                raise ValueError("device points must be a contiguous (n, 3) float64 tensor on the Green "
                                 "function's device")
            if not adjoint_double_layer:
                raise ValueError("device points need adjoint_double_layer=True")
            n = mesh1.shape[0]
            pts_d, dpn_d = mesh1, torch.zeros_like(mesh1)
        else:
            pts, dpn = self._get_colocation_points_and_normals(mesh1, mesh2, adjoint_double_layer)
            n = pts.shape[0]
            if mesh1 is mesh2:
                pts_d, dpn_d = md["centers"], md["normals"]
            else:
                pts_d = torch.as_tensor(np.ascontiguousarray(pts), dtype=torch.float64).to(dev)
                dpn_d = torch.as_tensor(np.ascontiguousarray(dpn), dtype=torch.float64).to(dev)
        nk = 1 if early_dot_product else 3
        if out is None:
            S = torch.empty((nb, m, n), dtype=torch.complex128, device=dev)
            K = torch.empty((nb, m, n) if nk == 1 else (nb, m, n, 3), dtype=torch.complex128, device=dev)
        else:
            S, K = out
            if S.dtype != torch.complex128 or K.dtype != torch.complex128:
                raise ValueError("out must hold complex128 matrices for FinGreen3D")
        ctx = _lib.Context.get(dev)
        for b, k in enumerate(ks):
            roots = self._roots(self.nb_dispersion_roots, float(k), float(water_depth))
            check(lib.cpt_build_matrices_fingreen3d(
                ctx.handle, n, ptr(pts_d), ptr(dpn_d), mesh2.nb_vertices, ptr(md["vertices"]), m, ptr(md["faces"]),
                ptr(md["centers"]), ptr(md["normals"]), ptr(md["areas"]), ptr(md["radii"]), md["nq"],
                ptr(md["quad_points"]), ptr(md["quad_weights"]), float(k), float(water_depth), len(roots),
                roots.ctypes.data, self.gf_singularities_index, int(bool(adjoint_double_layer)),
                int(bool(early_dot_product)), int(bool(diagonal_term_in_double_layer)), float(free_surface),
                ptr(S[b]), n, ptr(K[b]), n, stream_handle(dev)), "cpt_build_matrices_fingreen3d")
        return S, K

    def evaluate(self, mesh1, mesh2, *, free_surface=0.0, water_depth=np.inf, wavenumber=1.0,
                 adjoint_double_layer=True, early_dot_product=True, diagonal_term_in_double_layer=True):
        """hams.py:158-209: Fortran-ordered numpy S (n, m) and K (n, m) (or
        (3, n, m) if not early_dot_product)."""
        S, K = self.evaluate_on_device(mesh1, mesh2, free_surface=free_surface, water_depth=water_depth,
                                       wavenumber=float(wavenumber), adjoint_double_layer=adjoint_double_layer,
                                       early_dot_product=early_dot_product,
                                       diagonal_term_in_double_layer=diagonal_term_in_double_layer)
        S = S[0].cpu().numpy().T
        K = K[0].cpu().numpy().T if early_dot_product else K[0].cpu().numpy().transpose(2, 1, 0)
        if np.any(np.isnan(S)) or np.any(np.isnan(K)):
            raise GreenFunctionEvaluationError("Green function returned a NaN in the interaction matrix.\n"
                                               "It could be due to overlapping panels.")
        return S, K


def require_delhommeau(green_function, what):
    """The sweeps, graphs and the server assemble with the Delhommeau kernels
    (tables, Prony decompositions, batched launches); other Green functions
    go through BEMSolver.solve / solve_all."""
    if not isinstance(green_function, Delhommeau):
        raise NotImplementedError(f"{what} runs the Delhommeau Green function; with {green_function} use "
                                  "BEMSolver.solve or BEMSolver.solve_all")


class HAMS_GF(AbstractGreenFunction):
    """hams.py:212-234: FinGreen3D in finite depth.  Its infinite-depth half,
    the Liang-Wu-Noblesse wave term, is not available on the device."""

    floating_point_precision = "float64"

    def __init__(self, *, device=None):
        self.finite_depth_gf = FinGreen3D(nb_dispersion_roots=200, device=device)
        self.device = self.finite_depth_gf.device
        self.exportable_settings = {'green_function': "HAMS_GF"}

    def __str__(self):
        return "HAMS_GF()"

    __repr__ = __str__

    def _pick(self, water_depth):
        if water_depth == np.inf:
            raise NotImplementedError("HAMS_GF in infinite depth uses the Liang-Wu-Noblesse wave term, which is "
                                      "not available on the device; use Delhommeau()")
        return self.finite_depth_gf

    def evaluate_on_device(self, mesh1, mesh2, *, water_depth=np.inf, **kwargs):
        return self._pick(water_depth).evaluate_on_device(mesh1, mesh2, water_depth=water_depth, **kwargs)

    def evaluate(self, mesh1, mesh2, *, water_depth=np.inf, **kwargs):
        return self._pick(water_depth).evaluate(mesh1, mesh2, water_depth=water_depth, **kwargs)


class XieDelhommeau(Delhommeau):
    """Legacy way to call the gf_singularities="low_freq" variant."""

    def __init__(self, **kwargs):
        super().__init__(gf_singularities="low_freq", **kwargs)
