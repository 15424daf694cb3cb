"""ORACLE -- test infrastructure only.

ctypes wrappers of
  * liboracle.so: the C restatement of the reference libDelhommeau core
    (delhommeau_oracle.c), and
  * _ref/libdelhommeau_ref.so: the reference's own Fortran core, compiled
    from the reference tree by oracle/Makefile (present only where the
    reference tree was available at build time),
with identical Python signatures, plus a CPU pipeline (oracle assembly +
scipy LU, as capytaine's BEMSolver does) for end-to-end coefficients.

Only tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py
use this module.  The product package never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libdelhommeau_ref.so")

_P, _I, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double


def build(ref=True):
    """Compile liboracle.so (and the reference core when its sources exist)."""
    target = "all" if ref else "oracle"
    subprocess.run(["make", "-s", "-C", HERE, target], check=True)


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else ctypes.c_void_p(0)


def _f(a, dtype=np.float64):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(np.asarray(a, dtype=dtype))


class _Base:
    def r_spacing(self, nr, rmax, method=1):
        out = np.zeros(nr)
        self._r_spacing(nr, _D(rmax), method, _p(out))
        return out

    def z_spacing(self, nz, zmin, method=1):
        out = np.zeros(nz)
        self._z_spacing(nz, _D(zmin), method, _p(out))
        return out

    def numerical_integration(self, r, z, npts=1001):
        out = np.zeros(5)
        self._numint(_D(r), _D(z), npts, _p(out))
        return out

    def asymptotic(self, r, z):
        out = np.zeros(5)
        self._asym(_D(r), _D(z), _p(out))
        return out

    def tabulation(self, r_range, z_range, npts=1001):
        r_range, z_range = _f(r_range), _f(z_range)
        tab = np.zeros((len(r_range), len(z_range), 5), order="F")
        self._tab(r_range, z_range, npts, tab)
        return tab

    def wave_part(self, x0i, x0j, k, table, gfs=1, grid=1, npts=1001):
        """interface.f90:29-67 -> (g (np,), nablag (np, 3)) complex."""
        r_range, z_range, tab = table
        x0i, x0j = _f(x0i), _f(x0j)
        n = x0i.shape[0]
        g = np.zeros(n, dtype=np.complex128)
        ng = np.zeros((n, 3), dtype=np.complex128, order="F")
        self._wave(n, _p(x0i), _p(x0j), _D(k), npts, grid, len(r_range), _p(_f(r_range)), len(z_range),
                   _p(_f(z_range)), _p(_f(tab)), gfs, _p(g), _p(ng))
        return g, ng

    def rankine(self, kind, m, nodes, centers, normals, areas, radii, deriv_first=True, coefs=(-1.0, 0.0)):
        m, nodes, centers, normals = _f(m), _f(nodes), _f(centers), _f(normals)
        areas, radii, coefs = _f(areas), _f(radii), _f(coefs)
        n = m.shape[0]
        g = np.zeros(n)
        ng = np.zeros((n, 3), order="F")
        self._rankine(n, kind, _p(m), _p(nodes), _p(centers), _p(normals), _p(areas), _p(radii),
                      int(deriv_first), _p(coefs), _p(g), _p(ng))
        return g, ng


class Oracle(_Base):
    """The C restatement (delhommeau_oracle.c)."""

    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            build(ref=False)
        L = ctypes.CDLL(path)
        self.lib = L
        self._r_spacing = L.orc_default_r_spacing
        self._z_spacing = L.orc_default_z_spacing
        self._numint = L.orc_numerical_integration
        self._asym = L.orc_asymptotic_approximations
        self._wave = L.orc_wave_part_infinite_depth
        self._rankine = L.orc_rankine

    def _tab(self, r, z, npts, tab):
        self.lib.orc_construct_tabulation(_p(r), len(r), _p(z), len(z), npts, _p(tab))

    def build_matrices(self, colloc, dpn, vertices, faces, centers, normals, areas, radii, quad_points,
                       quad_weights, wavenumber, depth, table, *, grid=1, npts=1001, finite_depth_method=1,
                       prony=None, gf_singularities=1, adjoint=True, early_dot_product=True,
                       diagonal_term=True, free_surface=0.0):
        """matrices.f90:28-364 (+ add_diagonal_term).  Faces 0-based.
        Returns S (n, m) and K (n, m) or (3, n, m), Fortran-ordered."""
        colloc, dpn, vertices, centers, normals = map(_c, (colloc, dpn, vertices, centers, normals))
        areas, radii = _c(areas), _c(radii)
        faces = _c(faces, np.int32)
        qp, qw = _c(quad_points), _c(quad_weights)
        n, m = colloc.shape[0], faces.shape[0]
        nq = qw.shape[1]
        nk = 1 if early_dot_product else 3
        S = np.zeros((n, m), dtype=np.complex128, order="F")
        K = np.zeros((nk, n, m), dtype=np.complex128, order="F")
        if free_surface == np.inf:
            self.lib.orc_rankine_only_matrices(n, _p(colloc), _p(dpn), _p(vertices), m, _p(faces), _p(centers),
                                               _p(normals), _p(areas), _p(radii), int(adjoint), nk, _p(S), _p(K))
        else:
            r_range, z_range, tab = (_f(t) for t in table)
            if prony is None:
                lam, amp, nexp = np.zeros(1), np.zeros(1), 0
            else:
                prony = np.asarray(prony, float)
                lam, amp, nexp = _c(prony[0]), _c(prony[1]), prony.shape[1]
            self.lib.orc_build_matrices(n, _p(colloc), _p(dpn), vertices.shape[0], _p(vertices), m, _p(faces),
                                        _p(centers), _p(normals), _p(areas), _p(radii), nq, _p(qp), _p(qw),
                                        _D(wavenumber), _D(depth), npts, grid, len(r_range), _p(r_range),
                                        len(z_range), _p(z_range), _p(tab), finite_depth_method, nexp, _p(lam),
                                        _p(amp), gf_singularities, int(adjoint), nk, _p(S), _p(K))
        if diagonal_term:
            nd = min(n, m)
            self.lib.orc_add_diagonal_term(nd, _p(_c(centers[:nd])), _p(dpn), _D(free_surface), nk, n, _p(K))
        if early_dot_product:
            K = K.reshape((n, m))
        if free_surface == np.inf:
            S, K = np.real(S), np.real(K)
        return S, K

    def count_near_field(self, colloc, centers, radii):
        a, b = ctypes.c_long(), ctypes.c_long()
        colloc, centers, radii = _c(colloc), _c(centers), _c(radii)
        self.lib.orc_count_near_field(colloc.shape[0], _p(colloc), centers.shape[0], _p(centers), _p(radii),
                                      ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value


class Reference(_Base):
    """The reference Fortran core (amdflang build of libs/libDelhommeau/src)."""

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} was not built (the reference tree is needed to build it)")
        L = ctypes.CDLL(path)
        self.lib = L
        self._r_spacing = L.ref_default_r_spacing
        self._z_spacing = L.ref_default_z_spacing
        self._numint = L.ref_numerical_integration
        self._asym = L.ref_asymptotic_approximations
        self._wave = L.ref_wave_part_infinite_depth
        self._rankine = L.ref_rankine

    def _tab(self, r, z, npts, tab):
        self.lib.ref_construct_tabulation(len(r), _p(r), len(z), _p(z), npts, _p(tab))

    def build_matrices(self, colloc, dpn, vertices, faces, centers, normals, areas, radii, quad_points,
                       quad_weights, wavenumber, depth, table, *, grid=1, npts=1001, finite_depth_method=1,
                       prony=None, gf_singularities=1, adjoint=True, early_dot_product=True,
                       diagonal_term=True, free_surface=0.0):
        colloc, dpn, vertices, centers, normals = map(_f, (colloc, dpn, vertices, centers, normals))
        areas, radii = _f(areas), _f(radii)
        faces1 = _f(np.asarray(faces) + 1, np.int32)
        qp, qw = _f(quad_points), _f(quad_weights)
        n, m = colloc.shape[0], faces1.shape[0]
        nq = qw.shape[1]
        nk = 1 if early_dot_product else 3
        S = np.zeros((n, m), dtype=np.complex128, order="F")
        K = np.zeros((nk, n, m), dtype=np.complex128, order="F")
        r_range, z_range, tab = (_f(t) for t in table)
        prony = np.zeros((2, 1), order="F") if prony is None else _f(prony)
        self.lib.ref_build_matrices(n, _p(colloc), dpn.shape[0], _p(dpn), vertices.shape[0], m, _p(vertices),
                                    _p(faces1), _p(centers), _p(normals), _p(areas), _p(radii), nq, _p(qp), _p(qw),
                                    _D(wavenumber), _D(depth), npts, grid, len(r_range), _p(r_range), len(z_range),
                                    _p(z_range), _p(tab), finite_depth_method, prony.shape[1], _p(prony),
                                    gf_singularities, int(adjoint), nk, _p(S), _p(K))
        if diagonal_term:
            nd = min(n, m)
            self.lib.ref_add_diagonal_term(nd, _p(_f(centers[:nd])), dpn.shape[0], _p(dpn), _D(free_surface), nk,
                                           n, m, _p(K))
        if early_dot_product:
            K = K.reshape((n, m))
        return S, K

    # ---- FinGreen3D (FinGreen3D.f90, compiled in with Green_wave.f90's define)
    def fingreen3d(self, rr, zf, zp, v, wvn, h):
        """fingreen3d_routine (FinGreen3D.f90:71) per point: (np, 3) complex,
        [G, dG/dR, dG/dzf] of the wave part (1/r and 1/r1 taken out)."""
        rr, zf, zp = (_f(np.atleast_1d(a)) for a in (rr, zf, zp))
        wvn = _f(wvn)
        out = np.zeros((rr.shape[0], 3), dtype=np.complex128, order="F")
        self.lib.ref_fingreen3d(rr.shape[0], _p(rr), _p(zf), _p(zp), _D(v), wvn.shape[0], _p(wvn), _D(h), _p(out))
        return out

    def fg_dispersion_roots(self, nk, omega2_over_g, depth):
        """compute_dispersion_roots (Green_wave.f90:523-541): Newman's iteration."""
        out = np.zeros(nk)
        self.lib.ref_fg_dispersion_roots(nk, _D(omega2_over_g), _D(depth), _p(out))
        return out

    FG_SPECIAL = {"ik01a": 0, "jy01b": 1, "e1xa": 2, "eix": 3, "error": 4, "enxa": 5}

    def fg_special(self, name, x, n=0):
        """FinGreen3D's special functions per point: (np, 4) with the outputs
        in the routine's order (unused columns zero)."""
        x = _f(np.atleast_1d(x))
        out = np.zeros((x.shape[0], 4))
        outf = np.zeros((4, x.shape[0]), order="F")
        self.lib.ref_fg_special(x.shape[0], self.FG_SPECIAL[name], int(n), _p(x), _p(outf))
        out[:] = outf.T
        return out

    def fg_limes(self, partial_sums):
        """Limes (FinGreen3D.f90:3053) of the partial sums s(0:n): (limit, error)."""
        s = _f(partial_sums)
        out = np.zeros(2)
        self.lib.ref_fg_limes(s.shape[0] - 1, _p(s), _p(out))
        return out[0], out[1]

    def build_matrices_fingreen3d(self, colloc, dpn, vertices, faces, centers, normals, areas, radii, quad_points,
                                  quad_weights, wavenumber, depth, roots, *, adjoint=True, early_dot_product=True,
                                  diagonal_term=True, free_surface=0.0, gf_singularities=1):
        """build_matrices with finite_depth_method FINGREEN3D_METHOD (-1) and the
        dispersion roots (hams.py:158-209)."""
        colloc, dpn, vertices, centers, normals = map(_f, (colloc, dpn, vertices, centers, normals))
        areas, radii = _f(areas), _f(radii)
        faces1 = _f(np.asarray(faces) + 1, np.int32)
        qp, qw = _f(quad_points), _f(quad_weights)
        roots = _f(roots)
        n, m = colloc.shape[0], faces1.shape[0]
        nq = qw.shape[1]
        nk = 1 if early_dot_product else 3
        S = np.zeros((n, m), dtype=np.complex128, order="F")
        K = np.zeros((nk, n, m), dtype=np.complex128, order="F")
        self.lib.ref_build_matrices_roots(n, _p(colloc), dpn.shape[0], _p(dpn), vertices.shape[0], m, _p(vertices),
                                          _p(faces1), _p(centers), _p(normals), _p(areas), _p(radii), nq, _p(qp),
                                          _p(qw), _D(wavenumber), _D(depth), -1, roots.shape[0], _p(roots),
                                          gf_singularities, int(adjoint), nk, _p(S), _p(K))
        if diagonal_term:
            nd = min(n, m)
            self.lib.ref_add_diagonal_term(nd, _p(_f(centers[:nd])), dpn.shape[0], _p(dpn), _D(free_surface), nk,
                                           n, m, _p(K))
        if early_dot_product:
            K = K.reshape((n, m))
        return S, K


def mesh_arrays(mesh):
    """The arrays build_matrices takes, from a capytaine_amd Mesh."""
    qp, qw = mesh.quadrature_points
    return dict(colloc=mesh.faces_centers, dpn=mesh.faces_normals, vertices=mesh.vertices, faces=mesh.faces,
                centers=mesh.faces_centers, normals=mesh.faces_normals, areas=mesh.faces_areas,
                radii=mesh.faces_radiuses, quad_points=qp, quad_weights=qw)


def solve_coefficients(S, K, body, problems):
    """CPU pipeline of BEMSolver._solve (solver.py:188-207) with scipy LU
    (block_circulant_matrices.py:374-385): returns the forces of every
    problem (dict dof -> complex), all sharing S and K."""
    from scipy.linalg import lu_factor, lu_solve
    lu = lu_factor(K)
    out = []
    for pb in problems:
        sources = lu_solve(lu, pb.boundary_condition.astype(np.complex128))
        potential = S @ sources
        pressure = 1j * pb.omega * pb.rho * potential
        out.append(body.integrate_pressure(pressure))
    return out
