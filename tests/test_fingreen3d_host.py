"""FinGreen3D on the CPU: the oracle pinned, and the device code's host build
checked against it before any GPU run.

The oracle is the reference's own FinGreen3D.f90, compiled into
oracle/_ref by oracle/Makefile; its outputs are the fixtures of
tests/golden/fingreen3d.npz (make_golden.py fingreen3d).  They are pinned here
by what does not come from that source:
  * HAMS's published S matrices of two panels at k = 1 and 2, h = 10
    (test_consistency_with_HAMS.py:48-66, rtol 1e-3) and its dispersion roots
    (:39-45);
  * the dispersion relation itself (Brent's roots to 1e-12, Newman's
    iteration to the reference's tolerance, test_green_functions_finite_depth
    .py:196-215);
  * John's eigenfunction expansion of the finite-depth Green function summed
    here with scipy's Bessel functions and enough terms to converge, against
    every region of fingreen3d_routine (value and both derivatives);
  * the R = 0 convention (dG/dR = 0).
The device functions (capytaine_amd/csrc/fingreen3d.hpp, __host__
__device__) are built for the host with hipcc and compared with the fixtures:
the special functions bit for bit with the reference's where it is built, the
four regions to rounding, except where the reference's own epsilon-algorithm
limit is ill-conditioned (Pidcock's dG/dR and dG/dzf: a one-ulp change of one
partial sum moves the reference's limit by up to 1e-7 of the 2/(R h) term it
cancels against)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import load_golden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.dirname(os.path.abspath(__file__))
REGIONS = [("eigen", 0.5, np.inf), ("eigene", 0.05, 0.5), ("pidcock", 0.0005, 0.05), ("linton", -1.0, 0.0005)]
NCASES = 5


@pytest.fixture(scope="module")
def fg():
    return load_golden("fingreen3d.npz")


@pytest.fixture(scope="module")
def host_lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("fg_host") / "libfghost.so")
    cmd = ["/opt/rocm/bin/hipcc", "-std=c++17", "-O2", "-fPIC", "-shared", "--offload-host-only", "-x", "hip",
           os.path.join(HERE, "fg_host_check.cpp"), "-o", out]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        pytest.fail("host build of fingreen3d.hpp failed:\n" + res.stderr[-2000:])
    L = ctypes.CDLL(out)
    L.fgh_limes.restype = ctypes.c_double
    return L


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def host_points(L, rr, zf, zp, v, wvn, h):
    rr, zf, zp, wvn = (np.ascontiguousarray(a, dtype=np.float64) for a in (rr, zf, zp, wvn))
    out = np.zeros((rr.size, 3), dtype=np.complex128)
    L.fgh_points(rr.size, _p(rr), _p(zf), _p(zp), ctypes.c_double(v), _p(wvn), ctypes.c_double(h), _p(out))
    return out


def case(fg, c):
    g = lambda k: fg[f"points{c}_{k}"]
    return float(g("h")), float(g("k")), float(g("v")), g("roots"), g("rr"), g("zf"), g("zp"), g("grn")


def region_mask(rr, h, lo, hi):
    t = np.abs(rr / h)
    return (t >= lo) & (t < hi) if lo >= 0 else t < hi


# ---------------------------------------------------------------------------
# the oracle pinned

def john_series(R, zf, zp, h, k, roots, nterms):
    """John's (1950) eigenfunction expansion of the finite-depth Green
    function less 1/r and 1/r1, and its R and zf derivatives, with scipy's
    Bessel functions; the evanescent roots beyond those given are the
    m pi / h - asymptote refined by Newton on the dispersion relation."""
    from scipy.special import j0, j1, y0, y1, k0e, k1e
    nu = k * np.tanh(k * h)
    m = np.arange(1, nterms)
    km = np.empty(nterms - 1)
    given = min(len(roots) - 1, nterms - 1)
    km[:given] = roots[1:given + 1]
    if nterms - 1 > given:
        x = m[given:] * np.pi / h - 1e-3 / m[given:]
        for _ in range(50):  # nu + x tan(x h) = 0 on (m pi/h - pi/2h, m pi/h)
            f = nu + x * np.tan(x * h)
            df = np.tan(x * h) + x * h / np.cos(x * h) ** 2
            x = x - f / df
        km[given:] = x
    n0 = h / 2 * (1 + np.sinh(2 * k * h) / (2 * k * h))
    p0 = np.pi / n0 * np.cosh(k * (zf + h)) * np.cosh(k * (zp + h))
    d0 = np.pi / n0 * k * np.sinh(k * (zf + h)) * np.cosh(k * (zp + h))
    x = k * R
    G = p0 * (1j * j0(x) - y0(x))
    GR = -k * p0 * (1j * j1(x) - y1(x))
    GZ = d0 * (1j * j0(x) - y0(x))
    pm = 4 / (1 + np.sin(2 * km * h) / (2 * km * h)) / h
    xm = km * R
    ex = np.exp(-xm)  # k0e, k1e are scaled by exp(x)
    cf, cp, sf = np.cos(km * (zf + h)), np.cos(km * (zp + h)), np.sin(km * (zf + h))
    G = G + np.sum(pm * k0e(xm) * ex * cf * cp)
    GR = GR - np.sum(pm * km * k1e(xm) * ex * cf * cp)
    GZ = GZ - np.sum(pm * km * k0e(xm) * ex * sf * cp)
    r = np.hypot(R, zf - zp)
    r1 = np.hypot(R, zf + zp)
    G = G - (1 / r + 1 / r1)
    GR = GR + R / r ** 3 + R / r1 ** 3
    GZ = GZ + (zf - zp) / r ** 3 + (zf + zp) / r1 ** 3
    return np.array([G, GR, GZ])


# The reference's accuracy against the converged expansion, per region, as a
# fraction of the case's largest |G|, |dG/dR|, |dG/dzf| (measured on these
# fixtures: at most 7.5e-6 in the eigenfunction regions; (1.6e-4, 5.0e-2,
# 1.2e-3) in Pidcock's; in Linton's, where its Ewald parameter suits the case,
# (2.6e-5, 5.4e-7, 1.7e-4)).  Pidcock's gradient is the least accurate: its
# epsilon limits are as sensitive to the last bits of the partial sums (up to
# 4.7e-2 of the largest dG/dR under 2^-50 relative changes of the inputs,
# recorded with the fixtures).  Linton's form (R/h < 0.0005) misses the
# expansion by 8e-4 to 30 % of the largest value in three cases, (h, k) =
# (10, 2), (40, 0.3) and (2, 0.05): that is the reference's polynomial choice
# of the Ewald parameter in those regimes, kept for parity and reported
# here, not pinned.  The bounds keep a margin over these maxima; a wrong
# argument order, sign or image would miss them by orders of magnitude.
JOHN_TOL = {"eigen": (2e-5, 2e-5, 2e-5), "eigene": (2e-5, 2e-5, 2e-5), "pidcock": (1e-3, 1e-1, 1e-2),
            "linton": (1e-4, 1e-4, 1e-3)}
LINTON_PINNED = (0, 3)


@pytest.mark.parametrize("c", range(NCASES))
def test_oracle_regions_against_john_series(fg, c):
    """Every region of the reference's fingreen3d_routine against the
    eigenfunction expansion summed to convergence with scipy's Bessel
    functions (value and both derivatives), within JOHN_TOL."""
    h, k, v, roots, rr, zf, zp, grn = case(fg, c)
    assert np.isclose(v, k * np.tanh(k * h))
    worst = {}
    for name, lo, hi in REGIONS:
        sel = np.flatnonzero(region_mask(rr, h, lo, hi) & (rr > 1e-4 * h))  # the series needs R > 0
        assert sel.size >= 10
        for i in sel:
            nterms = int(min(400000, max(400, 40 * h / (np.pi * rr[i]))))
            ref = john_series(rr[i], zf[i], zp[i], h, k, roots, nterms)
            e = np.abs(grn[i] - ref) / np.abs(grn).max(axis=0)
            worst[name] = np.maximum(worst.get(name, 0), e)
    print(f"h {h} k {k}:", {n: np.array2string(w, precision=1) for n, w in worst.items()})
    for name, w in worst.items():
        if name == "linton" and c not in LINTON_PINNED:
            continue
        assert np.all(w < np.array(JOHN_TOL[name])), (name, w)


def test_oracle_r_zero_convention(fg):
    for c in range(NCASES):
        h, k, v, roots, rr, zf, zp, grn = case(fg, c)
        assert np.all(grn[rr == 0.0, 1] == 0)
        assert np.all(np.isfinite(grn))


def test_oracle_hams_published_S(fg, orc):
    """test_consistency_with_HAMS.py:48-66: -4 pi (S - S_Rankine) against
    HAMS's published matrices at rtol 1e-3 (S_Rankine from the C oracle's
    panel integral of 1/r)."""
    published = {
        "hams_k1": np.array([[0.73247863590752549 + 2.3114545333287739j, -0.43898677720635493 + 2.2883993524884838j],
                             [-0.43898677720635493 + 2.2883993524884838j, 0.73247863590752549 + 2.3114545333287739j]]),
        "hams_k2": np.array([[-0.36177831640996905 + 1.7007340213300215j, -1.7129415732152193 + 1.6333866990767409j],
                             [-1.7129415732152193 + 1.6333866990767409j, -0.36177831640996905 + 1.7007340213300215j]]),
    }
    for name, hams in published.items():
        g = lambda key: fg[f"{name}__{key}"]
        verts, faces = g("vertices"), g("faces")
        n = faces.shape[0]
        rank = np.zeros((n, n))
        for i in range(n):
            for j in range(n):
                gi, _ = orc.rankine(0, g("colloc")[i:i + 1], verts[faces[j]][None], g("centers")[j:j + 1],
                                    g("normals")[j:j + 1], g("areas")[j:j + 1], g("radii")[j:j + 1],
                                    deriv_first=False)
                rank[i, j] = -gi[0] / (4 * np.pi)
        got = -4 * np.pi * (g("S") - rank)
        np.testing.assert_allclose(got, hams * g("areas")[0], rtol=1e-3)


def test_dispersion_roots(fg):
    """HAMS's roots (test_consistency_with_HAMS.py:39-45) against the
    reference's Newman iteration, rtol 1e-3; Brent's roots of the package
    (FinGreen3D.compute_dispersion_relation_roots, hams.py:152-156) solve
    the dispersion relation to 1e-12 and agree with Newman's to 1e-4
    (test_green_functions_finite_depth.py:196-215)."""
    from capytaine_amd.green_functions import FinGreen3D
    hams = np.array([0.99996435, 0.17434080, 0.51912325, 0.85620867, 1.1866127, 1.5123763, 1.8350620, 2.1556817,
                     2.4748751, 2.7930533, 3.1104880, 3.4273639, 3.7438106, 4.0599209, 4.3757631, 4.6913883,
                     5.0068356, 5.3221352, 5.6373110, 5.9523820])
    np.testing.assert_allclose(fg["roots_hams_k1_h10"], hams, rtol=1e-3)
    for om in (0.1, 1.0, 10.0):
        for h in (0.1, 1.0, 10.0):
            newman = fg[f"roots_newman_om{om}_h{h}"]
            nu = om ** 2 / 9.81
            k = newman[0]
            assert abs(nu - k * np.tanh(k * h)) <= 1e-3 * nu
            brent = FinGreen3D._roots(100, float(k), h)
            nu_b = k * np.tanh(k * h)
            assert np.all(np.abs(nu_b + brent[1:] * np.tan(brent[1:] * h)) <= 1e-12 * (1 + brent[1:] * h) * 1e3)
            np.testing.assert_allclose(brent[1:], newman[1:], rtol=1e-4)


# ---------------------------------------------------------------------------
# the device code, built for the host

@pytest.mark.parametrize("c", range(NCASES))
def test_host_build_points_match_oracle(fg, host_lib, c):
    h, k, v, roots, rr, zf, zp, grn = case(fg, c)
    got = host_points(host_lib, rr, zf, zp, v, roots, h)
    for name, lo, hi in REGIONS:
        sel = region_mask(rr, h, lo, hi)
        a, b = got[sel], grn[sel]
        scale = np.abs(grn).max(axis=0)
        e = np.abs(a - b) / scale
        print(f"h {h} k {k} {name}: max err / max |component| {np.array2string(e.max(axis=0), precision=1)}, "
              f"bit-equal {np.mean(a == b):.2f}")
        if name == "pidcock":
            # the epsilon limits: G within 1e-8 of the largest G, the
            # gradient within 1e-6 of the 2/(R h) term its series cancels
            assert np.all(e[:, 0] <= 1e-8)
            tol = 1e-6 * 2.0 / (rr[sel] * h)
            assert np.all(np.abs(a[:, 1:] - b[:, 1:]) <= tol[:, None] + 1e-12 * scale[1:])
        else:
            assert np.all(e <= 1e-12)


def test_host_build_special_functions_bitwise(host_lib):
    """The special functions of the device code against the reference's
    routines, bit for bit (J/Y: all but rare last-bit differences of the
    elementary functions)."""
    try:
        from oracle import Reference
        R = Reference()
    except (ImportError, FileNotFoundError, OSError):
        pytest.skip("oracle/_ref is not built (needs the reference tree)")
    L = host_lib
    rng = np.random.default_rng(4)
    cases = {0: ("ik01a", np.concatenate([rng.uniform(1e-4, 40, 2000), np.geomspace(1e-5, 700, 300)])),
             1: ("jy01b", rng.uniform(1e-4, 300, 2000)), 2: ("e1xa", np.geomspace(1e-6, 60, 500)),
             3: ("eix", np.geomspace(1e-6, 60, 500)), 4: ("error", np.linspace(-8, 8, 1001)),
             5: ("enxa", np.geomspace(1e-6, 60, 500))}
    for kind, (name, x) in cases.items():
        x = np.ascontiguousarray(x)
        mine = np.zeros((x.size, 4))
        L.fgh_special(x.size, kind, 7, _p(x), _p(mine))
        ref = R.fg_special(name, x, 7)
        rel = np.abs(mine - ref) / np.maximum(np.abs(ref), 1e-300)
        print(f"{name}: bit-equal {np.mean(mine == ref):.4f}, max rel {rel.max():.1e}")
        assert np.mean(mine == ref) >= 0.999 and rel.max() <= 1e-15
    # the epsilon limit of partial sums of alternating and noisy series
    for trial in range(20):
        n = int(rng.integers(3, 99))
        terms = (-1.0) ** np.arange(n) / (1 + np.arange(n)) ** rng.uniform(0.5, 2.0) + 1e-3 * rng.normal(size=n)
        s = np.ascontiguousarray(np.concatenate([[0.0], np.cumsum(terms)]))
        assert L.fgh_limes(s.size - 1, _p(s)) == R.fg_limes(s)[0]
