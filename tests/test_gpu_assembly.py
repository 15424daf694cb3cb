"""Parity of the HIP assembly path (cpt_tabulate, cpt_build_matrices,
cpt_wave_part_infinite_depth) with the reference core's golden vectors and
with the oracle, on the GPU.

Tolerances (FP64 path): normwise relative error <= 1e-12 on S and K, and
elementwise |a - b| <= 1e-9 max(|b|, 1e-8 max|b|) -- the kernels use fused
multiply-adds and a different (but exactly rounded) sqrt/log/atan, so
agreement is to rounding, not bitwise."""
import numpy as np
import pytest
from conftest import elem_ok, rel_err

from gpu_helpers import build_raw, device, ref_table_gf

pytestmark = pytest.mark.gpu

RTOL_NORM = 1e-12


@pytest.fixture(scope="module")
def gf_ref(default_table):
    print("[gpu] Delhommeau with the reference table", flush=True)
    return ref_table_gf(default_table)


def _table_sha(tab):
    import hashlib
    return hashlib.sha256(np.asfortranarray(tab).tobytes(order="F")).hexdigest()


def test_tabulation_kernel_vs_reference(default_table, core_golden):
    """cpt_tabulate reproduces the reference table bit for bit: the device
    repeats the reference's arithmetic and takes the transcendental values
    (cos, exp, sin, the float complex log of exp_e1) from the host C
    library, as the reference does."""
    from capytaine_amd import Delhommeau
    print("[gpu] tabulation kernel", flush=True)
    gf = Delhommeau(tabulation_cache_dir=None, device=device())
    tab = gf.tabulated_integrals
    ref = default_table[2]
    assert np.array_equal(gf.tabulated_r_range, default_table[0])
    assert np.array_equal(gf.tabulated_z_range, default_table[1])
    frac = np.mean(tab == ref)
    print(f"[gpu] tabulation: {100 * frac:.2f}% of entries bit-identical to the reference", flush=True)
    assert _table_sha(tab) == str(core_golden["table_scaled_sha256"])
    assert np.array_equal(tab, ref)


def test_legacy_tabulation_kernel_vs_reference(legacy_table, core_golden):
    """The legacy grid (r = 0 first, signed zeros in the log branch) too."""
    from capytaine_amd import Delhommeau
    print("[gpu] legacy tabulation kernel", flush=True)
    gf = Delhommeau(tabulation_cache_dir=None, tabulation_grid_shape="legacy", tabulation_nr=328,
                    tabulation_nz=46, tabulation_zmin=-16.0, tabulation_nb_integration_points=251, device=device())
    assert _table_sha(gf.tabulated_integrals) == str(core_golden["table_legacy_sha256"])
    assert np.array_equal(gf.tabulated_integrals, legacy_table[2])


def test_default_delhommeau_vs_oracle(orc, core_golden, tmp_path, monkeypatch):
    """The path a user gets: Delhommeau() with its defaults and an empty
    cache computes its table on the device, bit-identical to the reference's
    (SHA-256 of the reference table); S and K of the 450-panel sphere and a
    160-row stripe of the 4352-panel barge then match the oracle on the
    reference table at 1e-12 normwise.  A second Delhommeau() loads the
    cached file and gets the same table."""
    from capytaine_amd import Delhommeau, mesh_parallelepiped, mesh_sphere
    from oracle import mesh_arrays
    monkeypatch.setenv("CAPYTAINE_CACHE_DIR", str(tmp_path))
    print("[gpu] default Delhommeau(): table on the device", flush=True)
    gf = Delhommeau()
    assert gf.tabulation_source == "computed"
    assert _table_sha(gf.tabulated_integrals) == str(core_golden["table_scaled_sha256"])
    table = (gf.tabulated_r_range, gf.tabulated_z_range, gf.tabulated_integrals)
    gf2 = Delhommeau()
    assert gf2.tabulation_source != "computed"
    assert _table_sha(gf2.tabulated_integrals) == str(core_golden["table_scaled_sha256"])

    print("[gpu] default Delhommeau(): sphere 450 vs oracle", flush=True)
    mesh = mesh_sphere(radius=1.0, resolution=(30, 30)).immersed_part()
    for k in (0.1, 1.0 / 9.81, 0.64, 1.0):
        S, K = gf.evaluate(mesh, mesh, wavenumber=k)
        So, Ko = orc.build_matrices(**mesh_arrays(mesh), wavenumber=k, depth=np.inf, table=table)
        print(f"[gpu]   k={k:.3f}: S {rel_err(S, So):.1e}, K {rel_err(K, Ko):.1e}", flush=True)
        assert rel_err(S, So) < RTOL_NORM and rel_err(K, Ko) < RTOL_NORM

    print("[gpu] default Delhommeau(): barge 4352 stripe vs oracle", flush=True)
    mesh = mesh_parallelepiped(size=(20, 8, 2), center=(0, 0, -1), resolution=(80, 32, 8), missing_sides={"top"})
    ks = [0.2 ** 2 / 9.81, 1.0 / 9.81, 2.5 ** 2 / 9.81]
    S, K = gf.evaluate_on_device(mesh, mesh, wavenumber=ks)
    arr = mesh_arrays(mesh)
    rows = 160
    sub = dict(arr, colloc=arr["colloc"][:rows], dpn=arr["dpn"][:rows])
    for b, k in enumerate(ks):
        So, Ko = orc.build_matrices(**sub, wavenumber=k, depth=np.inf, table=table)
        Sd = S[b].cpu().numpy().T[:rows]
        Kd = K[b].cpu().numpy().T[:rows]
        assert rel_err(Sd, So) < RTOL_NORM and rel_err(Kd, Ko) < RTOL_NORM
        assert elem_ok(Sd, So, 1e-9, 1e-8) and elem_ok(Kd, Ko, 1e-9, 1e-8)


@pytest.mark.parametrize("gfs", [0, 1, 2])
@pytest.mark.parametrize("k", [0.3, 2.0])
def test_wave_part_golden(gf_ref, core_golden, gfs, k):
    import torch
    from capytaine_amd import _lib
    from capytaine_amd._lib import check, ptr, stream_handle
    dev = device()
    x0i = torch.as_tensor(core_golden["wave_x0i"]).to(dev)
    x0j = torch.as_tensor(core_golden["wave_x0j"]).to(dev)
    n = x0i.shape[0]
    g = torch.empty(n, dtype=torch.complex128, device=dev)
    ng = torch.empty((n, 3), dtype=torch.complex128, device=dev)
    ctx = gf_ref._context()
    check(_lib.load().cpt_wave_part_infinite_depth(ctx.handle, n, ptr(x0i), ptr(x0j), k, gfs, ptr(g), ptr(ng),
                                                   stream_handle(dev)), "cpt_wave_part_infinite_depth")
    g, ng = g.cpu().numpy(), ng.cpu().numpy()
    gr, ngr = core_golden[f"wave_g_gfs{gfs}_k{k}"], core_golden[f"wave_ng_gfs{gfs}_k{k}"]
    assert rel_err(g, gr) < RTOL_NORM and rel_err(ng, ngr) < RTOL_NORM
    assert elem_ok(g, gr, 1e-9, 1e-8) and elem_ok(ng, ngr, 1e-9, 1e-8)


def _case(bm, name):
    keys = ["colloc", "dpn", "vertices", "faces", "centers", "normals", "areas", "radii", "quad_points",
            "quad_weights"]
    arr = {k: bm[f"{name}__{k}"] for k in keys}
    opts = {k[len(name) + 6:]: bm[k] for k in bm.files if k.startswith(f"{name}__opt_")}
    opts = {k: (v if v.ndim else v.item()) for k, v in opts.items()}
    kw = dict(depth=float(opts.pop("depth")))
    wavenumber = opts.pop("wavenumber")
    if "gf_singularities" in opts:
        kw["gfs"] = int(opts.pop("gf_singularities"))
    for key in ("adjoint", "early_dot_product", "finite_depth_method", "prony"):
        if key in opts:
            kw[key] = opts.pop(key)
    assert not opts
    return arr, wavenumber, kw


BM_CASES = ["sphere_k1", "sphere_k01_hf", "sphere_k3_rp", "sphere_k1_direct", "sphere_k1_full", "sphere_k0",
            "sphere_kinf", "barge_k06", "barge_lid_k06", "barge_fd_newer", "barge_fd_legacy", "sphere_gl2_k1"]


@pytest.mark.parametrize("name", BM_CASES)
def test_build_matrices_golden(gf_ref, bm_golden, name):
    print(f"[gpu] build_matrices golden case {name}", flush=True)
    arr, k, kw = _case(bm_golden, name)
    S, K = build_raw(gf_ref, arr, k, **kw)
    Sg, Kg = bm_golden[f"{name}__S"], bm_golden[f"{name}__K"]
    assert rel_err(S[0], Sg) < RTOL_NORM, rel_err(S[0], Sg)
    assert rel_err(K[0], Kg) < RTOL_NORM, rel_err(K[0], Kg)
    assert elem_ok(S[0], Sg, 1e-9, 1e-8)
    assert elem_ok(K[0], Kg, 1e-9, 1e-8)


@pytest.mark.parametrize("name", ["points_sphere_k07", "cylinder_faces_sphere_k07", "points_barge_fd"])
def test_post_processing_matrices_golden(gf_ref, pp_golden, name):
    """Points (and another mesh's faces) as collocation, full gradient."""
    print(f"[gpu] post-processing golden case {name}", flush=True)
    arr, k, kw = _case(pp_golden, name)
    S, K = build_raw(gf_ref, arr, k, **kw)
    Sg, Kg = pp_golden[f"{name}__S"], pp_golden[f"{name}__K"]
    assert rel_err(S[0], Sg) < RTOL_NORM and rel_err(K[0], Kg) < RTOL_NORM
    assert elem_ok(S[0], Sg, 1e-9, 1e-8) and elem_ok(K[0], Kg, 1e-9, 1e-8)


def test_sphere_450_vs_oracle(gf_ref, orc, default_table):
    """Sphere R = 1, 30 x 30 panels clipped at z = 0 (450 panels), k = 1/g."""
    from capytaine_amd import mesh_sphere
    from oracle import mesh_arrays
    print("[gpu] sphere 450 panels vs oracle", flush=True)
    mesh = mesh_sphere(radius=1.0, resolution=(30, 30)).immersed_part()
    assert mesh.nb_faces == 450
    k = 1.0 / 9.81
    S, K = gf_ref.evaluate(mesh, mesh, wavenumber=k)
    So, Ko = orc.build_matrices(**mesh_arrays(mesh), wavenumber=k, depth=np.inf, table=default_table)
    assert rel_err(S, So) < RTOL_NORM and rel_err(K, Ko) < RTOL_NORM
    assert elem_ok(S, So, 1e-9, 1e-8) and elem_ok(K, Ko, 1e-9, 1e-8)
    assert S.flags.f_contiguous and S.shape == (450, 450) and S.dtype == np.complex128


def test_barge_4352_stripe_vs_oracle(gf_ref, orc, default_table):
    """Barge 20 x 8 x 2 m, 0.25 m panels (4352 panels): full matrices on the
    GPU for the extreme wavenumbers of the 0.2-2.5 rad/s sweep, compared
    with the oracle on the first 160 rows and on the diagonal term."""
    from capytaine_amd import mesh_parallelepiped
    from oracle import mesh_arrays
    print("[gpu] barge 4352 panels stripe vs oracle", flush=True)
    mesh = mesh_parallelepiped(size=(20, 8, 2), center=(0, 0, -1), resolution=(80, 32, 8), missing_sides={"top"})
    assert mesh.nb_faces == 4352
    ks = [0.2 ** 2 / 9.81, 2.5 ** 2 / 9.81]
    S, K = gf_ref.evaluate_on_device(mesh, mesh, wavenumber=ks)
    arr = mesh_arrays(mesh)
    rows = 160
    sub = dict(arr, colloc=arr["colloc"][:rows], dpn=arr["dpn"][:rows])
    for b, k in enumerate(ks):
        So, Ko = orc.build_matrices(**sub, wavenumber=k, depth=np.inf, table=default_table)
        Sd = S[b].cpu().numpy().T[:rows]
        Kd = K[b].cpu().numpy().T[:rows]
        assert rel_err(Sd, So) < RTOL_NORM and rel_err(Kd, Ko) < RTOL_NORM
        assert elem_ok(Sd, So, 1e-9, 1e-8) and elem_ok(Kd, Ko, 1e-9, 1e-8)


def test_batched_wavenumbers_match_single(gf_ref):
    """One launch for several wavenumbers gives the same bits as one launch
    per wavenumber (the shared Rankine part is added in the same order)."""
    from capytaine_amd import mesh_sphere
    mesh = mesh_sphere(radius=1.0, resolution=(16, 16)).immersed_part()
    ks = [0.05, 0.4, 1.7, 0.0]
    S, K = gf_ref.evaluate_on_device(mesh, mesh, wavenumber=ks)
    for b, k in enumerate(ks):
        s1, k1 = gf_ref.evaluate_on_device(mesh, mesh, wavenumber=k)
        assert bool((s1[0] == S[b]).all()) and bool((k1[0] == K[b]).all())


def test_nemoh2_matrices_gpu(gf_ref, nemoh_golden):
    """pytest/test_consistency_with_Nemoh_2.py:61-99 on the device."""
    from capytaine_amd import Mesh
    mesh = Mesh(nemoh_golden["vertices"], nemoh_golden["faces"])
    gf_ref.gf_singularities = "high_freq"
    try:
        S0, _ = gf_ref.evaluate(mesh, mesh, water_depth=np.inf, wavenumber=0.0)
        Sinf, _ = gf_ref.evaluate(mesh, mesh, water_depth=np.inf, wavenumber=np.inf)
    finally:
        gf_ref.gf_singularities = "low_freq"
    assert np.allclose(S0, nemoh_golden["S_k0"])
    assert np.allclose(Sinf, nemoh_golden["S_kinf"])


def test_rankine_only_vs_oracle(gf_ref, orc):
    from capytaine_amd import mesh_sphere
    from oracle import mesh_arrays
    mesh = mesh_sphere(radius=1.0, resolution=(10, 20))
    S, K = gf_ref.evaluate(mesh, mesh, free_surface=np.inf, water_depth=np.inf)
    So, Ko = orc.build_matrices(**mesh_arrays(mesh), wavenumber=1.0, depth=np.inf, table=None,
                                free_surface=np.inf)
    assert S.dtype == np.float64
    assert rel_err(S, So) < RTOL_NORM and rel_err(K, Ko) < RTOL_NORM


def test_points_as_collocation(gf_ref, orc, default_table):
    """mesh1 given as an (n, 3) array of points (abstract_green_function.py:42-54)."""
    from capytaine_amd import mesh_sphere
    from oracle import mesh_arrays
    mesh = mesh_sphere(radius=1.0, resolution=(10, 20)).immersed_part()
    pts = np.random.default_rng(3).uniform(-2, 2, (37, 3))
    pts[:, 2] = -np.abs(pts[:, 2]) - 0.01
    S, _ = gf_ref.evaluate(pts, mesh, wavenumber=0.7, diagonal_term_in_double_layer=False)
    arr = dict(mesh_arrays(mesh), colloc=pts, dpn=np.zeros((37, 3)))
    So, _ = orc.build_matrices(**arr, wavenumber=0.7, depth=np.inf, table=default_table, diagonal_term=False)
    assert rel_err(S, So) < RTOL_NORM


def test_table_switch_between_green_functions(default_table, legacy_table):
    """Two Green functions with different tables share the device context:
    each evaluation runs on its own table, including after the other one
    was freed (the context tracks the uploaded tensor, not its id())."""
    import gc
    import torch
    import capytaine_amd as cpa
    print("[gpu] table switch", flush=True)
    mesh = cpa.mesh_sphere(radius=1.0, resolution=(6, 12)).immersed_part()

    def legacy_gf():
        gf = cpa.Delhommeau(tabulation_nr=328, tabulation_nz=46, tabulation_grid_shape="legacy",
                            tabulation_nb_integration_points=251, tabulation_cache_dir=None, device=device())
        gf._set_tabulation(legacy_table[0], legacy_table[1], np.asfortranarray(legacy_table[2]))
        return gf

    ref_legacy = [a.cpu() for a in legacy_gf().evaluate_on_device(mesh, mesh, wavenumber=0.5)]
    ref_default = [a.cpu() for a in ref_table_gf(default_table).evaluate_on_device(mesh, mesh, wavenumber=0.5)]
    assert not torch.equal(ref_legacy[0], ref_default[0])
    for _ in range(4):
        for make, ref in ((legacy_gf, ref_legacy), (lambda: ref_table_gf(default_table), ref_default)):
            gf = make()
            S, K = gf.evaluate_on_device(mesh, mesh, wavenumber=0.5)
            assert torch.equal(S.cpu(), ref[0]) and torch.equal(K.cpu(), ref[1])
            del gf, S, K
            gc.collect()
