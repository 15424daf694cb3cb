"""The benchmark workloads against the oracle pipeline (oracle S and K on the
same table, then scipy LU, as BEMSolver._solve does: solver.py:188-207,
block_circulant_matrices.py:374-385).

* Hull fleet (64 ellipsoid hulls, seed 0): the smallest and the largest hull
  of 300-1500 panels,
  three of the fleet's frequencies and its four headings, through the
  persistent multi-body assembly and the grouped LU (BatchedSweep).  FP64
  forces within 1e-10 of the oracle pipeline, mixed (complex64 factors
  refined against the FP64 matrix) within 1e-6.
* Serving stream: one request of >= 500 panels served among three others;
  its (mixed) coefficients within 1e-6 of the oracle pipeline.
* Cylinder array in 40 m of water: a 2 x 2 array of 800 panels at two
  frequencies, FP64 forces within 1e-10 of the oracle pipeline (finite
  depth: matrices.f90:212-304, Green_wave.f90:311-519); and the full 5 x 5
  array (20000 panels) at one frequency, where the oracle is too slow, by
  properties of the FP64 and mixed solves: backward error
  ||K x - b|| / (||K||_F ||x||) <= 50 n eps, relative residual <= 1e-12
  (FP64; the mixed refinement runs to 1e-11, about the floor of an FP64
  residual at this order) and the mixed solution within 1e-9 of the FP64
  one.

Forces are compared per problem, relative to the largest force of that
problem (coefficients of a problem share its scale)."""
import numpy as np
import pytest

from gpu_helpers import device, ref_table_gf

pytestmark = pytest.mark.gpu

RHO, G = 1025.0, 9.81


def _oracle_forces(orc, table, body, pbs_by_freq, depth=np.inf, gf=None):
    """(nfreq, nproblems, ndof) forces of the oracle pipeline."""
    from oracle import mesh_arrays, solve_coefficients
    arr = mesh_arrays(body.mesh)
    dofs = list(body.dofs)
    out = np.zeros((len(pbs_by_freq), len(pbs_by_freq[0]), len(dofs)), dtype=complex)
    for b, pbs in enumerate(pbs_by_freq):
        k = pbs[0].wavenumber
        kw = {}
        if depth < np.inf:
            kw = dict(prony=gf.find_best_exponential_decomposition(k * depth),
                      finite_depth_method=gf.finite_depth_method_index)
        S, K = orc.build_matrices(**arr, wavenumber=k, depth=depth, table=table, **kw)
        for r, f in enumerate(solve_coefficients(S, K, body, pbs)):
            out[b, r] = [f[d] for d in dofs]
    return out


def _per_problem_err(got, ref):
    """max over problems of max_dof |got - ref| / max_dof |ref|."""
    scale = np.abs(ref).max(axis=-1, keepdims=True)
    return float((np.abs(got - ref) / np.maximum(scale, 1e-300)).max())


def test_hull_fleet_two_hulls_vs_oracle(default_table, orc):
    from capytaine_amd import workloads
    from capytaine_amd.batch import BatchedSweep
    from capytaine_amd.sweep import boundary_conditions
    bodies, omegas, headings = workloads.hull_fleet(0)
    inside = sorted((b for b in bodies if 300 <= b.mesh.nb_faces <= 1500), key=lambda b: b.mesh.nb_faces)
    picked = [inside[0], inside[-1]]  # the smallest and the largest
    om = omegas[[0, 7, 15]]
    print(f"[gpu] fleet hulls {[b.mesh.nb_faces for b in picked]} panels vs oracle", flush=True)
    gf = ref_table_gf(default_table)
    pbs = [workloads.problems_by_frequency(b, om, headings, rho=RHO) for b in picked]
    nrhs = [len(p[0]) for p in pbs]
    bcs = [boundary_conditions(p, r, b.mesh.nb_faces, device()) for p, r, b in zip(pbs, nrhs, picked)]
    ks = [np.array([p[0].wavenumber for p in pp]) for pp in pbs]
    refs = [_oracle_forces(orc, default_table, b, p) for b, p in zip(picked, pbs)]
    for precision, tol in (("fp64", 1e-10), ("mixed", 1e-6)):
        bs = BatchedSweep(gf, picked, ks, nrhs, precision=precision)
        bs.run(bcs)
        got = bs.forces([om] * len(picked), RHO)
        for b, g, ref in zip(picked, got, refs):
            err = _per_problem_err(g, ref)
            print(f"[gpu]   {precision} {b.name} ({b.mesh.nb_faces} panels): {err:.2e}", flush=True)
            assert err <= tol, (precision, b.name, err)


def test_serving_request_vs_oracle(default_table, orc):
    from capytaine_amd import workloads
    from capytaine_amd.serving import BEMServer
    rng = np.random.default_rng(4)
    target = workloads.serving_request(rng, n_range=(500, 900))
    others = [workloads.serving_request(rng, n_range=(300, 700)) for _ in range(3)]
    body, om, hd = target
    om = om[:4]
    assert body.mesh.nb_faces >= 500
    print(f"[gpu] serving request of {body.mesh.nb_faces} panels, {len(om)} frequencies, {len(hd)} headings, "
          f"dofs {list(body.dofs)} vs oracle", flush=True)
    gf = ref_table_gf(default_table)
    with BEMServer(gf, slots=2, max_batch_jobs=2, rho=RHO, g=G) as srv:
        futs = [srv.submit(b, o[:3], h) for b, o, h in others[:2]]
        fut = srv.submit(body, om, hd)
        futs.append(srv.submit(*others[2]))
        res = fut.result(timeout=240)
        for f in futs:
            f.result(timeout=240)
    ref = _oracle_forces(orc, default_table, body, workloads.problems_by_frequency(body, om, hd, rho=RHO, g=G))
    err = _per_problem_err(res["forces"], ref)
    print(f"[gpu]   served (mixed) vs oracle: {err:.2e}, residual {res['final_rel_residual']:.1e}", flush=True)
    assert err <= 1e-6


def test_cylinder_array_2x2_finite_depth_vs_oracle(default_table, orc):
    from capytaine_amd import workloads
    from capytaine_amd.sweep import FrequencySweep, boundary_conditions
    body, _, headings, depth = workloads.cylinder_array(rows=2, cols=2, resolution=(4, 20, 6))
    om = np.array([0.3, 1.8])
    hd = headings[:2]
    print(f"[gpu] 2 x 2 cylinder array, {body.mesh.nb_faces} panels, depth {depth} m vs oracle", flush=True)
    gf = ref_table_gf(default_table)
    pbs = workloads.problems_by_frequency(body, om, hd, rho=RHO, water_depth=depth)
    ks = np.array([p[0].wavenumber for p in pbs])
    sw = FrequencySweep(gf, body, ks, len(pbs[0]), water_depth=depth)
    sw.run(boundary_conditions(pbs, len(pbs[0]), body.mesh.nb_faces, device()))
    got = sw.forces(om, RHO)
    ref = _oracle_forces(orc, default_table, body, pbs, depth=depth, gf=gf)
    err = _per_problem_err(got, ref)
    print(f"[gpu]   FP64 vs oracle: {err:.2e}", flush=True)
    assert err <= 1e-10


def test_cylinder_array_full_one_frequency_properties(default_table):
    """N = 20000 panels, omega = 1 rad/s, depth 40 m: the six dofs of the
    corner cylinder and two headings."""
    import torch
    from capytaine_amd import workloads
    from capytaine_amd.sweep import FrequencySweep, boundary_conditions
    body, _, headings, depth = workloads.cylinder_array()
    n = body.mesh.nb_faces
    assert n == 20000
    keep = [d for d in body.dofs if d.startswith("c0_")]
    for d in list(body.dofs):
        if d not in keep:
            del body.dofs[d]
    om = np.array([1.0])
    pbs = workloads.problems_by_frequency(body, om, headings[:2], rho=RHO, water_depth=depth)
    nrhs = len(pbs[0])
    print(f"[gpu] full cylinder array: {n} panels, {nrhs} right-hand sides, depth {depth} m", flush=True)
    gf = ref_table_gf(default_table)
    bc = boundary_conditions(pbs, nrhs, n, device())
    eps = np.finfo(float).eps
    xs = {}
    for precision in ("fp64", "mixed"):
        # at this order the FP64 residual itself is good to ~1e-12: the
        # refinement stops at 1e-11
        sw = FrequencySweep(gf, body, [pbs[0][0].wavenumber], nrhs, water_depth=depth, precision=precision,
                            tol=1e-11)
        sw.assemble()
        K0 = sw.K.clone() if precision == "fp64" else sw.K  # the mixed form keeps K unmodified
        sw.factor()
        sw.solve(bc)
        torch.cuda.synchronize()
        X = sw.X[0]
        # K (column-major blocks) acts as K[0]^T: K x_j = (X K[0])_j
        R = bc[0] - X @ K0[0]
        rel_res = (torch.linalg.vector_norm(R, dim=1) / torch.linalg.vector_norm(bc[0], dim=1)).max().item()
        bwd = (torch.linalg.vector_norm(R, dim=1)
               / (torch.linalg.matrix_norm(K0[0]) * torch.linalg.vector_norm(X, dim=1))).max().item()
        print(f"[gpu]   {precision}: relative residual {rel_res:.2e}, backward error {bwd:.2e} "
              f"(bound {50 * n * eps:.1e})", flush=True)
        assert bwd <= 50 * n * eps
        if precision == "fp64":
            assert rel_res <= 1e-12
        else:
            assert sw.residual <= sw.tol and rel_res <= 1e-11
        xs[precision] = X.cpu().numpy()
        del sw, K0, X, R
        torch.cuda.empty_cache()
    d = np.linalg.norm(xs["mixed"] - xs["fp64"], axis=1) / np.linalg.norm(xs["fp64"], axis=1)
    print(f"[gpu]   mixed vs fp64 solution: {d.max():.2e}", flush=True)
    assert d.max() <= 1e-9
