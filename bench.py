"""Benchmark of the BEM hot path on one MI355X.

Workload (one step): rectangular barge 20 m x 8 m, draft 2 m, 0.25 m panels
(4352 panels), infinite depth, rho = 1025, g = 9.81; 40 frequencies uniform
in 0.2-2.5 rad/s; 6 rigid-body radiation problems + diffraction at 8
headings (0-315 deg) per frequency: 560 problems.  The step runs, on the
device with the boundary conditions already in HBM: assembly of S and K for
the 40 wavenumbers (one launch), the 40 factorisations (batched), the 560
solves, the potentials S @ sources and the generalised forces.

Three forms are timed in the same run: the plain FP64 form (complex128 LU
on FP64 MFMA), the mixed-precision batched form (far panel pairs read from a
float copy of the table; complex64 LU with FP16-input tile products, refined
with FP64 residuals against the complex128 matrix), and the same mixed form
on the FP64 assembly.  `value` is the mixed form's throughput; the FP64
throughput, the ratio and the error of every coefficient family against the
FP64 form (and, at the CPU sample's frequencies, against the oracle) are
reported beside it.

Prints one JSON line (see the repository's bench contract).
    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline] [--no-fleet] [--no-serving]
                    [--no-post] [--no-symmetric] [--no-fingreen3d]
    python bench.py --workload cylinders     # 20000-panel finite-depth array instead
    python bench.py --workload serving       # the request-stream measurement alone
    python bench.py --workload post          # the post-processing maps alone
    python bench.py --workload symmetric     # the barge as two reflection planes alone
    python bench.py --workload fingreen3d    # FinGreen3D assembly of the 20000-panel array

The request-stream measurement runs in a child process (started after the
headline and fleet legs have released their device memory), so that a
failure there is reported in the line instead of ending the run.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "panel-pairs/s (S,K assembly) + BEM solves/s, mixed vs FP64 at 1e-6; p99 latency"
MI355X_FP64_MFMA_TFLOPS = 78.6
MI355X_FP32_MFMA_TFLOPS = 157.3
MI355X_FP16_MFMA_TFLOPS = 2500.0  # dense, FP16 inputs, FP32 accumulation
MI355X_HBM_GBS = 8000.0


def barge_workload():
    import capytaine_amd as cpa
    mesh = cpa.mesh_parallelepiped(size=(20, 8, 2), center=(0, 0, -1), resolution=(80, 32, 8),
                                   missing_sides={"top"})
    body = cpa.FloatingBody(mesh, name="barge").add_all_rigid_body_dofs(rotation_center=(0, 0, 0))
    omegas = np.linspace(0.2, 2.5, 40)
    headings = np.arange(8) * np.pi / 4
    return mesh, body, omegas, headings


def problems_by_frequency(body, omegas, headings, rho, g):
    import capytaine_amd as cpa
    out = []
    for om in omegas:
        pbs = [cpa.RadiationProblem(body=body, radiating_dof=d, omega=om, rho=rho, g=g) for d in body.dofs]
        pbs += [cpa.DiffractionProblem(body=body, omega=om, wave_direction=b, rho=rho, g=g) for b in headings]
        out.append(pbs)
    return out


def cpu_baseline(mesh, body, pbs_by_freq, table, n_freq=6):
    """The oracle (C restatement of the reference core, OpenMP on all host
    cores) + scipy LU, on a bounded sample: `n_freq` frequencies spread over
    the sweep, every problem of those frequencies.  Returns the baseline
    object and the oracle's forces (n_freq, nproblems, ndof) at the sample's
    frequency indices."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle, mesh_arrays, solve_coefficients
    orc = Oracle()
    arr = mesh_arrays(mesh)
    idx = list(np.linspace(0, len(pbs_by_freq) - 1, n_freq).round().astype(int))
    dofs = list(body.dofs)
    forces = np.zeros((len(idx), len(pbs_by_freq[0]), len(dofs)), dtype=complex)
    t0 = time.perf_counter()
    nprob = 0
    for s, i in enumerate(idx):
        pbs = pbs_by_freq[i]
        S, K = orc.build_matrices(**arr, wavenumber=pbs[0].wavenumber, depth=np.inf, table=table)
        for r, f in enumerate(solve_coefficients(S, K, body, pbs)):
            forces[s, r] = [f[d] for d in dofs]
        nprob += len(pbs)
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=nprob / dt, unit="problems/s", cores=threads, kind="port",
                sample=f"{len(idx)} of {len(pbs_by_freq)} frequencies ({nprob} problems), oracle assembly + "
                       f"scipy LU, {dt:.1f} s"), idx, forces


def coefficient_errors(f, ref, omegas, ndof):
    """Errors of every hydrodynamic coefficient of `f` against `ref`, both
    (frequencies, problems, ndof) forces with the ndof radiation problems
    first (radiating dof order) and the diffraction problems after them.
    Added mass Re F / omega^2, radiation damping Im F / omega and the
    diffraction force: each coefficient's |error| relative to the largest
    |coefficient| of its family at its frequency; `by_dof` is the largest
    over frequencies and problems for each influenced dof, `diagonal_own`
    the error of each diagonal added mass and damping relative to itself
    (where it is at least 1e-3 of its family's largest)."""
    w = np.asarray(omegas, float)[:, None, None]
    fams = {"added_mass": (f[:, :ndof].real / w ** 2, ref[:, :ndof].real / w ** 2),
            "radiation_damping": (f[:, :ndof].imag / w, ref[:, :ndof].imag / w),
            "diffraction_force": (f[:, ndof:], ref[:, ndof:])}
    out = {}
    for name, (x, x0) in fams.items():
        if x0.shape[1] == 0:
            continue
        scale = np.abs(x0).max(axis=(1, 2), keepdims=True)
        e = np.abs(x - x0) / np.maximum(scale, 1e-300)
        row = {"max": float(e.max()), "by_dof": [float(v) for v in e.max(axis=(0, 1))]}
        if name != "diffraction_force":
            # each diagonal coefficient against itself, where it is at least
            # 1e-3 of its family's largest (a vanishing coefficient, such as
            # a yaw damping, has no relative error to speak of)
            d = np.arange(ndof)
            big = np.abs(x0[:, d, d]) >= 1e-3 * scale[:, :, 0]
            own = np.where(big, np.abs(x[:, d, d] - x0[:, d, d]) / np.maximum(np.abs(x0[:, d, d]), 1e-300), 0.0)
            row["diagonal_own"] = [float(v) for v in own.max(axis=0)]
        out[name] = row
    return out


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (a line "traffic: <kernel> <bytes>"), or None."""
    path = os.path.join(ROOT, "profiles", "c2_mixed_hbm_traffic_counters.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic:" and parts[1] == kernel:
            return float(parts[2])
    return None


def assembly_error(gf, mesh, ks, _lib):
    """Largest |S_mixed - S_fp64| / max |S_fp64| (and the same for K) over
    the given wavenumbers: the FP32 far field against the FP64 assembly."""
    import torch
    ctx = gf._context()
    out = {}
    for fp32 in (0, 1):
        _lib.check(_lib.load().cpt_set_assembly_precision(ctx.handle, fp32), "assembly precision")
        out[fp32] = gf.evaluate_on_device(mesh, mesh, wavenumber=list(ks))
    _lib.check(_lib.load().cpt_set_assembly_precision(ctx.handle, 0), "assembly precision")
    torch.cuda.synchronize()
    err = {}
    for i, name in enumerate(("S", "K")):
        a64, a32 = out[0][i], out[1][i]
        err[name] = max(float((a32[b] - a64[b]).abs().max() / a64[b].abs().max()) for b in range(len(ks)))
    del out
    torch.cuda.empty_cache()
    return err


def fleet_benchmark(gf, dev, rho=1025.0, g=9.81, seed=0, budget_gb=110.0, log=None):
    """64 half-immersed ellipsoid hulls (panel counts log-uniform in
    300-6000), 16 frequencies, 6 dof + 4 headings: one BatchedSweep per
    memory chunk (persistent multi-body assembly + grouped LU), mixed and
    FP64.  Inputs (boundary conditions) are resident before timing."""
    import torch
    from capytaine_amd import workloads
    from capytaine_amd.batch import BatchedSweep, chunks_by_memory, footprint_bytes
    from capytaine_amd.sweep import boundary_conditions
    from capytaine_amd import _lib
    # library defaults (block width by matrix order) for the mix of sizes
    _lib.set_tuning(_lib.Context.get(dev), lu_block=0, panel_leaf=16, lookahead=1, assembly_chunk=0)
    bodies, omegas, headings = workloads.hull_fleet(seed)
    ks = omegas ** 2 / g
    nrhs = 6 + len(headings)
    ns = [b.mesh.nb_faces for b in bodies]
    bcs = []
    for b in bodies:
        pbs = workloads.problems_by_frequency(b, omegas, headings, rho=rho, g=g)
        bcs.append(boundary_conditions(pbs, nrhs, b.mesh.nb_faces, dev))
    out = {"workload": f"64 ellipsoid hulls (seed {seed}), panels {min(ns)}-{max(ns)} log-uniform, 16 frequencies "
                       f"0.3-3.0 rad/s, 6 dof + 4 headings, infinite depth",
           "problems": len(bodies) * len(ks) * nrhs, "panel_pairs": int(sum(len(ks) * n * n for n in ns))}
    forces = {}
    for precision in ("fp64", "mixed"):
        sizes = [footprint_bytes(n, len(ks), nrhs, precision) for n in ns]
        chunks = chunks_by_memory(sizes, budget_gb * 1e9)
        t_total, t_asm, sweeps, res = 0.0, 0.0, 0, 0.0
        forces[precision] = [None] * len(bodies)
        # untimed warm-up on the largest chunk: the library's workspaces and
        # lane streams, and the allocator's blocks, as a sweep meets them after
        # its first call (the device still holds the barge's cached blocks)
        torch.cuda.empty_cache()
        big = max(chunks, key=lambda c: sum(sizes[i] for i in c))
        warm = BatchedSweep(gf, [bodies[i] for i in big], ks, nrhs, precision=precision)
        warm.assemble()
        warm.solve([bcs[i] for i in big])
        warm.post()
        torch.cuda.synchronize()
        del warm
        for ch in chunks:
            bs = BatchedSweep(gf, [bodies[i] for i in ch], ks, nrhs, precision=precision)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            t0 = time.perf_counter()
            e0.record()
            bs.assemble()
            e1.record()
            bs.solve([bcs[i] for i in ch])
            bs.post()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            t_total += dt
            if log:
                log(f"[fleet] {precision}: {len(ch)} bodies, {dt:.3f} s (assembly {e0.elapsed_time(e1) / 1e3:.3f} s)")
            t_asm += e0.elapsed_time(e1) / 1e3
            if precision == "mixed":
                sweeps = max(sweeps, int(bs.iterations.max()))
                res = max(res, float(bs.residual.max()))
            for j, f in zip(ch, bs.forces([omegas] * len(ch), rho)):
                forces[precision][j] = f
            del bs  # its blocks are reused by the next chunk (same stream)
        torch.cuda.empty_cache()
        out[f"{precision}_problems_per_s"] = out["problems"] / t_total
        out[f"{precision}_seconds"] = t_total
        out[f"{precision}_assembly_panel_pairs_per_s"] = out["panel_pairs"] / t_asm
        out[f"{precision}_launches"] = len(chunks)
        if precision == "mixed":
            out["mixed_refinement_sweeps_max"] = sweeps
            out["mixed_final_rel_residual_max"] = res
        if log:
            log(f"[fleet] {precision}: {t_total:.2f} s in {len(chunks)} chunk(s)")
    err = 0.0
    for f64, fmx in zip(forces["fp64"], forces["mixed"]):
        scale = np.abs(f64).max(axis=2, keepdims=True)
        err = max(err, float(np.max(np.abs(fmx - f64) / scale)))
    out["mixed_over_fp64"] = out["mixed_problems_per_s"] / out["fp64_problems_per_s"]
    out["mixed_max_rel_err_vs_fp64"] = err
    return out


def serving_benchmark(gf, dev, seed=1, requests=40, loads=(0.5, 0.8, 1.0, 1.2), log=None):
    """Poisson request stream against the continuous-batching server.  One
    mix of `requests` requests is drawn once: its burst (all submitted at
    once) gives the saturation throughput, and the same mix, in the same
    order, arrives at Poisson rates of `loads` x saturation for the latency
    percentiles (p50, p90, p99 and the largest latency of each load point)."""
    import threading
    from capytaine_amd import workloads
    from capytaine_amd.serving import BEMServer

    def problems(reqs):
        return sum(len(om) * (len(b.dofs) + len(hd)) for b, om, hd in reqs)

    rng = np.random.default_rng(seed)
    mix = [workloads.serving_request(rng) for _ in range(requests)]
    sizes = np.array([b.mesh.nb_faces for b, _, _ in mix])
    out = {"workload": f"hull requests (seed {seed}): 500-12000 panels log-uniform, 1-32 frequencies in 0.2-3.0 "
                       "rad/s, random dof subset, 0-8 headings; mixed precision; Poisson arrivals",
           "requests_per_load_point": requests,
           "mix": {"panels_median": int(np.median(sizes)), "panels_max": int(sizes.max()),
                   "problems": problems(mix)}}
    with BEMServer(gf, rho=1025.0) as srv:
        # warm: first use of the slot contexts
        b, om, hd = workloads.serving_request(np.random.default_rng(seed + 1000), n_range=(500, 800))
        srv.submit(b, om[:1], hd).result()
        t0 = time.perf_counter()
        futs = [srv.submit(*r) for r in mix]
        for f in futs:
            f.result()
        span = time.perf_counter() - t0
        sat = requests / span
        out["saturation_requests_per_s"] = sat
        out["saturation_solves_per_s"] = problems(mix) / span
        if log:
            log(f"[serving] saturation {sat:.2f} requests/s ({span:.2f} s for {requests})")
        out["loads"] = []
        for load in loads:
            gaps = rng.exponential(1.0 / (load * sat), size=requests)
            arrivals = np.cumsum(gaps) - gaps[0]
            n_before = len(srv.latencies)
            futs = []

            def client():
                start = time.perf_counter()
                for r, t in zip(mix, arrivals):
                    delay = start + t - time.perf_counter()
                    if delay > 0:
                        time.sleep(delay)
                    futs.append(srv.submit(*r))

            t0 = time.perf_counter()
            th = threading.Thread(target=client)
            th.start()
            th.join()
            for f in futs:
                f.result()
            span = time.perf_counter() - t0
            lat = np.array(srv.latencies[n_before:])
            row = {"load": load, "requests": requests, "offered_requests_per_s": load * sat,
                   "p50_s": float(np.percentile(lat, 50)), "p90_s": float(np.percentile(lat, 90)),
                   "p99_s": float(np.percentile(lat, 99)), "max_s": float(lat.max()),
                   "requests_per_s": requests / span, "solves_per_s": problems(mix) / span}
            out["loads"].append(row)
            if log:
                log(f"[serving] load {load}: p50 {row['p50_s']:.3f} s p90 {row['p90_s']:.3f} s p99 "
                    f"{row['p99_s']:.3f} s max {row['max_s']:.3f} s, {row['solves_per_s']:.0f} solves/s")
        out["max_inflight_batches"] = srv.max_inflight_batches
    return out


def serving_child(timeout_s=900, requests=40):
    """serving_benchmark in a child process (same interpreter, this file
    with --workload serving): its JSON line, or the exit status when the
    child ends without one.  Its progress lines go to this process's stderr."""
    import subprocess
    cmd = [sys.executable, "-u", os.path.abspath(__file__), "--workload", "serving", "--requests", str(requests)]
    try:
        p = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=timeout_s)
    except subprocess.TimeoutExpired:
        return {"error": f"request-stream measurement exceeded {timeout_s} s"}
    lines = [ln for ln in p.stdout.splitlines() if ln.startswith("{")]
    if p.returncode != 0 or not lines:
        return {"error": f"request-stream measurement exited with status {p.returncode}"}
    return json.loads(lines[-1])


def post_processing_benchmark(gf, dev, mesh, body, headings, steps=3, rho=1025.0, g=9.81, omega=1.0,
                              cpu_points=16384, log=None):
    """Free-surface elevation map and velocity section of one frequency's
    problems (6 dof + 8 headings) around the barge, through BEMSolver's
    post-processing engine (points as collocation, products on the device).

    Elevation: 0.2 m grid over 80 m x 40 m at z = 0 minus the barge
    footprint, all 14 source vectors at once.  Velocity: 0.4 m grid on the
    vertical plane y = 6 m, x in [-40, 40], z in [-20, -0.2].  Times are per
    map, host points in and host values out; the CPU sample is the oracle's
    S and K for `cpu_points` points of the map (OpenMP on the host cores)."""
    import torch
    import capytaine_amd as cpa
    from capytaine_amd import _lib
    solver = cpa.BEMSolver(engine=cpa.DefaultMatrixEngine(green_function=gf))
    pbs = [cpa.RadiationProblem(body=body, radiating_dof=d, omega=omega, rho=rho, g=g) for d in body.dofs]
    pbs += [cpa.DiffractionProblem(body=body, omega=omega, wave_direction=b, rho=rho, g=g) for b in headings]
    results = solver.solve_all(pbs, keep_details=True)
    src = np.stack([r.sources for r in results], axis=1)
    gp = dict(free_surface=0.0, water_depth=np.inf, wavenumber=pbs[0].wavenumber)
    x, y = np.meshgrid(np.arange(-40.0, 40.0 + 1e-9, 0.2), np.arange(-20.0, 20.0 + 1e-9, 0.2), indexing="ij")
    keep = (np.abs(x) > 10.05) | (np.abs(y) > 4.05)
    fs_pts = np.stack([x[keep], y[keep], np.zeros(keep.sum())], axis=1)
    xs, zs = np.meshgrid(np.arange(-40.0, 40.0 + 1e-9, 0.4), np.arange(-20.0, -0.2 + 1e-9, 0.4), indexing="ij")
    vel_pts = np.stack([xs.ravel(), np.full(xs.size, 6.0), zs.ravel()], axis=1)
    n = mesh.nb_faces
    ctx = _lib.Context.get(dev)
    out = {}
    for name, pts, gradient in (("elevation_map", fs_pts, False), ("velocity_section", vel_pts, True)):
        solver.engine.field_at_points(pts, mesh, src, gradient=gradient, **gp)  # warm-up
        torch.cuda.synchronize()
        _lib.profile_begin(ctx)
        t0 = time.perf_counter()
        for _ in range(steps):
            val = solver.engine.field_at_points(pts, mesh, src, gradient=gradient, **gp)
        dt = (time.perf_counter() - t0) / steps
        kern = _lib.profile_end(ctx)
        if not gradient:
            val = 1j * omega / g * val  # compute_free_surface_elevation
        asm = kern["assembly"]
        prod = {k: v for k, v in kern.items() if k in ("zgemm", "skinny_z") and v["launches"]}
        out[name] = {"points": len(pts), "panels": n, "source_vectors": src.shape[1], "ms_per_map": dt * 1e3,
                     "points_per_s": len(pts) / dt, "panel_pairs_per_s": len(pts) * n / dt,
                     "assembly_ms_per_map": asm["ms"] / steps,
                     "assembly_panel_pairs_per_s": asm["work"] / (asm["ms"] / 1e3) if asm["ms"] else None,
                     "product_ms_per_map": sum(v["ms"] for v in prod.values()) / steps,
                     "finite": bool(np.all(np.isfinite(val)))}
        if log:
            log(f"[post] {name}: {len(pts)} points x {n} panels, {dt * 1e3:.1f} ms per map")
    out["elevation_map"]["grid"] = "0.2 m over [-40, 40] x [-20, 20] m at z = 0, barge footprint removed"
    out["velocity_section"]["grid"] = "0.4 m over x in [-40, 40], z in [-20, -0.2] m, y = 6 m"
    # CPU: the oracle's S, K at a bounded sample of the map's points
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle, mesh_arrays
    orc = Oracle()
    sample = fs_pts[np.linspace(0, len(fs_pts) - 1, cpu_points).round().astype(int)]
    arr = dict(mesh_arrays(mesh), colloc=sample, dpn=np.zeros_like(sample))
    table = (gf.tabulated_r_range, gf.tabulated_z_range, gf.tabulated_integrals)
    t0 = time.perf_counter()
    S, _ = orc.build_matrices(**arr, wavenumber=gp["wavenumber"], depth=np.inf, table=table)
    S @ src
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out["cpu_baseline"] = {"panel_pairs_per_s": cpu_points * n / dt, "cores": threads, "kind": "port",
                           "sample": f"{cpu_points} points of the elevation map x {n} panels, oracle S and K "
                                     f"+ product, {dt:.1f} s",
                           "ms_per_elevation_map_extrapolated": dt / cpu_points * len(fs_pts) * 1e3}
    del results, solver
    torch.cuda.empty_cache()
    return out


def symmetric_benchmark(gf, dev, steps=3, warmup=1, rho=1025.0, g=9.81, ref_forces=None, log=None):
    """The barge workload with both reflection planes (four blocks of 1088
    panels): per step, the first block column of S and K for the 40
    wavenumbers (one launch), the block transforms, one batched LU of the 160
    transformed blocks (FP64; then complex64 with FP64 refinement), the 560
    solves, S @ sources and the forces, inputs resident in HBM.  `ref_forces` (the plain FP64 form's) gives the
    largest relative difference of the coefficients."""
    import torch
    import capytaine_amd as cpa
    from capytaine_amd import _lib
    from capytaine_amd.sweep import boundary_conditions
    from capytaine_amd.symmetry import (BlockSymmetricDeviceMatrix, LUDecomposedSymmetricMatrix,
                                        MixedSymmetricSolver, symmetry_of)
    mesh = cpa.mesh_parallelepiped(size=(20, 8, 2), center=(0, 0, -1), resolution=(80, 32, 8), missing_sides={"top"},
                                   reflection_symmetry=True)
    body = cpa.FloatingBody(mesh, name="barge").add_all_rigid_body_dofs(rotation_center=(0, 0, 0))
    _, _, omegas, headings = barge_workload()
    pbs_by_freq = problems_by_frequency(body, omegas, headings, rho, g)
    sym = symmetry_of(mesh)
    n = mesh.nb_faces
    nrhs = len(pbs_by_freq[0])
    ks = [p[0].wavenumber for p in pbs_by_freq]
    nb = len(ks)
    bc = boundary_conditions(pbs_by_freq, nrhs, n, dev)
    W = -(body.dof_normal_matrix() * mesh.faces_areas[None, :]).T
    W_d = torch.as_tensor(np.ascontiguousarray(W), dtype=torch.complex128).to(dev)
    ndof = W.shape[1]
    ctx = _lib.Context.get(dev)
    merged = mesh.merged()
    G = torch.empty((nb, nrhs, ndof), dtype=torch.complex128, device=dev)

    def step(mixed):
        S, K = gf.evaluate_on_device(merged, sym.block_mesh, wavenumber=ks, adjoint_double_layer=True,
                                     diagonal_term_in_double_layer=True)
        if mixed:
            lu = MixedSymmetricSolver(BlockSymmetricDeviceMatrix(K, sym))
            X = lu.solve_device(bc)
        else:
            lu = LUDecomposedSymmetricMatrix(BlockSymmetricDeviceMatrix(K, sym))
            X = bc.clone()
            lu.solve_device(X)
        Phi = BlockSymmetricDeviceMatrix(S, sym).matmul_device(X)
        _lib.check(_lib.load().cpt_zgemm(ctx.handle, ndof, nrhs, n, 1.0, _lib.ptr(W_d), ndof, 0, _lib.ptr(Phi), n,
                                         n * nrhs, _lib.ptr(G), ndof, ndof * nrhs, 1, nb, _lib.stream_handle(dev)),
                   "cpt_zgemm")
        return lu

    out = {"workload": "barge 20x8x2 m, 4352 panels as two reflection planes (4 blocks of 1088), 40 frequencies, "
                       "560 problems; FP64 block-diagonal LU, and complex64 LU of the blocks with FP64 refinement",
           "lu_blocks": nb * sym.n, "block_order": sym.w}
    forces = {}
    for mixed in (False, True):
        for _ in range(warmup):
            step(mixed)
        torch.cuda.synchronize()
        _lib.profile_begin(ctx)
        t0 = time.perf_counter()
        for _ in range(steps):
            lu = step(mixed)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        kern = _lib.profile_end(ctx)
        forces[mixed] = 1j * np.asarray(omegas)[:, None, None] * rho * G.cpu().numpy()
        tag = "mixed_" if mixed else ""
        out[tag + "problems_per_s"] = nb * nrhs / dt
        out[tag + "ms_per_step"] = dt * 1e3
        if mixed:
            out["mixed_refinement_sweeps"], out["mixed_final_rel_residual"] = lu.iterations, lu.residual
        else:
            lu.check_info()
            out["assembly_ms_per_step"] = kern["assembly"]["ms"] / steps
            out["assembly_panel_pairs_per_s"] = kern["assembly"]["work"] / (kern["assembly"]["ms"] / 1e3)
        if log:
            log(f"[symmetric] {'mixed' if mixed else 'fp64'}: {dt * 1e3:.1f} ms per step, "
                f"{nb * nrhs / dt:.0f} problems/s")
        del lu
    scale = np.abs(forces[False]).max(axis=2, keepdims=True)
    out["mixed_max_rel_err_vs_fp64"] = float(np.max(np.abs(forces[True] - forces[False]) / scale))
    if ref_forces is not None:
        scale = np.abs(ref_forces).max(axis=2, keepdims=True)
        out["max_rel_diff_vs_plain_fp64"] = float(np.max(np.abs(forces[False] - ref_forces) / scale))
    torch.cuda.empty_cache()
    return out


def copy_calibration(nbytes=1 << 30, reps=4):
    """Element-wise copy of `nbytes` (float32, y = 1 * x) `reps` times: the
    HBM counters of this kernel (FETCH_SIZE, WRITE_SIZE) against the known
    bytes read and written give their scale factor."""
    import torch
    dev = torch.device("cuda", 0)
    x = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    print(json.dumps({"workload": "copy", "bytes_read_per_launch": nbytes, "bytes_written_per_launch": nbytes,
                      "launches": reps, "GB_per_s": 2 * nbytes / dt / 1e9}), flush=True)


def cylinders_benchmark(args, rho=1025.0, g=9.81, chunk=4):
    """5 x 5 truncated cylinders (20000 panels), depth 40 m, 20 frequencies,
    150 radiation dof + 8 headings: the finite-depth assembly and the
    N = 20000 factorisation, FP64 form and mixed form (complex64 factors
    refined against the complex128 K), `chunk` frequencies at a time
    (S + K are 12.8 GB per frequency).  One timed pass over the 20
    frequencies per form, after a warm pass over one chunk."""
    import torch
    import capytaine_amd as cpa
    from capytaine_amd import workloads
    from capytaine_amd.sweep import FrequencySweep, boundary_conditions
    dev = torch.device("cuda", 0)
    body, omegas, headings, depth = workloads.cylinder_array()
    n = body.mesh.nb_faces
    gf = cpa.Delhommeau(device=dev)
    pbs = workloads.problems_by_frequency(body, omegas, headings, rho=rho, g=g, water_depth=depth)
    nrhs = len(pbs[0])
    ks = np.array([p[0].wavenumber for p in pbs])
    out = {"metric": METRIC, "unit": "problems/s", "n_gpus": 1, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "data": "synthetic",
           "config": {"workload": "5 x 5 truncated vertical cylinders, radius 2 m, draft 4 m, 12 m apart, "
                                  "20000 panels, depth 40 m, 20 frequencies 0.3-1.8 rad/s, 150 dof + 8 headings",
                      "panels": n, "frequencies": len(ks), "rhs_per_frequency": nrhs, "frequency_chunk": chunk}}
    forces = {}
    for precision in ("fp64", "mixed"):
        t_asm = t_all = 0.0
        forces[precision] = []
        for warm in (True, False):
            for c0 in range(0, len(ks), chunk):
                sl = slice(c0, min(c0 + chunk, len(ks)))
                bc = boundary_conditions(pbs[sl], nrhs, n, dev)
                sw = FrequencySweep(gf, body, ks[sl], nrhs, water_depth=depth, precision=precision)
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                t0 = time.perf_counter()
                e0.record()
                sw.assemble()
                e1.record()
                sw.factor()
                sw.solve(bc)
                sw.post()
                torch.cuda.synchronize()
                if not warm:
                    t_all += time.perf_counter() - t0
                    t_asm += e0.elapsed_time(e1) / 1e3
                    forces[precision].append(sw.forces(omegas[sl], rho))
                del sw, bc
                print(f"[cylinders] {precision} {'warm' if warm else 'timed'} frequencies {sl.start}-{sl.stop - 1}",
                      file=sys.stderr, flush=True)
                if warm:
                    break
        torch.cuda.empty_cache()
        nprob = len(ks) * nrhs
        out[f"{precision}_problems_per_s"] = nprob / t_all
        out[f"{precision}_seconds"] = t_all
        out[f"{precision}_assembly_panel_pairs_per_s"] = len(ks) * n * n / t_asm
        out[f"{precision}_lu_tflops"] = len(ks) * 8.0 * n ** 3 / 3.0 / (t_all - t_asm) / 1e12
    f64 = np.concatenate(forces["fp64"])
    fmx = np.concatenate(forces["mixed"])
    scale = np.abs(f64).max(axis=2, keepdims=True)
    out["value"] = out["mixed_problems_per_s"]
    out["mixed_over_fp64"] = out["mixed_problems_per_s"] / out["fp64_problems_per_s"]
    out["mixed_max_rel_err_vs_fp64"] = float(np.max(np.abs(fmx - f64) / scale))
    out["dtype"] = "c64 factor + f64 refinement (mixed); f64 reported beside"
    out["steps"], out["warmup"] = 1, 1
    print(json.dumps(out), flush=True)


def fingreen3d_benchmark(no_cpu_baseline=False, nfreq=4, cpu_points=100, log=None):
    """FinGreen3D (the HAMS finite-depth Green function) assembly of S and K
    for the 5 x 5 cylinder array (20000 panels, depth 40 m) at `nfreq`
    wavenumbers, one launch pair per wavenumber, the panels resident in HBM;
    timed after one untimed evaluation.  The CPU baseline is the reference's
    own FinGreen3D.f90 (oracle/_ref, OpenMP over the host cores) on the
    first `cpu_points` collocation rows at the first wavenumber, whose rows
    are also compared with the device's.  Returns the measurement as a dict
    (the `--workload fingreen3d` line, and the default line's "fingreen3d")."""
    import torch
    import capytaine_amd as cpa
    from capytaine_amd import workloads
    log = log or (lambda m: print(m, file=sys.stderr, flush=True))
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    body, omegas, _, depth = workloads.cylinder_array()
    mesh = body.mesh
    n = mesh.nb_faces
    gf = cpa.FinGreen3D(device=dev)
    sel = np.linspace(0, len(omegas) - 1, nfreq).round().astype(int)
    from capytaine_amd.problems import wavenumber_from_omega
    ks = [float(wavenumber_from_omega(float(omegas[i]), 9.81, depth)) for i in sel]
    S = torch.empty((1, n, n), dtype=torch.complex128, device=dev)
    K = torch.empty((1, n, n), dtype=torch.complex128, device=dev)
    gf.evaluate_on_device(mesh, mesh, water_depth=depth, wavenumber=ks[0], out=(S, K))  # untimed
    torch.cuda.synchronize()
    times = []
    for k in ks:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        gf.evaluate_on_device(mesh, mesh, water_depth=depth, wavenumber=k, out=(S, K))
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        log(f"[fingreen3d] k {k:.4f}: {times[-1]:.3f} s")
    t = float(np.sum(times))
    out = {"metric": "FinGreen3D panel-pairs/s (S and K assembly)", "value": nfreq * n * n / t,
           "unit": "panel-pairs/s", "n_gpus": 1, "steps": nfreq, "warmup": 1, "ms_per_step": 1e3 * t / nfreq,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "FinGreen3D S and K of 5 x 5 truncated vertical cylinders, radius 2 m, draft "
                                  "4 m, 12 m apart, 20000 panels, depth 40 m, adjoint double layer",
                      "panels": n, "wavenumbers": [float(k) for k in ks], "nb_dispersion_roots": 200}}
    # the first rows at the first wavenumber, for the comparison with the
    # reference (S[0] holds the column-major (n, n) matrix: rows are columns)
    gf.evaluate_on_device(mesh, mesh, water_depth=depth, wavenumber=ks[0], out=(S, K))
    S_dev = S[0][:, :cpu_points].T.cpu().numpy()
    K_dev = K[0][:, :cpu_points].T.cpu().numpy()
    del S, K
    torch.cuda.empty_cache()
    if no_cpu_baseline:
        return out
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        from oracle import Reference, mesh_arrays
        R = Reference()
    except (ImportError, OSError, FileNotFoundError) as e:
        out["cpu_baseline"] = {"error": f"oracle/_ref not available: {e}"}
        return out
    cores = min(os.cpu_count() or 1, 16)
    arr = mesh_arrays(mesh)
    arr["colloc"] = arr["colloc"][:cpu_points]
    arr["dpn"] = arr["dpn"][:cpu_points]
    roots = cpa.FinGreen3D._roots(200, float(ks[0]), float(depth))
    t0 = time.perf_counter()
    S_ref, K_ref = R.build_matrices_fingreen3d(**arr, wavenumber=float(ks[0]), depth=float(depth),
                                               roots=np.array(roots), adjoint=True, free_surface=0.0)
    t_cpu = time.perf_counter() - t0
    out["cpu_baseline"] = {"value": cpu_points * n / t_cpu, "unit": "panel-pairs/s", "cores": cores,
                           "kind": "reference",
                           "sample": f"the reference's FinGreen3D.f90 build_matrices (OpenMP, OMP_NUM_THREADS="
                                     f"{os.environ.get('OMP_NUM_THREADS', 'unset')}) on {cpu_points} collocation "
                                     f"rows x {n} panels at k = {ks[0]:.4f}: {t_cpu:.2f} s"}
    out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    out["max_diff_vs_reference"] = {"S": float(np.abs(S_dev - S_ref).max() / np.abs(S_ref).max()),
                                    "K": float(np.abs(K_dev - K_ref).max() / np.abs(K_ref).max())}
    # the reference's own sensitivity on the same rows: its change when the
    # geometry moves by 2^-50 relative (its series limits are not smooth)
    rng = np.random.default_rng(0)
    pa = dict(arr)
    for key in ("colloc", "vertices", "centers", "quad_points"):
        a = np.asarray(arr[key], dtype=np.float64)
        pa[key] = a * (1 + rng.choice([-1.0, 1.0], size=a.shape) * 2.0 ** -50)
    S_p, K_p = R.build_matrices_fingreen3d(**pa, wavenumber=float(ks[0]), depth=float(depth), roots=np.array(roots),
                                           adjoint=True, free_surface=0.0)
    out["reference_sensitivity"] = {"S": float(np.abs(S_p - S_ref).max() / np.abs(S_ref).max()),
                                    "K": float(np.abs(K_p - K_ref).max() / np.abs(K_ref).max())}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fleet", action="store_true", help="skip the 64-hull batched measurement")
    ap.add_argument("--no-serving", action="store_true", help="skip the request-stream latency measurement")
    ap.add_argument("--no-post", action="store_true", help="skip the post-processing (elevation, velocity) maps")
    ap.add_argument("--no-symmetric", action="store_true", help="skip the reflection-symmetric barge measurement")
    ap.add_argument("--no-fingreen3d", action="store_true",
                    help="skip the FinGreen3D assembly of the 20000-panel cylinder array")
    ap.add_argument("--gemm-tiles", type=int, default=None,
                    help="force the complex64 tile product variant of the mixed form (cpt_set_tuning "
                         "gemm_tiles, 0-4) after tuning, for profiler passes of one variant")
    ap.add_argument("--lu-block", type=int, default=None, choices=(128, 256, 512, 1024),
                    help="force the LU block width of both forms (cpt_set_tuning lu_block) after tuning, for "
                         "profiler passes of one variant")
    ap.add_argument("--panel-threads", type=int, default=None, choices=(256, 512),
                    help="force the panel leaf work-group size of both forms (cpt_set_tuning "
                         "panel_threads) after tuning, for profiler passes of one variant")
    ap.add_argument("--requests", type=int, default=40,
                    help="requests of the serving mix (drawn once; the saturation burst and every load point "
                         "use the same mix)")
    ap.add_argument("--loads", type=str, default="0.5,0.8,1.0,1.2",
                    help="arrival rates of the serving measurement, as fractions of its saturation throughput")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (a stored tuning of this size class is still applied)")
    ap.add_argument("--workload", choices=("barge", "cylinders", "copy", "serving", "post", "symmetric",
                                           "fingreen3d"),
                    default="barge",
                    help="barge: the headline line (plus fleet and serving); cylinders: the 20000-panel "
                         "finite-depth array; copy: a 1 GiB device copy, the calibration of the HBM "
                         "byte counters; serving: the request-stream measurement alone")
    ap.add_argument("--only", choices=("fp64", "mixed"), default=None,
                    help="run one form of the barge sweep only (for counter passes); prints no JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the frequency sweep shards by "
              "frequency across GPUs by construction but this build measures a single GPU.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "cylinders":
        cylinders_benchmark(args)
        return
    if args.workload == "copy":
        copy_calibration()
        return
    if args.workload == "fingreen3d":
        print(json.dumps(fingreen3d_benchmark(no_cpu_baseline=args.no_cpu_baseline)), flush=True)
        return
    if args.workload == "serving":
        import torch
        import capytaine_amd as cpa
        dev = torch.device("cuda", 0)
        torch.cuda.set_device(dev)
        res = serving_benchmark(cpa.Delhommeau(device=dev), dev, requests=args.requests,
                                loads=tuple(float(v) for v in args.loads.split(",")),
                                log=lambda m: print(m, file=sys.stderr, flush=True))
        print(json.dumps(res), flush=True)
        return

    if args.workload == "symmetric":
        import torch
        import capytaine_amd as cpa
        dev = torch.device("cuda", 0)
        torch.cuda.set_device(dev)
        res = symmetric_benchmark(cpa.Delhommeau(device=dev), dev, steps=args.steps, warmup=args.warmup,
                                  log=lambda m: print(m, file=sys.stderr, flush=True))
        print(json.dumps(res), flush=True)
        return
    if args.workload == "post":
        import torch
        import capytaine_amd as cpa
        dev = torch.device("cuda", 0)
        torch.cuda.set_device(dev)
        mesh, body, _, headings = barge_workload()
        res = post_processing_benchmark(cpa.Delhommeau(device=dev), dev, mesh, body, headings, steps=args.steps,
                                        log=lambda m: print(m, file=sys.stderr, flush=True))
        print(json.dumps(res), flush=True)
        return

    import torch
    import capytaine_amd as cpa
    from capytaine_amd.sweep import FrequencySweep, boundary_conditions
    from capytaine_amd import _lib

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    rho, g = 1025.0, 9.81
    mesh, body, omegas, headings = barge_workload()
    pbs_by_freq = problems_by_frequency(body, omegas, headings, rho, g)
    n = mesh.nb_faces
    nrhs = len(pbs_by_freq[0])
    ks = [p[0].wavenumber for p in pbs_by_freq]
    gf = cpa.Delhommeau(device=dev)
    bc = boundary_conditions(pbs_by_freq, nrhs, n, dev)
    stream = torch.cuda.current_stream(dev)
    omegas_arr = np.asarray(omegas)

    from capytaine_amd import autotune
    ctx = _lib.Context.get(dev)
    tuning = {}
    if not args.no_tune:
        def log(msg):
            print(msg, file=sys.stderr, flush=True)
        for prec in ("fp64", "mixed"):
            tuning[prec] = autotune.tune_lu(n, len(ks), prec, device=dev, nrhs=nrhs, log=log)
        tuning["assembly"] = autotune.tune_assembly(gf, mesh, ks, log=log)
        torch.cuda.empty_cache()

    def run(precision, far_field_fp32=None):
        applied = autotune.apply(ctx, n, len(ks), precision, nb_wavenumbers=len(ks))
        if args.gemm_tiles is not None and precision == "mixed":
            _lib.set_tuning(ctx, gemm_tiles=args.gemm_tiles)
            applied = dict(applied or {}, gemm_tiles=args.gemm_tiles)
        if args.panel_threads is not None:
            _lib.set_tuning(ctx, panel_threads=args.panel_threads)
            applied = dict(applied or {}, panel_threads=args.panel_threads)
        if args.lu_block is not None:
            _lib.set_tuning(ctx, lu_block=args.lu_block)
            applied = dict(applied or {}, lu_block=args.lu_block)
        if applied:
            print(f"[bench] {precision}: tuning {applied}", file=sys.stderr, flush=True)
        sweep = FrequencySweep(gf, body, ks, nrhs, precision=precision, max_iter=10, tol=1e-12,
                               far_field_fp32=far_field_fp32)
        for _ in range(args.warmup):
            sweep.run(bc)
        torch.cuda.synchronize()
        phases = {"assembly": 0.0, "lu_factor": 0.0, "lu_solve": 0.0, "post": 0.0}
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(5 * args.steps)]
        torch.cuda.synchronize()
        _lib.profile_begin(sweep.ctx)
        t0 = time.perf_counter()
        for s in range(args.steps):
            e = ev[5 * s:5 * s + 5]
            e[0].record(stream)
            sweep.assemble()
            e[1].record(stream)
            sweep.factor()
            e[2].record(stream)
            sweep.solve(bc)
            e[3].record(stream)
            sweep.post()
            e[4].record(stream)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        kernels = _lib.profile_end(sweep.ctx)
        for s in range(args.steps):
            e = ev[5 * s:5 * s + 5]
            for name, (a, b) in zip(phases, [(0, 1), (1, 2), (2, 3), (3, 4)]):
                phases[name] += e[a].elapsed_time(e[b]) / 1e3
        forces = sweep.forces(omegas_arr, rho)
        out = dict(seconds=t1 - t0, phases={k: v / args.steps for k, v in phases.items()}, forces=forces,
                   kernels=kernels,
                   iterations=sweep.iterations, residual=sweep.residual,
                   info_nonzero=int(np.count_nonzero(sweep.info.cpu().numpy())) if precision == "fp64" else 0)
        del sweep
        torch.cuda.empty_cache()
        return out

    if args.only:
        # one form, for profiler counter passes: same kernels and shapes
        r = run(args.only)
        print(f"[bench] {args.only}: {r['seconds'] / args.steps * 1e3:.1f} ms per step", file=sys.stderr, flush=True)
        return
    r64 = run("fp64")
    rmx = run("mixed")  # far pairs from the float table: the mixed form's default
    rfa = run("mixed", far_field_fp32=False)  # the same on the FP64 assembly
    asm_err = assembly_error(gf, mesh, [ks[0], ks[len(ks) // 2], ks[-1]], _lib)
    nb = len(ks)
    nprob = nb * nrhs
    value64 = nprob * args.steps / r64["seconds"]
    valuemx = nprob * args.steps / rmx["seconds"]
    # per-coefficient agreement of the mixed form with the FP64 form
    f64, fmx = r64["forces"], rmx["forces"]
    scale = np.abs(f64).max(axis=2, keepdims=True)  # per problem
    mixed_err = float(np.max(np.abs(fmx - f64) / scale))
    fa_err = float(np.max(np.abs(rfa["forces"] - f64) / scale))
    t_asm = r64["phases"]["assembly"]
    t_lu = r64["phases"]["lu_factor"]
    pairs_per_s = nb * n * n / t_asm
    lu_tflops = nb * 8.0 * n ** 3 / 3.0 / t_lu / 1e12
    # roofline of the dominant kernel of the reported (mixed) form, from the
    # live HIP-event timing of every launch of it in the timed region
    kern = rmx["kernels"]
    dom = max(("zgemm", "cgemm", "skinny_z", "skinny_c", "cgemm_packed"), key=lambda k: kern[k]["ms"])
    kd = kern[dom]
    # the complex64 product the tuner chose for the mixed form
    ctiles = _lib.get_tuning(ctx)["gemm_tiles"]
    cname = {0: "gemm_tiled_kernel<float,64>", 1: "gemm_tiled_kernel<float,128>",
             2: "gemm_tiled_gauss_w1_kernel<float,128>", 3: "gemm_tiled_kernel<float,64,true>",
             4: "gemm_h16_kernel<16>"}[ctiles]
    cnote = (", three real products per complex one" if ctiles in (2, 3) else "") + \
        (", FP16 inputs on two planes, four real products per complex one, FP32 accumulation" if ctiles == 4 else "")
    names = {"zgemm": "gemm_tiled_kernel<double,64> (complex128 MFMA tile product)",
             "cgemm": f"{cname} (complex64 MFMA tile product{cnote})",
             "skinny_z": "gemm_skinny_kernel<double> (complex128 product, <= 32 columns)",
             "skinny_c": "gemm_skinny_kernel<float> (complex64 product, <= 32 columns)",
             "cgemm_packed": "gemm_h16p_kernel (complex64 trailing update on FP16 operands packed once per block "
                             "step as two planes, v_mfma_f32_16x16x32_f16, four real products per complex one, "
                             "FP32 accumulation)"}
    if dom.startswith("skinny"):
        achieved = kd["work"] / (kd["ms"] / 1e3) / 1e9
        roofline = {"bound": "hbm", "kernel": names[dom], "achieved": achieved, "peak": MI355X_HBM_GBS,
                    "unit": "GB/s", "frac": achieved / MI355X_HBM_GBS, "traffic": None,
                    "work_per_launch": "bytes of A, B, C per product, summed over the batch"}
    else:
        achieved = kd["work"] / (kd["ms"] / 1e3) / 1e12
        peak = MI355X_FP64_MFMA_TFLOPS if dom == "zgemm" else MI355X_FP16_MFMA_TFLOPS \
            if dom == "cgemm_packed" or ctiles == 4 else MI355X_FP32_MFMA_TFLOPS
        roofline = {"bound": "mfma", "kernel": names[dom], "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                    "frac": achieved / peak, "traffic": None,
                    "work_per_launch": "8 M N K flop per complex M x N x K product, summed over the batch"}
    roofline.update(launches=kd["launches"], avg_launch_us=kd["ms"] / kd["launches"] * 1e3,
                    work_per_launch_avg=kd["work"] / kd["launches"],
                    algorithmic_bytes_per_launch=kd["bytes"] / kd["launches"])
    if dom == "cgemm" and ctiles in (2, 3):
        roofline["mfma_flops_note"] = ("the Gauss form executes 6 M N K real MFMA flops for the 8 M N K counted in "
                                       "`achieved`; its MFMA pipes run at 3/4 of `achieved`")
    traffic = measured_traffic(names[dom].split(" ")[0])
    if traffic:
        roofline["traffic"] = traffic
        roofline["traffic_over_algorithmic"] = traffic / roofline["algorithmic_bytes_per_launch"]
        roofline["traffic_source"] = "profiles/c2_mixed_hbm_traffic_counters.md (rocprofv3 FETCH_SIZE, " \
                                     "WRITE_SIZE, scaled by a copy kernel)"

    def kernel_table(r):
        out = {}
        for k, v in r["kernels"].items():
            if not v["launches"]:
                continue
            rate = v["work"] / (v["ms"] / 1e3) if v["ms"] else 0.0
            unit = {"assembly": ("gpairs_per_s", 1e9)}.get(k, ("gb_per_s", 1e9) if k.startswith("skinny")
                                                          else ("tflops", 1e12))
            out[k] = {"ms_per_step": v["ms"] / args.steps, "launches_per_step": v["launches"] / args.steps,
                      unit[0]: rate / unit[1]}
        return out
    out = {
        "metric": METRIC,
        "value": valuemx,
        "unit": "problems/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": rmx["seconds"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 factor" + (" (FP16-input tile products)" if ctiles == 4 else "") +
                 " + f64 refinement, far pairs from an f32 table (mixed); f64 reported beside",
        "data": "synthetic",
        "config": {"workload": "barge 20x8x2 m, 4352 panels, 40 frequencies 0.2-2.5 rad/s, 6 dof + 8 headings "
                               "(560 problems), infinite depth; mixed-precision batched form vs FP64 plain form",
                   "panels": n, "frequencies": nb, "rhs_per_frequency": nrhs},
        "fp64_problems_per_s": value64,
        "fp64_ms_per_step": r64["seconds"] / args.steps * 1e3,
        "mixed_over_fp64": valuemx / value64,
        "mixed_max_rel_err_vs_fp64": mixed_err,
        "coefficient_errors_mixed_vs_fp64": coefficient_errors(fmx, f64, omegas_arr, len(body.dofs)),
        "far_field_assembly_max_rel_err": {"S": asm_err["S"], "K": asm_err["K"],
                                           "wavenumbers": "first, middle and last of the sweep",
                                           "assembly_panel_pairs_per_s": nb * n * n / rmx["phases"]["assembly"]},
        "mixed_fp64_assembly": {
            "problems_per_s": nprob * args.steps / rfa["seconds"],
            "phase_ms_per_step": {k: v * 1e3 for k, v in rfa["phases"].items()},
            "assembly_panel_pairs_per_s": nb * n * n / rfa["phases"]["assembly"],
            "max_rel_err_vs_fp64": fa_err},
        "mixed_refinement_sweeps": rmx["iterations"],
        "mixed_final_rel_residual": rmx["residual"],
        "assembly_panel_pairs_per_s": pairs_per_s,
        "phase_ms_per_step_fp64": {k: v * 1e3 for k, v in r64["phases"].items()},
        "phase_ms_per_step_mixed": {k: v * 1e3 for k, v in rmx["phases"].items()},
        "lu_info_nonzero": r64["info_nonzero"],
        "fp64_lu_factor_tflops": lu_tflops,
        "kernels_fp64": kernel_table(r64),
        "kernels_mixed": kernel_table(rmx),
        "roofline": roofline,
        "tuning": {k: {kk: vv for kk, vv in v.items() if kk != "timings_ms"} for k, v in tuning.items()},
    }
    if not args.no_symmetric:
        out["symmetric"] = symmetric_benchmark(gf, dev, steps=args.steps, warmup=args.warmup, rho=rho, g=g,
                                               ref_forces=r64["forces"],
                                               log=lambda m: print(m, file=sys.stderr, flush=True))
    if not args.no_post:
        out["post_processing"] = post_processing_benchmark(gf, dev, mesh, body, headings, rho=rho, g=g,
                                                           log=lambda m: print(m, file=sys.stderr, flush=True))
    if not args.no_fleet:
        out["fleet"] = fleet_benchmark(gf, dev, log=lambda m: print(m, file=sys.stderr, flush=True))
    if not args.no_fingreen3d:
        # the FinGreen3D finite-depth Green function on the 20000-panel array
        out["fingreen3d"] = fingreen3d_benchmark(no_cpu_baseline=args.no_cpu_baseline)
    if not args.no_serving:
        _lib.check(_lib.load().cpt_release_workspace(ctx.handle), "cpt_release_workspace")
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        out["serving"] = serving_child(requests=args.requests)
        loads = out["serving"].get("loads", [])
        out["p99_latency_s_at_saturation"] = next((r["p99_s"] for r in loads if r["load"] == 1.0), None)
    if not args.no_cpu_baseline:
        out["cpu_baseline"], idx, f_orc = cpu_baseline(
            mesh, body, pbs_by_freq, (gf.tabulated_r_range, gf.tabulated_z_range, gf.tabulated_integrals))
        out["coefficient_errors_vs_oracle"] = {
            "frequencies_rad_s": [float(omegas_arr[i]) for i in idx],
            "fp64": coefficient_errors(f64[idx], f_orc, omegas_arr[idx], len(body.dofs)),
            "mixed": coefficient_errors(fmx[idx], f_orc, omegas_arr[idx], len(body.dofs))}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
