"""Benchmark of the BEM hot path on one MI355X.

Workload (one step): rectangular barge 20 m x 8 m, draft 2 m, 0.25 m panels
(4352 panels), infinite depth, rho = 1025, g = 9.81; 40 frequencies uniform
in 0.2-2.5 rad/s; 6 rigid-body radiation problems + diffraction at 8
headings (0-315 deg) per frequency: 560 problems.  The step runs, on the
device with the boundary conditions already in HBM: assembly of S and K for
the 40 wavenumbers (one launch), 40 complex128 LU factorisations (batched),
the 560 solves, the potentials S @ sources and the generalised forces.

Prints one JSON line (see the repository's bench contract).
    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
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


def cpu_baseline(mesh, body, pbs_by_freq, table, n_freq=2):
    """The oracle (C restatement of the reference core, OpenMP on all host
    cores) + scipy LU, on a bounded sample: the first and last frequencies
    of the sweep, every problem of those frequencies."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle, mesh_arrays, solve_coefficients
    orc = Oracle()
    arr = mesh_arrays(mesh)
    idx = [0, len(pbs_by_freq) - 1][:n_freq]
    t0 = time.perf_counter()
    nprob = 0
    for i in idx:
        pbs = pbs_by_freq[i]
        S, K = orc.build_matrices(**arr, wavenumber=pbs[0].wavenumber, depth=np.inf, table=table)
        solve_coefficients(S, K, body, pbs)
        nprob += len(pbs)
    dt = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=nprob / dt, unit="problems/s", cores=threads, kind="port",
                sample=f"{len(idx)} of 40 frequencies ({nprob} problems), oracle assembly + scipy LU, "
                       f"{dt:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the frequency sweep shards by "
              "frequency across GPUs by construction but this build measures a single GPU.", file=sys.stderr)
        sys.exit(2)

    import torch
    import capytaine_amd as cpa
    from capytaine_amd.sweep import FrequencySweep, boundary_conditions

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    rho, g = 1025.0, 9.81
    mesh, body, omegas, headings = barge_workload()
    pbs_by_freq = problems_by_frequency(body, omegas, headings, rho, g)
    n = mesh.nb_faces
    nrhs = len(pbs_by_freq[0])
    ks = [p[0].wavenumber for p in pbs_by_freq]
    gf = cpa.Delhommeau(device=dev)
    sweep = FrequencySweep(gf, body, ks, nrhs)
    bc = boundary_conditions(pbs_by_freq, nrhs, n, dev)
    stream = torch.cuda.current_stream(dev)

    for _ in range(args.warmup):
        sweep.run(bc)
    torch.cuda.synchronize()

    phases = {"assembly": 0.0, "lu_factor": 0.0, "lu_solve": 0.0, "post": 0.0}
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(5 * args.steps)]
    torch.cuda.synchronize()
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
    for s in range(args.steps):
        e = ev[5 * s:5 * s + 5]
        for name, (a, b) in zip(phases, [(0, 1), (1, 2), (2, 3), (3, 4)]):
            phases[name] += e[a].elapsed_time(e[b]) / 1e3
    info = sweep.info.cpu().numpy()
    ms_per_step = (t1 - t0) / args.steps * 1e3
    nprob = len(ks) * nrhs
    value = nprob * args.steps / (t1 - t0)
    nb = len(ks)
    t_asm = phases["assembly"] / args.steps
    t_lu = phases["lu_factor"] / args.steps
    pairs_per_s = nb * n * n / t_asm
    lu_flops = nb * 8.0 * n ** 3 / 3.0
    lu_tflops = lu_flops / t_lu / 1e12
    dominant = max(phases, key=phases.get)
    roofline = {"bound": "mfma", "kernel": "lu_factor (panel + laswp + trsm + zgemm, batched)",
                "achieved": lu_tflops, "peak": MI355X_FP64_MFMA_TFLOPS, "unit": "TFLOP/s",
                "frac": lu_tflops / MI355X_FP64_MFMA_TFLOPS, "traffic": None,
                "work_per_launch": f"{nb} x 8 n^3/3 flop, n = {n}"}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "problems/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "barge 20x8x2 m, 4352 panels, 40 frequencies 0.2-2.5 rad/s, 6 dof + 8 headings "
                               "(560 problems), infinite depth, FP64 plain form",
                   "panels": n, "frequencies": nb, "rhs_per_frequency": nrhs, "precision": "fp64"},
        "assembly_panel_pairs_per_s": pairs_per_s,
        "phase_ms_per_step": {k: v / args.steps * 1e3 for k, v in phases.items()},
        "dominant_phase": dominant,
        "lu_info_nonzero": int(np.count_nonzero(info)),
        "roofline": roofline,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(mesh, body, pbs_by_freq,
                                           (gf.tabulated_r_range, gf.tabulated_z_range, gf.tabulated_integrals))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
