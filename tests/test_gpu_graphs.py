"""Per-frequency-batch HIP graphs (capytaine_amd.graphs.SweepGraph): the
FP64 sweep step captured once and replayed for new wavenumbers and boundary
conditions must give the eager FrequencySweep's forces (the same kernels on
the same inputs: equal to rounding, checked at 1e-13 of the largest force)
on every replay, with eager work and synchronisations between the replays;
infinite and finite depth.  The replay and eager step times of a small body
are printed (the graph saves the launches of the blocked LU and solves)."""
import time

import numpy as np
import pytest

from gpu_helpers import device

pytestmark = pytest.mark.gpu


def _body(resolution):
    import capytaine_amd as cpa
    mesh = cpa.mesh_parallelepiped(size=(8.0, 4.0, 2.0), center=(0, 0, -1.0), resolution=resolution,
                                   missing_sides={"top"})
    return cpa.FloatingBody(mesh).add_all_rigid_body_dofs(rotation_center=(0, 0, 0))


@pytest.mark.parametrize("depth", [np.inf, 6.0])
def test_sweep_graph_replays_match_eager(depth):
    import torch
    import capytaine_amd as cpa
    from capytaine_amd import workloads
    from capytaine_amd.graphs import SweepGraph
    from capytaine_amd.sweep import FrequencySweep, boundary_conditions
    body = _body((12, 6, 4))
    n = body.mesh.nb_faces
    headings = [0.0, 0.6]
    nrhs = len(body.dofs) + len(headings)
    print(f"[gpu] sweep graph, {n} panels, depth {depth}", flush=True)
    gf = cpa.Delhommeau(tabulation_cache_dir=None, device=device())
    graph = SweepGraph(gf, body, 3, nrhs, water_depth=depth)
    for trial, omegas in enumerate(([0.5, 1.0, 1.5], [0.7, 1.2, 2.1], [0.5, 1.0, 1.5], [0.9, 0.3, 2.6])):
        pbs = workloads.problems_by_frequency(body, np.array(omegas), headings, rho=1025.0, water_depth=depth)
        ks = np.array([p[0].wavenumber for p in pbs])
        bc = boundary_conditions(pbs, nrhs, n, device())
        G = graph.run(ks, bc).clone()
        assert np.all(graph.check_info() == 0)
        # eager work between the replays
        sw = FrequencySweep(gf, body, ks, nrhs, water_depth=depth)
        sw.run(bc)
        torch.cuda.synchronize()
        Ge = sw.G
        err = (torch.abs(G - Ge).max() / torch.abs(Ge).max()).item()
        same = bool(torch.equal(G, Ge))
        print(f"[gpu]   replay {trial}: max |graph - eager| / max |eager| = {err:.1e} (bitwise equal: {same})",
              flush=True)
        assert err <= 1e-13
        del sw


def test_sweep_graph_step_time():
    """A 216-panel body, 4 frequencies, 10 right-hand sides: replay time
    against the eager step (printed; the replay must not be slower than the
    eager step by more than its own noise)."""
    import torch
    import capytaine_amd as cpa
    from capytaine_amd import workloads
    from capytaine_amd.graphs import SweepGraph
    from capytaine_amd.sweep import FrequencySweep, boundary_conditions
    body = _body((12, 6, 4))
    n = body.mesh.nb_faces
    headings = [0.0, 0.5, 1.0, 1.5]
    nrhs = len(body.dofs) + len(headings)
    omegas = np.array([0.4, 0.9, 1.4, 2.0])
    pbs = workloads.problems_by_frequency(body, omegas, headings, rho=1025.0)
    ks = np.array([p[0].wavenumber for p in pbs])
    bc = boundary_conditions(pbs, nrhs, n, device())
    gf = cpa.Delhommeau(tabulation_cache_dir=None, device=device())
    graph = SweepGraph(gf, body, len(ks), nrhs)
    sw = FrequencySweep(gf, body, ks, nrhs)
    reps = 20
    for name, step in (("eager", lambda: sw.run(bc)), ("graph", lambda: graph.run(ks, bc))):
        step()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            step()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / reps
        print(f"[gpu]   {name}: {dt * 1e3:.3f} ms per step ({n} panels, {len(ks)} frequencies, {nrhs} rhs)",
              flush=True)
        if name == "eager":
            t_eager = dt
    err = (torch.abs(graph.G - sw.G).max() / torch.abs(sw.G).max()).item()
    assert err <= 1e-13
    assert dt <= 1.2 * t_eager
