"""Admission and batching rule of the serving loop and a request's
incident-wave terms (host logic, no GPU)."""
import collections
from types import SimpleNamespace

import numpy as np

from capytaine_amd.serving import select_batch, size_class


def _jobs(*spec):
    return collections.deque(SimpleNamespace(nbytes=b, cls=c, name=i) for i, (b, c) in enumerate(spec))


def test_size_classes():
    assert [size_class(n) for n in (300, 999, 1000, 2499, 2500, 5999, 6000, 12000)] == [0, 0, 1, 1, 2, 2, 3, 3]


def test_batch_takes_same_class_in_order_within_memory():
    w = _jobs((10, 0), (50, 1), (20, 0), (30, 0), (5, 0))
    batch, nbytes = select_batch(w, used=0, budget=60, inflight=0, max_jobs=16)
    assert [j.name for j in batch] == [0, 2, 3] and nbytes == 60  # job 4 would exceed the budget


def test_batch_size_limit():
    w = _jobs(*[(1, 2)] * 10)
    batch, _ = select_batch(w, used=0, budget=100, inflight=0, max_jobs=4)
    assert [j.name for j in batch] == [0, 1, 2, 3]


def test_head_waits_for_memory_no_overtaking():
    w = _jobs((80, 1), (5, 1), (5, 0))
    assert select_batch(w, used=30, budget=100, inflight=1, max_jobs=16) is None


def test_oversized_head_runs_alone_when_idle():
    w = _jobs((150, 3), (5, 3))
    batch, nbytes = select_batch(w, used=0, budget=100, inflight=0, max_jobs=16)
    assert [j.name for j in batch] == [0] and nbytes == 150


def test_empty_queue():
    assert select_batch(collections.deque(), used=0, budget=1, inflight=0, max_jobs=1) is None


def test_request_terms_match_problem_objects():
    """A request's boundary conditions and Froude-Krylov forces, formed for
    all of its frequencies and headings at once, equal those of the
    RadiationProblem / DiffractionProblem objects one by one."""
    from capytaine_amd import workloads
    from capytaine_amd.problems import DiffractionProblem, froude_krylov_force
    from capytaine_amd.serving import _Request
    rng = np.random.default_rng(5)
    body, om, _ = workloads.serving_request(rng, n_range=(150, 250))
    hd = np.array([0.0, 0.7, np.pi])
    req = _Request(0, body, om, hd, 1025.0, 9.81, "mixed")
    bc = req.boundary_conditions(np.arange(len(om)))
    nd = len(req.dofs)
    dn = body.dof_normal_matrix()
    req.forces[:] = 0
    fk = req.result()["froude_krylov_force"]
    for i, w in enumerate(om):
        np.testing.assert_array_equal(bc[i, :nd], -1j * w * dn)
        for h, beta in enumerate(hd):
            pb = DiffractionProblem(body=body, omega=w, wave_direction=beta, rho=1025.0, g=9.81)
            ref = pb.boundary_condition
            assert np.abs(bc[i, nd + h] - ref).max() <= 1e-14 * np.abs(ref).max()
            f = froude_krylov_force(pb)
            ref = np.array([f[d] for d in req.dofs])
            assert np.abs(fk[i, h] - ref).max() <= 1e-12 * np.abs(ref).max()


def test_first_chunk_of_a_new_request_goes_ahead_of_later_chunks():
    """Round robin over requests: the rank-0 chunk of a later request opens
    the next batch before the rank-1 and rank-2 chunks of an earlier one."""
    w = collections.deque([SimpleNamespace(nbytes=10, cls=3, name="big1", rank=1),
                           SimpleNamespace(nbytes=10, cls=3, name="big2", rank=2),
                           SimpleNamespace(nbytes=5, cls=0, name="small0", rank=0)])
    batch, nbytes = select_batch(w, used=0, budget=100, inflight=0, max_jobs=16)
    assert [j.name for j in batch] == ["small0"] and nbytes == 5
    w.remove(batch[0])
    batch, _ = select_batch(w, used=0, budget=100, inflight=0, max_jobs=16)
    assert [j.name for j in batch] == ["big1", "big2"]
