"""Admission and batching rule of the serving loop (host logic, no GPU)."""
import collections
from types import SimpleNamespace

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
