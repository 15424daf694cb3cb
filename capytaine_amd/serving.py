"""Serving loop: continuous batching of BEM requests on one GPU.

A request is a body (mesh + dofs), a list of frequencies and a list of wave
headings; its answer is what capytaine's ``BEMSolver.solve_all`` returns
for the radiation problems of every dof and the diffraction problems of
every heading at every frequency (solver.py:268-372): added mass,
radiation damping, diffraction and Froude-Krylov forces.

Design
  * Requests become jobs of at most `max_job_bytes` of device memory
    (frequency chunks of a request whose matrices do not fit at once).
  * Admission is by device-memory footprint (S, K, the complex64 copy and
    the right-hand sides: about N^2 x 40 bytes per frequency in mixed
    precision) against `memory_budget`; a job waits until in-flight work has
    released enough memory.  Jobs start in arrival order (no overtaking of a
    waiting job, so large requests are not starved).
  * Jobs are batched by size class of their panel count: the oldest waiting
    job opens a batch, which takes the waiting jobs of the same class in
    arrival order while memory and `max_batch_jobs` allow.
  * A batch runs on one of `slots` workers, each with its own HIP stream and
    its own library context (workspace): one persistent assembly launch for
    all its bodies and frequencies, one grouped LU, the forces.  Batches on
    different slots overlap on the device; a batch admitted while others run
    joins them as soon as a slot and memory are free (continuous batching,
    not a queue of whole-device jobs).
  * The Green function's tabulation is resident in every context; meshes
    keep their device arrays; each factorisation is computed once per
    (body, frequency) and serves all of its right-hand sides.

Every job's arithmetic is independent of what it is batched with (the
assembly tile and grouped-LU arithmetic is per body), so a result is the
same alone and in a mix.
"""
import collections
import concurrent.futures
import threading
import time

import numpy as np

from . import _lib
from .batch import BatchedSweep, footprint_bytes
from .problems import DiffractionProblem, froude_krylov_force, wavenumber_from_omega

SIZE_CLASSES = (1000, 2500, 6000)


def size_class(n):
    for i, edge in enumerate(SIZE_CLASSES):
        if n < edge:
            return i
    return len(SIZE_CLASSES)


class _Request:
    def __init__(self, rid, body, omegas, headings, rho, g, precision):
        self.id = rid
        self.body = body
        self.omegas = np.asarray(omegas, dtype=float)
        self.headings = np.asarray(headings, dtype=float)
        self.dofs = list(body.dofs)
        self.ks = np.array([wavenumber_from_omega(om, g, np.inf) for om in self.omegas])
        self.n = body.mesh.nb_faces
        self.nrhs = len(self.dofs) + len(self.headings)
        self.rho, self.g = rho, g
        self.future = concurrent.futures.Future()
        self.t_submit = time.perf_counter()
        self.t_start = None
        self.forces = np.empty((len(self.omegas), self.nrhs, len(self.dofs)), dtype=complex)
        self.pending_jobs = 0
        self.iterations = 0
        self.residual = 0.0
        self.precision = precision

    def boundary_conditions(self, idx):
        """(len(idx), nrhs, n) host array: radiation then diffraction."""
        dn = self.body.dof_normal_matrix()  # (ndof, n)
        out = np.empty((len(idx), self.nrhs, self.n), dtype=complex)
        for a, i in enumerate(idx):
            om = self.omegas[i]
            out[a, :len(self.dofs)] = -1j * om * dn
            for h, beta in enumerate(self.headings):
                pb = DiffractionProblem(body=self.body, omega=om, wave_direction=beta, rho=self.rho, g=self.g)
                out[a, len(self.dofs) + h] = pb.boundary_condition
        return out

    def result(self):
        """Coefficients per frequency; F = i omega rho W^T S sigma."""
        om = self.omegas
        nd = len(self.dofs)
        F = self.forces
        rad = F[:, :nd, :]  # [omega, radiating dof, influenced dof]
        added_mass = np.real(rad) / om[:, None, None] ** 2
        damping = np.imag(rad) / om[:, None, None]
        diffraction = F[:, nd:, :]
        fk = np.zeros_like(diffraction)
        for i, w in enumerate(om):
            for h, beta in enumerate(self.headings):
                pb = DiffractionProblem(body=self.body, omega=w, wave_direction=beta, rho=self.rho, g=self.g)
                f = froude_krylov_force(pb)
                fk[i, h] = [f[d] for d in self.dofs]
        return {"omegas": om, "dofs": self.dofs, "headings": self.headings, "added_mass": added_mass,
                "radiation_damping": damping, "diffraction_force": diffraction, "froude_krylov_force": fk,
                "excitation_force": diffraction + fk, "forces": F, "refinement_sweeps": self.iterations,
                "final_rel_residual": self.residual,
                "latency_s": time.perf_counter() - self.t_submit}


class _Job:
    def __init__(self, req, idx, nbytes):
        self.req, self.idx, self.nbytes = req, np.asarray(idx), nbytes
        self.cls = size_class(req.n)


def select_batch(waiting, used, budget, inflight, max_jobs):
    """Admission and batching rule: the oldest waiting job opens the batch
    if its bytes fit beside the `used` in-flight bytes (or nothing is in
    flight: a job larger than the free memory runs alone); later jobs of the
    same size class join in arrival order while they fit, up to `max_jobs`.
    Returns (jobs, bytes) or None (the head waits: no overtaking)."""
    if not waiting:
        return None
    head = waiting[0]
    if used + head.nbytes > budget and inflight > 0:
        return None
    batch, nbytes = [head], head.nbytes
    for job in list(waiting)[1:]:
        if len(batch) >= max_jobs:
            break
        if job.cls == head.cls and used + nbytes + job.nbytes <= budget:
            batch.append(job)
            nbytes += job.nbytes
    return batch, nbytes


class BEMServer:
    """Continuous-batching BEM server on one GPU (see the module docstring).

        with BEMServer(Delhommeau(device=dev)) as srv:
            fut = srv.submit(body, omegas, headings)
            res = fut.result()   # dict of coefficients, latency_s
    """

    def __init__(self, green_function, *, precision="mixed", slots=3, memory_budget=None, max_job_bytes=None,
                 max_batch_jobs=16, rho=1000.0, g=9.81, max_iter=10, tol=1e-12):
        import torch
        self.gf = green_function
        self.dev = green_function.device
        if memory_budget is None:
            memory_budget = int(0.5 * torch.cuda.get_device_properties(self.dev).total_memory)
        self.memory_budget = int(memory_budget)
        self.max_job_bytes = int(max_job_bytes or self.memory_budget // 2)
        self.max_batch_jobs = max_batch_jobs
        self.precision = precision
        self.rho, self.g = rho, g
        self.max_iter, self.tol = max_iter, tol
        self._cv = threading.Condition()
        self._waiting = collections.deque()
        self._free_slots = collections.deque(range(slots))
        self._used = 0
        self._closed = False
        self._next_id = 0
        self.latencies = []
        self.completed = 0
        self.t0 = time.perf_counter()
        self.max_inflight_batches = 0
        self._inflight = 0
        self._contexts = [_lib.Context(self.dev.index if self.dev.index is not None else 0) for _ in range(slots)]
        self._streams = [torch.cuda.Stream(device=self.dev) for _ in range(slots)]
        self._pools = [concurrent.futures.ThreadPoolExecutor(max_workers=1) for _ in range(slots)]
        self._scheduler = threading.Thread(target=self._schedule, name="bem-scheduler", daemon=True)
        self._scheduler.start()

    # -- client side ------------------------------------------------------
    def submit(self, body, omegas, headings=()):
        req = _Request(self._next_id, body, omegas, headings, self.rho, self.g, self.precision)
        self._next_id += 1
        per_freq = footprint_bytes(req.n, 1, req.nrhs, self.precision)
        if per_freq > self.memory_budget:
            req.future.set_exception(MemoryError(f"one frequency of a {req.n}-panel body needs {per_freq} bytes, "
                                                 f"over the memory budget {self.memory_budget}"))
            return req.future
        per_job = max(1, min(len(req.omegas), self.max_job_bytes // per_freq))
        chunks = [list(range(i, min(i + per_job, len(req.omegas)))) for i in range(0, len(req.omegas), per_job)]
        req.pending_jobs = len(chunks)
        with self._cv:
            if self._closed:
                raise RuntimeError("server is closed")
            for c in chunks:
                self._waiting.append(_Job(req, c, per_freq * len(c)))
            self._cv.notify_all()
        return req.future

    def close(self):
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        self._scheduler.join()
        for p in self._pools:
            p.shutdown(wait=True)
        for c in self._contexts:
            c.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def stats(self):
        lat = np.array(self.latencies) if self.latencies else np.zeros(1)
        return {"completed": self.completed, "p50_s": float(np.percentile(lat, 50)),
                "p99_s": float(np.percentile(lat, 99)), "max_inflight_batches": self.max_inflight_batches}

    # -- scheduler ----------------------------------------------------------
    def _take_batch(self):
        """Called with the lock held: the next batch that fits, or None."""
        if not self._free_slots:
            return None
        picked = select_batch(self._waiting, self._used, self.memory_budget, self._inflight, self.max_batch_jobs)
        if picked is None:
            return None
        batch, nbytes = picked
        for job in batch:
            self._waiting.remove(job)
        self._used += nbytes
        return batch, nbytes, self._free_slots.popleft()

    def _schedule(self):
        while True:
            with self._cv:
                while True:
                    if self._closed and not self._waiting and self._inflight == 0:
                        return
                    taken = self._take_batch()
                    if taken is not None:
                        break
                    self._cv.wait()
                batch, nbytes, slot = taken
                self._inflight += 1
                self.max_inflight_batches = max(self.max_inflight_batches, self._inflight)
            self._pools[slot].submit(self._run_batch, batch, nbytes, slot)

    def _run_batch(self, batch, nbytes, slot):
        import torch
        try:
            with torch.cuda.device(self.dev), torch.cuda.stream(self._streams[slot]):
                for job in batch:
                    if job.req.t_start is None:
                        job.req.t_start = time.perf_counter()
                bodies = [job.req.body for job in batch]
                ks = [job.req.ks[job.idx] for job in batch]
                nrhs = [job.req.nrhs for job in batch]
                bcs = [torch.as_tensor(job.req.boundary_conditions(job.idx)).to(self.dev) for job in batch]
                sweep = BatchedSweep(self.gf, bodies, ks, nrhs, precision=self.precision, max_iter=self.max_iter,
                                     tol=self.tol, context=self._contexts[slot])
                sweep.run(bcs)
                G = [g.cpu().numpy() for g in sweep.G]  # synchronises this slot's stream
                its, res = sweep.iterations, sweep.residual
                del sweep, bcs
            for b, job in enumerate(batch):
                om = job.req.omegas[job.idx]
                job.req.forces[job.idx] = 1j * om[:, None, None] * job.req.rho * G[b]
                if self.precision == "mixed":
                    job.req.iterations = max(job.req.iterations, int(its[b]))
                    job.req.residual = max(job.req.residual, float(res[b]))
            error = None
        except Exception as exc:  # reported through the futures
            error = exc
        finished = []
        with self._cv:
            self._used -= nbytes
            self._free_slots.append(slot)
            self._inflight -= 1
            for job in batch:
                job.req.pending_jobs -= 1
                if error is not None or job.req.pending_jobs == 0:
                    finished.append(job.req)
            self._cv.notify_all()
        for req in finished:
            if req.future.done():
                continue
            if error is not None:
                req.future.set_exception(error)
                continue
            res = req.result()
            with self._cv:
                self.latencies.append(res["latency_s"])
                self.completed += 1
            req.future.set_result(res)
