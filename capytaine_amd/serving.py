"""Serving loop: continuous batching of BEM requests on one GPU.

A request is a body (mesh + dofs), a list of frequencies and a list of wave
headings; its answer is what capytaine's ``BEMSolver.solve_all`` returns
for the radiation problems of every dof and the diffraction problems of
every heading at every frequency (solver.py:268-372): added mass,
radiation damping, diffraction and Froude-Krylov forces.

Design
  * Requests become jobs of at most `max_job_bytes` of device memory
    (frequency chunks; by default an eighth of the budget, so that a large
    request is served in chunks that take turns with other requests).
  * Admission is by device-memory footprint (S, K, the complex64 copy and
    the right-hand sides: about N^2 x 40 bytes per frequency in mixed
    precision) against `memory_budget`; a job waits until in-flight work has
    released enough memory.  Service order is round robin over requests (a
    job's rank is its index among its request's chunks), then arrival; the
    first job in that order is never overtaken, so nothing is starved.
  * Jobs are batched by size class of their panel count: the first waiting
    job in service order opens a batch, which takes the waiting jobs of the
    same class in that order while memory and `max_batch_jobs` allow.
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
from .problems import wavenumber_from_omega

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
        self.n = body.mesh_including_lid.nb_faces
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

    def _incident(self, idx):
        """Deep-water Airy wave at the panel centres for the frequencies idx
        and every heading, (F, H, n) each: exp(k z) exp(i k (x cos b +
        y sin b)) and the normal factor cos b n_x + sin b n_y - i n_z, the
        terms of DiffractionProblem.boundary_condition and of the
        Froude-Krylov pressure (problems.airy_waves_velocity / _potential,
        airy_waves.py:19-94, no forward speed) for all problems at once."""
        mesh = self.body.mesh
        x, y, z = np.asarray(mesh.faces_centers).T
        nx, ny, nz = np.asarray(mesh.faces_normals).T
        k = self.ks[idx][:, None, None]
        cb = np.cos(self.headings)[None, :, None]
        sb = np.sin(self.headings)[None, :, None]
        wave = np.exp(k * z) * np.exp(1j * k * (x * cb + y * sb))
        return wave, cb * nx + sb * ny - 1j * nz

    def boundary_conditions(self, idx):
        """(len(idx), nrhs, n) host array: radiation then diffraction."""
        idx = np.asarray(idx)
        om = self.omegas[idx]
        dn = self.body.dof_normal_matrix(including_lid=True)  # (ndof, n), zero on a lid
        nd = len(self.dofs)
        out = np.empty((len(idx), self.nrhs, self.n), dtype=complex)
        out[:, :nd] = -1j * om[:, None, None] * dn[None]
        if len(self.headings):
            wave, normal = self._incident(idx)
            out[:, nd:] = 0.0
            out[:, nd:, self.body.hull_mask] = -(self.g * self.ks[idx] / om)[:, None, None] * wave * normal
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
        if len(self.headings):
            # pressure rho g exp(kz) exp(i k w); force on dof d: -sum p (d . n) A
            wave, _ = self._incident(np.arange(len(om)))
            dna = self.body.dof_normal_matrix() * np.asarray(self.body.mesh.faces_areas)[None]
            fk[:] = -(self.rho * self.g) * (wave @ dna.T)
        return {"omegas": om, "dofs": self.dofs, "headings": self.headings, "added_mass": added_mass,
                "radiation_damping": damping, "diffraction_force": diffraction, "froude_krylov_force": fk,
                "excitation_force": diffraction + fk, "forces": F, "refinement_sweeps": self.iterations,
                "final_rel_residual": self.residual,
                "latency_s": time.perf_counter() - self.t_submit}


class _Job:
    def __init__(self, req, idx, nbytes, rank=0):
        self.req, self.idx, self.nbytes, self.rank = req, np.asarray(idx), nbytes, rank
        self.cls = size_class(req.n)


def _order(job, pos):
    """Service order: round robin over requests (the rank of a job is its
    index among its request's frequency chunks), then arrival."""
    return (getattr(job, "rank", 0), pos)


def select_batch(waiting, used, budget, inflight, max_jobs):
    """Admission and batching rule: the first waiting job in service order
    (the lowest chunk rank, then the oldest) opens the batch if its bytes fit
    beside the `used` in-flight bytes (or nothing is in flight: a job larger
    than the free memory runs alone); other jobs of the same size class join
    in service order while they fit, up to `max_jobs`.  A new request's
    first chunk thus goes ahead of the remaining chunks of a large one.
    Returns (jobs, bytes) or None (the head waits: no overtaking)."""
    if not waiting:
        return None
    ordered = [job for _, job in sorted(((_order(j, i), j) for i, j in enumerate(waiting)), key=lambda t: t[0])]
    head = ordered[0]
    if used + head.nbytes > budget and inflight > 0:
        return None
    batch, nbytes = [head], head.nbytes
    for job in ordered[1:]:
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
                 max_batch_jobs=16, rho=1000.0, g=9.81, max_iter=10, tol=1e-10):
        from .green_functions import require_delhommeau
        require_delhommeau(green_function, "BEMServer")
        # tol: relative residual of the mixed-precision refinement.  Requests
        # reach 12000 panels, where the FP64 residual itself is only good to
        # about 1e-12; 1e-10 keeps each solution within ~1e-9 relative of the
        # FP64 direct solve for these well-conditioned matrices.
        import torch
        self.gf = green_function
        self.dev = green_function.device
        if memory_budget is None:
            memory_budget = int(0.5 * torch.cuda.get_device_properties(self.dev).total_memory)
        self.memory_budget = int(memory_budget)
        # torch keeps the blocks of finished batches cached; batches of other
        # sizes cannot reuse them all, so the cache would grow until the
        # device is full and the runtime could not place a kernel's scratch
        # or the library's workspaces.  Above this many reserved bytes the
        # cache is returned to the driver after a batch.
        total = torch.cuda.get_device_properties(self.dev).total_memory
        self.cache_limit = self.memory_budget + max(0, total - self.memory_budget) // 2
        # jobs of at most an eighth of the budget: a large request is served
        # in frequency chunks that take turns with other requests
        self.max_job_bytes = int(max_job_bytes or self.memory_budget // 8)
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
            for rank, c in enumerate(chunks):
                self._waiting.append(_Job(req, c, per_freq * len(c), rank))
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
                if torch.cuda.memory_reserved(self.dev) > self.cache_limit:
                    torch.cuda.empty_cache()
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
