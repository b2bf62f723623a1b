"""Serving loop: continuous batching of diffusion-solve requests on one GPU.

Requests (a mesh, a diffusivity and a source) arrive over time.  The dispatcher admits
waiting requests whenever a job slot and enough of the HBM budget are free: the admitted
requests of one dimension are packed into one context (``batch.DGBatch``), so each
operator apply, smoothing step and V-cycle level of the job is a single launch over all
of its solves.  Up to ``max_jobs`` jobs run at once, each on its own stream and host
thread (library calls release the GIL), so requests that arrive while others are on the
GPU start as soon as a slot and memory free up and their work overlaps on the device.  A
request of at least ``solo_elements`` elements runs as a job of its own, so that small
requests are never held back by a large one packed with them, and is admitted ahead of
the small ones (the longest job first: it then overlaps them); packing stops at
``job_elements`` elements, so that packed jobs stay short.  A job's geometry,
multigrid hierarchy and Krylov state stay device-resident until the job ends; its slot
frees when its results are published, its memory returns to the budget once the library
has taken it back (into its block cache: no device-wide synchronisation).  Each request's
latency runs from its arrival to the end of the job that solved it.

Failures stay with the request that caused them:
* a request with invalid data (non-finite or non-positive diffusivity, non-finite source
  or boundary data, a malformed mesh) is answered at once with an error and never packed;
* a request whose estimated memory exceeds what the device holds is rejected;
* inside a job, a component whose data turn non-finite is stopped by the solver while the
  other components finish (packed contexts isolate components);
* a job whose device allocation fails is split in two and both halves are queued again; a
  single request that still cannot be allocated is answered with an error;
* any other error of a job is reported on each of its requests; the loop keeps serving.
A solve that stops at ``max_its`` above the tolerance is returned with
``converged = False``.
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass

import numpy as np
import torch

from ._lib import DgmiError
from .batch import DGBatch


@dataclass
class Request:
    rid: int
    mesh: object          # meshgen.Mesh-like (D, x_ref_in, cells, verts)
    kappa: np.ndarray
    f: np.ndarray
    u_bc: np.ndarray
    arrival: float        # seconds after the start of the run


@dataclass
class Result:
    rid: int
    arrival: float
    admitted: float
    done: float
    its: int
    relres: float
    job: int
    x: torch.Tensor | None = None
    converged: bool = True
    error: str | None = None
    # the job's wall seconds of set-up (context, multigrid, right-hand side) and solve, its
    # element count and the library's set-up phase timings (shared by the job's results)
    job_info: dict | None = None

    @property
    def latency(self):
        return self.done - self.arrival

    @property
    def ok(self):
        return self.error is None and self.converged


# device bytes per element of a job (geometry, smoother blocks in FP64 and FP32, the block
# diagonal, traces, multigrid and solver vectors, an FP32 Krylov basis of 31 + 30 vectors);
# refined from measured jobs as the run proceeds
_BYTES_PER_ELEM = {2: 5.0e3, 3: 17.0e3}


def validate(req) -> str | None:
    """Why a request cannot be solved, or None."""
    m = req.mesh
    try:
        D = int(m.D)
        nele = m.cells.shape[0]
        nloc = 10 if D == 2 else 20
        if D not in (2, 3) or m.cells.shape != (nele, D + 1) or nele < 1:
            return "malformed mesh"
        if m.x_ref_in.shape != (nele, D, nloc) or m.verts.ndim != 2 or m.verts.shape[1] != D:
            return "malformed mesh"
        if m.cells.min() < 0 or m.cells.max() >= m.verts.shape[0]:
            return "cell vertex id out of range"
        k = np.asarray(req.kappa, dtype=np.float64).reshape(-1)
        if k.size != nele or not np.all(np.isfinite(k)) or not np.all(k > 0):
            return "diffusivity must be finite and positive on every element"
        for name, a in (("source", req.f), ("boundary data", req.u_bc)):
            a = np.asarray(a, dtype=np.float64).reshape(-1)
            if a.size != nele * nloc:
                return f"{name} must have nele*nloc values"
            if not np.all(np.isfinite(a)):
                return f"non-finite {name}"
    except (AttributeError, TypeError, ValueError) as e:
        return f"malformed request: {e}"
    return None


@dataclass
class _Job:
    jid: int
    reqs: list
    bytes: int
    admitted: float = 0.0
    thread: threading.Thread | None = None
    alloc_failed: bool = False
    published: bool = False  # results handed out: the job no longer holds a slot
    stream: torch.cuda.Stream | None = None


class Server:
    """``hbm_budget`` bounds the estimated bytes of the running jobs (a request larger than
    the budget still runs when nothing else does); ``hbm_capacity`` (default: 90% of the
    device's memory) rejects requests that could not run even alone."""

    def __init__(self, hbm_budget=200e9, max_jobs=4, max_batch=64, inner=torch.float32,
                 rtol=1e-10, max_its=400, device="cuda:0", keep_solutions=False,
                 hbm_capacity=None, job_byte_limit=0, solo_elements=250_000,
                 job_elements=400_000, params=None):
        self.budget = float(hbm_budget)
        self.solo_elements = int(solo_elements)
        self.job_elements = int(job_elements)  # packing stops once a job holds this many
        self.capacity = hbm_capacity
        self.max_jobs = max_jobs
        self.max_batch = max_batch
        self.inner = inner
        self.rtol = rtol
        self.max_its = max_its
        self.device = device
        self.keep = keep_solutions
        self.job_byte_limit = int(job_byte_limit)  # device bytes one job may hold (0: any)
        self.params = dict(params or {})  # solver parameters of every job's context
        self.bpe = dict(_BYTES_PER_ELEM)
        self._lock = threading.Lock()
        self._cv = threading.Condition(self._lock)
        # one stream per job slot, created together: HIP spreads streams over its hardware
        # queues in creation order, so streams created one after another get different
        # queues, and work on streams that share a queue runs in queue order (a job's
        # set-up would wait behind another job's queued solve)
        self._streams = []

    def _estimate(self, req):
        return int(self.bpe[req.mesh.D] * req.mesh.cells.shape[0]) + (64 << 20)

    def _run_job(self, job, t0, results):
        out = []
        used, ne = 0.0, 0
        try:
            torch.cuda.set_device(self.device)
            stream = job.stream
            with torch.cuda.stream(stream):
                bt = None
                try:
                    # jobs set up and solve concurrently: no stream capture while other
                    # threads allocate and copy
                    ta = time.perf_counter()
                    bt = DGBatch([r.mesh for r in job.reqs], [r.kappa for r in job.reqs],
                                 device=self.device, graph_vcycle=False,
                                 alloc_limit=self.job_byte_limit, params=self.params)
                    b = bt.rhs([r.f for r in job.reqs], [r.u_bc for r in job.reqs])
                    stream.synchronize()
                    tb = time.perf_counter()
                    x, st = bt.solve(b, rtol=self.rtol, inner=self.inner, max_its=self.max_its)
                    stream.synchronize()
                    done = time.perf_counter() - t0
                    info = {"setup_s": tb - ta, "solve_s": time.perf_counter() - tb,
                            "start": ta - t0, "setup_end": tb - t0,
                            "elements": int(bt.space.nele), "solves": len(job.reqs),
                            "phases": bt.space.setup_timing()}
                    used = float(bt.space.device_bytes())
                    ne = sum(r.mesh.cells.shape[0] for r in job.reqs)
                    parts = bt.split(x) if self.keep else [None] * len(job.reqs)
                    for r, s, p in zip(job.reqs, st, parts):
                        rel = float(s["relres"])
                        err = None if np.isfinite(rel) else "non-finite data in the solve"
                        out.append(Result(r.rid, r.arrival, job.admitted, done, s["its"], rel,
                                          job.jid, p.clone() if p is not None and err is None
                                          else None, converged=err is None and rel <= self.rtol,
                                          error=err, job_info=info))
                    stream.synchronize()
                    with self._cv:  # publish now: the slot frees before the memory goes back
                        results.extend(out)
                        job.published = True
                        # the next job may queue its work behind this job's last host steps
                        self._streams.append(stream)
                        self._cv.notify_all()
                finally:
                    if bt is not None:
                        try:
                            stream.synchronize()
                            synced = True
                        except Exception:  # noqa: BLE001 - then the library synchronises
                            synced = False
                        bt.close(synced=synced)
                        if out:  # when the job's device memory went back
                            out[0].job_info["released"] = time.perf_counter() - t0
        except DgmiError as e:
            if "allocation failed" in str(e):
                job.alloc_failed = True
                out = []
            else:
                out = self._failed(job, t0, f"job failed: {e}")
        except Exception as e:  # noqa: BLE001 - reported on the job's requests
            out = self._failed(job, t0, f"job failed: {type(e).__name__}: {e}")
        finally:
            with self._cv:
                if ne:
                    D = job.reqs[0].mesh.D
                    # keep the per-element estimate at or above what jobs actually held
                    self.bpe[D] = max(self.bpe[D], used / ne)
                if job.alloc_failed and len(job.reqs) == 1:
                    job.alloc_failed = False
                    out = self._failed(job, t0, "device allocation failed for this request")
                if not job.published:
                    results.extend(out)
                    job.published = not job.alloc_failed
                    self._streams.append(job.stream)
                self._cv.notify_all()

    def _failed(self, job, t0, msg):
        now = time.perf_counter() - t0
        return [Result(r.rid, r.arrival, job.admitted, now, 0, float("nan"), job.jid,
                       converged=False, error=msg) for r in job.reqs]

    def run(self, requests):
        """Serve ``requests`` (arrival times relative to the start); returns one result per
        request, in completion order."""
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting, jobs, results = [], [], []
        in_use = 0
        cap = self.capacity
        if cap is None:  # what one job could hold at most (the library caches freed blocks)
            cap = 0.9 * float(torch.cuda.mem_get_info(torch.device(self.device))[1])
        if not self._streams:
            self._streams = [torch.cuda.Stream(device=self.device) for _ in range(self.max_jobs)]
        t0 = time.perf_counter()
        jid = 0
        with self._cv:
            while pending or waiting or jobs:
                now = time.perf_counter() - t0
                while pending and pending[0].arrival <= now:
                    r = pending.pop(0)
                    why = validate(r)
                    if why is None and self._estimate(r) > cap:
                        why = (f"estimated {self._estimate(r) / 1e9:.1f} GB exceeds the device's "
                               f"{cap / 1e9:.1f} GB")
                    if why is not None:
                        results.append(Result(r.rid, r.arrival, now, now, 0, float("nan"), -1,
                                              converged=False, error=why))
                    else:
                        waiting.append(r)
                # retire finished jobs; a job that could not allocate goes back in two halves
                for j in [j for j in jobs if not j.thread.is_alive()]:
                    jobs.remove(j)
                    in_use -= j.bytes
                    if j.alloc_failed:
                        h = len(j.reqs) // 2
                        for part in (j.reqs[:h], j.reqs[h:]):
                            self._admit(part, jid, t0, jobs, results)
                            jid += 1
                            in_use += jobs[-1].bytes
                # admit: oldest waiting request's dimension, FIFO, within the budget; a
                # large request runs alone.  Jobs that published their results hold memory
                # until the library has it back, but no slot.
                while waiting and sum(not j.published for j in jobs) < self.max_jobs:
                    # the oldest large request first: it is the longest job, started early
                    # it overlaps the small ones
                    solo_first = [r for r in waiting
                                  if r.mesh.cells.shape[0] >= self.solo_elements][:1]
                    D = (solo_first or waiting)[0].mesh.D
                    take, need, nel = [], 0, 0
                    for r in solo_first or waiting:
                        if r.mesh.D != D or len(take) >= self.max_batch:
                            continue
                        if take and nel + r.mesh.cells.shape[0] > self.job_elements:
                            continue
                        solo = r.mesh.cells.shape[0] >= self.solo_elements
                        if solo and take:
                            continue
                        e = self._estimate(r)
                        if in_use + need + e > self.budget and take:
                            break
                        if in_use + need + e > self.budget and jobs:
                            break  # wait for memory to free up
                        take.append(r)
                        need += e
                        nel += r.mesh.cells.shape[0]
                        if solo:
                            break
                    if not take:
                        break
                    for r in take:
                        waiting.remove(r)
                    self._admit(take, jid, t0, jobs, results)
                    jid += 1
                    in_use += need
                # sleep until the next arrival or a job end
                timeout = None
                if pending:
                    timeout = max(0.0, pending[0].arrival - (time.perf_counter() - t0))
                if jobs or pending:
                    self._cv.wait(timeout=0.05 if timeout is None else min(timeout, 0.05))
        return results

    def _admit(self, reqs, jid, t0, jobs, results):
        job = _Job(jid, list(reqs), sum(self._estimate(r) for r in reqs))
        # a free slot's stream (a job re-queued after an allocation failure may find none
        # free: then a stream of its own)
        job.stream = (self._streams.pop(0) if self._streams
                      else torch.cuda.Stream(device=self.device))
        job.admitted = time.perf_counter() - t0
        job.thread = threading.Thread(target=self._run_job, args=(job, t0, results), daemon=True)
        jobs.append(job)
        job.thread.start()


def summarize(results, t_end=None):
    good = [r for r in results if r.error is None]
    lat = np.array([r.latency for r in good]) if good else np.array([np.nan])
    t_end = max(r.done for r in results) if t_end is None else t_end
    t_first = min(r.arrival for r in results)
    return {"requests": len(results), "p50_s": float(np.percentile(lat, 50)),
            "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean()),
            "solves_per_s": len(good) / max(t_end - t_first, 1e-9),
            "max_relres": float(max(r.relres for r in good)) if good else float("nan"),
            "not_converged": sum(1 for r in good if not r.converged),
            "errors": sum(1 for r in results if r.error is not None),
            "jobs": len({r.job for r in results if r.job >= 0})}
