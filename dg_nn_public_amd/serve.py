"""Serving loop: continuous batching of diffusion-solve requests on one GPU.

Requests (a mesh, a diffusivity and a source) arrive over time.  The dispatcher admits
waiting requests whenever a job slot and enough of the HBM budget are free: the admitted
requests of one dimension are packed into one context (``batch.DGBatch``), so each
operator apply, smoothing step and V-cycle level of the job is a single launch over all
of its solves.  Up to ``max_jobs`` jobs run at once, each on its own stream and host
thread (library calls release the GIL), so requests that arrive while others are on the
GPU start as soon as a slot and memory free up and their work overlaps on the device.  A
job's geometry, multigrid hierarchy and Krylov state stay device-resident until the job
ends; its memory then returns to the budget.  Each request's latency runs from its
arrival to the end of the job that solved it.
"""
from __future__ import annotations

import threading
import time
from dataclasses import dataclass, field

import numpy as np
import torch

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

    @property
    def latency(self):
        return self.done - self.arrival


# device bytes per element of a job (geometry, smoother blocks in FP64 and FP32, the block
# diagonal, traces, multigrid and solver vectors, an FP32 Krylov basis of 31 + 30 vectors);
# refined from measured jobs as the run proceeds
_BYTES_PER_ELEM = {2: 5.0e3, 3: 17.0e3}


@dataclass
class _Job:
    jid: int
    reqs: list
    bytes: int
    thread: threading.Thread | None = None
    error: BaseException | None = None


class Server:
    def __init__(self, hbm_budget=200e9, max_jobs=4, max_batch=64, inner=torch.float32,
                 rtol=1e-10, device="cuda:0", keep_solutions=False):
        self.budget = float(hbm_budget)
        self.max_jobs = max_jobs
        self.max_batch = max_batch
        self.inner = inner
        self.rtol = rtol
        self.device = device
        self.keep = keep_solutions
        self.bpe = dict(_BYTES_PER_ELEM)
        self._lock = threading.Lock()
        self._cv = threading.Condition(self._lock)

    def _estimate(self, req):
        return int(self.bpe[req.mesh.D] * req.mesh.cells.shape[0]) + (64 << 20)

    def _run_job(self, job, t0, results):
        try:
            torch.cuda.set_device(self.device)
            stream = torch.cuda.Stream(device=self.device)
            with torch.cuda.stream(stream):
                # jobs set up and solve concurrently: no stream capture while other
                # threads allocate and copy
                bt = DGBatch([r.mesh for r in job.reqs], [r.kappa for r in job.reqs],
                             device=self.device, graph_vcycle=False)
                b = bt.rhs([r.f for r in job.reqs], [r.u_bc for r in job.reqs])
                x, st = bt.solve(b, rtol=self.rtol, inner=self.inner)
                stream.synchronize()
                done = time.perf_counter() - t0
                info = (bt.space.lib, bt.space.ctx)
                used = _device_bytes(*info)
                parts = bt.split(x) if self.keep else [None] * len(job.reqs)
                out = [Result(r.rid, r.arrival, 0.0, done, s["its"], s["relres"], job.jid,
                              p.clone() if p is not None else None)
                       for r, s, p in zip(job.reqs, st, parts)]
                bt.close()
            with self._cv:
                ne = sum(r.mesh.cells.shape[0] for r in job.reqs)
                D = job.reqs[0].mesh.D
                # keep the per-element estimate at or above what jobs actually held
                self.bpe[D] = max(self.bpe[D], used / ne)
                for o in out:
                    o.admitted = job.admitted
                results.extend(out)
        except BaseException as e:  # surfaced by run()
            job.error = e
        finally:
            with self._cv:
                self._cv.notify_all()

    def run(self, requests):
        """Serve `requests` (arrival times relative to the start); returns the results in
        completion order."""
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting, jobs, results = [], [], []
        in_use = 0
        t0 = time.perf_counter()
        jid = 0
        with self._cv:
            while pending or waiting or jobs:
                now = time.perf_counter() - t0
                while pending and pending[0].arrival <= now:
                    waiting.append(pending.pop(0))
                # retire finished jobs
                for j in [j for j in jobs if not j.thread.is_alive()]:
                    jobs.remove(j)
                    in_use -= j.bytes
                    if j.error is not None:
                        raise RuntimeError(f"job {j.jid} failed") from j.error
                # admit: oldest waiting request's dimension, FIFO, within the budget
                while waiting and len(jobs) < self.max_jobs:
                    D = waiting[0].mesh.D
                    take, need = [], 0
                    for r in waiting:
                        if r.mesh.D != D or len(take) >= self.max_batch:
                            continue
                        e = self._estimate(r)
                        if in_use + need + e > self.budget and take:
                            break
                        if in_use + need + e > self.budget and jobs:
                            break  # wait for memory to free up
                        take.append(r)
                        need += e
                    if not take:
                        break
                    for r in take:
                        waiting.remove(r)
                    job = _Job(jid, take, need)
                    job.admitted = time.perf_counter() - t0
                    jid += 1
                    in_use += need
                    job.thread = threading.Thread(target=self._run_job, args=(job, t0, results),
                                                  daemon=True)
                    jobs.append(job)
                    job.thread.start()
                # sleep until the next arrival or a job end
                timeout = None
                if pending:
                    timeout = max(0.0, pending[0].arrival - (time.perf_counter() - t0))
                if jobs or pending:
                    self._cv.wait(timeout=0.05 if timeout is None else min(timeout, 0.05))
        return results


def _device_bytes(lib, ctx):
    import ctypes as C
    info = (C.c_int64 * 8)()
    lib.dgmi_ctx_info(ctx, info)
    return float(info[7])


def summarize(results, t_end=None):
    lat = np.array([r.latency for r in results])
    t_end = max(r.done for r in results) if t_end is None else t_end
    t_first = min(r.arrival for r in results)
    return {"requests": len(results), "p50_s": float(np.percentile(lat, 50)),
            "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean()),
            "solves_per_s": len(results) / max(t_end - t_first, 1e-9),
            "max_relres": float(max(r.relres for r in results)),
            "jobs": len({r.job for r in results})}
