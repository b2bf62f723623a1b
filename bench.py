"""Benchmark of the IP-DG diffusion solve on one MI355X.

Default line (``python bench.py``): the serving workload the metric is quoted on.  Poisson
arrivals of diffusion-solve requests (half 2D P3 triangle meshes, half 3D P3 tet meshes,
2e3..2e5 elements log-uniform, plus 5 % large 3D jobs of 55^3 x 6 = 998,250 tets; jittered
structured grids, lognormal diffusivity sigma U[0.5, 2], sine source of amplitude
U[0.5, 2]; seeded) are served by the continuous-batching loop (``dg_nn_public_amd.serve``)
under an HBM budget, every request solved to ||b - A x|| <= 1e-10 ||b||.  A step is one
stream of ``--nreq`` requests arriving at ``--rate`` req/s; ``value`` is the served solves/s
of the mixed-precision solver (FP32 flexible GMRES cycles inside FP64 iterative
refinement) over the timed steps, with p50/p99 latency; the plain FP64 solver serves the
same streams beside it.  Inputs (meshes, diffusivity, sources) are in host memory when a
request arrives, as a served request's are; each job's set-up (context, multigrid
hierarchy, smoother) is part of its latency.

Beside it: the 48^3 x 6 tet mesh (663,552 elements, 13.3M DOFs, sigma 2) with the operator
apply in GDoF/s per kernel family and the roofline of the FP32 smoothing sweep (the
dominant kernel of the mixed solve, HBM-bound) and of the FP32 / FP64 applies; the
24^3 x 6 single-solve line (FP64, FP32-IR, bf16-IR); and the CPU baseline: the FP64 NumPy
oracle timed on this host's cores (a 32 x 32 x 2 triangle solve end to end, and one
24^3 x 6 apply), 5 warm-ups and the median of 5 timed runs each.

Other workloads (``--workload``), each printing its own JSON line:
  solve      the 24^3x6 single-solve line alone (the previous default)
  precision  48^3 x 6 tets, sigma 2: operator error of the FP32 and bf16-input applies
             against FP64, iteration counts of FP64 / FP32 / bf16 solves
  batch      256 independent solves packed per dimension into one context each
  sweep      the serving workload at several arrival rates
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6e12   # MI355X FP64 vector/matrix, FLOP/s
PEAK_FP32 = 157.3e12  # MI355X FP32 matrix, FLOP/s
PEAK_HBM = 8.0e12     # B/s
METRIC = "solves/s + p50/p99 latency at rtol 1e-10; DG apply GDoF/s, mixed vs FP64"
# calibrated HBM bytes per launch of the FP32 single-pass smoothing sweep with the FP16
# smoother copy (k_staged<3,5,true>, the 3D default): FETCH_SIZE x 2 + WRITE_SIZE from two rocprofv3
# --pmc passes (profiles/), by element count
PMC_SWEEP = {663552: 577.1e6}


def flops_per_elem(D, nloc, ngi, sngi, nface):
    """Flops of one operator apply per element in the reference's point-by-point form
    (volume GEMMs + metric, face traces, lifts and pointwise flux).  The vector kernels
    evaluate the face terms through a quadratic interpolant and perform fewer, so TFLOP/s
    derived from this count are rates of the reference's formulation."""
    f_vol = 4 * D * ngi * nloc + ngi * (2 * D * D + 1)
    f_face = nface * sngi * (4 * (1 + D) * nloc + 4 * D * D + 10)
    return f_vol + f_face


def bytes_per_elem(D, nloc, nface, s, read_b=True):
    return 2 * nloc * s + s + (D * D + 1) * s + 4 * nface + (nloc * s if read_b else 0)


def sweep_bytes_per_elem(D, nloc, nface, sngi, single_pass=True, half=False):
    """Algorithmic HBM bytes of one FP32 smoothing sweep x_new = (1-w)x + S(b - A_off x),
    per element.  Single pass (the default kernel): packed symmetric S (FP32, or with
    ``half`` its FP16 copy in 32-bit pairs plus one FP32 scale), x read once (the element's
    own row and, through L2, its neighbours' rows), b read, x_new written, and per face
    c = J^-T n on both sides, wf, pen, code and nbface.  Two-pass form: the same with the
    neighbours' traces read and this element's traces written (value and normal derivative
    at every face point) and c on one side only."""
    ns = nloc * (nloc + 1) // 2
    s_packed = 4 * ((ns + 1) // 2) + 4 if half else 4 * ns
    vectors = 4 * 3 * nloc
    if single_pass:
        return s_packed + vectors + nface * (8 * D + 16)
    traces = 2 * nface * sngi * 8
    geometry = nface * (4 * D + 16)
    return s_packed + vectors + traces + geometry


def sweep_flops_per_elem(D, nloc, nface, sngi):
    """Flops of one such sweep per element as the kernel performs them: per face the
    neighbour's normal derivative at the face's nk quadratic nodes and the flux in node
    form (face mass and mixed face matrices over the nfn face nodes); the faces' lifts
    summed at the np2 P2 nodes and applied once; then the packed smoother product."""
    nfn, nk, np2 = (10, 6, 10) if D == 3 else (4, 3, 6)
    dk = nk * (D * nloc + D)
    nodes = nfn * nfn + 2 * nfn * nk
    lift = nface * nk * D + nloc * np2 * D
    return 2 * (nface * (dk + nodes) + lift) + 2 * nloc * nloc


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="serve",
                    choices=["serve", "solve", "precision", "batch", "sweep", "setup"])
    # serving
    ap.add_argument("--nreq", type=int, default=48, help="requests per stream")
    ap.add_argument("--rate", type=float, default=500.0, help="arrival rate of a stream, req/s")
    ap.add_argument("--rates", default="10,50,200,500", help="sweep: arrival rates, req/s")
    ap.add_argument("--large-n", type=int, default=55, help="hex cells of the large jobs")
    ap.add_argument("--emax", type=float, default=2e5, help="largest element count of the mix")
    ap.add_argument("--jobs", type=int, default=4, help="concurrent job slots")
    ap.add_argument("--budget", type=float, default=200e9, help="HBM budget, bytes")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 serving streams")
    ap.add_argument("--no-c3", action="store_true")
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--cpu-warmup", type=int, default=5)
    # single-solve / apply settings
    ap.add_argument("--n", type=int, default=24, help="cells per direction (solve workload)")
    ap.add_argument("--dim", type=int, default=3, choices=[2, 3])
    ap.add_argument("--sigma", type=float, default=1.0)
    ap.add_argument("--applies", type=int, default=50)
    ap.add_argument("--sweeps", type=int, default=40)
    ap.add_argument("--face-kappa", default="harmonic", choices=["harmonic", "mean"],
                    help="interior-face diffusivity; 'mean' is the reference's arithmetic "
                         "mean, which is not coercive at sigma=1 and stagnates")
    ap.add_argument("--apply-only", action="store_true",
                    help="solve workload: only applies, sweeps and a calibration copy "
                         "(for rocprofv3 --pmc passes)")
    ap.add_argument("--smoother", default="direct", choices=["direct", "jacobian"])
    ap.add_argument("--graph", action="store_true")
    ap.add_argument("--cpu-applies", type=int, default=3)
    ap.add_argument("--nsolves", type=int, default=256, help="batch workload size")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not "
                                   "supported (multi-GPU sharding is design-only)"}))
        sys.exit(2)
    {"serve": run_headline, "solve": run_solve, "precision": run_precision,
     "batch": run_batch, "sweep": run_serve, "setup": run_setup}[args.workload](args)


# ----------------------------------------------------------------------------------------
def build_stream(nreq, large_n, emax, seed=11):
    """The serving mix: ``nreq`` requests, half 2D / half 3D with element counts
    log-uniform in [2e3, emax], 5 % (at least one) large 3D jobs of large_n^3 x 6 tets;
    sigma and source amplitude U[0.5, 2]; shuffled, seeded."""
    import numpy as np

    from dg_nn_public_amd import meshgen
    rng = np.random.default_rng(7)
    specs = _batch_requests(nreq, emax, seed=seed)
    rng.shuffle(specs)
    nlarge = max(1, int(round(0.05 * nreq)))
    for i in range(nlarge):
        specs[i * (len(specs) // nlarge)] = dict(D=3, n=large_n, mesh_seed=5000 + i,
                                                 kappa_seed=6000 + i, sigma=1.0)
    payload = []
    for i, r in enumerate(specs):
        m = meshgen.square_mesh(r["n"], seed=r["mesh_seed"]) if r["D"] == 2 else \
            meshgen.cube_mesh(r["n"], seed=r["mesh_seed"])
        sig = float(rng.uniform(0.5, 2.0))
        amp = float(rng.uniform(0.5, 2.0))
        payload.append((m, meshgen.sine_problem(m, sigma=sig, kappa_seed=r["kappa_seed"],
                                                amplitude=amp)))
        if i % 12 == 11:
            log(f"generated {i + 1}/{len(specs)} request meshes")
    return payload, nlarge


# the FP64 solver's V-cycle (the reference's): sweeps x += S (b - A x) with the whole
# operator, every SFC level swept; run in FP32 refinement it separates the effect of
# precision from that of the FP32 path's cycle (neighbour-form sweeps, exact coarse level)
REFERENCE_CYCLE = {"smooth_offdiag": 0, "coarse_exact": 0}


def serve_stream(payload, rate, inner, args, seed, params=None):
    import numpy as np

    from dg_nn_public_amd.serve import Request, Server
    ar = np.cumsum(np.random.default_rng(seed).exponential(1.0 / rate, size=len(payload)))
    reqs = [Request(i, m, p.kappa, p.f, p.u_bc, float(a))
            for i, ((m, p), a) in enumerate(zip(payload, ar))]
    srv = Server(hbm_budget=args.budget, max_jobs=args.jobs, max_batch=64, inner=inner,
                 params=params)
    t0 = time.perf_counter()
    res = srv.run(reqs)
    return res, time.perf_counter() - t0


def run_headline(args):
    import numpy as np
    import torch

    from dg_nn_public_amd.serve import summarize
    torch.zeros(1, device="cuda:0")
    t0 = time.time()
    payload, nlarge = build_stream(args.nreq, args.large_n, args.emax)
    nele = sum(m.nele for m, _ in payload)
    log(f"{len(payload)} requests, {nele} elements ({nlarge} of {args.large_n}^3x6), meshes in "
        f"{time.time() - t0:.1f}s")
    # warm-up: a short stream (kernel loads, allocator, host thread pools)
    for w in range(args.warmup):
        serve_stream(payload[: max(4, len(payload) // 6)], args.rate, torch.float32, args, 99 + w)
    log("warm-up done")
    out = {}
    variants = [("f32_ir", torch.float32, None)]
    if not args.no_fp64:
        variants += [("f64", torch.float64, None),
                     ("f32_ir_reference_cycle", torch.float32, REFERENCE_CYCLE)]
    for name, inner, params in variants:
        results, walls = [], []
        for k in range(args.steps):
            torch.cuda.synchronize()
            res, wall = serve_stream(payload, args.rate, inner, args, seed=1000 + k,
                                     params=params)
            torch.cuda.synchronize()
            results += res
            walls.append(wall)
            s = summarize(res)
            log(f"{name} stream {k}: {s}")
        lat = np.array([r.latency for r in results if r.error is None])
        good = [r for r in results if r.error is None]
        jobs = {}
        for i, r in enumerate(results):  # one job_info per (stream, job)
            if r.job_info is not None:
                jobs[(i // len(payload), r.job)] = r.job_info
        phases = {}
        for j in jobs.values():
            for k, v in j["phases"].items():
                phases[k] = phases.get(k, 0.0) + v / args.steps
        out[name] = {"solves_per_s": len(good) / sum(walls), "stream_s": walls,
                     "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
                     "max_relres": float(max(r.relres for r in good)),
                     "its_mean": float(np.mean([r.its for r in good])),
                     "not_converged": sum(1 for r in good if not r.converged),
                     "errors": len(results) - len(good),
                     "jobs_per_stream": len(jobs) / args.steps,
                     "job_setup_s_per_stream": sum(j["setup_s"] for j in jobs.values()) / args.steps,
                     "job_solve_s_per_stream": sum(j["solve_s"] for j in jobs.values()) / args.steps,
                     "setup_phases_s_per_stream": phases}
    head = out["f32_ir"]
    line = {
        "metric": METRIC,
        "value": head["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * float(np.mean(head["stream_s"])),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"serving: {len(payload)} diffusion-solve requests per step, "
                               f"Poisson arrivals at {args.rate:g} req/s; half 2D P3 tris / half "
                               f"3D P3 tets, 2e3..{args.emax:.0e} elements log-uniform, "
                               f"{nlarge} job(s) of {args.large_n}^3x6 tets; lognormal kappa "
                               "sigma U[0.5,2], harmonic face kappa, exact block smoother, "
                               "GMRES(30)+SFC-MG, FP32 flexible cycles in FP64 iterative "
                               f"refinement to rtol 1e-10; {args.jobs} job slots, "
                               f"{args.budget / 1e9:g} GB HBM budget",
                   "requests_per_step": len(payload), "elements_per_step": nele,
                   "parallelism": "single GPU"},
        "latency_s": {"p50": head["p50_s"], "p99": head["p99_s"]},
        "serve": out,
        "speedup_vs_fp64": (head["solves_per_s"] / out["f64"]["solves_per_s"]
                            if "f64" in out else None),
    }
    if "f64" in out:
        # the speed-up split: FP32 refinement with the FP64 solver's own cycle against FP64
        # (precision alone), and the FP32 path's cycle against that (algorithm)
        ref = out["f32_ir_reference_cycle"]["solves_per_s"]
        line["speedup_split"] = {"precision": ref / out["f64"]["solves_per_s"],
                                 "cycle": head["solves_per_s"] / ref}
    if not args.no_c3:
        c3 = c3_block(args)
        line["c3_48cube"] = c3
        line["roofline"] = c3.pop("roofline")
    if not args.no_c2:
        line["c2_24cube"] = c2_block(args)
    line["cpu_baseline"] = None if args.no_cpu else cpu_block(args)
    print(json.dumps(line))


def _events():
    import torch
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def c3_block(args):
    """48^3 x 6 tets, sigma 2: apply GDoF/s per kernel family and the rooflines, all timed
    with HIP events on the stream the library launches on."""
    import torch

    from dg_nn_public_amd import meshgen
    from dg_nn_public_amd.diffusion import DGDiffusion
    t0 = time.time()
    mesh = meshgen.cube_mesh(48, jitter=0.15, seed=0)
    pb = meshgen.sine_problem(mesh, sigma=2.0, kappa_seed=1)
    g = DGDiffusion.from_mesh(mesh, pb.kappa, face_kappa="harmonic", blk_solver="direct")
    g._get_RAR_and_sfc_data_Um()
    tuned = g.autotune(reps=5)
    b = g._get_rhs(g.new_vector(), pb.f, pb.u_bc)
    torch.cuda.synchronize()
    log(f"C3 set up in {time.time() - t0:.1f}s, autotune {tuned}")
    stream = torch.cuda.current_stream()
    nele, nloc = g.nele, g.nloc
    ta = {"f64": _time_apply(g, torch.float64, args.applies),
          "f32_vector": _time_apply(g, torch.float32, args.applies, "vector"),
          "f32_mfma": _time_apply(g, torch.float32, args.applies, "mfma"),
          "bf16_mfma": _time_apply(g, torch.float32, args.applies, bf16=True)}
    g.set_f32_kernel(tuned["f32_kernel"])
    b32 = b.float()
    xs = torch.zeros(g.ndof, dtype=torch.float32, device="cuda:0")
    g.smooth_n(xs, b32, 4)
    e0, e1 = _events()
    e0.record(stream)
    g.smooth_n(xs, b32, args.sweeps)
    e1.record(stream)
    torch.cuda.synchronize()
    t_sweep = e0.elapsed_time(e1) / 1e3 / args.sweeps
    log(f"C3 apply us {({k: round(v * 1e6, 1) for k, v in ta.items()})}, sweep {t_sweep * 1e6:.1f} us")
    F = flops_per_elem(3, nloc, g.ngi, g.sngi, g.nface)
    SB = sweep_bytes_per_elem(3, nloc, g.nface, g.sngi, half=g.smoother_half)
    SF = sweep_flops_per_elem(3, nloc, g.nface, g.sngi)
    vec = tuned["f32_kernel"] == "vector"
    t32 = ta["f32_vector"] if vec else ta["f32_mfma"]
    gbps = SB * nele / t_sweep / 1e9
    out = {
        "workload": f"3D P3 tets 48^3x6 = {nele} elements, {nele * nloc} DOFs, lognormal kappa "
                    "sigma 2, harmonic face kappa",
        "apply_us": {k: v * 1e6 for k, v in ta.items()},
        "apply_gdofs": {k: nele * nloc / v / 1e9 for k, v in ta.items()},
        "autotune": tuned,
        "roofline": {"bound": "hbm", "achieved": gbps, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                     "frac": gbps * 1e9 / PEAK_HBM, "traffic": PMC_SWEEP.get(nele),
                     "kernel": "FP32 smoothing sweep x_new = (1-w)x + S(b - A_off x), the "
                               "neighbours' face traces formed from their rows of x, the "
                               "smoother read " + ("from its FP16 copy" if g.smoother_half
                                                   else "in FP32") + " (" +
                               ("k_staged<3,5>, one launch" if vec
                                else "MFMA trace + residual") +
                               "), 48^3x6 tets; the largest share of the mixed solve's kernel "
                               "time",
                     "bytes_per_elem": SB, "flop_per_elem": SF, "elements": nele,
                     "avg_launch_us": t_sweep * 1e6, "launches_timed": args.sweeps,
                     "achieved_tflops": SF * nele / t_sweep / 1e12},
        "roofline_apply_f32": {"bound": "mfma", "achieved": F * nele / t32 / 1e12,
                               "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                               "frac": F * nele / t32 / PEAK_FP32, "flop_per_elem": F,
                               "kernel": "FP32 operator apply (" + ("vector" if vec else "MFMA")
                                         + " kernels, autotuned)",
                               "achieved_GBps_min_traffic":
                                   bytes_per_elem(3, nloc, g.nface, 4) * nele / t32 / 1e9},
        "roofline_apply_f64": {"bound": "mfma", "achieved": F * nele / ta["f64"] / 1e12,
                               "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                               "frac": F * nele / ta["f64"] / PEAK_FP64,
                               "achieved_GBps_min_traffic":
                                   bytes_per_elem(3, nloc, g.nface, 8) * nele / ta["f64"] / 1e9},
    }
    g.close()
    return out


def c2_block(args):
    """24^3 x 6 tets, sigma 1: one solve to 1e-10 with each inner precision."""
    import numpy as np
    import torch

    from dg_nn_public_amd import meshgen
    from dg_nn_public_amd.diffusion import DGDiffusion
    mesh = meshgen.cube_mesh(24, jitter=0.15, seed=0)
    pb = meshgen.sine_problem(mesh, sigma=1.0, kappa_seed=1)
    g = DGDiffusion.from_mesh(mesh, pb.kappa, face_kappa="harmonic", blk_solver="direct")
    g._get_RAR_and_sfc_data_Um()
    g.autotune(reps=5)
    b = g._get_rhs(g.new_vector(), pb.f, pb.u_bc)
    out = {"workload": f"3D P3 tets 24^3x6 = {g.nele} elements, {g.ndof} DOFs, lognormal "
                       "kappa sigma 1, harmonic face kappa, one solve to rtol 1e-10"}
    for name, inner, params in (("f64", torch.float64, None), ("f32_ir", torch.float32, None),
                                ("f32_ir_reference_cycle", torch.float32, REFERENCE_CYCLE),
                                ("bf16_ir", torch.bfloat16, None)):
        for k, v in (params or {}).items():
            g.set_param(k, v)
        lat = []
        for i in range(1 + args.steps):
            x = g.new_vector()
            torch.cuda.synchronize()
            ts = time.perf_counter()
            g._gmres_mg_solver(x, b, rtol=1e-10, inner=inner)
            torch.cuda.synchronize()
            if i:
                lat.append(time.perf_counter() - ts)
        for k in (params or {}):
            g.set_param(k, 1)  # both parameters default to 1
        out[name] = {"solves_per_s": len(lat) / sum(lat), "ms": 1e3 * float(np.median(lat)),
                     **g.last_stats}
    out["speedup_f32_vs_f64"] = out["f32_ir"]["solves_per_s"] / out["f64"]["solves_per_s"]
    out["speedup_split"] = {
        "precision": out["f32_ir_reference_cycle"]["solves_per_s"] / out["f64"]["solves_per_s"],
        "cycle": out["f32_ir"]["solves_per_s"] / out["f32_ir_reference_cycle"]["solves_per_s"]}
    log(f"C2 {out}")
    g.close()
    return out


def _cpu_info():
    import platform
    model = platform.processor()
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model


def cpu_block(args):
    """The FP64 NumPy oracle on this host: a 32 x 32 x 2 P3 triangle solve end to end
    (space, multigrid set-up, right-hand side, GMRES(30) to 1e-10) and one 24^3 x 6 apply;
    BLAS threads = the cores this process may use; warm-ups, then the median of the timed
    runs."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    from dg_nn_public_amd import meshgen
    from oracle import dg, krylov, mg
    cores = len(os.sched_getaffinity(0))
    with threadpool_limits(cores):
        m = meshgen.square_mesh(32, jitter=0.2, seed=0)
        pb = meshgen.sine_problem(m)

        def solve():
            sp = dg.make_space(m, pb.kappa)
            mgd = mg.build_mg(sp, m.verts)
            b = dg.get_rhs(sp, pb.f, pb.u_bc)
            return krylov.gmres_mg_solver(sp, mgd, np.zeros(sp.ndof), b, rtol=1e-10, m=30)

        ts = []
        for i in range(args.cpu_warmup + args.cpu_reps):
            t = time.perf_counter()
            _, its = solve()
            if i >= args.cpu_warmup:
                ts.append(time.perf_counter() - t)
        t_c1 = float(np.median(ts))
        log(f"CPU C1 solve {t_c1:.2f}s ({its} its)")
        m2 = meshgen.cube_mesh(24, jitter=0.15, seed=0)
        pb2 = meshgen.sine_problem(m2, sigma=1.0, kappa_seed=1)
        sp2 = dg.make_space(m2, pb2.kappa, face_kappa="harmonic")
        x = np.random.default_rng(0).standard_normal(sp2.ndof)
        ta = []
        for i in range(args.cpu_warmup + args.cpu_reps):
            t = time.perf_counter()
            dg.apply_A(sp2, x)
            if i >= args.cpu_warmup:
                ta.append(time.perf_counter() - t)
        t_c2 = float(np.median(ta))
        log(f"CPU C2 apply {t_c2:.2f}s")
    return {"value": 1.0 / t_c1, "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": f"FP64 NumPy oracle: 2D P3 32x32x2 triangles (20,480 DOFs) solved end to end "
                      f"to rtol 1e-10 ({its} GMRES(30) its), median of {args.cpu_reps} after "
                      f"{args.cpu_warmup} warm-ups; and one 24^3x6 tet apply (1.66M DOFs)",
            "c1_solve_s": t_c1, "c2_apply_s": t_c2,
            "c2_apply_gdofs": m2.nele * m2.nloc / t_c2 / 1e9,
            "os_cpu_count": os.cpu_count(), "threads": cores, "cpu_model": _cpu_info()}


def run_solve(args):
    """The 24^3x6 (or --n/--dim) single-solve line: FP64, FP32-IR and bf16-IR solves, apply
    GDoF/s per kernel family, the FP32 sweep roofline, the oracle apply baseline."""
    import numpy as np
    import torch

    from dg_nn_public_amd import meshgen
    from dg_nn_public_amd.diffusion import DGDiffusion

    dev = "cuda:0"
    t0 = time.time()
    mesh = meshgen.cube_mesh(args.n, jitter=0.15, seed=0) if args.dim == 3 else \
        meshgen.square_mesh(args.n, jitter=0.2, seed=0)
    pb = meshgen.sine_problem(mesh, sigma=args.sigma, kappa_seed=1)
    g = DGDiffusion.from_mesh(mesh, pb.kappa, device=dev, face_kappa=args.face_kappa,
                              blk_solver=args.smoother)
    g._get_RAR_and_sfc_data_Um()
    if args.graph:
        g.lib.dgmi_set_param(g.ctx, b"graph_vcycle", 1.0)
    tuned = g.autotune(reps=5)  # FP32 kernel family and tile shape for this mesh
    torch.cuda.synchronize()
    setup_s = time.time() - t0
    print(f"[bench] autotune: {tuned}", file=sys.stderr, flush=True)
    print(f"[bench] setup done in {setup_s:.1f}s: nele={g.nele} ndof={g.ndof} "
          f"levels={g.nlevel} colours={g.ncolor}", file=sys.stderr, flush=True)

    b = g._get_rhs(g.new_vector(), pb.f, pb.u_bc)
    stream = torch.cuda.current_stream()

    # ---- operator apply throughput (HIP events on the stream the kernels run on)
    def time_applies(dtype, k, kernel="mfma", bf16=False):
        if dtype == torch.float32:
            g.set_f32_kernel(kernel)
        x = torch.randn(g.ndof, dtype=dtype, device=dev)
        bb = torch.randn(g.ndof, dtype=dtype, device=dev)
        r = torch.empty_like(x)
        for _ in range(3):
            g.apply(x, bb, out=r, bf16=bf16)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(k):
            g.apply(x, bb, out=r, bf16=bf16)
        e1.record(stream)
        torch.cuda.synchronize()
        g.set_f32_kernel(tuned["f32_kernel"])  # back to the autotuner's choice
        return e0.elapsed_time(e1) / 1e3 / k

    t_apply64 = time_applies(torch.float64, args.applies)
    t_apply32v = time_applies(torch.float32, args.applies, kernel="vector")
    t_apply32 = time_applies(torch.float32, args.applies, kernel="mfma")
    t_apply16 = time_applies(torch.float32, args.applies, bf16=True)
    print(f"[bench] apply f64 {t_apply64*1e6:.1f} us, f32 vector {t_apply32v*1e6:.1f} us, "
          f"f32 mfma {t_apply32*1e6:.1f} us, bf16 mfma {t_apply16*1e6:.1f} us",
          file=sys.stderr, flush=True)
    # ---- the dominant kernel of the mixed solve: one FP32 smoothing sweep (autotuned
    # family), timed with HIP events on the stream the kernels run on
    g.set_f32_kernel(tuned["f32_kernel"])
    b32 = b.float()
    xs32 = torch.zeros(g.ndof, dtype=torch.float32, device=dev)
    sweeps = 40
    g.smooth_n(xs32, b32, 4)
    es0, es1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    es0.record(stream)
    g.smooth_n(xs32, b32, sweeps)
    es1.record(stream)
    torch.cuda.synchronize()
    t_sweep = es0.elapsed_time(es1) / 1e3 / sweeps
    print(f"[bench] FP32 smoothing sweep ({tuned['f32_kernel']}): {t_sweep*1e6:.1f} us",
          file=sys.stderr, flush=True)

    if args.apply_only:
        # FP32 preconditioner V-cycles with each kernel family (smoothing kernels for the
        # counter passes)
        for kern in ("vector", "mfma", "mfma_bf16"):
            g.set_f32_kernel(kern)
            for _ in range(3):
                g._um_left_precond(g.new_vector(torch.float32), b32)
        g.set_f32_kernel(tuned["f32_kernel"])
        torch.cuda.synchronize()
        # calibration: elementwise copy of 256 MiB (reads 256 MiB, writes 256 MiB)
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            torch.add(src, 0.0, out=dst)
        torch.cuda.synchronize()
        print(json.dumps({"apply_us": {"f64": t_apply64 * 1e6, "f32_vector": t_apply32v * 1e6,
                                       "f32_mfma": t_apply32 * 1e6, "bf16_mfma": t_apply16 * 1e6},
                          "calibration_copy_bytes": 2 * src.numel() * 8}))
        return

    # ---- solves
    def run_solves(inner, k, w):
        lat, stats = [], []
        for i in range(w + k):
            x = g.new_vector()
            torch.cuda.synchronize()
            ts = time.perf_counter()
            g._gmres_mg_solver(x, b, rtol=1e-10, inner=inner)
            torch.cuda.synchronize()
            te = time.perf_counter()
            if i >= w:
                lat.append(te - ts)
                stats.append(dict(g.last_stats))
        return lat, stats

    torch.cuda.synchronize()
    T0 = time.perf_counter()
    lat64, st64 = run_solves(torch.float64, args.steps, args.warmup)
    T1 = time.perf_counter()
    print(f"[bench] fp64 solves: {[round(v, 3) for v in lat64]} s, stats {st64[-1]}",
          file=sys.stderr, flush=True)
    lat32, st32 = run_solves(torch.float32, args.steps, args.warmup)
    print(f"[bench] mixed solves: {[round(v, 3) for v in lat32]} s, stats {st32[-1]}",
          file=sys.stderr, flush=True)
    lat16, st16 = run_solves(torch.bfloat16, args.steps, args.warmup)
    print(f"[bench] bf16-inner solves: {[round(v, 3) for v in lat16]} s, stats {st16[-1]}",
          file=sys.stderr, flush=True)
    total64 = sum(lat64)

    # the reference's arithmetic-mean face diffusivity on the same problem, 60 iterations
    ref_form = None
    if args.face_kappa != "mean" and args.sigma > 0:
        gm = DGDiffusion.from_mesh(mesh, pb.kappa, device=dev, face_kappa="mean",
                                   blk_solver=args.smoother)
        gm._get_RAR_and_sfc_data_Um()
        bm = gm._get_rhs(gm.new_vector(), pb.f, pb.u_bc)
        gm._gmres_mg_solver(gm.new_vector(), bm, rtol=1e-10, max_its=60)
        ref_form = {"face_kappa": "mean", "iterations": gm.last_stats["its"],
                    "relres": gm.last_stats["relres"]}
        gm.close()
        print(f"[bench] reference mean-kappa form after 60 its: {ref_form}", file=sys.stderr,
              flush=True)

    nele, nloc = g.nele, g.nloc
    F = flops_per_elem(args.dim, nloc, g.ngi, g.sngi, g.nface)
    achieved = F * nele / t_apply64
    gdofs64 = nele * nloc / t_apply64 / 1e9
    gdofs32 = nele * nloc / t_apply32 / 1e9

    # ---- CPU baseline: the oracle's FP64 apply on this host, bounded sample
    cpu = None
    if not args.no_cpu:
        try:
            from threadpoolctl import threadpool_limits
        except Exception:  # pragma: no cover
            threadpool_limits = None
        from oracle import dg as odg
        ctx = threadpool_limits(1) if threadpool_limits else None
        if ctx:
            ctx.__enter__()
        sp = odg.make_space(mesh, pb.kappa, face_kappa=args.face_kappa)
        xs = np.random.default_rng(0).standard_normal(sp.ndof)
        odg.apply_A(sp, xs)  # builds the per-space face cache
        tc = time.perf_counter()
        for _ in range(args.cpu_applies):
            odg.apply_A(sp, xs)
        t_cpu_apply = (time.perf_counter() - tc) / args.cpu_applies
        if ctx:
            ctx.__exit__(None, None, None)
        its = np.mean([s["its"] for s in st64])
        cycles = np.mean([s["cycles"] for s in st64])
        applies_per_solve = 7 * its + 7 * cycles + 1
        cpu = {"value": 1.0 / (applies_per_solve * t_cpu_apply), "unit": "solves/s",
               "cores": 1, "kind": "port",
               "sample": f"{args.cpu_applies} FP64 operator applies of the oracle on this "
                         f"mesh ({t_cpu_apply:.2f} s each, {nele*nloc/t_cpu_apply/1e9:.5f} "
                         f"GDoF/s), scaled by the {applies_per_solve:.0f} applies one GPU "
                         f"solve performs"}

    # headline: the mixed-precision solve (FP32 flexible GMRES cycles inside FP64 iterative
    # refinement, FP64 true residual to 1e-10); the plain FP64 solver is reported beside it
    sps32 = len(lat32) / sum(lat32)
    sps64 = len(lat64) / total64
    vec32 = tuned["f32_kernel"] == "vector"
    t32 = t_apply32v if vec32 else t_apply32
    ach32 = F * nele / t32
    # calibrated HBM bytes per launch of this workload (FETCH_SIZE x 2 + WRITE_SIZE, two
    # rocprofv3 --pmc passes, profiles/pmc_bytes_apply_tet24.md)
    pmc = {"f64": 240e6, "f32_vector": 113e6, "f32_mfma": 125e6, "f32_sweep": PMC_SWEEP.get(82944)}
    on_profiled = args.dim == 3 and args.n == 24
    SB = sweep_bytes_per_elem(args.dim, nloc, g.nface, g.sngi, half=g.smoother_half)
    SF = sweep_flops_per_elem(args.dim, nloc, g.nface, g.sngi)
    sweep_gbps = SB * nele / t_sweep / 1e9
    out = {
        "metric": "solves/s + p50/p99 latency at rtol 1e-10; DG apply GDoF/s, mixed vs FP64",
        "value": sps32,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * sum(lat32) / len(lat32),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": (f"3D IP-DG P3 tets, {args.n}^3x6" if args.dim == 3 else
                                f"2D IP-DG P3 triangles, {args.n}^2x2") + f" = {nele} elements, "
                               f"{nele*nloc} DOFs, lognormal kappa sigma={args.sigma}, "
                               f"{args.face_kappa}-mean face kappa, {args.smoother} block "
                               "smoother, GMRES(30)+SFC-MG, FP32 flexible cycles in FP64 "
                               "iterative refinement, rtol 1e-10",
                   "global_batch": 1, "parallelism": "single GPU"},
        "latency_s": {"p50": float(np.percentile(lat32, 50)), "p99": float(np.percentile(lat32, 99))},
        "iterations": [s["its"] for s in st32], "outer": [s["outer"] for s in st32],
        "relres": max(s["relres"] for s in st32),
        "fp64": {"solves_per_s": sps64,
                 "p50": float(np.percentile(lat64, 50)), "p99": float(np.percentile(lat64, 99)),
                 "iterations": [s["its"] for s in st64], "relres": max(s["relres"] for s in st64)},
        "speedup_vs_fp64": sps32 / sps64,
        "mixed_bf16": {"solves_per_s": len(lat16) / sum(lat16),
                       "p50": float(np.percentile(lat16, 50)), "p99": float(np.percentile(lat16, 99)),
                       "speedup_vs_fp64": (len(lat16) / sum(lat16)) / sps64,
                       "iterations": [s["its"] for s in st16], "outer": [s["outer"] for s in st16],
                       "relres": max(s["relres"] for s in st16)},
        "reference_face_kappa_check": ref_form,
        "autotune": tuned,
        "apply_gdofs": {"f64": gdofs64, "f32_vector": nele * nloc / t_apply32v / 1e9,
                        "f32_mfma": gdofs32, "bf16_mfma": nele * nloc / t_apply16 / 1e9},
        "setup_s": setup_s,
        "roofline": {"bound": "hbm", "achieved": sweep_gbps, "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": sweep_gbps * 1e9 / PEAK_HBM,
                     "traffic": (pmc["f32_sweep"] if on_profiled and vec32 else None),
                     "kernel": "FP32 smoothing sweep x_new = (1-w)x + S(b - A_off x), the "
                               "neighbours' face traces formed from their rows of x (" +
                               ("k_staged<D,5>, one launch" if vec32
                                else "trace + residual MFMA launches") + ", autotuned family); "
                               "the largest share of the mixed solve's kernel time",
                     "bytes_per_elem": SB, "flop_per_elem": SF, "elements": nele,
                     "avg_launch_us": t_sweep * 1e6,
                     "achieved_tflops": SF * nele / t_sweep / 1e12},
        "roofline_apply": {"bound": "mfma", "achieved": ach32 / 1e12, "peak": PEAK_FP32 / 1e12,
                           "unit": "TFLOP/s", "frac": ach32 / PEAK_FP32,
                           "traffic": (pmc["f32_vector" if vec32 else "f32_mfma"] if on_profiled else None),
                           "kernel": "FP32 operator apply (trace + residual, "
                                     + ("vector" if vec32 else "MFMA") + " kernels, autotuned)",
                           "flop_per_elem": F, "elements": nele, "avg_launch_us": t32 * 1e6},
        "roofline_fp64": {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12,
                          "unit": "TFLOP/s", "frac": achieved / PEAK_FP64,
                          "traffic": pmc["f64"] if on_profiled else None,
                          "avg_launch_us": t_apply64 * 1e6,
                          "achieved_GBps_min_traffic": bytes_per_elem(args.dim, nloc, g.nface, 8) * nele / t_apply64 / 1e9},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))




def _time_apply(g, dtype, k, kernel="mfma", bf16=False):
    import torch
    if dtype == torch.float32:
        g.set_f32_kernel(kernel)
    x = torch.randn(g.ndof, dtype=dtype, device="cuda:0")
    bb = torch.randn(g.ndof, dtype=dtype, device="cuda:0")
    r = torch.empty_like(x)
    stream = torch.cuda.current_stream()
    for _ in range(3):
        g.apply(x, bb, out=r, bf16=bf16)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(k):
        g.apply(x, bb, out=r, bf16=bf16)
    e1.record(stream)
    torch.cuda.synchronize()
    g.set_f32_kernel("vector")
    return e0.elapsed_time(e1) / 1e3 / k


def run_precision(args):
    """48^3 x 6 tets, lognormal sigma 2: operator error of the FP32 / bf16-input applies
    against the FP64 apply, and FP64 / FP32-IR / bf16-IR solves to 1e-10."""
    import numpy as np
    import torch

    from dg_nn_public_amd import meshgen
    from dg_nn_public_amd.diffusion import DGDiffusion

    n = 48 if args.n == 24 else args.n
    sigma = 2.0
    t0 = time.time()
    mesh = meshgen.cube_mesh(n, jitter=0.15, seed=0)
    pb = meshgen.sine_problem(mesh, sigma=sigma, kappa_seed=1)
    g = DGDiffusion.from_mesh(mesh, pb.kappa, face_kappa="harmonic", blk_solver=args.smoother)
    g._get_RAR_and_sfc_data_Um()
    b = g._get_rhs(g.new_vector(), pb.f, pb.u_bc)
    torch.cuda.synchronize()
    setup = time.time() - t0
    print(f"[precision] setup {setup:.1f}s nele={g.nele}", file=sys.stderr, flush=True)
    solves = {}
    xh = None
    for name, inner in (("f64", torch.float64), ("f32_ir", torch.float32),
                        ("bf16_ir", torch.bfloat16)):
        x = g.new_vector()
        torch.cuda.synchronize()
        ts = time.perf_counter()
        g._gmres_mg_solver(x, b, rtol=1e-10, inner=inner)
        torch.cuda.synchronize()
        solves[name] = dict(g.last_stats, seconds=time.perf_counter() - ts)
        if name == "f64":
            xh = x
        print(f"[precision] {name}: {solves[name]}", file=sys.stderr, flush=True)

    def op_err(x64):
        ref = g.apply(x64)
        out = {}
        for kname, kern, bf in (("f32_vector", "vector", False), ("f32_mfma", "mfma", False),
                                ("bf16_mfma", "mfma", True)):
            g.set_f32_kernel(kern)
            y = g.apply(x64.float(), bf16=bf).double()
            out[kname] = float(torch.linalg.vector_norm(y - ref) / torch.linalg.vector_norm(ref))
        g.set_f32_kernel("vector")
        return out

    gen = torch.Generator(device="cpu").manual_seed(2)
    xr = (torch.rand(g.ndof, generator=gen, dtype=torch.float64) * 2 - 1).to("cuda:0")
    errs = {"x_uniform": op_err(xr), "x_solution": op_err(xh)}
    ap = {"f64": _time_apply(g, torch.float64, args.applies),
          "f32_vector": _time_apply(g, torch.float32, args.applies, "vector"),
          "f32_mfma": _time_apply(g, torch.float32, args.applies, "mfma"),
          "bf16_mfma": _time_apply(g, torch.float32, args.applies, bf16=True)}
    nd = g.nele * g.nloc
    print(json.dumps({
        "metric": "operator rel. L2 error vs FP64 and iteration counts (precision workload)",
        "config": {"workload": f"3D IP-DG P3 tets, {n}^3x6 = {g.nele} elements, {nd} DOFs, "
                               f"lognormal kappa sigma={sigma}, harmonic face kappa, "
                               f"{args.smoother} block smoother, GMRES(30)+SFC-MG, rtol 1e-10"},
        "operator_rel_err": errs,
        "solves": solves,
        "iteration_change_vs_f64": {k: v["its"] - solves["f64"]["its"] for k, v in solves.items()},
        "apply_us": {k: v * 1e6 for k, v in ap.items()},
        "apply_gdofs": {k: nd / v / 1e9 for k, v in ap.items()},
        "setup_s": setup,
    }))


def _batch_requests(nsolves, emax, seed=0):
    """The batch mix: half 2D, half 3D; element counts log-uniform in [2e3, emax];
    sigma U[0.5, 2]; grid n from the element count."""
    import numpy as np
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(nsolves):
        D = 2 if i < nsolves // 2 else 3
        E = float(np.exp(rng.uniform(np.log(2e3), np.log(emax))))
        n = max(2, int(round((E / 2) ** 0.5))) if D == 2 else max(2, int(round((E / 6) ** (1 / 3))))
        reqs.append(dict(D=D, n=n, mesh_seed=i, kappa_seed=1000 + i,
                         sigma=float(rng.uniform(0.5, 2.0))))
    return reqs


def run_setup(args):
    """Set-up of one large serving job (--large-n^3 x 6 tets, as the serving mix's large
    jobs): DGBatch creation (context, smoother, multigrid) and the right-hand side,
    --steps times after --warmup untimed ones; per-phase host times from the library."""
    import numpy as np
    import torch

    from dg_nn_public_amd import meshgen
    from dg_nn_public_amd.batch import DGBatch
    m = meshgen.cube_mesh(args.large_n, seed=5000)
    pb = meshgen.sine_problem(m, sigma=1.0, kappa_seed=6000)
    torch.zeros(1, device="cuda:0")
    walls, rhs, phases = [], [], []
    for k in range(args.warmup + args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bt = DGBatch([m], [pb.kappa], graph_vcycle=False)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        bt.rhs([pb.f], [pb.u_bc])
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        ph = bt.space.setup_timing()
        bt.close()
        log(f"set-up {k}: DGBatch {t1 - t0:.3f}s rhs {t2 - t1:.3f}s "
            f"{ {q: round(v, 4) for q, v in ph.items()} }")
        if k >= args.warmup:
            walls.append(t1 - t0)
            rhs.append(t2 - t1)
            phases.append(ph)
    print(json.dumps({"workload": f"set-up of one {args.large_n}^3x6 tet job ({m.nele} "
                                  "elements)",
                      "dgbatch_s": float(np.median(walls)), "rhs_s": float(np.median(rhs)),
                      "phases_s": {q: float(np.median([p[q] for p in phases]))
                                   for q in phases[0]}}))


def run_batch(args):
    """256 independent solves, packed per dimension into one context each (one launch per
    kernel covers every solve of that dimension); FP32-IR and FP64 solves."""
    import numpy as np
    import torch

    from dg_nn_public_amd import meshgen
    from dg_nn_public_amd.batch import DGBatch

    reqs = _batch_requests(args.nsolves, args.emax)
    out = {"groups": {}}
    total_solve = {"f32_ir": 0.0, "f64": 0.0}
    nsol = 0
    t_setup = 0.0
    for D in (2, 3):
        rs = [r for r in reqs if r["D"] == D]
        if not rs:
            continue
        t0 = time.time()
        pbs = []
        for r in rs:
            m = meshgen.square_mesh(r["n"], seed=r["mesh_seed"]) if D == 2 else \
                meshgen.cube_mesh(r["n"], seed=r["mesh_seed"])
            pbs.append(meshgen.sine_problem(m, sigma=r["sigma"], kappa_seed=r["kappa_seed"]))
        tm = time.time() - t0
        bt = DGBatch([p.mesh for p in pbs], [p.kappa for p in pbs], blk_solver=args.smoother)
        b = bt.rhs([p.f for p in pbs], [p.u_bc for p in pbs])
        torch.cuda.synchronize()
        ts = time.time() - t0 - tm
        t_setup += ts
        print(f"[batch] D={D}: {len(rs)} solves, {bt.space.nele} elements, mesh gen {tm:.1f}s, "
              f"context+multigrid setup {ts:.1f}s", file=sys.stderr, flush=True)
        grp = {"solves": len(rs), "elements": int(bt.space.nele), "setup_s": ts}
        for name, inner in (("f32_ir", torch.float32), ("f64", torch.float64)):
            lat = []
            for rep in range(args.warmup + args.steps):
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                x, st = bt.solve(b, rtol=1e-10, inner=inner)
                torch.cuda.synchronize()
                if rep >= args.warmup:
                    lat.append(time.perf_counter() - t1)
            its = [s["its"] for s in st]
            grp[name] = {"seconds": float(np.median(lat)), "max_relres": max(s["relres"] for s in st),
                         "its_min": min(its), "its_max": max(its), "its_mean": float(np.mean(its))}
            total_solve[name] += float(np.median(lat))
            print(f"[batch] D={D} {name}: {grp[name]}", file=sys.stderr, flush=True)
        nsol += len(rs)
        out["groups"][D] = grp
        bt.close()
    sps = {k: nsol / v for k, v in total_solve.items()}
    print(json.dumps({
        "metric": "solves/s at rtol 1e-10, packed batch (batch workload)",
        "value": sps["f32_ir"], "unit": "solves/s", "higher_is_better": True,
        "config": {"workload": f"{nsol} independent solves, half 2D P3 tris / half 3D P3 tets, "
                               f"2e3..{args.emax:.0e} elements log-uniform, sigma U[0.5,2], "
                               "packed per dimension, FP32 inner + FP64 refinement"},
        "fp64_solves_per_s": sps["f64"], "speedup_vs_fp64": sps["f32_ir"] / sps["f64"],
        "setup_s_total": t_setup, "groups": out["groups"],
    }))


def run_serve(args):
    """Poisson request streams at several rates, served by the continuous-batching loop;
    FP32-IR and FP64 inner solves on the same streams."""
    import numpy as np
    import torch

    from dg_nn_public_amd import meshgen
    from dg_nn_public_amd.serve import Request, Server, summarize

    rng = np.random.default_rng(7)
    specs = _batch_requests(args.nreq, args.emax, seed=11)
    rng.shuffle(specs)
    nlarge = max(1, int(round(0.05 * args.nreq)))
    for i in range(nlarge):  # 5 % large 3D jobs
        specs[i * (len(specs) // nlarge)] = dict(D=3, n=args.large_n, mesh_seed=5000 + i,
                                                 kappa_seed=6000 + i, sigma=1.0)
    t0 = time.time()
    payload = []
    for i, r in enumerate(specs):
        m = meshgen.square_mesh(r["n"], seed=r["mesh_seed"]) if r["D"] == 2 else \
            meshgen.cube_mesh(r["n"], seed=r["mesh_seed"])
        sig = float(rng.uniform(0.5, 2.0))
        amp = float(rng.uniform(0.5, 2.0))
        p = meshgen.sine_problem(m, sigma=sig, kappa_seed=r["kappa_seed"], amplitude=amp)
        payload.append((m, p))
        if i % 8 == 7:
            print(f"[serve] generated {i + 1}/{len(specs)} request meshes", file=sys.stderr, flush=True)
    print(f"[serve] {len(payload)} requests, {sum(m.nele for m, _ in payload)} elements, "
          f"meshes in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    rates = [float(v) for v in args.rates.split(",")]
    out = {}
    for lam in rates:
        ar = np.cumsum(np.random.default_rng(int(lam)).exponential(1.0 / lam, size=len(payload)))
        reqs = [Request(i, m, p.kappa, p.f, p.u_bc, float(a))
                for i, ((m, p), a) in enumerate(zip(payload, ar))]
        row = {}
        for name, inner in (("f32_ir", torch.float32), ("f64", torch.float64)):
            res = Server(hbm_budget=200e9, max_jobs=4, max_batch=64, inner=inner).run(reqs)
            row[name] = summarize(res)
            print(f"[serve] rate {lam:g}/s {name}: {row[name]}", file=sys.stderr, flush=True)
        row["speedup_vs_fp64"] = row["f32_ir"]["solves_per_s"] / row["f64"]["solves_per_s"]
        out[f"{lam:g}"] = row
    best = max(out.values(), key=lambda r: r["f32_ir"]["solves_per_s"])
    print(json.dumps({
        "metric": "solves/s + p50/p99 latency at rtol 1e-10 (serving workload)",
        "value": best["f32_ir"]["solves_per_s"], "unit": "solves/s", "higher_is_better": True,
        "config": {"workload": f"Poisson arrivals, {len(payload)} requests per rate: half 2D "
                               f"P3 tris / half 3D P3 tets 2e3..{args.emax:.0e} elements, "
                               f"{nlarge} jobs of {args.large_n}^3x6 tets, sigma and source "
                               "amplitude U[0.5,2], 200 GB HBM budget, 4 job slots"},
        "rates": out,
    }))


if __name__ == "__main__":
    main()
