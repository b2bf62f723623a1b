"""Benchmark of the IP-DG diffusion solve on one MI355X.

Workload (3D, the configuration the metric is quoted on): P3 tetrahedra in the unit
cube from a 24^3 hex grid, 6 Kuhn tets per hex (82,944 elements, 1,658,880 DG DOFs),
interior vertices jittered +-0.15h (seed 0), lognormal diffusivity per element
(mu 0, sigma 1, seed 1), Dirichlet boundary, manufactured sine source.  A step is one
full solve ||b - A x|| <= 1e-10 ||b|| from x = 0 with GMRES(30) + SFC multigrid, inputs
resident in HBM.

Prints one JSON line: solves/s (FP64 solver) as `value`, latency p50/p99, the
mixed-precision solver's solves/s, operator-apply GDoF/s in FP64 and FP32, the roofline
of the apply, and the CPU oracle baseline.
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


def flops_per_elem(D, nloc, ngi, sngi, nface):
    """Algorithmic flops of one operator apply per element (volume GEMMs + metric,
    face traces, lifts and pointwise flux)."""
    f_vol = 4 * D * ngi * nloc + ngi * (2 * D * D + 1)
    f_face = nface * sngi * (4 * (1 + D) * nloc + 4 * D * D + 10)
    return f_vol + f_face


def bytes_per_elem(D, nloc, nface, s, read_b=True):
    return 2 * nloc * s + s + (D * D + 1) * s + 4 * nface + (nloc * s if read_b else 0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=24, help="hex cells per direction")
    ap.add_argument("--sigma", type=float, default=1.0)
    ap.add_argument("--applies", type=int, default=50)
    ap.add_argument("--face-kappa", default="harmonic", choices=["harmonic", "mean"],
                    help="interior-face diffusivity; 'mean' is the reference's arithmetic "
                         "mean, which is not coercive at sigma=1 and stagnates")
    ap.add_argument("--apply-only", action="store_true",
                    help="only operator applies (FP64, FP32) and a 256 MiB elementwise copy "
                         "that calibrates HBM byte counters; for rocprofv3 --pmc passes")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-applies", type=int, default=3)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not "
                                   "supported (multi-GPU sharding is design-only)"}))
        sys.exit(2)

    import numpy as np
    import torch

    from dg_nn_public_amd import meshgen
    from dg_nn_public_amd.diffusion import DGDiffusion

    dev = "cuda:0"
    t0 = time.time()
    mesh = meshgen.cube_mesh(args.n, jitter=0.15, seed=0)
    pb = meshgen.sine_problem(mesh, sigma=args.sigma, kappa_seed=1)
    g = DGDiffusion.from_mesh(mesh, pb.kappa, device=dev, face_kappa=args.face_kappa)
    g._get_RAR_and_sfc_data_Um()
    torch.cuda.synchronize()
    setup_s = time.time() - t0
    print(f"[bench] setup done in {setup_s:.1f}s: nele={g.nele} ndof={g.ndof} "
          f"levels={g.nlevel} colours={g.ncolor}", file=sys.stderr, flush=True)

    b = g._get_rhs(g.new_vector(), pb.f, pb.u_bc)
    stream = torch.cuda.current_stream()

    # ---- operator apply throughput (HIP events on the stream the kernels run on)
    def time_applies(dtype, k):
        x = torch.randn(g.ndof, dtype=dtype, device=dev)
        bb = torch.randn(g.ndof, dtype=dtype, device=dev)
        r = torch.empty_like(x)
        for _ in range(3):
            g.apply(x, bb, out=r)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(k):
            g.apply(x, bb, out=r)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3 / k

    t_apply64 = time_applies(torch.float64, args.applies)
    t_apply32 = time_applies(torch.float32, args.applies)
    print(f"[bench] apply f64 {t_apply64*1e6:.1f} us, f32 {t_apply32*1e6:.1f} us",
          file=sys.stderr, flush=True)
    if args.apply_only:
        # calibration: elementwise copy of 256 MiB (reads 256 MiB, writes 256 MiB)
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            torch.add(src, 0.0, out=dst)
        torch.cuda.synchronize()
        print(json.dumps({"apply_us": {"f64": t_apply64 * 1e6, "f32": t_apply32 * 1e6},
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
    total64 = sum(lat64)

    # the reference's arithmetic-mean face diffusivity on the same problem, 60 iterations
    ref_form = None
    if args.face_kappa != "mean" and args.sigma > 0:
        gm = DGDiffusion.from_mesh(mesh, pb.kappa, device=dev, face_kappa="mean")
        gm._get_RAR_and_sfc_data_Um()
        bm = gm._get_rhs(gm.new_vector(), pb.f, pb.u_bc)
        gm._gmres_mg_solver(gm.new_vector(), bm, rtol=1e-10, max_its=60)
        ref_form = {"face_kappa": "mean", "iterations": gm.last_stats["its"],
                    "relres": gm.last_stats["relres"]}
        gm.close()
        print(f"[bench] reference mean-kappa form after 60 its: {ref_form}", file=sys.stderr,
              flush=True)

    nele, nloc = g.nele, g.nloc
    F = flops_per_elem(3, nloc, g.ngi, g.sngi, g.nface)
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

    out = {
        "metric": "solves/s + p50/p99 latency at rtol 1e-10; DG apply GDoF/s, mixed vs FP64",
        "value": len(lat64) / total64,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total64 / len(lat64),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3D IP-DG P3 tets, {args.n}^3x6 = {nele} elements, "
                               f"{nele*nloc} DOFs, lognormal kappa sigma={args.sigma}, "
                               f"{args.face_kappa}-mean face kappa, GMRES(30)+SFC-MG, FP64, "
                               f"rtol 1e-10",
                   "global_batch": 1, "parallelism": "single GPU"},
        "latency_s": {"p50": float(np.percentile(lat64, 50)), "p99": float(np.percentile(lat64, 99))},
        "mixed": {"solves_per_s": len(lat32) / sum(lat32),
                  "p50": float(np.percentile(lat32, 50)), "p99": float(np.percentile(lat32, 99)),
                  "speedup_vs_fp64": (len(lat32) / sum(lat32)) / (len(lat64) / total64),
                  "iterations": [s["its"] for s in st32], "outer": [s["outer"] for s in st32],
                  "relres": max(s["relres"] for s in st32)},
        "reference_face_kappa_check": ref_form,
        "fp64_iterations": [s["its"] for s in st64],
        "fp64_relres": max(s["relres"] for s in st64),
        "apply_gdofs": {"f64": gdofs64, "f32": gdofs32},
        "setup_s": setup_s,
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64, "traffic": None,
                     "kernel": "operator apply (k_trace + k_residual), FP64",
                     "flop_per_elem": F, "elements": nele, "avg_launch_us": t_apply64 * 1e6,
                     "achieved_GBps_min_traffic": bytes_per_elem(3, nloc, g.nface, 8) * nele / t_apply64 / 1e9},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
