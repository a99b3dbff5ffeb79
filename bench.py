#!/usr/bin/env python3
"""Benchmark of the MI355X Walk-on-Spheres engine.

Workload (one "step" = one full solve of this batch, inputs resident in HBM):
  benchmark problem 2, plain GPU Walk-on-Spheres at real size -- UV sphere
  stacks=128, slices=256, r = 1 + 0.1 sin(6 theta) sin(6 phi) (65,024
  triangles, normalised), the interior cell centres of a 512 x 512 grid on
  [-1, 1]^2 at z = 0, 256 walks per point, eps = 1e-4, step scale 0.999, at
  most 256 steps, Philox4x32-10 seeded per (point, walk, step), 8-charge
  harmonic Dirichlet data, per-point FP64 mean/variance on the device.
  The BVH is built with leaves of at most 4 triangles.

Prints ONE JSON line.  `value` is walk-steps/s of the whole job.  `roofline`
prices the walker kernel's algorithmic bytes (node lines x 112 B + triangle
records x 40 B + 12 B per walk start + 24 B per work unit, visit counts from a
counting run with the same seeds) against MI355X HBM peak, with the kernel's
duration measured by HIP events on its own stream.  `cpu_baseline` times the
oracle's CPU walker (same walks, same Philox streams) on a sample of the
points.  Extra figures (closest-point queries/s, RMSE against the FP64
solution) sit under "extra".
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "walk-steps/s + time-to-RMSE≤1e-2 (NC+2LMC vs plain WoS); p99 req latency"
HBM_PEAK_GBS = 8000.0
NODE_BYTES = 112   # four child boxes (96 B) + four child refs (16 B)
TRI_BYTES = 40     # three vertices (36 B) + original index (4 B)


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--wpp", type=int, default=256, help="walks per point")
    ap.add_argument("--grid", type=int, default=512)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target length of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="run a 1 GiB device copy (counter calibration) and one walker solve, then exit; "
                         "used under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE")
    return ap.parse_args()


def cpu_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch

    import __graft_entry__ as entry
    eng = entry.load_engine()
    from n2wos_amd import hiptime, problems as P, wos as W

    torch.cuda.set_device(0)
    verts, tris = P.make_mesh("mesh2")
    bvh = eng.Bvh(verts, tris, leaf_size=4)
    pts, _ = P.grid_slice_points(bvh, res=args.grid)
    n = len(pts)
    bc = P.charges_boundary(0)
    params = W.WalkParams(max_steps=256, epsilon=1e-4, step_scale=0.999, seed=2024)
    solver = W.Solver(bvh, bc, params)
    tp = torch.from_numpy(pts).cuda()
    stats = [W.PointStats.zeros(n) for _ in range(args.steps)]
    events = [hiptime.EventPair() for _ in range(args.steps)]

    if args.traffic_probe:
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        solver.solve(tp, args.wpp, seed=params.seed)
        torch.cuda.synchronize()
        print(json.dumps({"traffic_probe": "done", "copy_bytes": 1 << 30, "points": n}))
        return

    for w in range(args.warmup):
        solver.solve(tp, args.wpp, seed=1000 + w)
    torch.cuda.synchronize()

    t0 = time.perf_counter()
    for k in range(args.steps):
        solver.solve(tp, args.wpp, stats=stats[k], seed=params.seed + k, events=events[k].handles)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0

    steps_total = sum(int(s.steps.sum().item()) for s in stats)
    walks_total = n * args.wpp * args.steps
    kernel_ms = [e.elapsed() for e in events]
    value = steps_total / elapsed

    # exact visit counts of the same walks (counting build of the same kernel)
    visits = torch.zeros(2, dtype=torch.int64, device="cuda")
    solver.solve(tp, args.wpp, seed=params.seed, visit_counts=visits)
    torch.cuda.synchronize()
    nodes, tri_tests = (int(x) for x in visits.cpu().tolist())
    units = n * ((args.wpp + 63) // 64)
    bytes_launch = nodes * NODE_BYTES + tri_tests * TRI_BYTES + 12 * n * args.wpp + 24 * units
    steps0 = int(stats[0].steps.sum().item())
    avg_ms = sum(kernel_ms) / len(kernel_ms)
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "wos_walker_mesh2_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    truth = P.charges_solution_f64(bc.charges, pts)
    mean = stats[0].mean.cpu().numpy()
    rmse = float(np.sqrt(np.mean((mean - truth) ** 2)))

    # closest-point queries/s: benchmark problem 1 (65,536 uniform queries)
    v1, t1 = P.make_mesh("mesh1")
    b1 = eng.Bvh(v1, t1)
    q1 = torch.from_numpy(P.uniform_box_queries(65536, seed=0)).cuda()
    for _ in range(2):
        b1.query_tensors(q1)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        b1.query_tensors(q1)
    e1.record()
    torch.cuda.synchronize()
    cpq_ms = e0.elapsed_time(e1) / 20

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "walk-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "problem 2: plain GPU WoS, mesh2 (UV sphere 128x256, r=1+0.1 sin6t sin6p, 65,024 tris), "
                        f"{args.grid}x{args.grid} z=0 grid interior points, {args.wpp} walks/point, eps 1e-4, "
                        "8-charge Laplace, FP64 per-point Welford/Chan",
            "points": n,
            "walks_per_point": args.wpp,
            "triangles": int(len(tris)),
            "walks_per_step": n * args.wpp,
            "mean_steps_per_walk": steps0 / (n * args.wpp),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "k_wos_persistent",
            "kernel_ms_avg": avg_ms,
            "algorithmic_bytes_per_launch": bytes_launch,
            "node_visits_per_launch": nodes,
            "triangle_tests_per_launch": tri_tests,
        },
        "extra": {
            "rmse_vs_fp64": rmse,
            "walks_per_s": walks_total / elapsed,
            "closest_point_queries_per_s_problem1": 65536 / (cpq_ms * 1e-3),
            "closest_point_ms_problem1": cpq_ms,
        },
    }

    if not args.no_cpu:
        import oracle
        oracle.build()
        threads = cpu_threads()
        b2 = oracle.Bvh2(verts, tris)
        # bounded sample: every k-th grid point, sized for about cpu_seconds
        pstride = max(1, n // 64)
        probe_ids = np.arange(0, n, pstride)[:64]
        t0 = time.perf_counter()
        _, st_probe, _, _ = b2.walks(pts[probe_ids], 16, max_steps=256, eps=1e-4, step_scale=0.999, bmode=bc.mode,
                                     bparams=bc.oracle_params(), seed=params.seed, point_ids=probe_ids,
                                     nthreads=threads)
        rate = st_probe.sum() / max(1e-9, time.perf_counter() - t0)
        want_steps = rate * args.cpu_seconds
        n_sample = int(max(16, min(n, want_steps / (25.0 * args.wpp))))
        stride = max(1, n // n_sample)
        ids = np.arange(0, n, stride)[:n_sample]
        sample = pts[ids]
        t0 = time.perf_counter()
        _, st_cpu, _, _ = b2.walks(sample, args.wpp, max_steps=256, eps=1e-4, step_scale=0.999, bmode=bc.mode,
                                   bparams=bc.oracle_params(), seed=params.seed, point_ids=ids, nthreads=threads)
        dt = time.perf_counter() - t0
        line["cpu_baseline"] = {
            "value": float(st_cpu.sum() / dt),
            "unit": "walk-steps/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{len(sample)} of {n} grid points (every {stride}th) x {args.wpp} walks, same walks as the "
                      f"GPU run, {dt:.1f} s on {threads} host threads",
        }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
