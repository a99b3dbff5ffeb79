#!/usr/bin/env python3
"""Benchmark of the MI355X Walk-on-Spheres engine.

Workload (one "step" = one full solve of this batch, inputs resident in HBM):
  benchmark problem 2, plain GPU Walk-on-Spheres at real size -- UV sphere
  stacks=128, slices=256, r = 1 + 0.1 sin(6 theta) sin(6 phi) (65,024
  triangles, normalised), the interior cell centres of a 512 x 512 grid on
  [-1, 1]^2 at z = 0, 256 walks per point, eps = 1e-4, step scale 0.999, at
  most 256 steps, Philox4x32-10 seeded per (point, walk, step), 8-charge
  harmonic Dirichlet data, per-point FP64 mean/variance on the device.
  The BVH is built by the binned surface-area heuristic with leaves of at
  most 4 triangles (the reference's default builder is cuBQL SAH).

Prints ONE JSON line.  `value` is walk-steps/s of the whole job.  `roofline`
prices the walker kernel's algorithmic bytes (node lines x 112 B + triangle
records x 40 B + 12 B per walk start + 24 B per work unit, visit counts from a
counting run with the same seeds) against MI355X HBM peak, with the kernel's
duration measured by HIP events on its own stream.  The roof that binds the
walker is VALU issue: `roofline.valu` gives its counters (VALU instructions
per walk step, busy and lane-utilisation fractions) and `roofline.traffic` the
HBM bytes per launch, both read from profiles/ only when the profile was
taken of this build's walker sources (`source_sha256`), else null.  `cpu_baseline` times the
oracle's CPU walker (same walks, same Philox streams) on a sample of the
points.  Extra figures sit under "extra":
  closest_point_problem1   65,536 uniform queries on the 64 x 128 sphere
  closest_point_problem1_cpu  the reference's CPU brute force (compiled from
                           its sources) and the oracle's brute force / binary
                           BVH on 1 and on all host threads, bounded samples
  closest_point_stress     262,144 queries each from uniform box, near-
                           boundary shell (|d| < 0.02) and WoS prefixes (1-20
                           steps) on the 128 x 256 sphere: queries/s and HBM
                           roofline fraction of the query kernel
  wos_problem1             4,096 interior points x 128 walks, eps 1e-3, on the
                           64 x 128 sphere; the oracle's CPU walker on the
                           same walks in full beside it
  poisson_problem3         262,144 interior points x 512 walks on the
                           512 x 1024 multi-frequency sphere (1,046,528 tris),
                           Green's-function source sampling, RMSE vs FP64 u*
  time_to_rmse             plain WoS vs neural cache (NC) and NC + 2LMC to
                           RMSE <= 1e-2 on the headline problem, cache
                           training included in the NC times
  progressive_2lmc         NC + 2LMC in blocks while a second cache trains on
                           another stream (snapshot overlap) vs the same
                           schedule on one stream: wall times, RMSE, identity
  cache                    hash-grid + MLP training step and inference rates
  serving                  request server: three resident meshes, seeded
                           Poisson arrivals of log-uniform request sizes,
                           p50 / p99 latency and throughput
--quick skips the heavy sections (problem 3, time-to-RMSE, progressive,
serving); --section runs one section's GPU work alone (counter profiles).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "walk-steps/s + time-to-RMSE≤1e-2 (NC+2LMC vs plain WoS); p99 req latency"
HBM_PEAK_GBS = 8000.0
NODE_BYTES = 112   # four child boxes (96 B) + four child refs (16 B)
TRI_BYTES = 40     # three vertices (36 B) + original index (4 B)
# tree of the walker workloads: SAH, 4 triangles per leaf (measured fastest
# of leaf sizes 2-8 for the walker, whose triangle tests are spread over the
# wave, on mesh 2; results are identical for any tree, since every query is
# exact)
WALKER_TREE = dict(leaf_size=4, build_method="sah")


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--wpp", type=int, default=256, help="walks per point")
    ap.add_argument("--grid", type=int, default=512)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target length of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="skip problem 3, time-to-RMSE and serving")
    ap.add_argument("--serve-requests", type=int, default=240)
    ap.add_argument("--serve-rate", type=float, default=32.0, help="request arrivals per second")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="run a 1 GiB device copy (counter calibration) and one walker solve, then exit; "
                         "used under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--source-hash", action="store_true", help="print the walker sources' hash and exit")
    ap.add_argument("--section", choices=["walker", "cache", "closest_stress", "poisson", "cv"], default=None,
                    help="run one section's GPU work only (one headline solve; the cache's training steps and "
                         "inference; the closest-point stress sets; problem 3; the control variate's training "
                         "pipeline and its 2-walk estimate, three times), print a short JSON line and exit: the "
                         "workload of the profiles")
    return ap.parse_args()


WALKER_SOURCES = ("walker_kernel.hpp", "traverse.hpp", "geom.hpp", "cvfield.hpp", "wave.hpp", "bvh.hpp", "wos.hpp",
                  "wos.hip", "bvh_build.cpp")


def walker_source_hash() -> str:
    """SHA-256 of the walker's sources and build flags: a counter profile
    applies to the build whose hash it records."""
    import hashlib
    h = hashlib.sha256()
    for f in WALKER_SOURCES:
        with open(os.path.join(ROOT, "n2wos-cuda_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    with open(os.path.join(ROOT, "n2wos-cuda_amd", "Makefile"), "rb") as fh:
        h.update(fh.read())
    return h.hexdigest()


def profile_for_build(name: str):
    """profiles/<name> when it was measured on this build's walker, else None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        doc = json.load(f)
    return doc if doc.get("source_sha256") == walker_source_hash() else None


def host_cpu() -> dict:
    """Host CPU facts for the CPU baseline: online CPUs, this process's
    affinity, the model string."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "affinity": aff, "model": model}


def cpu_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed_events(torch, fn, reps: int) -> float:
    """Average ms of fn() over reps launches on torch's current stream."""
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def achievable_hbm_gbs(torch, nbytes: int = 1 << 30, reps: int = 10) -> float:
    """Achievable HBM bandwidth: a device-to-device copy of 1 GiB (read +
    write bytes / time), the STREAM-copy denominator of the roofline."""
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    ms = timed_events(torch, lambda: y.copy_(x), reps)
    del x, y
    return 2 * nbytes / (ms * 1e-3) / 1e9


def closest_point_stress(eng, torch, P, n: int = 262144) -> dict:
    """Queries/s and the query kernel's HBM roofline fraction per distribution.
    Algorithmic bytes per query: 12 (point) + 20 (d2, closest, id) + 112 per
    4-wide node loaded + 40 per triangle record tested, counts from the
    counting build of the same traversal."""
    v, t = P.make_mesh("mesh2")
    bvh = eng.Bvh(v, t)
    interior = P.interior_points(bvh, v, 65536, seed=3)
    sets = {
        "uniform_box": P.uniform_box_queries(n, seed=11),
        "near_boundary_shell": P.surface_shell_queries(v, t, n, seed=12),
        "wos_prefixes": P.wos_prefix_queries(bvh, interior, n, seed=13),
    }
    out = {}
    for name, q in sets.items():
        qt = torch.from_numpy(q).cuda()
        bvh.query_tensors(qt)
        ms = timed_events(torch, lambda: bvh.query_tensors(qt), 10)
        _, _, _, ovf, nn, nt = bvh.query_tensors(qt, counted=True)
        torch.cuda.synchronize()
        nodes, tris = int(nn.sum().item()), int(nt.sum().item())
        byts = 32 * n + NODE_BYTES * nodes + TRI_BYTES * tris
        gbs = byts / (ms * 1e-3) / 1e9
        # what binds: the critical path of the slowest query (its node visits
        # and 4-record leaf rounds are one chain of dependent loads), timed
        # alone, and the bulk without the 0.1% costliest queries
        cost = nn.double() + torch.ceil(nt.double() / 4.0)
        worst = int(torch.argmax(cost).item())
        q1 = qt[worst:worst + 1].contiguous()
        bvh.query_tensors(q1)
        crit_ms = timed_events(torch, lambda: bvh.query_tensors(q1), 10)
        keep = torch.sort(torch.argsort(cost, descending=True)[n // 1000:]).values
        qb = qt[keep].contiguous()
        bvh.query_tensors(qb)
        bulk_ms = timed_events(torch, lambda: bvh.query_tensors(qb), 10)
        bulk_bytes = 32 * len(keep) + NODE_BYTES * int(nn[keep].sum().item()) + TRI_BYTES * int(nt[keep].sum().item())
        out[name] = {"queries": n, "kernel_ms": ms, "queries_per_s": n / (ms * 1e-3),
                     "node_lines_per_query": nodes / n, "triangle_tests_per_query": tris / n,
                     "algorithmic_GBps": gbs, "hbm_frac": gbs / HBM_PEAK_GBS,
                     "overflow": int(ovf.sum().item()),
                     "critical_path": {"query_cost_iterations": float(cost[worst].item()),
                                       "mean_cost_iterations": float(cost.mean().item()),
                                       "alone_ms": crit_ms, "share_of_kernel": crit_ms / ms},
                     "bulk_without_costliest_0.1pct": {"queries": len(keep), "kernel_ms": bulk_ms,
                                                       "algorithmic_GBps": bulk_bytes / (bulk_ms * 1e-3) / 1e9}}
        log(f"closest point {name}: {n / (ms * 1e-3) / 1e6:.1f} M q/s, {gbs:.0f} GB/s; slowest query alone "
            f"{crit_ms:.3f} ms, bulk {bulk_ms:.3f} ms")
    return out


def cpu_closest_point(P, threads: int, gpu_qps: float, sample: int = 2048) -> dict:
    """The reference's CPU path for closest points -- closest_point_bruteforce
    (src/geometry/closest_point.cpp:50-68), single-threaded as in its probe
    (src/tools/probe_cuda_bvh.cpp:153-167) -- timed on a bounded sample of the
    problem-1 queries, beside the oracle's brute force and its binary-BVH
    traversal on 1 and on all host threads.  The reference's and the BVH's
    results are checked equal to the oracle brute force on the sample."""
    import oracle
    oracle.build()
    v1, t1 = P.make_mesh("mesh1")
    q = P.uniform_box_queries(65536, seed=0)[:sample]
    out = {"queries": sample, "of": 65536, "threads_all": threads, "gpu_queries_per_s": gpu_qps}

    def rate(fn, n):
        t0 = time.perf_counter()
        r = fn()
        return n / (time.perf_counter() - t0), r

    d2o, _, trio = oracle.bruteforce(v1, t1, q, 1)
    if oracle.ref_available():
        R = oracle.ref()
        d2 = np.zeros(sample, np.float32)
        cp = np.zeros((sample, 3), np.float32)
        tri = np.zeros(sample, np.int32)
        qps, _ = rate(lambda: R.ref_bruteforce(np.ascontiguousarray(v1), len(v1), np.ascontiguousarray(t1),
                                               len(t1), q, sample, d2, cp, tri), sample)
        out["reference_bruteforce_1_thread"] = {"queries_per_s": qps, "kind": "reference",
                                                "equal_to_oracle": bool(np.array_equal(tri, trio))}
    qps1, _ = rate(lambda: oracle.bruteforce(v1, t1, q, 1), sample)
    qpsn, _ = rate(lambda: oracle.bruteforce(v1, t1, q, threads), sample)
    out["oracle_bruteforce"] = {"queries_per_s_1_thread": qps1, "queries_per_s_all_threads": qpsn, "kind": "port"}
    b2 = oracle.Bvh2(v1, t1)
    big = P.uniform_box_queries(65536, seed=0)
    qb1, r1 = rate(lambda: b2.query(big, 1), 65536)
    qbn, _ = rate(lambda: b2.query(big, threads), 65536)
    out["oracle_bvh2"] = {"queries": 65536, "queries_per_s_1_thread": qb1, "queries_per_s_all_threads": qbn,
                          "kind": "port", "equal_to_bruteforce_on_sample": bool(np.array_equal(r1[2][:sample], trio))}
    log(f"cpu closest point: reference brute force "
        f"{out.get('reference_bruteforce_1_thread', {}).get('queries_per_s', float('nan')):.0f} q/s, "
        f"oracle bvh {qbn / 1e6:.2f} M q/s on {threads} threads")
    return out


def wos_problem1(eng, torch, P, W, cpu: bool, threads: int) -> dict:
    v, t = P.make_mesh("mesh1")
    bvh = eng.Bvh(v, t)
    pts = P.interior_points(bvh, v, 4096, seed=0)
    bc = P.charges_boundary(0)
    wp = W.WalkParams(max_steps=256, epsilon=1e-3, step_scale=0.999, seed=7)
    solver = W.Solver(bvh, bc, wp)
    tp = torch.from_numpy(pts).cuda()
    solver.solve(tp, 128, seed=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = solver.solve(tp, 128, seed=wp.seed)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    steps = int(st.steps.sum().item())
    truth = P.charges_solution_f64(bc.charges, pts)
    mean = st.mean.cpu().numpy()
    out = {"points": 4096, "walks_per_point": 128, "seconds": dt, "walk_steps_per_s": steps / dt,
           "rmse_vs_fp64": float(np.sqrt(np.mean((mean - truth) ** 2))),
           "mean_steps_per_walk": steps / (4096 * 128)}
    if cpu:
        import oracle
        b2 = oracle.Bvh2(v, t)
        t0 = time.perf_counter()
        _, st_cpu, _, _ = b2.walks(pts, 128, max_steps=256, eps=1e-3, step_scale=0.999, bmode=bc.mode,
                                   bparams=bc.oracle_params(), seed=wp.seed, nthreads=threads)
        dtc = time.perf_counter() - t0
        out["cpu_oracle"] = {"seconds": dtc, "walk_steps_per_s": float(st_cpu.sum() / dtc), "threads": threads,
                             "same_total_steps": bool(int(st_cpu.sum()) == steps)}
    log(f"wos problem 1: {steps / dt / 1e9:.3f} G steps/s, rmse {out['rmse_vs_fp64']:.4g}")
    return out


def poisson_problem3(eng, torch, P, W) -> dict:
    t0 = time.perf_counter()
    v, t = P.make_mesh("mesh3")
    bvh = eng.Bvh(v, t, leaf_size=4, build_method="sah")
    build_s = time.perf_counter() - t0
    pts = P.interior_points(bvh, v, 262144, seed=0)
    bc = P.poisson_boundary()
    wp = W.WalkParams(max_steps=256, epsilon=1e-4, step_scale=0.999, seed=31, source=W.SOURCE_POISSON_SINES)
    solver = W.Solver(bvh, bc, wp)
    tp = torch.from_numpy(pts).cuda()
    solver.solve(tp[:4096].contiguous(), 16, seed=1)
    torch.cuda.synchronize()
    from n2wos_amd import hiptime
    ev = hiptime.EventPair()
    st = W.PointStats.zeros(len(pts))
    t0 = time.perf_counter()
    solver.solve(tp, 512, stats=st, seed=wp.seed, events=ev.handles)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    kernel_ms = ev.elapsed()
    steps = int(st.steps.sum().item())
    truth = P.poisson_solution_f64(pts)
    mean = st.mean.cpu().numpy()
    var = (st.m2 / (st.n - 1)).cpu().numpy()
    err = mean - truth
    z = err / np.sqrt(np.maximum(var / 512, 1e-300))
    out = {"triangles": int(len(t)), "points": 262144, "walks_per_point": 512, "seconds": dt,
           "walk_steps_per_s": steps / dt, "walker_kernel_ms": kernel_ms, "mean_steps_per_walk": steps / (262144 * 512),
           "rmse_vs_fp64": float(np.sqrt(np.mean(err ** 2))), "mean_error": float(err.mean()),
           "frac_within_3_sigma": float(np.mean(np.abs(z[var > 0]) <= 3.0)), "host_mesh_and_bvh_s": build_s,
           "forced_walks": int(st.forced.sum().item())}
    log(f"poisson problem 3: {steps / dt / 1e9:.3f} G steps/s, rmse {out['rmse_vs_fp64']:.4g}")
    return out


def cache_rates(eng, torch) -> dict:
    """Training step (batch 2^18) and inference (2^20 points) of the benchmark
    cache; MFMA work per point: 28,800 FLOP forward, 86,400 train."""
    from n2wos_amd import cache as K
    c = K.Cache(K.CacheConfig.for_bounds(np.full(3, -1.0), np.full(3, 1.0), seed=1))
    g = torch.Generator(device="cuda").manual_seed(0)
    x = (torch.rand((1 << 20, 3), device="cuda", generator=g) * 2 - 1).contiguous()
    y = torch.sin(3 * x[:, 0]) * x[:, 1]
    xb, yb = x[: 1 << 18].contiguous(), y[: 1 << 18].contiguous()
    c.train_step(xb, yb)
    train_ms = timed_events(torch, lambda: c.train_step(xb, yb), 10)
    c.infer(x)
    infer_ms = timed_events(torch, lambda: c.infer(x), 10)
    tflops_train = 86400.0 * (1 << 18) / (train_ms * 1e-3) / 1e12
    tflops_inf = 28800.0 * (1 << 20) / (infer_ms * 1e-3) / 1e12
    return {"train_step_ms_batch_2^18": train_ms, "infer_ms_2^20": infer_ms,
            "train_points_per_s": (1 << 18) / (train_ms * 1e-3), "infer_points_per_s": (1 << 20) / (infer_ms * 1e-3),
            "train_mfma_TFLOPs": tflops_train, "infer_mfma_TFLOPs": tflops_inf,
            "train_mfma_frac_fp16_dense": tflops_train / 2500.0, "infer_mfma_frac_fp16_dense": tflops_inf / 2500.0}


def cv_pipeline(torch, bvh, verts, tp, bc, params, pts, reps: int = 3):
    """The control variate's first time-to-RMSE setup on problem 2: the
    training pipeline (capped label solves, Adam steps, bakes) and the
    estimate at its first walk count, `reps` times; host-clock seconds of
    each part of the last repetition and the estimate's RMSE against FP64."""
    from n2wos_amd import cv as CV, problems as P, ttrmse
    su = ttrmse.DEFAULT_CV_SETUPS[0]
    cs = CV.CvSolver(bvh, bc, params)
    truth = P.charges_solution_f64(bc.charges, pts)
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fld, cache, _ = CV.train_field(bvh, verts, bc, params, tp, su.cache_kw, su.label_stride, su.path_len,
                                       su.rounds, su.train_steps, su.batch, su.res, su.points, 11, solver=cs,
                                       label_max_steps=su.label_max_steps)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        st = cs.solve(tp, su.wpp[0], fld, seed=1000 + su.wpp[0])
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        del cache
    rmse = float(np.sqrt(np.mean((st.mean.cpu().numpy() - truth) ** 2)))
    return {"setup": su.name, "train_seconds": t1 - t0, "estimate_seconds": t2 - t1, "walks_per_point": su.wpp[0],
            "rmse": rmse}


def main():
    args = parse()
    if args.source_hash:
        print(walker_source_hash())
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    import __graft_entry__ as entry
    eng = entry.load_engine()
    from n2wos_amd import hiptime, problems as P, wos as W

    torch.cuda.set_device(0)
    verts, tris = P.make_mesh("mesh2")
    bvh = eng.Bvh(verts, tris, **WALKER_TREE)
    pts, _ = P.grid_slice_points(bvh, res=args.grid)
    n = len(pts)
    bc = P.charges_boundary(0)
    params = W.WalkParams(max_steps=256, epsilon=1e-4, step_scale=0.999, seed=2024)
    solver = W.Solver(bvh, bc, params)
    tp = torch.from_numpy(pts).cuda()
    stats = [W.PointStats.zeros(n) for _ in range(args.steps)]
    events = [hiptime.EventPair() for _ in range(args.steps)]

    if args.section == "walker":
        st = solver.solve(tp, args.wpp, seed=params.seed)
        torch.cuda.synchronize()
        print(json.dumps({"section": "walker", "points": n, "walks_per_point": args.wpp,
                          "walk_steps": int(st.steps.sum().item()), "source_sha256": walker_source_hash()}))
        return
    if args.section == "cache":
        print(json.dumps({"section": "cache", **cache_rates(eng, torch)}))
        return
    if args.section == "closest_stress":
        print(json.dumps({"section": "closest_stress", **closest_point_stress(eng, torch, P)}))
        return
    if args.section == "poisson":
        print(json.dumps({"section": "poisson", **poisson_problem3(eng, torch, P, W)}))
        return
    if args.section == "cv":
        print(json.dumps({"section": "cv", **cv_pipeline(torch, bvh, verts, tp, bc, params, pts)}))
        return

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

    tprof = profile_for_build("wos_walker_mesh2_traffic.json")
    traffic = tprof.get("hbm_bytes_per_launch") if tprof else None
    vprof = profile_for_build("wos_walker_mesh2_valu.json")
    valu = None
    if vprof:
        valu = {k: vprof.get(k) for k in ("valu_insts_per_step", "valu_busy", "valu_utilization", "waves_per_cu",
                                          "kernel_ms", "source")}
        valu["frac"] = vprof.get("valu_busy")

    truth = P.charges_solution_f64(bc.charges, pts)
    mean = stats[0].mean.cpu().numpy()
    rmse = float(np.sqrt(np.mean((mean - truth) ** 2)))
    achievable = achievable_hbm_gbs(torch)

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
                        "8-charge Laplace, FP64 per-point Welford/Chan, SAH BVH leaf 4",
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
            "achievable_GBps": achievable,
            "frac_of_achievable": achieved / achievable,
            "traffic": traffic,
            "binding_roof": "valu_issue",
            "valu": valu,
            "source_sha256": walker_source_hash(),
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

    threads = cpu_threads()
    extra = line["extra"]
    extra["closest_point_stress"] = closest_point_stress(eng, torch, P)
    for v in extra["closest_point_stress"].values():
        v["frac_of_achievable"] = v["algorithmic_GBps"] / achievable
        v["bulk_without_costliest_0.1pct"]["frac_of_achievable"] = (
            v["bulk_without_costliest_0.1pct"]["algorithmic_GBps"] / achievable)
    if not args.no_cpu:
        extra["closest_point_problem1_cpu"] = cpu_closest_point(P, threads, 65536 / (cpq_ms * 1e-3))
    extra["wos_problem1"] = wos_problem1(eng, torch, P, W, not args.no_cpu, threads)
    extra["cache"] = cache_rates(eng, torch)
    if not args.quick:
        extra["poisson_problem3"] = poisson_problem3(eng, torch, P, W)
        from n2wos_amd import ttrmse
        ttr_params = W.WalkParams(max_steps=256, epsilon=1e-4, step_scale=0.999, seed=2024)
        r = ttrmse.run(bvh, verts, pts, bc, truth, ttr_params, plain_wpp=(320, 384, 448, 512), log=log)
        extra["time_to_rmse"] = r
        from n2wos_amd import progressive as PR
        extra["progressive_2lmc"] = PR.progressive_vs_serial(bvh, verts, pts, bc, truth, ttr_params, log=log)
        from n2wos_amd import serve as S
        extra["serving"] = S.serving_workload(eng, n_requests=args.serve_requests, rate_per_s=args.serve_rate,
                                              config=S.ServerConfig(hbm_budget_bytes=1 << 30),
                                              log=log)
        sv = extra["serving"]
        log(f"serving: p50 {sv['latency_p50_s']:.3f} s p99 {sv['latency_p99_s']:.3f} s over {sv['latency_samples']} "
            f"requests, device busy {sv['gpu_busy_fraction']:.2f}, {sv['points_per_s']:.0f} points/s")

    if not args.no_cpu:
        import oracle
        oracle.build()
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
        hc = host_cpu()
        line["cpu_baseline"] = {
            "value": float(st_cpu.sum() / dt),
            "unit": "walk-steps/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{len(sample)} of {n} grid points (every {stride}th) x {args.wpp} walks, same walks as the "
                      f"GPU run, {dt:.1f} s on {threads} host threads",
            "threads": threads,
            "nproc": hc["nproc"],
            "affinity": hc["affinity"],
            "cpu_model": hc["model"],
            "threads_note": "all the host CPUs one GPU is allotted (16 per GPU on the GPU box), "
                            "capped at this process's affinity",
        }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
