"""The wavefront-engine evaluation tool at one point x0: plain WoS and the
two-level estimator with the analytic solution as the cache (the counterpart
of the reference's n2wos_eval_wavefront_wos, src/tools/eval_wavefront_wos.cu).

Methods (boundary data g = x^2 - y^2, harmonic, so u = x^2 - y^2 inside):
  pure_wos         `samples` plain walks from x0;
  oracle_coarse    walks stopped at depth m, plus the exact u at X_m (the
                   "cache" is the analytic solution): unbiased, since u is
                   harmonic;
  oracle_residual  coupled walks W - S_m - u(X_m): zero mean;
  oracle_2lmc      coarse (coarse_samples) + residual (residual_samples,
                   seed xor 0x517cc1b727220a95), as the reference (:327-342).
Same flags (:96-181) and the JSON "n2wos_wavefront_wos_eval_v1" (:356-432)
with truthful implementation_mode labels: every engine here is the
persistent wave-scheduled walker, whatever --engine names.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from datetime import datetime, timezone

import numpy as np

from . import eval_nc as EV, geometry as G, nc as N, wos as W
from .hiptime import EventPair

METHODS = ("pure_wos", "oracle_coarse", "oracle_residual", "oracle_2lmc")


def parse_args(argv=None):
    ap = argparse.ArgumentParser(description="Plain WoS and the analytic-oracle two-level estimator at x0.")
    a = ap.add_argument
    a("--mesh", default="procedural_bumpy_sphere", help="procedural_bumpy_sphere | obj | ply | mesh1 | mesh2 | mesh3")
    a("--mesh-path", default="")
    a("--normalize", type=int, default=1)
    a("--bumpy-stacks", type=int, default=128)
    a("--bumpy-slices", type=int, default=256)
    a("--bumpy-amplitude", type=float, default=0.15)
    a("--method", default="pure_wos", choices=METHODS)
    a("--engine", default="wavefront", choices=["wavefront", "persistent"])
    a("--samples", type=int, default=65536)
    a("--coarse-samples", type=int, default=0)
    a("--residual-samples", type=int, default=0)
    a("--depth-m", type=int, default=8)
    a("--max-steps", type=int, default=256)
    a("--epsilon", type=float, default=1e-4)
    a("--step-scale", type=float, default=0.999)
    a("--x0", default="0.1,0.05,0.0")
    a("--seed", type=int, default=12345)
    a("--block-size", type=int, default=128)
    a("--cubql-leaf-size", type=int, default=8)
    a("--cubql-build-method", default="sah")
    a("--output", default="results/eval_wavefront_wos.json")
    opt = ap.parse_args(argv)
    x0 = [float(v) for v in opt.x0.split(",")]
    if len(x0) != 3:
        ap.error("--x0 expects x,y,z")
    opt.x0 = x0
    if opt.samples <= 0 or opt.max_steps <= 0 or opt.block_size <= 0:
        ap.error("samples, max-steps, and block-size must be positive")
    if opt.depth_m < 0:
        ap.error("--depth-m must be >= 0")
    # the BVH builders of this engine (the reference's cuBQL names map onto them)
    opt.bc = "harmonic_x2_minus_y2"
    return opt


def _stats(values: np.ndarray, steps_total: int, samples: int, elapsed_ms: float, forced: int, overflow: int):
    v = values.astype(np.float64)
    mean = float(v.mean())
    var = float(v.var(ddof=1)) if samples > 1 else 0.0
    return {"samples": samples, "mean": mean, "sample_variance": var, "estimator_variance": var / samples,
            "stderr": float(np.sqrt(var / samples)), "mean_steps": steps_total / samples, "elapsed_ms": elapsed_ms,
            "us_per_sample": elapsed_ms * 1e3 / samples, "launched_query_slots": int(steps_total + samples),
            "mean_launched_query_slots_per_sample": (steps_total + samples) / samples,
            "scheduled_query_rounds": None, "forced_max_steps": int(forced), "overflow_count": int(overflow)}


def _run_json(s: dict, exact: float) -> dict:
    err = s["mean"] - exact
    out = {"samples": s["samples"], "mean": s["mean"], "exact_target": exact, "error": err, "abs_error": abs(err)}
    out.update({k: s[k] for k in ("sample_variance", "estimator_variance", "stderr", "mean_steps", "elapsed_ms",
                                  "us_per_sample", "launched_query_slots", "mean_launched_query_slots_per_sample",
                                  "scheduled_query_rounds", "forced_max_steps", "overflow_count")})
    return out


def run_method(bvh, bc, opt, method: str, samples: int, seed: int) -> dict:
    """One method's samples from x0 on the device; per-sample values come
    back for the statistics (FP64 on the host, as the reference's finish)."""
    import torch
    x0 = torch.tensor([opt.x0], dtype=torch.float32, device="cuda")
    params = W.WalkParams(max_steps=opt.max_steps, epsilon=opt.epsilon, step_scale=opt.step_scale, seed=seed)
    ev = EventPair()
    if method == "pure_wos":
        st = W.Solver(bvh, bc, params).solve(x0, samples, per_walk=True, events=ev.handles)
        ms = ev.elapsed()
        vals = st.extra["walk_values"].cpu().numpy()
        return _stats(vals, int(st.steps.sum().item()), samples, ms, int(st.forced.sum().item()), 0)
    ns = N.NcSolver(bvh, bc, params)
    mode = N.NC_PREFIX if method == "oracle_coarse" else N.NC_COUPLED
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b = ns.walks(x0, samples, mode, opt.depth_m, stream_id=0, seed=seed)
    xm = b.xm
    cache = (xm[:, 0] * xm[:, 0] - xm[:, 1] * xm[:, 1]).contiguous()  # the analytic solution at X_m
    need = b.need.bool()
    if mode == N.NC_PREFIX:
        vals = torch.where(need, b.prefix + cache, b.prefix)
    else:
        vals = torch.where(need, b.full - (b.prefix + cache), torch.zeros_like(b.prefix))
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    return _stats(vals.cpu().numpy(), int(b.steps.sum().item()), samples, ms, 0, 0)


def main(argv=None) -> int:
    import torch
    opt = parse_args(argv)
    cmd = " ".join([os.path.basename(sys.argv[0])] + list(argv if argv is not None else sys.argv[1:]))
    eopt = argparse.Namespace(mesh=opt.mesh, mesh_path=opt.mesh_path, normalize=opt.normalize,
                              bumpy_stacks=opt.bumpy_stacks, bumpy_slices=opt.bumpy_slices,
                              bumpy_amplitude=opt.bumpy_amplitude, seed=opt.seed)
    v, t, name, center, scale = EV.make_mesh(eopt)
    bvh = G.Bvh(v, t, leaf_size=opt.cubql_leaf_size, build_method=opt.cubql_build_method)
    bc = W.Boundary(W.BC_X2_MINUS_Y2)
    exact = float(np.float64(opt.x0[0]) ** 2 - np.float64(opt.x0[1]) ** 2)
    runs = {}
    if opt.method == "oracle_2lmc":
        cs = opt.coarse_samples if opt.coarse_samples > 0 else opt.samples
        rs = opt.residual_samples if opt.residual_samples > 0 else max(1, opt.samples // 8)
        coarse = run_method(bvh, bc, opt, "oracle_coarse", cs, opt.seed)
        resid = run_method(bvh, bc, opt, "oracle_residual", rs, opt.seed ^ 0x517cc1b727220a95)
        runs["coarse"] = _run_json(coarse, exact)
        runs["residual"] = _run_json(resid, 0.0)
        mean = coarse["mean"] + resid["mean"]
        evar = coarse["estimator_variance"] + resid["estimator_variance"]
        elapsed = coarse["elapsed_ms"] + resid["elapsed_ms"]
    else:
        s = run_method(bvh, bc, opt, opt.method, opt.samples, opt.seed)
        runs[opt.method] = _run_json(s, 0.0 if opt.method == "oracle_residual" else exact)
        mean, evar, elapsed = s["mean"], s["estimator_variance"], s["elapsed_ms"]
    st = bvh.stats()
    lo, hi = v.min(0), v.max(0)
    dev = torch.cuda.get_device_properties(0)
    doc = {
        "schema": "n2wos_wavefront_wos_eval_v1",
        "patch": "n2wos_amd_gfx950",
        "generated_at_utc": datetime.now(timezone.utc).strftime("%Y-%m-%dT%H:%M:%SZ"),
        "command_line": cmd,
        "cuda": {"runtime": f"HIP {torch.version.hip}", "device": f"{dev.name} {dev.gcnArchName}"},
        "implementation_mode": {
            "engine": "persistent_wave_scheduled_walker",
            "geometry_backend": "hip_bvh4_gfx950",
            "geometry_build_method": opt.cubql_build_method,
            "gpu_resident_walker_state": True,
            "gpu_rng": "philox4x32_10_counter_based",
            "cpu_gpu_transfer_inside_sampling_loop": False,
            "host_controls_global_step_loop": False,
            "per_walk_kernel_launch": False,
            "active_compaction": True,
            "inactive_slots_skip_bvh_traversal": True,
            "coarse_query_rounds_capped_at_depth_m": opt.method in ("oracle_coarse", "oracle_2lmc"),
            "cache_backend": "analytic_oracle_debug_only",
            "tcnn_in_solver": False,
            "timing_scope": "hip_events_walker_kernel_only" if opt.method == "pure_wos"
            else "host_clock_walks_and_oracle_combine_exclude_bvh_build_and_final_readback",
            "production_status": "evaluation_tool",
        },
        "options": {"mesh": opt.mesh, "mesh_path": opt.mesh_path, "normalize": bool(opt.normalize),
                    "method": opt.method, "engine": opt.engine, "samples": opt.samples,
                    "coarse_samples": opt.coarse_samples, "residual_samples": opt.residual_samples,
                    "depth_m": opt.depth_m, "max_steps": opt.max_steps, "epsilon": opt.epsilon,
                    "step_scale": opt.step_scale, "x0": opt.x0, "seed": opt.seed, "block_size": opt.block_size,
                    "cubql_leaf_size": opt.cubql_leaf_size, "cubql_build_method": opt.cubql_build_method},
        "mesh_stats": {"name": name, "vertices": int(len(v)), "triangles": int(len(t)),
                       "degenerate_triangles": int(G.count_degenerate_triangles(v, t, 1e-24)),
                       "bounds_min": lo.tolist(), "bounds_max": hi.tolist(),
                       "normalization": {"center": np.asarray(center).tolist(), "scale": float(scale)}},
        "bvh_stats": {"triangles": st["triangle_count"], "nodes": st["node_count"], "prim_ids": st["index_count"],
                      "leaf_size": st["leaf_size"], "build_method": st["build_method"],
                      "build_milliseconds": st["build_milliseconds"]},
        "target": {"boundary_condition": "harmonic_x2_minus_y2", "exact_u_x0": exact},
        "runs": runs,
        "estimator": {"method": opt.method, "engine": opt.engine, "mean": mean, "exact": exact,
                      "error": mean - exact, "abs_error": abs(mean - exact), "estimator_variance": evar,
                      "stderr": float(np.sqrt(evar)), "elapsed_ms_total": elapsed},
    }
    d = os.path.dirname(opt.output)
    if d:
        os.makedirs(d, exist_ok=True)
    text = json.dumps(doc, indent=2)
    with open(opt.output, "w") as f:
        f.write(text + "\n")
    print(text)
    return 0
