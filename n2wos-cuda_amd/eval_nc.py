"""Neural-cache evaluation tool: plain WoS, NC and NC + 2LMC on one problem,
written with the reference evaluation tool's outputs so its analysis scripts
read them unchanged (src/tools/eval_tcnn_nc_wos.cu):

  <output>.json                 schema n2wos_tcnn_nc_wos_eval_v4 (:802-849),
                                implementation_mode labels stating what ran
  <prefix>_estimates.csv        per-point estimates, same columns (:626-645)
  <slice>_mask.ppm, _points.csv slice interior mask and points (:403-404)
  <prefix>_train_points.csv     training points (point_id,x,y,z)

Flags keep the reference names (--pure-walks-per-point, --depth-m,
--label-refreshes, --slice-width, ...).  Training follows the reference's
schedule: label_refreshes rounds, each adding walks_per_label_refresh WoS
walks per training point to a running label (the per-point mean over all
rounds, nc_update_labels_kernel, src/cuda/cubql_bvh.cu:1372-1420) or using
the analytic solution, then train_steps_per_refresh Adam steps on the whole
training set.  The cache is this engine's hash grid + ReLU MLP on MFMA with
the preset's (or the flags') grid and network shape; `--network` and
`--jit` are reported, not used.  Exact values are the FP64 analytic
solutions.

Run as: python n2wos-cuda_amd/tools/n2wos_eval_nc_wos.py [flags]
"""
from __future__ import annotations

import argparse
import datetime
import json
import os
import shlex
import sys
import time

import numpy as np

from . import cache as K, geometry as G, nc as N, problems as P, wos as W

SCHEMA = "n2wos_tcnn_nc_wos_eval_v4"

# boundary names of the reference (src/cuda/cubql_bvh.cu:1700-1730) this
# engine evaluates, plus the benchmark's 8-charge data
BOUNDARIES = {
    "harmonic_x2_minus_y2": W.BC_X2_MINUS_Y2,
    "harmonic_zebra_k4": W.BC_ZEBRA_K4,
    "harmonic_zebra_k8": W.BC_ZEBRA_K8,
    "harmonic_zebra_k12": W.BC_ZEBRA_K12,
    "constant_one": W.BC_CONSTANT_ONE,
    "external_charges8": W.BC_CHARGES,
}
ALIASES = {"x2_minus_y2": "harmonic_x2_minus_y2", "zebra_k4": "harmonic_zebra_k4", "zebra_k8": "harmonic_zebra_k8",
           "zebra_k12": "harmonic_zebra_k12", "constant": "constant_one", "one": "constant_one",
           "const1": "constant_one", "charges8": "external_charges8"}

# the reference's cache presets (eval_tcnn_nc_wos.cu:249-266): n_levels,
# log2_hashmap_size, base_resolution, per_level_scale, n_neurons,
# n_hidden_layers, n_features_per_level
PRESETS = {
    "baseline": (12, 18, 8, 1.5, 32, 2, 2), "paper_like": (12, 18, 8, 1.5, 32, 2, 2),
    "light": (8, 15, 8, 1.5, 16, 1, 2), "nano": (6, 13, 8, 1.5, 16, 1, 2), "heavy": (16, 19, 8, 1.4, 64, 3, 2),
    "xlarge": (16, 20, 8, 1.35, 64, 4, 4),
}
PRESET_FIELDS = ("n_levels", "log2_hashmap_size", "base_resolution", "per_level_scale", "n_neurons",
                 "n_hidden_layers", "n_features_per_level")


def parse_args(argv=None):
    ap = argparse.ArgumentParser(description="Plain WoS, neural-cache WoS and NC + 2LMC on one problem.")
    a = ap.add_argument
    a("--mesh", default="procedural_bumpy_sphere",
      help="procedural_bumpy_sphere | obj | ply | mesh1 | mesh2 | mesh3 | blob")
    a("--mesh-path", default="")
    a("--normalize", type=int, default=1)
    a("--bumpy-stacks", type=int, default=128)
    a("--bumpy-slices", type=int, default=256)
    a("--bumpy-amplitude", type=float, default=0.15)
    a("--cubql-build-method", default="sah", help="sah | median (BVH builder)")
    a("--cubql-leaf-size", type=int, default=8)
    a("--bc", "--boundary", dest="bc", default="external_charges8")
    a("--charges-seed", type=int, default=0)
    a("--label-source", default="wos_supervision", choices=["wos_supervision", "exact_analytic"])
    a("--train-points", type=int, default=20000)
    a("--train-sampler", default="rejection", choices=["rejection"])
    a("--save-train-points-prefix", default="")
    a("--eval-points", type=int, default=8192)
    a("--eval-mode", "--eval-sampler", dest="eval_mode", default="ball", choices=["ball", "slice", "train_points"])
    a("--eval-use-train-points", type=int, default=0)
    a("--slice-width", type=int, default=512)
    a("--slice-height", type=int, default=512)
    a("--slice-view", default="xy", choices=["xy", "xz", "yz"])
    a("--slice-plane", type=float, default=0.0)
    a("--slice-frame", default="mesh_bounds", help="mesh_bounds | u_min,u_max,v_min,v_max")
    a("--slice-padding-fraction", type=float, default=0.02)
    a("--slice-preserve-world-aspect", type=int, default=1)
    a("--slice-output-prefix", default="")
    a("--label-refreshes", type=int, default=4)
    a("--walks-per-label-refresh", type=int, default=50)
    a("--train-steps-per-refresh", type=int, default=5000)
    a("--skip-training", type=int, default=0)
    a("--pure-walks-per-point", type=int, default=64)
    a("--hybrid-walks-per-point", type=int, default=4)
    a("--coarse-walks-per-point", type=int, default=64)
    a("--residual-walks-per-point", type=int, default=4)
    a("--enable-2lmc", type=int, default=1)
    a("--depth-m", type=int, default=1)
    a("--max-steps", type=int, default=256)
    a("--epsilon", type=float, default=1e-4)
    a("--step-scale", type=float, default=0.999)
    a("--seed", type=int, default=12345)
    # default: the reference's "baseline" preset; an explicit encoding or
    # network flag overrides the preset's value and makes the run "custom"
    a("--cache-preset", default="baseline")
    a("--network", default="hashgrid_mlp_mfma")
    a("--n-levels", type=int, default=None)
    a("--n-features-per-level", type=int, default=None)
    a("--log2-hashmap-size", type=int, default=None)
    a("--base-resolution", type=int, default=None)
    a("--per-level-scale", type=float, default=None)
    a("--n-neurons", type=int, default=None)
    a("--n-hidden-layers", type=int, default=None)
    a("--learning-rate", type=float, default=1e-2)
    a("--batch-size", type=int, default=0, help="Adam batch (0 = the whole training set, as the reference)")
    a("--antithetic", type=int, default=0, help="mirrored first steps in NC walk pairs (this engine's option)")
    a("--jit", type=int, default=0)
    a("--block-size", type=int, default=128)
    a("--output", default="results/eval_nc_wos.json")
    a("--save-estimates-prefix", default="")
    opt = ap.parse_args(argv)
    opt.bc = ALIASES.get(opt.bc, opt.bc)
    if opt.bc not in BOUNDARIES:
        ap.error(f"unsupported boundary {opt.bc!r}; this engine evaluates {sorted(BOUNDARIES)}")
    preset = "baseline" if opt.cache_preset == "custom" else opt.cache_preset
    if preset not in PRESETS:
        ap.error(f"unknown cache preset {opt.cache_preset!r} (expected {sorted(PRESETS)} or custom)")
    explicit = False
    for name, value in zip(PRESET_FIELDS, PRESETS[preset]):
        if getattr(opt, name) is None:
            setattr(opt, name, value)
        else:
            explicit = True
    if explicit:
        opt.cache_preset = "custom"
    if opt.n_features_per_level != 2:
        ap.error("this engine's hash grid has 2 features per level (preset xlarge needs 4: use "
                 "--cache-preset heavy --n-hidden-layers 4 --log2-hashmap-size 20 --per-level-scale 1.35)")
    if not 1 <= opt.n_levels <= 16:
        ap.error("--n-levels must be in [1, 16]")
    if opt.n_neurons not in (16, 32, 64):
        ap.error("--n-neurons must be 16, 32 or 64")
    if not 1 <= opt.n_hidden_layers <= 4:
        ap.error("--n-hidden-layers must be in [1, 4]")
    return opt


# ------------------------------------------------------------------ problem
def make_mesh(opt):
    if opt.mesh == "procedural_bumpy_sphere":
        v, t = G.uv_sphere(opt.bumpy_stacks, opt.bumpy_slices, G.SPHERE_REFERENCE_BUMP, [opt.bumpy_amplitude])
        name = "procedural_bumpy_sphere"
    elif opt.mesh in ("obj", "ply"):
        if not opt.mesh_path:
            raise SystemExit(f"--mesh-path required for {opt.mesh}")
        v, t = G.load_mesh(opt.mesh_path, opt.mesh)
        name = os.path.basename(opt.mesh_path)
    elif opt.mesh in P.MESHES:
        stacks, slices, kind, params = P.MESHES[opt.mesh]
        v, t = G.uv_sphere(stacks, slices, kind, params if params is not None else P.sh_blob_coefficients(opt.seed))
        name = opt.mesh
    else:
        raise SystemExit(f"unknown mesh: {opt.mesh}")
    center, scale = np.zeros(3, np.float32), 1.0
    if opt.normalize:
        v, center, scale = G.normalize_to_unit_radius(v)
    return v, t, name, center, scale


def boundary(opt) -> W.Boundary:
    mode = BOUNDARIES[opt.bc]
    if mode == W.BC_CHARGES:
        return W.Boundary(mode, P.charges(opt.charges_seed))
    return W.Boundary(mode)


def exact_f64(b: W.Boundary, pts) -> np.ndarray:
    """FP64 solution at interior points (each boundary datum here is the
    trace of a harmonic function, so the solution is that function)."""
    p = np.asarray(pts, np.float64).reshape(-1, 3)
    if b.mode == W.BC_CHARGES:
        return P.charges_solution_f64(b.charges, p)
    if b.mode == W.BC_X2_MINUS_Y2:
        return p[:, 0] ** 2 - p[:, 1] ** 2
    if b.mode in (W.BC_ZEBRA_K4, W.BC_ZEBRA_K8, W.BC_ZEBRA_K12):
        k = {W.BC_ZEBRA_K4: 4, W.BC_ZEBRA_K8: 8, W.BC_ZEBRA_K12: 12}[b.mode]
        return np.real((p[:, 0] + 1j * p[:, 1]) ** k)
    if b.mode == W.BC_CONSTANT_ONE:
        return np.ones(len(p))
    raise ValueError("no analytic solution for this boundary")


AXES = {"xy": (0, 1, 2), "xz": (0, 2, 1), "yz": (1, 2, 0)}


def slice_frame(opt, lo, hi):
    """The reference's slice frame: mesh bounds padded by a fraction, widened
    to the pixel aspect (make_slice_eval_points, eval_tcnn_nc_wos.cu:401)."""
    ua, va, _ = AXES[opt.slice_view]
    if opt.slice_frame == "mesh_bounds":
        du, dv = float(hi[ua] - lo[ua]), float(hi[va] - lo[va])
        f = opt.slice_padding_fraction
        u0, u1, v0, v1 = lo[ua] - f * du, hi[ua] + f * du, lo[va] - f * dv, hi[va] + f * dv
    else:
        u0, u1, v0, v1 = (float(x) for x in opt.slice_frame.split(","))
    u0, u1, v0, v1 = (np.float32(x) for x in (u0, u1, v0, v1))
    if opt.slice_preserve_world_aspect:
        du, dv = u1 - u0, v1 - v0
        target = np.float32(opt.slice_width) / np.float32(opt.slice_height)
        cur = du / dv
        if cur < target:
            nd, c = target * dv, np.float32(0.5) * (u0 + u1)
            u0, u1 = c - np.float32(0.5) * nd, c + np.float32(0.5) * nd
        elif cur > target:
            nd, c = du / target, np.float32(0.5) * (v0 + v1)
            v0, v1 = c - np.float32(0.5) * nd, c + np.float32(0.5) * nd
    if not (u1 > u0 and v1 > v0):
        raise SystemExit("invalid slice frame")
    return float(u0), float(u1), float(v0), float(v1)


def slice_points(opt, bvh, lo, hi):
    """Pixel (ix, iy) at u = u0 + (u1 - u0) ix / (W - 1), v likewise; kept when
    inside the mesh (this engine's ray-parity test).  Returns points, mask, frame."""
    u0, u1, v0, v1 = slice_frame(opt, lo, hi)
    w, h = opt.slice_width, opt.slice_height
    uu = np.float32(u0) + (np.float32(u1) - np.float32(u0)) * (np.arange(w, dtype=np.float32) / np.float32(max(1, w - 1)))
    vv = np.float32(v0) + (np.float32(v1) - np.float32(v0)) * (np.arange(h, dtype=np.float32) / np.float32(max(1, h - 1)))
    if w == 1:
        uu = np.full(1, 0.5 * (u0 + u1), np.float32)
    if h == 1:
        vv = np.full(1, 0.5 * (v0 + v1), np.float32)
    ua, va, fa = AXES[opt.slice_view]
    U, V = np.meshgrid(uu, vv, indexing="xy")
    cand = np.zeros((w * h, 3), np.float32)
    cand[:, ua], cand[:, va], cand[:, fa] = U.ravel(), V.ravel(), np.float32(opt.slice_plane)
    mask = P.inside_mask(bvh, cand)
    if not mask.any():
        raise SystemExit("slice contains no interior pixels; adjust --slice-view, --slice-plane, or --slice-frame")
    return cand[mask], mask.astype(np.uint8), (u0, u1, v0, v1)


def ball_points(n: int, lo, hi, seed: int) -> np.ndarray:
    """Uniform points in the ball of radius 0.42 min(extent) at the bounds'
    centre (make_ball_points, eval_tcnn_nc_wos.cu:383-389; numpy PCG64 here)."""
    rng = np.random.default_rng(seed)
    c = 0.5 * (lo + hi)
    r = 0.42 * float(np.min(hi - lo))
    z = 1.0 - 2.0 * rng.uniform(size=n)
    phi = 2 * np.pi * rng.uniform(size=n)
    rr = r * np.cbrt(rng.uniform(size=n))
    s = np.sqrt(np.maximum(0.0, 1 - z * z))
    return np.stack([c[0] + rr * s * np.cos(phi), c[1] + rr * s * np.sin(phi), c[2] + rr * z], 1).astype(np.float32)


# ------------------------------------------------------------------ outputs
def write_points_csv(path, pts):
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    with open(path, "w") as f:
        f.write("point_id,x,y,z\n")
        for i, p in enumerate(pts):
            f.write(f"{i},{p[0]:.10g},{p[1]:.10g},{p[2]:.10g}\n")


def write_mask_ppm(path, mask, w, h):
    """P6, rows top to bottom = v descending; inside 235, outside 20."""
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    img = np.where(mask.reshape(h, w)[::-1], 235, 20).astype(np.uint8)
    with open(path, "wb") as f:
        f.write(f"P6\n{w} {h}\n255\n".encode())
        f.write(np.repeat(img[:, :, None], 3, axis=2).tobytes())


def write_estimates_csv(path, pts, exact, pure, nc, two=None):
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    cols = "point_id,x,y,z,analytic_value,pure_mean,pure_sample_variance,nc_wos_mean,nc_wos_sample_variance"
    if two is not None:
        cols += (",nc_2lmc_mean,nc_2lmc_coarse_mean,nc_2lmc_residual_mean,nc_2lmc_coarse_sample_variance,"
                 "nc_2lmc_residual_sample_variance")
    with open(path, "w") as f:
        f.write(cols + "\n")
        for i in range(len(pts)):
            row = [str(i)] + [f"{x:.10g}" for x in (pts[i, 0], pts[i, 1], pts[i, 2], exact[i], pure[0][i],
                                                    pure[1][i], nc[0][i], nc[1][i])]
            if two is not None:
                row += [f"{x:.10g}" for x in (two[0][i] + two[2][i], two[0][i], two[2][i], two[1][i], two[3][i])]
            f.write(",".join(row) + "\n")


def _mv(st):
    n = st.n.cpu().numpy()
    mean = st.mean.cpu().numpy()
    var = np.where(n > 1, st.m2.cpu().numpy() / np.maximum(n - 1, 1), 0.0)
    return mean, np.maximum(var, 0.0)


def run_stats(mean, var, exact, steps_total, wpp, elapsed_ms, training_ms=0.0, cache_queries=0, forced=0):
    n = len(mean)
    err = mean - exact
    samples = n * wpp
    return {
        "eval_points": n, "walks_per_point": wpp, "samples": samples,
        "rmse": float(np.sqrt(np.mean(err ** 2))), "mae": float(np.mean(np.abs(err))),
        "max_abs_error": float(np.max(np.abs(err))), "mean_estimate": float(mean.mean()),
        "mean_exact": float(exact.mean()), "mean_bias": float(mean.mean() - exact.mean()),
        "mean_sample_variance": float(var.mean()), "mean_steps": float(steps_total) / samples,
        "cache_queries": int(cache_queries), "cache_query_fraction": cache_queries / samples if samples else 0.0,
        "forced_max_steps": int(forced), "overflow_count": 0, "elapsed_ms": elapsed_ms,
        "training_plus_elapsed_ms": training_ms + elapsed_ms, "us_per_point": 1000.0 * elapsed_ms / n,
        "us_per_sample": 1000.0 * elapsed_ms / samples,
    }


# ------------------------------------------------------------------ main
def main(argv=None) -> int:
    import torch

    opt = parse_args(argv)
    if not torch.cuda.is_available():
        print("error: no GPU", file=sys.stderr)
        return 1
    torch.cuda.set_device(0)

    def sync_ms(t0):
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3

    v, t, mesh_name, center, scale = make_mesh(opt)
    lo, hi = v.min(0), v.max(0)
    bvh = G.Bvh(v, t, leaf_size=opt.cubql_leaf_size, build_method=opt.cubql_build_method)
    bst = bvh.stats()
    bc = boundary(opt)
    params = W.WalkParams(max_steps=opt.max_steps, epsilon=opt.epsilon, step_scale=opt.step_scale, seed=opt.seed)

    # training set: rejection sampling in the bounding box with the inside test
    train = P.interior_points(bvh, v, opt.train_points, seed=opt.seed)
    if opt.save_train_points_prefix:
        write_points_csv(opt.save_train_points_prefix + "_train_points.csv", train)

    slice_info = None
    if opt.eval_mode == "slice":
        pts, mask, frame = slice_points(opt, bvh, lo, hi)
        sp = opt.slice_output_prefix or (os.path.splitext(opt.output)[0] + "_slice")
        write_mask_ppm(sp + "_mask.ppm", mask, opt.slice_width, opt.slice_height)
        write_points_csv(sp + "_points.csv", pts)
        slice_info = {"width": opt.slice_width, "height": opt.slice_height,
                      "candidate_pixels": opt.slice_width * opt.slice_height, "inside_pixels": int(len(pts)),
                      "view": opt.slice_view, "plane": opt.slice_plane, "frame_u_min": frame[0],
                      "frame_u_max": frame[1], "frame_v_min": frame[2], "frame_v_max": frame[3],
                      "mask_ppm": sp + "_mask.ppm", "points_csv": sp + "_points.csv"}
    elif opt.eval_mode == "train_points" or opt.eval_use_train_points:
        pts = train
    else:
        pts = ball_points(opt.eval_points, lo, hi, opt.seed + 1)
    exact = exact_f64(bc, pts)
    tp = torch.from_numpy(np.ascontiguousarray(pts, np.float32)).cuda()
    ttrain = torch.from_numpy(np.ascontiguousarray(train, np.float32)).cuda()

    # cache and training
    cfg = K.CacheConfig.for_bounds(lo, hi, n_levels=opt.n_levels, log2_table_size=opt.log2_hashmap_size,
                                   base_resolution=opt.base_resolution, growth=opt.per_level_scale,
                                   learning_rate=opt.learning_rate, seed=opt.seed,
                                   max_batch=opt.batch_size if opt.batch_size > 0 else len(train),
                                   n_neurons=opt.n_neurons, n_hidden_layers=opt.n_hidden_layers)
    cache = K.Cache(cfg)
    solver = W.Solver(bvh, bc, params)
    label_ms = train_ms = 0.0
    if not opt.skip_training:
        labels = W.PointStats.zeros(len(train))
        gen = torch.Generator(device="cuda").manual_seed(opt.seed)
        batch = opt.batch_size if opt.batch_size > 0 else len(train)
        for r in range(opt.label_refreshes):
            t0 = time.perf_counter()
            if opt.label_source == "exact_analytic":
                y = torch.from_numpy(exact_f64(bc, train).astype(np.float32)).cuda()
            else:
                solver.solve(ttrain, opt.walks_per_label_refresh, stats=labels, seed=opt.seed + 7919 * (r + 1),
                             accumulate=r > 0)
                y = labels.mean.to(torch.float32)
            label_ms += sync_ms(t0)
            t0 = time.perf_counter()
            for _ in range(opt.train_steps_per_refresh):
                if batch >= len(train):
                    cache.train_step(ttrain, y)
                else:
                    idx = torch.randint(0, len(train), (batch,), device="cuda", generator=gen)
                    cache.train_step(ttrain[idx].contiguous(), y[idx].contiguous())
            train_ms += sync_ms(t0)

    # runs
    solver.solve(tp[:64].contiguous(), 4)
    t0 = time.perf_counter()
    pure = solver.solve(tp, opt.pure_walks_per_point, seed=opt.seed + 1)
    pure_ms = sync_ms(t0)
    ns = N.NcSolver(bvh, bc, params)
    t0 = time.perf_counter()
    nc, nb = ns.nc(cache, tp, opt.hybrid_walks_per_point, opt.depth_m, stream_id=1, seed=opt.seed + 2,
                   antithetic=bool(opt.antithetic))
    nc_ms = sync_ms(t0)
    pm, pv = _mv(pure)
    nm, nv = _mv(nc)
    training_ms = label_ms + train_ms
    runs = {
        "pure_wos": run_stats(pm, pv, exact, int(pure.steps.sum().item()), opt.pure_walks_per_point, pure_ms,
                              forced=int(pure.forced.sum().item())),
        "nc_wos": run_stats(nm, nv, exact, int(nb.steps.sum().item()), opt.hybrid_walks_per_point, nc_ms,
                            training_ms, cache_queries=int(nb.need.sum().item())),
    }
    two = None
    if opt.enable_2lmc:
        t0 = time.perf_counter()
        coarse, cb = ns.nc(cache, tp, opt.coarse_walks_per_point, opt.depth_m, stream_id=3, seed=opt.seed + 3,
                           antithetic=bool(opt.antithetic))
        coarse_ms = sync_ms(t0)
        t0 = time.perf_counter()
        pb = ns.walks(tp, opt.residual_walks_per_point, N.NC_COUPLED, opt.depth_m, 4, opt.seed + 4)
        resid = ns.combine(pb, cache.infer(pb.xm), len(pts), opt.residual_walks_per_point, N.NC_COUPLED)
        resid_ms = sync_ms(t0)
        cm, cv = _mv(coarse)
        rm, rv = _mv(resid)
        two = (cm, cv, rm, rv)
        est = cm + rm
        err = est - exact
        cs, rs = len(pts) * opt.coarse_walks_per_point, len(pts) * opt.residual_walks_per_point
        tl = {
            "eval_points": len(pts), "coarse_walks_per_point": opt.coarse_walks_per_point,
            "residual_walks_per_point": opt.residual_walks_per_point, "coarse_samples": cs, "residual_samples": rs,
            "rmse": float(np.sqrt(np.mean(err ** 2))), "mae": float(np.mean(np.abs(err))),
            "max_abs_error": float(np.max(np.abs(err))), "mean_estimate": float(est.mean()),
            "mean_exact": float(exact.mean()), "mean_bias": float(est.mean() - exact.mean()),
            "mean_coarse": float(cm.mean()), "mean_residual": float(rm.mean()),
            "mean_coarse_sample_variance": float(cv.mean()), "mean_residual_sample_variance": float(rv.mean()),
            "mean_coarse_steps": float(cb.steps.sum().item()) / cs, "mean_residual_steps": float(pb.steps.sum().item()) / rs,
            "coarse_cache_queries": int(cb.need.sum().item()), "residual_cache_queries": int(pb.need.sum().item()),
            "coarse_cache_query_fraction": float(cb.need.sum().item()) / cs,
            "residual_cache_query_fraction": float(pb.need.sum().item()) / rs,
            "coarse_forced_max_steps": 0, "residual_forced_max_steps": 0,
            "coarse_overflow_count": 0, "residual_overflow_count": 0,
            "coarse_elapsed_ms": coarse_ms, "residual_elapsed_ms": resid_ms, "elapsed_ms": coarse_ms + resid_ms,
            "training_plus_elapsed_ms": training_ms + coarse_ms + resid_ms,
        }
        runs["nc_2lmc"] = tl
        runs[f"nc_2lmc_m{opt.depth_m}"] = tl
    if opt.save_estimates_prefix:
        write_estimates_csv(opt.save_estimates_prefix + "_estimates.csv", pts, exact, (pm, pv), (nm, nv), two)

    pr, nr = runs["pure_wos"]["rmse"], runs["nc_wos"]["rmse"]
    comp = {"nc_wos_rmse_div_pure_wos_rmse": nr / pr if pr > 0 else 0.0}
    if two is not None:
        tl = runs["nc_2lmc"]
        comp.update({
            "nc_2lmc_rmse_div_pure_wos_rmse": tl["rmse"] / pr if pr > 0 else 0.0,
            "nc_2lmc_rmse_div_nc_wos_rmse": tl["rmse"] / nr if nr > 0 else 0.0,
            "nc_wos_mean_bias": runs["nc_wos"]["mean_bias"], "nc_2lmc_mean_bias": tl["mean_bias"],
            "pure_elapsed_div_nc_2lmc_inference_elapsed": pure_ms / tl["elapsed_ms"] if tl["elapsed_ms"] else 0.0,
            "pure_elapsed_div_nc_2lmc_total_with_training":
                pure_ms / (tl["elapsed_ms"] + training_ms) if tl["elapsed_ms"] + training_ms else 0.0,
        })
    comp.update({"pure_elapsed_div_nc_inference_elapsed": pure_ms / nc_ms if nc_ms else 0.0,
                 "pure_elapsed_div_nc_total_with_training": pure_ms / (nc_ms + training_ms),
                 "training_cost_counted_in_total_speedup": True})

    dev = torch.cuda.get_device_properties(0)
    doc = {
        "schema": SCHEMA,
        "patch": "0009-add-slice-eval-sampler",
        "generated_at_utc": datetime.datetime.now(datetime.timezone.utc).strftime("%Y-%m-%dT%H:%M:%SZ"),
        "command_line": " ".join(shlex.quote(a) for a in ([sys.argv[0]] + list(argv if argv is not None else sys.argv[1:]))),
        "cuda": {"device": dev.name, "runtime": "hip " + str(torch.version.hip), "gcn_arch": getattr(dev, "gcnArchName", "")},
        "implementation_mode": {
            "solver": "neural_cache_wos_and_nc_2lmc_persistent_walker",
            "geometry_backend": "hip_bvh4_gfx950_" + bst["build_method"],
            "cache_backend": "hip_hashgrid_mlp_mfma_gfx950",
            "training_schedule": "fixed_refresh_count_and_fixed_train_steps",
            "training_labels": opt.label_source,
            "rng": "philox4x32_10",
            "host_transfer_between_prefix_and_tcnn": False,
            "host_transfer_between_tcnn_and_residual_consumer": False,
            "csv_postprocess": False,
            "python_bindings": True,
            "nc_wos_biased": True,
            "nc_2lmc_correction_enabled": bool(opt.enable_2lmc),
            "nc_and_2lmc_share_depth_m": True,
            "nc_antithetic_first_step": bool(opt.antithetic),
            "point_sampler": {"slice": "fixed_slice_interior_pixels",
                              "train_points": "training_points_reused_for_eval_sanity"}.get(
                                  opt.eval_mode, "inscribed_ball_screening_sampler"),
            "exact_values": "fp64_analytic",
            "production_status": "screening_diagnostic_not_final_wall_clock",
        },
        "options": {
            "mesh": opt.mesh, "mesh_path": opt.mesh_path, "boundary_condition": opt.bc,
            "label_source": opt.label_source, "train_points_requested": opt.train_points,
            "train_points_padded": len(train), "train_sampler": opt.train_sampler,
            "eval_points_requested": opt.eval_points, "eval_points": len(pts), "eval_mode": opt.eval_mode,
            "eval_use_train_points": bool(opt.eval_mode == "train_points" or opt.eval_use_train_points),
            "label_refreshes": opt.label_refreshes, "walks_per_label_refresh": opt.walks_per_label_refresh,
            "train_steps_per_refresh": opt.train_steps_per_refresh,
            "pure_walks_per_point": opt.pure_walks_per_point, "hybrid_walks_per_point": opt.hybrid_walks_per_point,
            "coarse_walks_per_point": opt.coarse_walks_per_point,
            "residual_walks_per_point": opt.residual_walks_per_point, "depth_m": opt.depth_m,
            "max_steps": opt.max_steps, "epsilon": opt.epsilon, "cache_preset": opt.cache_preset,
            "network": "hashgrid_relu_mlp_%dx%d_fp16_mfma" % (opt.n_hidden_layers, opt.n_neurons),
            "n_levels": opt.n_levels,
            "n_features_per_level": 2, "log2_hashmap_size": opt.log2_hashmap_size,
            "base_resolution": opt.base_resolution, "per_level_scale": opt.per_level_scale,
            "n_neurons": opt.n_neurons, "n_hidden_layers": opt.n_hidden_layers, "jit_requested": bool(opt.jit),
            "jit_supported": False,
            "jit_enabled": False,
        },
    }
    if slice_info is not None:
        doc["slice_eval"] = slice_info
    if opt.save_estimates_prefix:
        doc["estimate_outputs"] = {"estimates_csv": opt.save_estimates_prefix + "_estimates.csv"}
    doc["training_point_sampler"] = {
        "mode": opt.train_sampler, "requested_points": opt.train_points, "padded_points": len(train),
        "unique_points": len(train), "inside_test": "hip_bvh_ray_parity_majority_of_3",
        "padding_reuses_accepted_points": False,
        "train_points_csv": opt.save_train_points_prefix + "_train_points.csv" if opt.save_train_points_prefix else "",
    }
    doc["mesh_stats"] = {
        "name": mesh_name, "vertices": int(len(v)), "triangles": int(len(t)),
        "degenerate_triangles": int(G.count_degenerate_triangles(v, t)),
        "bounds_min": [float(x) for x in lo], "bounds_max": [float(x) for x in hi],
        "normalization": {"center": [float(x) for x in center], "scale": float(scale)},
    }
    doc["bvh_stats"] = {"triangles": bst["triangle_count"], "nodes": bst["node_count"],
                        "prim_ids": bst["index_count"], "leaf_size": bst["leaf_size"],
                        "build_method": bst["build_method"], "build_milliseconds": bst["build_milliseconds"]}
    doc["training"] = {"label_update_ms": label_ms, "tcnn_training_ms": train_ms, "total_training_ms": training_ms,
                       "host_readback_in_training_loop": False}
    doc["runs"] = runs
    doc["comparison"] = comp
    os.makedirs(os.path.dirname(opt.output) or ".", exist_ok=True)
    with open(opt.output, "w") as f:
        json.dump(doc, f, indent=2)
    print(f"wrote {opt.output}")
    print(f"pure RMSE {pr} in {pure_ms} ms")
    print(f"NC+WoS RMSE {nr} in {nc_ms} ms")
    if two is not None:
        print(f"NC+2LMC RMSE {runs['nc_2lmc']['rmse']} in {runs['nc_2lmc']['elapsed_ms']} ms")
    return 0
