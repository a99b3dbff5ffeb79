"""Time-to-RMSE of plain WoS, neural-cache WoS (NC) and NC + 2LMC on one
problem (the harness formula of the reference: time-to-threshold = the least
wall time over runs whose RMSE against the exact solution is <= tau,
scripts/run_nc_slice_time_to_mse.py:213-246).

Cache training is online and counted in the NC / 2LMC wall time: coupled walks
from the evaluation points (their own Philox stream, never reused for an
estimate) give (X_m, W - S_m) pairs, unbiased labels of u(X_m); the cache
takes Adam steps on random batches of them, then is frozen for the estimate.

Several cache setups can be compared in one run (depth m, grid levels,
labels per point, Adam steps); each reports its own training time, the cache
error at held-out walk positions X_m, and the NC and NC + 2LMC runs.  The
control-variate setups (cv.py: the cache subtracted along the whole walk,
unbiased for any cache) report their pipeline time and estimates the same way.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np


def _rmse(mean, truth):
    return float(np.sqrt(np.mean((mean - truth) ** 2)))


@dataclass
class CacheSetup:
    name: str
    depth: int = 4
    cache_kw: dict = field(default_factory=dict)  # CacheConfig overrides (n_levels, ...)
    train_pairs: int = 32                          # coupled label walks per labelled point
    label_stride: int = 1                          # label every k-th evaluation point
    train_steps: int = 200
    batch: int = 1 << 18
    normalize_targets: bool = False                # cache output offset = mean label
    antithetic: bool = False                       # mirrored first steps (labels and NC samples)
    nc_wpp: tuple = (64, 128, 256)
    lmc: tuple = ((256, 16), (512, 16))            # (coarse samples, coupled pairs) per point


# The learned-cache configuration of the benchmark problem (16 levels x 2
# features, T = 2^19, N_min 16, growth 1.5; 4 x 64 ReLU MLP; depth 4; 32
# label walks per point; 16 coupled pairs per point for 2LMC), and shallow
# setups found on the same problem: depth 1, a 2-level grid (coarse grids
# generalise best from single-walk labels), targets normalised by their mean,
# 16 antithetic label pairs on every 8th (or 4th) point, 60 (100) Adam steps
# at batch 2^15, and NC samples in antithetic pairs (the pair mean cancels the
# odd part of the cache around each point, so few samples reach the cache's
# bias floor).  The 2LMC run of the every-4th-point setup with 384 coupled
# pairs per point reaches RMSE 1e-2 without bias, at about the cost of plain
# WoS.
DEFAULT_SETUPS = (
    # (1024, 384): the optimal allocation n_c / n_r = sqrt(V_c c_r / (V_r c_c))
    # (docs/wos_wavefront_policy.md:66-83) with V_c 0.031, V_r 0.026 and a
    # depth-4 prefix at about 1/6 of a full walk's cost: about 2.7
    CacheSetup("depth4_levels16", depth=4, train_steps=200, nc_wpp=(64, 256), lmc=((256, 16), (1024, 384))),
    CacheSetup("depth1_levels2_anti", depth=1, cache_kw={"n_levels": 2}, train_pairs=16, label_stride=8,
               train_steps=60, batch=1 << 15, normalize_targets=True, antithetic=True, nc_wpp=(24, 32), lmc=()),
    CacheSetup("depth1_levels2_anti_more_labels", depth=1, cache_kw={"n_levels": 2}, train_pairs=16,
               label_stride=4, train_steps=100, batch=1 << 15, normalize_targets=True, antithetic=True,
               nc_wpp=(32, 64), lmc=((64, 64), (64, 384))),
    CacheSetup("depth1_levels4", depth=1, cache_kw={"n_levels": 4}, train_steps=100, batch=1 << 16,
               nc_wpp=(96, 128, 192, 256), lmc=()),
)


@dataclass
class CvSetup:
    """The learned cache as an along-walk control variate (unbiased): labels
    from the path records of one walk per labelled point, Adam steps, a bake
    of the cache onto a grid, then `rounds` bootstrap rounds (labels from
    control-variate walks with the current field, which are far less noisy,
    and more Adam steps), then the estimate with wpp walks per point."""
    name: str
    cache_kw: dict = field(default_factory=dict)
    label_stride: int = 2
    path_len: int = 32
    rounds: int = 1
    train_steps: tuple = (40, 40)  # Adam steps per round (round 0 = plain labels)
    batch: int = 1 << 15
    res: int = 64                  # grid cells per axis of the baked field
    points: int = 6                # design: 2 (antipodal) or 6 (turned octahedron)
    wpp: tuple = (1, 2, 4)
    label_max_steps: int | None = None  # step cap of the label walks only


# A 3-level dense grid (N_min 8, growth 2) with the 4 x 64 MLP; labels on every
# second grid point from walks capped at 48 steps (the label solves are then
# shorter; the estimator stays unbiased for any field); one bootstrap round
# with 60 Adam steps.  Measured on the benchmark problem: RMSE 0.0094 with 2
# walks per point in 0.055 s, training included.  The uncapped pipeline with
# 40 bootstrap steps follows for comparison (RMSE 0.0105 at 2 walks per point,
# so it needs 4).
DEFAULT_CV_SETUPS = (
    CvSetup("cv_levels3_boot1_cap48", cache_kw={"n_levels": 3, "base_resolution": 8, "growth": 2.0},
            label_max_steps=48, train_steps=(40, 60), wpp=(2, 4)),
    CvSetup("cv_levels3_boot1", cache_kw={"n_levels": 3, "base_resolution": 8, "growth": 2.0}),
    CvSetup("cv_levels3_boot2", cache_kw={"n_levels": 3, "base_resolution": 8, "growth": 2.0}, rounds=2,
            train_steps=(40, 40, 40), wpp=(1, 2)),
    # the antipodal pair instead of the octahedron, for comparison
    CvSetup("cv_levels3_boot1_pair", cache_kw={"n_levels": 3, "base_resolution": 8, "growth": 2.0}, points=2,
            wpp=(2, 8)),
)


def run_cv(bvh, verts, points, boundary, truth, params, plain_time, tau=1e-2, setups=DEFAULT_CV_SETUPS, seed=11,
           log=print):
    """Time-to-RMSE of the control-variate estimator: each setup's training
    pipeline is timed once; each estimate adds its own solve time."""
    import torch

    from . import cv as CV
    tp = torch.from_numpy(points).cuda()
    cs = CV.CvSolver(bvh, boundary, params)
    cs.solve(tp[:1024].contiguous(), 1, None, path_len=4)  # warm the extended walker
    # warm every kernel of the pipeline (first launches load code objects):
    # a tiny run on 2,048 points, not reused
    su0 = setups[0]
    CV.train_field(bvh, verts, boundary, params, tp[:2048].contiguous(), su0.cache_kw, 1, su0.path_len, 1, (2, 2),
                   1024, 8, su0.points, seed + 7, solver=cs)
    torch.cuda.synchronize()
    out = []
    best = None
    for su in setups:
        res = {"name": su.name, "cache": dict(su.cache_kw), "label_stride": su.label_stride,
               "path_len": su.path_len, "rounds": su.rounds, "train_steps": list(su.train_steps),
               "batch": su.batch, "field_res": su.res, "design_points": su.points,
               "label_max_steps": su.label_max_steps}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fld, cache, nlab = CV.train_field(bvh, verts, boundary, params, tp, su.cache_kw, su.label_stride,
                                          su.path_len, su.rounds, su.train_steps, su.batch, su.res, su.points,
                                          seed, solver=cs, label_max_steps=su.label_max_steps)
        torch.cuda.synchronize()
        t_train = time.perf_counter() - t0
        res["train_seconds"] = t_train
        res["labels_per_round"] = nlab
        log(f"[{su.name}] pipeline (labels, {su.rounds} bootstrap rounds, training, bake) {t_train:.4f}s")
        runs = []
        for wpp in su.wpp:
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            st = cs.solve(tp, wpp, fld, seed=seed + 1000 + wpp)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t1
            mean = st.mean.cpu().numpy()
            var = float(st.variance().mean().item()) if wpp > 1 else None
            r_ = _rmse(mean, truth)
            runs.append({"wpp": wpp, "seconds": dt, "seconds_with_training": dt + t_train, "rmse": r_,
                         "mean_walk_variance": var, "mean_error": float(np.mean(mean - truth))})
            log(f"[{su.name}] cv wpp={wpp} t={dt:.4f}s (+train {t_train:.4f}s) rmse={r_:.4g} var={var}")
        res["runs"] = runs
        ok = [p["seconds_with_training"] for p in runs if p["rmse"] <= tau]
        v = min(ok) if ok else None
        res["time_to_rmse"] = v
        res["speedup_vs_plain"] = (plain_time / v) if (plain_time and v) else None
        if v is not None and (best is None or v < best[0]):
            best = (v, su.name)
        out.append(res)
        del cache
    return {"setups": out, "best_time_to_rmse": best[0] if best else None,
            "best_setup": best[1] if best else None,
            "best_speedup_vs_plain": (plain_time / best[0]) if (plain_time and best) else None}


def run(bvh, verts, points, boundary, truth, params, tau=1e-2, plain_wpp=(256, 320, 384, 448, 512),
        setups=DEFAULT_SETUPS, seed=11, log=print, cv_setups=DEFAULT_CV_SETUPS):
    import torch

    from . import cache as K, nc as N, wos as W
    tp = torch.from_numpy(points).cuda()
    out = {"tau": tau, "points": int(len(points))}

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t0

    # ---------------------------------------------------------------- plain
    solver = W.Solver(bvh, boundary, params)
    solver.solve(tp, 16)
    plain = []
    for wpp in plain_wpp:
        st, dt = timed(lambda: solver.solve(tp, wpp, seed=seed + wpp))
        r = _rmse(st.mean.cpu().numpy(), truth)
        plain.append({"wpp": wpp, "seconds": dt, "rmse": r})
        log(f"plain wpp={wpp} t={dt:.3f}s rmse={r:.4g}")
    out["plain"] = plain
    ok = [p["seconds"] for p in plain if p["rmse"] <= tau]
    pt = min(ok) if ok else None
    out["plain_time_to_rmse"] = pt

    ns = N.NcSolver(bvh, boundary, params)
    best = {}
    out["setups"] = []
    for su in setups:
        res = {"name": su.name, "depth": su.depth, "cache": dict(su.cache_kw), "train_pairs": su.train_pairs,
               "label_stride": su.label_stride, "train_steps": su.train_steps, "batch": su.batch,
               "normalize_targets": su.normalize_targets, "antithetic": su.antithetic}
        cfg = K.CacheConfig.for_bounds(verts.min(0), verts.max(0), seed=seed, **su.cache_kw)
        cache = K.Cache(cfg)

        def train(cache=cache):
            lp = tp[:: su.label_stride].contiguous() if su.label_stride > 1 else tp
            b = ns.walks(lp, su.train_pairs, N.NC_COUPLED, su.depth, stream_id=7, seed=seed,
                         antithetic=su.antithetic)
            x, y = ns.training_pairs(b)
            if su.normalize_targets:
                off = float(y.mean().item())
                cache.set_output_offset(off)
            g = torch.Generator(device="cuda").manual_seed(seed)
            n = x.shape[0]
            for _ in range(su.train_steps):
                idx = torch.randint(0, n, (min(su.batch, n),), device="cuda", generator=g)
                cache.train_step(x[idx].contiguous(), y[idx].contiguous())
            return n

        ntrain, t_train = timed(train)
        res["train_label_walks"] = int(ntrain)
        res["train_seconds"] = t_train
        log(f"[{su.name}] train labels={ntrain} steps={su.train_steps} t={t_train:.3f}s")

        # cache quality at held-out walk positions X_m (exact u there)
        b = ns.walks(tp[:: max(1, len(points) // 4096)].contiguous(), 4, N.NC_PREFIX, su.depth, stream_id=9,
                     seed=seed)
        sel = b.need.bool()
        xq = b.xm[sel].contiguous()
        if xq.shape[0] > 0 and boundary.charges is not None:
            from . import problems as P
            cq = cache.infer(xq).cpu().numpy()
            res["cache_rmse_at_xm"] = _rmse(cq, P.charges_solution_f64(boundary.charges, xq.cpu().numpy()))
            log(f"[{su.name}] cache rmse at X_m = {res['cache_rmse_at_xm']:.4g}")

        ncr = []
        for wpp in su.nc_wpp:
            (st, _), dt = timed(lambda: ns.nc(cache, tp, wpp, su.depth, stream_id=1, seed=seed + wpp,
                                              antithetic=su.antithetic))
            r = _rmse(st.mean.cpu().numpy(), truth)
            ncr.append({"wpp": wpp, "seconds": dt, "seconds_with_training": dt + t_train, "rmse": r})
            log(f"[{su.name}] nc wpp={wpp} t={dt:.3f}s (+train {t_train:.3f}s) rmse={r:.4g}")
        res["nc"] = ncr
        # pipelined: the estimate's prefix walks do not depend on the cache,
        # so they run on a second stream while the labels are walked and the
        # cache trains (they fill the CUs the label walks' tail leaves idle);
        # only the inference and the combine wait for the training.  Same
        # walks and training as above, timed as one pipeline.
        ok_wpp = [r["wpp"] for r in ncr if r["rmse"] <= tau]
        if ok_wpp:
            wpp = min(ok_wpp)
            cache_p = K.Cache(cfg)
            ns_est = N.NcSolver(bvh, boundary, params)
            s_est = torch.cuda.Stream()
            main = torch.cuda.current_stream()

            def piped():
                s_est.wait_stream(main)
                with torch.cuda.stream(s_est):
                    bp = ns_est.walks(tp, wpp, N.NC_PREFIX, su.depth, stream_id=1, seed=seed + wpp,
                                      antithetic=su.antithetic)
                train(cache_p)
                main.wait_stream(s_est)
                return ns.combine(bp, cache_p.infer(bp.xm), len(points), wpp, N.NC_PREFIX)

            st, dt = timed(piped)
            r = _rmse(st.mean.cpu().numpy(), truth)
            res["nc_pipelined"] = {"wpp": wpp, "seconds_with_training": dt, "rmse": r}
            log(f"[{su.name}] nc pipelined wpp={wpp} t={dt:.3f}s (training included) rmse={r:.4g}")
            del cache_p
        tl = []
        for nc_, nr in su.lmc:
            (r2, dt) = timed(lambda: ns.two_level(cache, tp, nc_, nr, su.depth, seed=seed + nc_ + 1000 * nr,
                                                  antithetic=su.antithetic))
            est, co, re, _ = r2
            r = _rmse(est.cpu().numpy(), truth)
            vc = float((co.m2 / (co.n - 1)).mean().item())
            vr = float((re.m2 / (re.n - 1)).mean().item())
            tl.append({"coarse_wpp": nc_, "pairs": nr, "seconds": dt, "seconds_with_training": dt + t_train,
                       "rmse": r, "coarse_var": vc, "residual_var": vr})
            log(f"[{su.name}] 2lmc coarse={nc_} pairs={nr} t={dt:.3f}s (+train) rmse={r:.4g} Vc={vc:.4g} "
                f"Vr={vr:.4g}")
        res["nc_2lmc"] = tl
        for k, runs in (("nc", ncr + ([res["nc_pipelined"]] if "nc_pipelined" in res else [])), ("nc_2lmc", tl)):
            ok = [p["seconds_with_training"] for p in runs if p["rmse"] <= tau]
            v = min(ok) if ok else None
            res[f"{k}_time_to_rmse"] = v
            res[f"{k}_speedup_vs_plain"] = (pt / v) if (pt and v) else None
            if v is not None and (k not in best or v < best[k][0]):
                best[k] = (v, su.name)
        out["setups"].append(res)
        del cache
    for k in ("nc", "nc_2lmc"):
        v = best.get(k)
        out[f"best_{k}_time_to_rmse"] = v[0] if v else None
        out[f"best_{k}_setup"] = v[1] if v else None
        out[f"best_{k}_speedup_vs_plain"] = (pt / v[0]) if (pt and v) else None
    if cv_setups:
        # the learned cache as an along-walk control variate: unbiased
        cvr = run_cv(bvh, verts, points, boundary, truth, params, pt, tau, cv_setups, seed, log)
        out["cv"] = cvr
        out["best_unbiased_time_to_rmse"] = cvr["best_time_to_rmse"]
        out["best_unbiased_setup"] = cvr["best_setup"]
        out["best_unbiased_speedup_vs_plain"] = cvr["best_speedup_vs_plain"]
    return out
