"""Time-to-RMSE of plain WoS, neural-cache WoS (NC) and NC + 2LMC on one
problem (the harness formula of the reference: time-to-threshold = the least
wall time over runs whose RMSE against the exact solution is <= tau,
scripts/run_nc_slice_time_to_mse.py:213-246).

Cache training is online and counted in the NC / 2LMC wall time: coupled walks
from the evaluation points (their own Philox stream, never reused for an
estimate) give (X_m, W - S_m) pairs, unbiased labels of u(X_m); the cache
takes Adam steps on random batches of them, then is frozen for the estimate.
"""
from __future__ import annotations

import time

import numpy as np


def _rmse(mean, truth):
    return float(np.sqrt(np.mean((mean - truth) ** 2)))


def run(bvh, verts, points, boundary, truth, params, tau=1e-2, plain_wpp=(256, 320, 384, 448, 512),
        depth=4, train_pairs=4, train_steps=200, batch=1 << 18, nc_wpp=(32, 64, 128, 256),
        lmc=((128, 16), (192, 16), (256, 16), (256, 32), (384, 32), (512, 64)), cache_config=None, seed=11,
        log=print):
    import torch

    from . import cache as K, nc as N, wos as W
    tp = torch.from_numpy(points).cuda()
    out = {"tau": tau, "points": int(len(points)), "depth": depth}

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
    out["plain_time_to_rmse"] = min(ok) if ok else None

    # ---------------------------------------------------------------- train
    ns = N.NcSolver(bvh, boundary, params)
    cfg = cache_config or K.CacheConfig.for_bounds(verts.min(0), verts.max(0), seed=seed)
    cache = K.Cache(cfg)

    def train():
        b = ns.walks(tp, train_pairs, N.NC_COUPLED, depth, stream_id=7, seed=seed)
        x, y = ns.training_pairs(b)
        g = torch.Generator(device="cuda").manual_seed(seed)
        n = x.shape[0]
        for _ in range(train_steps):
            idx = torch.randint(0, n, (min(batch, n),), device="cuda", generator=g)
            cache.train_step(x[idx].contiguous(), y[idx].contiguous())
        return n

    ntrain, t_train = timed(train)
    out["train"] = {"pairs": int(ntrain), "steps": train_steps, "batch": batch, "seconds": t_train}
    log(f"train pairs={ntrain} steps={train_steps} t={t_train:.3f}s")

    # cache quality at walk positions: compare C(X_m) with the exact u(X_m)
    b = ns.walks(tp[:: max(1, len(points) // 4096)].contiguous(), 4, N.NC_PREFIX, depth, stream_id=9, seed=seed)
    sel = b.need.bool()
    xq = b.xm[sel].contiguous()
    if xq.shape[0] > 0:
        from . import problems as P
        cq = cache.infer(xq).cpu().numpy()
        uq = P.charges_solution_f64(boundary.charges, xq.cpu().numpy()) if boundary.charges is not None else None
        if uq is not None:
            out["cache_rmse_at_xm"] = _rmse(cq, uq)
            log(f"cache rmse at X_m = {out['cache_rmse_at_xm']:.4g}")

    # ---------------------------------------------------------------- NC
    ncr = []
    for wpp in nc_wpp:
        (st, _), dt = timed(lambda: ns.nc(cache, tp, wpp, depth, stream_id=1, seed=seed + wpp))
        r = _rmse(st.mean.cpu().numpy(), truth)
        ncr.append({"wpp": wpp, "seconds": dt, "seconds_with_training": dt + t_train, "rmse": r})
        log(f"nc wpp={wpp} t={dt:.3f}s (+train {t_train:.3f}s) rmse={r:.4g}")
    out["nc"] = ncr
    ok = [p["seconds_with_training"] for p in ncr if p["rmse"] <= tau]
    out["nc_time_to_rmse"] = min(ok) if ok else None

    # ---------------------------------------------------------------- 2LMC
    tl = []
    for nc_, nr in lmc:
        (res, dt) = timed(lambda: ns.two_level(cache, tp, nc_, nr, depth, seed=seed + nc_ + 1000 * nr))
        est, co, re, _ = res
        r = _rmse(est.cpu().numpy(), truth)
        vc = float((co.m2 / (co.n - 1)).mean().item())
        vr = float((re.m2 / (re.n - 1)).mean().item())
        tl.append({"coarse_wpp": nc_, "pairs": nr, "seconds": dt, "seconds_with_training": dt + t_train, "rmse": r,
                   "coarse_var": vc, "residual_var": vr})
        log(f"2lmc coarse={nc_} pairs={nr} t={dt:.3f}s (+train) rmse={r:.4g} Vc={vc:.4g} Vr={vr:.4g}")
    out["nc_2lmc"] = tl
    ok = [p["seconds_with_training"] for p in tl if p["rmse"] <= tau]
    out["nc_2lmc_time_to_rmse"] = min(ok) if ok else None
    pt = out["plain_time_to_rmse"]
    for k in ("nc", "nc_2lmc"):
        v = out[f"{k}_time_to_rmse"]
        out[f"{k}_speedup_vs_plain"] = (pt / v) if (pt and v) else None
    return out
