"""Progressive online cache training with snapshot overlap: NC + 2LMC
evaluated in blocks while the cache keeps training (the two-cache snapshot
discipline of docs/progressive_snapshot_accounting_policy.md:28-42 in the
reference, which describes it and measures only its accounting model;
docs/unfinished_cache_policy.md:34-48).

Two caches of one layout:
  train   mutable, takes Adam steps toward theta_{k+1} on the training stream;
  infer   the frozen snapshot theta_k that estimator block k reads.
Block k runs its NC + 2LMC samples (coarse NC samples and coupled residual
pairs, merged into per-point FP64 statistics) on the estimator stream with
theta_k while the training stream generates label walks and takes Adam steps.
At the block boundary both streams join and infer <- train (device copy).

Every block's coarse + residual estimate is unbiased whatever its snapshot,
so the merged estimate is unbiased.  Residual walks are never used as
labels: labels come from their own coupled walks (own Philox stream), as the
reference policy requires.  Given its snapshots, the estimate is a fixed
function of the seeds (each block's walks and merges run in a fixed order),
so the overlap cannot change it: replay() recomputes it from recorded
snapshots on one stream, bit for bit.  The snapshots themselves vary in
their last bits from run to run, because the grid gradients are accumulated
with FP32 atomics.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from . import cache as K, nc as N, wos as W


@dataclass
class ProgressiveConfig:
    blocks: int = 4                 # estimator blocks (snapshots theta_0 .. theta_{K-1})
    warmup_steps: int = 50          # Adam steps before block 0 (theta_0), not overlapped
    steps_per_block: int = 50       # Adam steps per block, overlapped with the estimator
    batch: int = 1 << 15
    label_pairs: int = 16           # coupled label walks per labelled point, per label round
    label_stride: int = 4           # label every k-th evaluation point
    coarse_wpp: int = 32            # coarse NC samples per point per block
    residual_wpp: int = 16          # coupled residual pairs per point per block
    depth: int = 1
    antithetic: bool = True
    normalize_targets: bool = True
    # share of the walker's full grid the estimator may take when overlapping
    # (measured on the benchmark problem: 1.0 -> 1.15x over one stream, 0.75 ->
    # 1.09x, 0.875 -> 0.95x; the training kernels fill CUs as walker blocks end)
    estimator_cu_fraction: float = 1.0
    seed: int = 101


class ProgressiveTwoLevel:
    def __init__(self, bvh, boundary: W.Boundary, params: W.WalkParams, cache_config: K.CacheConfig,
                 cfg: ProgressiveConfig | None = None):
        self.bvh = bvh
        self.boundary = boundary
        self.params = params
        self.cfg = cfg or ProgressiveConfig()
        self.train_cache = K.Cache(cache_config)
        self.infer_cache = K.Cache(cache_config)

    def _labels(self, ns, tp, rnd: int):
        c = self.cfg
        lp = tp[:: c.label_stride].contiguous() if c.label_stride > 1 else tp
        b = ns.walks(lp, c.label_pairs, N.NC_COUPLED, c.depth, stream_id=7, seed=c.seed + 7919 * rnd,
                     antithetic=c.antithetic)
        return ns.training_pairs(b)

    def _adam(self, x, y, steps: int, gen):
        import torch
        n = x.shape[0]
        for _ in range(steps):
            idx = torch.randint(0, n, (min(self.cfg.batch, n),), device="cuda", generator=gen)
            self.train_cache.train_step(x[idx].contiguous(), y[idx].contiguous())

    def _estimate_block(self, ns, points, k: int, coarse, resid):
        c = self.cfg
        n = points.shape[0]
        cb = ns.walks(points, c.coarse_wpp, N.NC_PREFIX, c.depth, stream_id=1, seed=c.seed + 31 * k,
                      antithetic=c.antithetic)
        ns.combine(cb, self.infer_cache.infer(cb.xm), n, c.coarse_wpp, N.NC_PREFIX, stats=coarse, accumulate=k > 0)
        pb = ns.walks(points, c.residual_wpp, N.NC_COUPLED, c.depth, stream_id=2, seed=c.seed + 31 * k)
        ns.combine(pb, self.infer_cache.infer(pb.xm), n, c.residual_wpp, N.NC_COUPLED, stats=resid, accumulate=k > 0)

    def replay(self, points, snapshots):
        """The estimator blocks alone, on the current stream, with recorded
        snapshots (grid, weights, offset) in place of the training."""
        n = points.shape[0]
        coarse = W.PointStats.zeros(n, points.device)
        resid = W.PointStats.zeros(n, points.device)
        ns = N.NcSolver(self.bvh, self.boundary, self.params)
        for k, (grid, weights, off) in enumerate(snapshots):
            self.infer_cache.set_params(grid, weights)
            self.infer_cache.set_output_offset(off)
            self._estimate_block(ns, points, k, coarse, resid)
        return coarse.mean + resid.mean, coarse, resid

    def run(self, points, overlap: bool = True, record_snapshots: bool = False):
        """Returns (estimate per point, coarse stats, residual stats, info);
        with record_snapshots, info["snapshots"] holds each block's snapshot
        (host copies taken between blocks, outside the overlapped work)."""
        import torch
        c = self.cfg
        n = points.shape[0]
        dev = points.device
        main = torch.cuda.current_stream()
        s_est = torch.cuda.Stream() if overlap else main
        s_train = torch.cuda.Stream() if overlap else main
        est_params = W.WalkParams(**{**self.params.__dict__})
        if overlap and c.estimator_cu_fraction < 1.0:
            full = W.full_grid_blocks(self.params.source)
            est_params.max_blocks = max(1, int(c.estimator_cu_fraction * full))
        ns_est = N.NcSolver(self.bvh, self.boundary, est_params)
        ns_train = N.NcSolver(self.bvh, self.boundary, self.params)
        coarse = W.PointStats.zeros(n, dev)
        resid = W.PointStats.zeros(n, dev)
        gen = torch.Generator(device="cuda").manual_seed(c.seed)
        info = {"blocks": [], "overlap": overlap}
        torch.cuda.synchronize()
        t_all = time.perf_counter()

        # warm-up: labels, target offset and theta_0
        with torch.cuda.stream(s_train):
            x, y = self._labels(ns_train, points, 0)
            if c.normalize_targets:
                self.train_cache.set_output_offset(float(y.mean().item()))
            self._adam(x, y, c.warmup_steps, gen)
            self.infer_cache.copy_from(self.train_cache)
        torch.cuda.synchronize()
        info["warmup_seconds"] = time.perf_counter() - t_all
        snaps = []
        paused = 0.0

        def record():
            if record_snapshots:
                t = time.perf_counter()
                g, w = self.infer_cache.get_params()
                snaps.append((g, w, self.infer_cache.output_offset))
                return time.perf_counter() - t
            return 0.0

        paused += record()
        for k in range(c.blocks):
            t0 = time.perf_counter()
            # estimator block k with the frozen snapshot theta_k
            with torch.cuda.stream(s_est):
                self._estimate_block(ns_est, points, k, coarse, resid)
            # meanwhile: the training stream moves theta toward theta_{k+1}
            if k + 1 < c.blocks:
                with torch.cuda.stream(s_train):
                    x, y = self._labels(ns_train, points, k + 1)
                    self._adam(x, y, c.steps_per_block, gen)
                    ev_train = torch.cuda.Event()
                    ev_train.record()
                # block boundary: the snapshot is replaced after block k's
                # estimator work (same stream) and theta_{k+1}'s last step
                with torch.cuda.stream(s_est):
                    s_est.wait_event(ev_train)
                    self.infer_cache.copy_from(self.train_cache)
            # both streams drain before the next block starts
            torch.cuda.synchronize()
            info["blocks"].append({"block": k, "seconds": time.perf_counter() - t0})
            if k + 1 < c.blocks:
                paused += record()
        torch.cuda.synchronize()
        info["seconds"] = time.perf_counter() - t_all - paused
        if record_snapshots:
            info["snapshots"] = snaps
        est = coarse.mean + resid.mean
        return est, coarse, resid, info


def progressive_vs_serial(bvh, verts, points_np, boundary, truth, params, cfg: ProgressiveConfig | None = None,
                          cache_kw: dict | None = None, log=print) -> dict:
    """Runs the same progressive schedule overlapped and on one stream; reports
    both wall times and the RMSE of each against the FP64 solution."""
    import torch
    cfg = cfg or ProgressiveConfig()
    tp = torch.from_numpy(np.ascontiguousarray(points_np, np.float32)).cuda()
    out = {"config": dict(cfg.__dict__)}
    for overlap in (False, True):
        kcfg = K.CacheConfig.for_bounds(verts.min(0), verts.max(0), seed=cfg.seed, **(cache_kw or {"n_levels": 2}))
        pr = ProgressiveTwoLevel(bvh, boundary, params, kcfg, cfg)
        est, co, re, info = pr.run(tp, overlap=overlap)
        e = est.cpu().numpy()
        r = float(np.sqrt(np.mean((e - truth) ** 2)))
        key = "overlapped" if overlap else "serial"
        out[key] = {"seconds": info["seconds"], "warmup_seconds": info["warmup_seconds"], "rmse": r,
                    "blocks": info["blocks"], "mean_residual": float(re.mean.mean().item())}
        log(f"progressive 2lmc {key}: {info['seconds']:.3f} s, rmse {r:.4g}")
    out["overlap_speedup"] = out["serial"]["seconds"] / out["overlapped"]["seconds"]
    return out
