"""Neural-cache Walk-on-Spheres (NC) and the two-level Monte Carlo estimator
(NC + 2LMC) over the C ABI, with online cache training on the solver's own
walks.

Estimators (per evaluation point x, depth m):
  NC (biased):   v = S_m + C(X_m)  (S_m: prefix source; walks that reach the
                 boundary before depth m return their boundary value)
  2LMC:          u(x) ~ mean_coarse[S_m + C(X_m)] + mean_pairs[W - S_m - C(X_m)]
                 where each pair shares its first m steps (same Philox stream)
                 and the two levels use independent streams, as the reference
                 does (src/tools/eval_tcnn_nc_wos.cu:767-783).  Unbiased for
                 any fixed cache C.
Walk values and prefixes come from the persistent walker; C is evaluated on
MFMA; the combine step reduces per-point FP64 statistics on the device.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import check, lib
from .wos import Boundary, PointStats, WalkParams

NC_PREFIX = 1
NC_COUPLED = 2
NC_ANTITHETIC = 0x100  # OR-ed into the mode: mirrored first steps in walk pairs (2j, 2j+1)


@dataclass
class WalkBatch:
    prefix: object
    xm: object
    need: object
    full: object
    steps: object


class NcSolver:
    def __init__(self, bvh, boundary: Boundary, params: WalkParams | None = None):
        self.bvh = bvh
        self.boundary = boundary
        self.params = params or WalkParams()
        self._bc = boundary.to_c()
        self._ws = None

    def _workspace(self, n, wpp):
        import torch
        need = int(lib().n2wos_nc_workspace_size(n, wpp))
        if need < 0:
            check(need)
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(need, dtype=torch.uint8, device="cuda")
        return self._ws

    def walks(self, points, wpp: int, mode: int, depth: int, stream_id: int, seed: int | None = None,
              point_offset: int = 0, antithetic: bool = False) -> WalkBatch:
        import torch
        if antithetic:
            mode |= NC_ANTITHETIC
        n = points.shape[0]
        dev = points.device
        total = n * wpp
        prefix = torch.empty(total, dtype=torch.float32, device=dev)
        xm = torch.empty((total, 3), dtype=torch.float32, device=dev)
        need = torch.empty(total, dtype=torch.uint8, device=dev)
        full = torch.empty(total, dtype=torch.float32, device=dev) if (mode & ~NC_ANTITHETIC) == NC_COUPLED else None
        steps = torch.empty(n, dtype=torch.int64, device=dev)
        wp = self.params.to_c()
        wp.rng_stream = stream_id
        if seed is not None:
            wp.seed = seed
        ws = self._workspace(n, wpp)
        check(lib().n2wos_nc_walks(self.bvh.handle, C.byref(self._bc), C.byref(wp), points.data_ptr(), n, wpp,
                                   point_offset, mode, depth, prefix.data_ptr(), xm.data_ptr(), need.data_ptr(),
                                   full.data_ptr() if full is not None else None, steps.data_ptr(), ws.data_ptr(),
                                   ws.numel(), torch.cuda.current_stream().cuda_stream))
        return WalkBatch(prefix, xm, need, full, steps)

    @staticmethod
    def combine(batch: WalkBatch, cache_values, n_points: int, wpp: int, mode: int, stats: PointStats | None = None,
                accumulate: bool = False) -> PointStats:
        import torch
        if stats is None:
            stats = PointStats.zeros(n_points, batch.prefix.device)
        check(lib().n2wos_nc_combine(n_points, wpp, mode, batch.prefix.data_ptr(), batch.need.data_ptr(),
                                     cache_values.data_ptr(), batch.full.data_ptr() if batch.full is not None else None,
                                     int(accumulate), stats.n.data_ptr(), stats.mean.data_ptr(), stats.m2.data_ptr(),
                                     torch.cuda.current_stream().cuda_stream))
        return stats

    # ------------------------------------------------------------ estimators
    def nc(self, cache, points, wpp: int, depth: int, stream_id: int = 1, seed: int | None = None,
           antithetic: bool = False):
        """Neural-cache (biased) estimate: per-point stats of S_m + C(X_m).
        With antithetic=True the walks come in pairs with mirrored first steps
        (same expectation; the pair mean cancels the odd part of C around x)."""
        b = self.walks(points, wpp, NC_PREFIX, depth, stream_id, seed, antithetic=antithetic)
        cv = cache.infer(b.xm)
        st = self.combine(b, cv, points.shape[0], wpp, NC_PREFIX)
        st.steps = b.steps
        return st, b

    def two_level(self, cache, points, n_coarse: int, n_pairs: int, depth: int, seed: int | None = None,
                  coarse_stream: int = 1, pair_stream: int = 2, antithetic: bool = False):
        """NC + 2LMC: returns (estimate per point, coarse stats, residual stats, pair batch).
        antithetic applies to the coarse level's walks."""
        coarse, _ = self.nc(cache, points, n_coarse, depth, coarse_stream, seed, antithetic=antithetic)
        pb = self.walks(points, n_pairs, NC_COUPLED, depth, pair_stream, seed)
        cv = cache.infer(pb.xm)
        resid = self.combine(pb, cv, points.shape[0], n_pairs, NC_COUPLED)
        resid.steps = pb.steps
        return coarse.mean + resid.mean, coarse, resid, pb

    @staticmethod
    def training_pairs(batch: WalkBatch):
        """(X_m, W - S_m) for the coupled walks that reached depth m: unbiased
        labels of the solution at X_m (the continuation's value)."""
        sel = batch.need.bool()
        return batch.xm[sel].contiguous(), (batch.full[sel] - batch.prefix[sel]).contiguous()
