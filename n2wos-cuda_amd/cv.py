"""The learned cache as an along-walk control variate (unbiased).

Each Walk-on-Spheres step moves from X_k to X_k + R_k d_k, d_k uniform on the
sphere.  For a field C, the term
    t_k = C(X_k + R_k d_k) - (1/m) sum_j C(X_k + R_k s_j)
over a spherical design s_j (antipodal pair, m = 2, or the octahedron turned
by a uniform angle about d_k, m = 6) has zero mean whatever C is, so a walk's
value W - beta sum_k t_k is an unbiased estimate of the solution.  When C is
close to the solution u, sum_k t_k follows W's fluctuation step by step and
cancels most of it.

Here C is the trained hash-grid + MLP cache, baked onto a grid (its own
inference kernel evaluated at the grid vertices) and read by trilinear
interpolation inside the walker.  Training labels come from the walker's
path record: every visited position X_k of a walk, labelled with the value
the walk went on to collect from there (an unbiased sample of u(X_k)).  No
counterpart in the reference: its two-level estimator
(src/tools/eval_tcnn_nc_wos.cu:767-783) uses the cache at one depth only.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import check, lib
from .wos import Boundary, PointStats, WalkParams, workspace_bytes


class CvFieldC(C.Structure):
    _fields_ = [("d_values", C.c_void_p), ("res", C.c_int), ("points", C.c_int), ("origin", C.c_float * 3),
                ("cell_inverse", C.c_float * 3), ("beta", C.c_float)]


@dataclass
class CvField:
    """Field values at the (res+1)^3 vertices (x fastest) of the box with
    the given origin and cells per unit length, as a CUDA float32 tensor."""
    values: object
    res: int
    origin: tuple
    cell_inverse: tuple
    points: int = 6
    beta: float = 1.0

    def to_c(self) -> CvFieldC:
        f = CvFieldC()
        f.d_values = self.values.data_ptr()
        f.res, f.points, f.beta = int(self.res), int(self.points), float(self.beta)
        for a in range(3):
            f.origin[a] = float(self.origin[a])
            f.cell_inverse[a] = float(self.cell_inverse[a])
        return f

    def vertices(self):
        """Vertex positions (n, 3), x fastest, as a CUDA float32 tensor."""
        import torch
        n = self.res + 1
        axes = [torch.arange(n, dtype=torch.float32, device="cuda") / float(self.cell_inverse[a]) +
                float(self.origin[a]) for a in range(3)]
        z, y, x = torch.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
        return torch.stack([x.reshape(-1), y.reshape(-1), z.reshape(-1)], dim=1).contiguous()


def field_box(lo, hi, res: int, pad: float = 0.05):
    """Box of the mesh bounds padded by pad per side: (origin, cells per unit length)."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    ext = (hi - lo) * (1 + 2 * pad)
    origin = lo - pad * (hi - lo)
    return tuple(np.float32(origin)), tuple(np.float32(res / ext))


def bake(cache, lo, hi, res: int = 64, points: int = 6, beta: float = 1.0, pad: float = 0.05, stream=None) -> CvField:
    """The cache evaluated at the vertices of a res^3 grid over the padded
    mesh bounds (one inference launch)."""
    origin, inv = field_box(lo, hi, res, pad)
    f = CvField(None, res, origin, inv, points, beta)
    f.values = cache.infer(f.vertices(), stream=stream)
    return f


@dataclass
class PathRecord:
    pos: object      # (walks, L, 3)
    q: object        # (walks, L)
    length: int


class CvSolver:
    """Plain walks with the along-walk control variate over a field (or
    none), optionally recording each walk's path for training labels."""

    def __init__(self, bvh, boundary: Boundary, params: WalkParams | None = None):
        self.bvh = bvh
        self.boundary = boundary
        self.params = params or WalkParams()
        self._bc = boundary.to_c()
        self._ws = None

    def _workspace(self, n, wpp):
        import torch
        need = workspace_bytes(n, wpp)
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(need, dtype=torch.uint8, device="cuda")
        return self._ws

    def solve(self, points, wpp: int, field: CvField | None = None, stats: PointStats | None = None,
              accumulate: bool = False, point_offset: int = 0, per_walk: bool = False, path_len: int = 0,
              seed: int | None = None, stream_id: int | None = None, events=None, max_steps: int | None = None):
        import torch
        n = points.shape[0]
        dev = points.device
        if stats is None:
            stats = PointStats.zeros(n, dev)
        ws = self._workspace(n, wpp)
        wv = wst = None
        if per_walk or path_len > 0:
            wv = torch.empty(n * wpp, dtype=torch.float32, device=dev)
            wst = torch.empty(n * wpp, dtype=torch.int32, device=dev)
        rec = None
        if path_len > 0:
            rec = PathRecord(torch.empty((n * wpp, path_len, 3), dtype=torch.float32, device=dev),
                             torch.empty((n * wpp, path_len), dtype=torch.float32, device=dev), path_len)
        wp = self.params.to_c()
        if seed is not None:
            wp.seed = seed
        if max_steps is not None:
            wp.max_steps = max_steps
        if stream_id is not None:
            wp.rng_stream = stream_id
        if events is not None:
            wp.event_begin, wp.event_end = events
        fc = field.to_c() if field is not None else None
        check(lib().n2wos_cv_solve(self.bvh.handle, C.byref(self._bc), C.byref(wp), points.data_ptr(), n, wpp,
                                   point_offset, C.byref(fc) if fc is not None else None, stats.n.data_ptr(),
                                   stats.mean.data_ptr(), stats.m2.data_ptr(), stats.steps.data_ptr(),
                                   stats.forced.data_ptr(), int(accumulate), ws.data_ptr(), ws.numel(),
                                   wv.data_ptr() if wv is not None else None,
                                   wst.data_ptr() if wst is not None else None,
                                   rec.pos.data_ptr() if rec is not None else None,
                                   rec.q.data_ptr() if rec is not None else None, path_len,
                                   torch.cuda.current_stream().cuda_stream))
        if wv is not None:
            stats.extra["walk_values"] = wv
            stats.extra["walk_steps"] = wst
        if rec is not None:
            stats.extra["path"] = rec
        return stats

    @staticmethod
    def path_labels(stats: PointStats):
        """(X_k, W - q_k) for every recorded step k <= steps of every walk:
        unbiased samples of the solution at the positions the walks visited."""
        import torch
        rec = stats.extra["path"]
        wv, wst = stats.extra["walk_values"], stats.extra["walk_steps"]
        k = torch.arange(rec.length, device=wv.device)
        sel = k[None, :] <= wst[:, None].to(torch.int64)
        y = (wv[:, None] - rec.q)[sel]
        return rec.pos[sel].contiguous(), y.contiguous()


def train_field(bvh, verts, boundary, params, points, cache_kw=None, label_stride: int = 2, path_len: int = 32,
                rounds: int = 1, train_steps=(40, 40), batch: int = 1 << 15, res: int = 64, design: int = 6,
                seed: int = 11, solver: "CvSolver | None" = None, label_max_steps: int | None = None):
    """The training pipeline of the control variate: labels from the path
    records of one plain walk per labelled point, Adam steps, a bake of the
    cache onto a res^3 grid over the padded mesh bounds, then `rounds`
    bootstrap rounds (labels from control-variate walks with the current
    field, far less noisy, and more Adam steps).  points: a CUDA (n, 3)
    tensor of interior points (every label_stride-th is labelled).
    label_max_steps caps the label walks only (a capped walk ends on the
    boundary value at its closest point, as at max_steps): the labels then
    carry a small bias, which the estimator does not inherit, since the
    correction has zero mean for any field, while the label solves no longer
    wait for the longest walks.  Returns (field, cache, labels per round)."""
    import torch

    from . import cache as K
    lo, hi = verts.min(0), verts.max(0)
    cs = solver or CvSolver(bvh, boundary, params)
    cache = K.Cache(K.CacheConfig.for_bounds(lo, hi, seed=seed, max_batch=batch, **(cache_kw or {})))
    lp = points[::label_stride].contiguous() if label_stride > 1 else points
    st = cs.solve(lp, 1, None, path_len=path_len, seed=seed + 101, stream_id=5, max_steps=label_max_steps)
    x, y = CvSolver.path_labels(st)
    cache.set_output_offset(float(y.mean().item()))
    gen = torch.Generator(device="cuda").manual_seed(seed)
    nlab = []
    fld = None
    for r in range(rounds + 1):
        nlab.append(int(x.shape[0]))
        for _ in range(train_steps[r]):
            idx = torch.randint(0, x.shape[0], (batch,), device="cuda", generator=gen)
            cache.train_step(x[idx].contiguous(), y[idx].contiguous())
        fld = bake(cache, lo, hi, res=res, points=design)
        if r < rounds:
            st = cs.solve(lp, 1, fld, path_len=path_len, seed=seed + 200 + r, stream_id=6 + r,
                          max_steps=label_max_steps)
            x, y = CvSolver.path_labels(st)
    return fld, cache, nlab


def measure_gain(bvh, boundary, params, field: CvField, points, wpp: int = 16, seed: int = 3) -> dict:
    """Plain walks against control-variate walks on the same points: mean
    walk variance, walker time, and the gain (variance reduction / cost ratio
    per walk) that decides whether a mesh is served with the field."""
    import time

    import torch

    from .wos import Solver
    plain, cs = Solver(bvh, boundary, params), CvSolver(bvh, boundary, params)
    out = {}
    for name, fn in (("plain", lambda: plain.solve(points, wpp, seed=seed)),
                     ("cv", lambda: cs.solve(points, wpp, field, seed=seed))):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = fn()
        torch.cuda.synchronize()
        out[name + "_seconds"] = time.perf_counter() - t0
        out[name + "_variance"] = float(st.variance().mean().item())
    out["variance_reduction"] = out["plain_variance"] / max(out["cv_variance"], 1e-300)
    out["cost_ratio"] = out["cv_seconds"] / out["plain_seconds"]
    out["gain"] = out["variance_reduction"] / out["cost_ratio"]
    return out
