"""Walk-on-Spheres solves over the C ABI (device-resident, torch for memory)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._lib import check, lib

BC_X2_MINUS_Y2 = 0
BC_ZEBRA_K4 = 3
BC_ZEBRA_K8 = 4
BC_ZEBRA_K12 = 5
BC_CONSTANT_ONE = 12
BC_CHARGES = 100
BC_POISSON_SINES = 101

SOURCE_NONE = 0
SOURCE_POISSON_SINES = 1


class BoundaryC(C.Structure):
    _fields_ = [("mode", C.c_int), ("n_charges", C.c_int), ("charges", (C.c_float * 4) * 16)]


class WalkParamsC(C.Structure):
    _fields_ = [("max_steps", C.c_int), ("epsilon", C.c_float), ("step_scale", C.c_float), ("source", C.c_int),
                ("seed", C.c_uint64), ("rng_stream", C.c_uint32), ("max_blocks", C.c_int),
                ("d_visit_counts", C.c_void_p), ("event_begin", C.c_void_p), ("event_end", C.c_void_p)]


@dataclass
class Boundary:
    mode: int = BC_CONSTANT_ONE
    charges: np.ndarray | None = None  # (n, 4): cx, cy, cz, q

    def to_c(self) -> BoundaryC:
        b = BoundaryC()
        b.mode = self.mode
        if self.charges is not None:
            ch = np.asarray(self.charges, np.float32).reshape(-1, 4)
            b.n_charges = len(ch)
            for i, row in enumerate(ch):
                for k in range(4):
                    b.charges[i][k] = float(row[k])
        return b

    def oracle_params(self) -> np.ndarray:
        """Flat parameter array in the oracle's layout (n, then cx cy cz q ...)."""
        if self.charges is None:
            return np.zeros(1, np.float32)
        ch = np.asarray(self.charges, np.float32).reshape(-1, 4)
        return np.concatenate([[len(ch)], ch.ravel()]).astype(np.float32)


@dataclass
class WalkParams:
    max_steps: int = 256
    epsilon: float = 1e-4
    step_scale: float = 0.999
    source: int = SOURCE_NONE
    seed: int = 12345
    rng_stream: int = 0
    max_blocks: int = 0

    def to_c(self) -> WalkParamsC:
        return WalkParamsC(self.max_steps, self.epsilon, self.step_scale, self.source, self.seed, self.rng_stream,
                           self.max_blocks, None, None, None)


@dataclass
class PointStats:
    """Per-point FP64 statistics resident on the device (torch tensors)."""
    n: object
    mean: object
    m2: object
    steps: object
    forced: object
    extra: dict = field(default_factory=dict)

    @classmethod
    def zeros(cls, n_points: int, device="cuda"):
        import torch
        return cls(torch.zeros(n_points, dtype=torch.float64, device=device),
                   torch.zeros(n_points, dtype=torch.float64, device=device),
                   torch.zeros(n_points, dtype=torch.float64, device=device),
                   torch.zeros(n_points, dtype=torch.int64, device=device),
                   torch.zeros(n_points, dtype=torch.int32, device=device))

    def variance(self):
        import torch
        return torch.where(self.n > 1, self.m2 / (self.n - 1).clamp(min=1), torch.zeros_like(self.m2))


def full_grid_blocks(source: int = SOURCE_NONE) -> int:
    """Blocks of the walker's full-device grid on the current device (what
    max_blocks = 0 launches)."""
    out = C.c_int(0)
    check(lib().n2wos_wos_full_grid_blocks(source, C.byref(out)))
    return out.value


def workspace_bytes(n_points: int, walks_per_point: int) -> int:
    r = int(lib().n2wos_wos_workspace_size(n_points, walks_per_point))
    if r < 0:
        check(r)
    return r


class Solver:
    """Plain WoS on one resident mesh: the drop-in for run_persistent_harmonic /
    launch_nc_pure_wos_samples plus the host summarize step."""

    def __init__(self, bvh, boundary: Boundary, params: WalkParams | None = None):
        self.bvh = bvh
        self.boundary = boundary
        self.params = params or WalkParams()
        self._bc = boundary.to_c()
        self._wp = self.params.to_c()
        self._ws = None

    def workspace(self, n_points: int, wpp: int):
        import torch
        need = workspace_bytes(n_points, wpp)
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(need, dtype=torch.uint8, device="cuda")
        return self._ws

    def solve(self, points, walks_per_point: int, stats: PointStats | None = None, accumulate: bool = False,
              point_offset: int = 0, per_walk: bool = False, stream=None, seed: int | None = None,
              visit_counts=None, events=None):
        """Runs walks_per_point walks from each point of a (n, 3) CUDA float32 tensor."""
        import torch
        n = points.shape[0]
        if stats is None:
            stats = PointStats.zeros(n, points.device)
        ws = self.workspace(n, walks_per_point)
        wv = ws_steps = None
        if per_walk:
            wv = torch.empty(n * walks_per_point, dtype=torch.float32, device=points.device)
            ws_steps = torch.empty(n * walks_per_point, dtype=torch.int32, device=points.device)
        wp = self._wp
        if seed is not None or visit_counts is not None or events is not None:
            wp = self.params.to_c()
            if seed is not None:
                wp.seed = seed
            if visit_counts is not None:
                wp.d_visit_counts = visit_counts.data_ptr()
            if events is not None:
                wp.event_begin, wp.event_end = events
        s = torch.cuda.current_stream().cuda_stream if stream is None else getattr(stream, "cuda_stream", stream)
        check(lib().n2wos_wos_solve(self.bvh.handle, C.byref(self._bc), C.byref(wp), points.data_ptr(), n,
                                    walks_per_point, point_offset, stats.n.data_ptr(), stats.mean.data_ptr(),
                                    stats.m2.data_ptr(), stats.steps.data_ptr(), stats.forced.data_ptr(),
                                    int(accumulate), ws.data_ptr(), ws.numel(),
                                    wv.data_ptr() if wv is not None else None,
                                    ws_steps.data_ptr() if ws_steps is not None else None, s))
        if per_walk:
            stats.extra["walk_values"] = wv
            stats.extra["walk_steps"] = ws_steps
        return stats
