"""Request server over resident meshes (n2wos_server_* of the C ABI).

The server itself -- admission by HBM budget, continuous batching of walker
pools across requests, per-request rounds sized on the device by the
estimated RMSE, the learned cache as an along-walk control variate on meshes
that carry a field -- is native code; this module wraps it, chooses per mesh
between the control variate and plain walks by their measured gain, and
replays a request trace against it.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from ._lib import check, lib
from .wos import Boundary, WalkParams

REQ_QUEUED, REQ_ACTIVE, REQ_DONE = 0, 1, 2


class ServerConfigC(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("tick_walks", C.c_int64), ("first_round_walks", C.c_int),
                ("max_round_walks", C.c_int), ("max_walks_per_point", C.c_int)]


class RequestInfoC(C.Structure):
    _fields_ = [("state", C.c_int), ("mesh", C.c_int), ("n_points", C.c_int), ("rounds", C.c_int),
                ("walks_per_point", C.c_int), ("estimated_rmse", C.c_double), ("walk_steps", C.c_uint64),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
                ("control_variate", C.c_int)]


class ServerStatsC(C.Structure):
    _fields_ = [("ticks", C.c_int64), ("requests_submitted", C.c_int64), ("requests_done", C.c_int64),
                ("walks", C.c_int64), ("walk_steps", C.c_uint64), ("arena_bytes", C.c_int64),
                ("arena_used_bytes", C.c_int64), ("fixed_bytes", C.c_int64), ("busy_seconds", C.c_double),
                ("launches", C.c_int64), ("gpu_busy_seconds", C.c_double), ("gpu_span_seconds", C.c_double)]


_SIGS = {
    "n2wos_server_create": ([C.c_void_p, C.POINTER(C.c_void_p)], C.c_int),
    "n2wos_server_destroy": ([C.c_void_p], C.c_int),
    "n2wos_server_add_mesh": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int)], C.c_int),
    "n2wos_server_submit": ([C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_double, C.c_uint32,
                             C.POINTER(C.c_int64)], C.c_int),
    "n2wos_server_step": ([C.c_void_p, C.POINTER(C.c_int)], C.c_int),
    "n2wos_server_set_mesh_field": ([C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int], C.c_int),
    "n2wos_server_request_info": ([C.c_void_p, C.c_int64, C.POINTER(RequestInfoC)], C.c_int),
    "n2wos_server_take_result": ([C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p], C.c_int),
    "n2wos_server_get_stats": ([C.c_void_p, C.POINTER(ServerStatsC)], C.c_int),
}


def _L():
    L = lib()
    if not getattr(L, "_serve_sigs", False):
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        L._serve_sigs = True
    return L


@dataclass
class ServerConfig:
    hbm_budget_bytes: int = 1 << 30
    # walks per slot launch: larger launches keep the walker grid full longer
    # (measured on the serving benchmark: 2^22 -> 1.16 G steps/s while busy,
    # p99 4.9 s; 2^25 -> 1.29 G steps/s, p99 4.2 s, p50 also lower)
    tick_walks: int = 1 << 25
    first_round_walks: int = 64
    max_round_walks: int = 2048
    max_walks_per_point: int = 4096

    def to_c(self) -> ServerConfigC:
        return ServerConfigC(self.hbm_budget_bytes, self.tick_walks, self.first_round_walks, self.max_round_walks,
                             self.max_walks_per_point)


@dataclass
class RequestResult:
    n: np.ndarray
    mean: np.ndarray
    m2: np.ndarray
    steps: np.ndarray
    forced: np.ndarray


class Server:
    def __init__(self, config: ServerConfig | None = None):
        self.config = config or ServerConfig()
        h = C.c_void_p()
        cfg = self.config.to_c()
        check(_L().n2wos_server_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._meshes = []  # keep BVH handles and C structs alive
        self._next_rng = 0

    def close(self):
        if self._h:
            _L().n2wos_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_mesh(self, bvh, boundary: Boundary, params: WalkParams) -> int:
        bc, wp = boundary.to_c(), params.to_c()
        mid = C.c_int()
        check(_L().n2wos_server_add_mesh(self._h, bvh.handle, C.byref(bc), C.byref(wp), C.byref(mid)))
        self._meshes.append((bvh, bc, wp))
        return mid.value

    def set_mesh_field(self, mesh: int, field, unit_walks: int = 8, first_round_walks: int = 8) -> None:
        """Serve the mesh with the along-walk control variate over `field` (a
        cv.CvField, kept alive here) in rounds of unit_walks walks; None
        restores plain walks.  Before any request to the mesh."""
        fc = field.to_c() if field is not None else None
        check(_L().n2wos_server_set_mesh_field(self._h, mesh, C.byref(fc) if fc is not None else None, unit_walks,
                                               first_round_walks))
        self._meshes.append((field, fc))

    def submit(self, mesh: int, points: np.ndarray, target_rmse: float = 1e-2, rng_base: int | None = None) -> int:
        pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        if rng_base is None:
            rng_base = self._next_rng
        self._next_rng = max(self._next_rng, rng_base + len(pts))
        rid = C.c_int64()
        check(_L().n2wos_server_submit(self._h, mesh, pts.ctypes.data, len(pts), target_rmse, rng_base,
                                       C.byref(rid)))
        return rid.value

    def step(self) -> int:
        n = C.c_int()
        check(_L().n2wos_server_step(self._h, C.byref(n)))
        return n.value

    def info(self, rid: int) -> RequestInfoC:
        inf = RequestInfoC()
        check(_L().n2wos_server_request_info(self._h, rid, C.byref(inf)))
        return inf

    def take_result(self, rid: int) -> RequestResult:
        n = self.info(rid).n_points
        out = RequestResult(np.empty(n), np.empty(n), np.empty(n), np.empty(n, np.uint64), np.empty(n, np.uint32))
        check(_L().n2wos_server_take_result(self._h, rid, out.n.ctypes.data, out.mean.ctypes.data,
                                            out.m2.ctypes.data, out.steps.ctypes.data, out.forced.ctypes.data))
        return out

    def stats(self) -> dict:
        st = ServerStatsC()
        check(_L().n2wos_server_get_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStatsC._fields_}

    def drain(self, ids) -> None:
        """Steps until every request in ids has finished."""
        ids = list(ids)
        while any(self.info(i).state != REQ_DONE for i in ids):
            self.step()


def request_trace(n_requests: int, rate_per_s: float, n_meshes: int, seed: int = 0, min_points: int = 1000,
                  max_points: int = 262144):
    """Seeded request stream: Poisson arrivals (exponential gaps at rate_per_s),
    log-uniform sizes in [min_points, max_points], meshes uniform."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, n_requests)
    arrivals = np.cumsum(gaps) - gaps[0]
    sizes = np.exp(rng.uniform(np.log(min_points), np.log(max_points), n_requests)).astype(np.int64)
    meshes = rng.integers(0, n_meshes, n_requests)
    return arrivals, sizes, meshes


def replay(server: Server, arrivals, point_sets, meshes, target_rmse: float = 1e-2, clock=time.perf_counter):
    """Submits request k when the wall clock passes arrivals[k] (seconds from
    the start) and steps the server until every request has finished.  A
    request that arrives during a step is submitted right after it, and its
    latency counts from its arrival: latency = (submit - arrival) + (done -
    submit on the server's clock).  Returns the request ids, their infos, the
    latencies (s) and the wall time of the whole replay."""
    t0 = clock()
    k = 0
    n = len(arrivals)
    ids = [None] * n
    lag = np.zeros(n)
    pending = set()
    while k < n or pending:
        now = clock() - t0
        while k < n and arrivals[k] <= now:
            ids[k] = server.submit(int(meshes[k]), point_sets[k], target_rmse)
            lag[k] = (clock() - t0) - arrivals[k]
            pending.add(ids[k])
            k += 1
        if not pending:
            time.sleep(max(0.0, min(0.01, arrivals[k] - (clock() - t0))))
            continue
        if server.step():
            for rid in list(pending):
                if server.info(rid).state == REQ_DONE:
                    pending.discard(rid)
    wall = clock() - t0
    infos = [server.info(i) for i in ids]
    lat = np.array([lag[q] + infos[q].t_done - infos[q].t_submit for q in range(n)])
    return ids, infos, lat, wall


# the control-variate cache of the serving meshes: a 3-level dense grid
# (N_min 8, growth 2) with the 4 x 64 MLP, the shape that fits single-walk
# labels best on the benchmark problem (finer grids overfit them)
SERVE_FIELD_CACHE = {"n_levels": 3, "base_resolution": 8, "growth": 2.0}


def mesh_fields(eng, meshes, log=None, n_train: int = 65536, n_probe: int = 4096, seed: int = 0, **train_kw):
    """Per mesh: the control-variate field (cv.train_field on uniform interior
    points) and its measured gain over plain walks on other interior points;
    the policy serves the mesh with the field when the gain exceeds 1.
    Returns one dict per mesh (field or None, gain measurement, seconds)."""
    import time

    import torch

    from . import cv as CV, problems as P
    out = []
    for k, (name, v, bvh, bc, wp) in enumerate(meshes):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tr = torch.from_numpy(P.interior_points(bvh, v, n_train, seed=seed + 7000 + k)).cuda()
        fld, _, nlab = CV.train_field(bvh, v, bc, wp, tr, **{"label_stride": 1, "cache_kw": SERVE_FIELD_CACHE,
                                                             **train_kw})
        torch.cuda.synchronize()
        secs = time.perf_counter() - t0
        probe = torch.from_numpy(P.interior_points(bvh, v, n_probe, seed=seed + 8000 + k)).cuda()
        g = CV.measure_gain(bvh, bc, wp, fld, probe)
        use = g["gain"] > 1.0
        out.append({"mesh": name, "field": fld if use else None, "use_control_variate": use, "train_seconds": secs,
                    "labels": nlab, **g})
        if log:
            log(f"serving field {name}: variance reduction {g['variance_reduction']:.1f}x, cost {g['cost_ratio']:.2f}x,"
                f" gain {g['gain']:.1f} -> {'control variate' if use else 'plain walks'}")
    return out


def serving_workload(eng, n_requests: int = 40, rate_per_s: float = 1.5, seed: int = 0, target_rmse: float = 1e-2,
                     config: ServerConfig | None = None, min_points: int = 1000, max_points: int = 262144, log=None,
                     estimator: str = "auto", unit_walks: int = 8):
    """The serving benchmark: three resident meshes (UV sphere 64x128 and
    128x256 with r = 1 + 0.1 sin 6t sin 6p, and the 256x512 random-harmonic
    blob), each with its own 8-charge Dirichlet data; a seeded Poisson stream
    of requests of log-uniform size, WoS at eps 1e-4 to the target RMSE.
    estimator "auto": each mesh gets a learned control-variate field when its
    measured gain over plain walks exceeds 1 (plain walks otherwise); "plain":
    plain walks everywhere.  Returns throughput, latency percentiles, the
    device busy fraction and the true RMSE of every request against the FP64
    solution."""
    from . import problems as P
    from .wos import WalkParams
    meshes = []
    for k, name in enumerate(("mesh1", "mesh2", "blob")):
        v, t = P.make_mesh(name)
        bvh = eng.Bvh(v, t, leaf_size=4, build_method="sah")  # the walker's fastest tree
        meshes.append((name, v, bvh, P.charges_boundary(k), WalkParams(max_steps=256, epsilon=1e-4, seed=500 + k)))
    arrivals, sizes, mesh_of = request_trace(n_requests, rate_per_s, len(meshes), seed, min_points, max_points)
    point_sets = [P.interior_points(meshes[m][2], meshes[m][1], int(n), seed=seed * 1000 + q)
                  for q, (m, n) in enumerate(zip(mesh_of, sizes))]
    server = Server(config or ServerConfig(hbm_budget_bytes=1 << 30))
    for _, _, bvh, bc, wp in meshes:
        server.add_mesh(bvh, bc, wp)
    policy = []
    if estimator == "auto":
        policy = mesh_fields(eng, meshes, log=log)
        for k, pm in enumerate(policy):
            if pm["field"] is not None:
                server.set_mesh_field(k, pm["field"], unit_walks, unit_walks)
    if log:
        log(f"serving: {n_requests} requests, {int(sizes.sum())} points, rate {rate_per_s}/s")
    ids, infos, lat, wall = replay(server, arrivals, point_sets, mesh_of, target_rmse)
    st = server.stats()
    rmse = []
    for q, rid in enumerate(ids):
        r = server.take_result(rid)
        truth = P.charges_solution_f64(meshes[mesh_of[q]][3].charges, point_sets[q])
        rmse.append(float(np.sqrt(np.mean((r.mean - truth) ** 2))))
    server.close()
    steps = sum(int(i.walk_steps) for i in infos)
    return {
        "requests": n_requests,
        "points": int(sizes.sum()),
        "rate_per_s": rate_per_s,
        "target_rmse": target_rmse,
        "wall_s": wall,
        "busy_s": st["busy_seconds"],
        "points_per_s": float(sizes.sum()) / wall,
        "walk_steps_per_s": steps / wall,
        "walk_steps_per_busy_s": steps / st["busy_seconds"],
        "latency_p50_s": float(np.percentile(lat, 50)),
        "latency_p99_s": float(np.percentile(lat, 99)),
        "latency_max_s": float(lat.max()),
        "ticks": st["ticks"],
        "launches": st["launches"],
        "gpu_busy_s": st["gpu_busy_seconds"],
        "gpu_busy_fraction": st["gpu_busy_seconds"] / wall,
        "latency_samples": int(len(lat)),
        "estimator": estimator,
        "mesh_policy": [{k: v for k, v in pm.items() if k != "field"} for pm in policy],
        "control_variate_requests": int(sum(i.control_variate for i in infos)),
        "true_rmse_max": max(rmse),
        "true_rmse_mean": float(np.mean(rmse)),
        "walks_per_point_mean": float(np.mean([i.walks_per_point for i in infos])),
        "fixed_bytes": st["fixed_bytes"],
        "arena_bytes": st["arena_bytes"],
    }
