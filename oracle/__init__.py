"""CPU oracle for the Walk-on-Spheres hot path (test infrastructure only).

The checker: ctypes bindings over ``_build/liboracle.so`` (built from
``oracle.c``) and, where it was built, over ``_ref/libn2wos_ref.so`` (the
reference's own CPU geometry core, compiled from its sources).  Only the tests,
the smoke entry and the CPU-baseline leg of ``bench.py`` import this package;
the product library never calls it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libn2wos_ref.so")

# boundary modes (ids 0..12 are the reference's NcBoundaryMode values)
BC_X2_MINUS_Y2 = 0
BC_ZEBRA_K4 = 3
BC_ZEBRA_K8 = 4
BC_ZEBRA_K12 = 5
BC_CONSTANT_ONE = 12
BC_CHARGES = 100
BC_POISSON_SINES = 101

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")

_lib = None
_ref = None


def build(force: bool = False) -> None:
    """Compile the oracle (and the reference core when its tree is present);
    make rebuilds only what is older than its sources unless force is set."""
    subprocess.run(["make", "-s", "-C", HERE] + (["-B"] if force else []), check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_closest_point_on_triangle.argtypes = [_f32p] * 5
        L.orc_bruteforce.argtypes = [_f32p, _u32p, C.c_int, _f32p, C.c_long, _f32p, _f32p, _i32p, C.c_int]
        L.orc_count_degenerate.argtypes = [_f32p, _u32p, C.c_int, C.c_float]
        L.orc_count_degenerate.restype = C.c_long
        L.orc_uv_sphere_counts.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_uv_sphere.argtypes = [C.c_int, C.c_int, C.c_int, _f32p, _f32p, _u32p]
        L.orc_sh_real.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double]
        L.orc_sh_real.restype = C.c_double
        L.orc_normalize.argtypes = [_f32p, C.c_int, _f32p, C.POINTER(C.c_float)]
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_u01.argtypes = [C.c_uint32]
        L.orc_u01.restype = C.c_float
        L.orc_sincos2pi.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_boundary_f32.argtypes = [C.c_int, _f32p, _f32p]
        L.orc_boundary_f32.restype = C.c_float
        L.orc_poisson_source_f32.argtypes = [_f32p]
        L.orc_poisson_source_f32.restype = C.c_float
        L.orc_solution_f64.argtypes = [C.c_int, _f32p, _f64p]
        L.orc_solution_f64.restype = C.c_double
        L.orc_poisson_source_f64.argtypes = [_f64p]
        L.orc_poisson_source_f64.restype = C.c_double
        L.orc_bvh2_build.argtypes = [_f32p, C.c_int, _u32p, C.c_int, C.c_int]
        L.orc_bvh2_build.restype = C.c_void_p
        L.orc_bvh2_free.argtypes = [C.c_void_p]
        L.orc_bvh2_info.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_bvh2_query.argtypes = [C.c_void_p, _f32p, C.c_long, _f32p, _f32p, _i32p, C.c_int]
        L.orc_prune_threshold.argtypes = [C.c_float, C.c_float]
        L.orc_prune_threshold.restype = C.c_float
        L.orc_walks.argtypes = [C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_int, _f32p, C.c_int, C.c_uint64,
                                C.c_uint32, _f32p, C.c_void_p, C.c_long, C.c_int, C.c_int, _f32p, _i32p,
                                C.c_void_p, C.c_void_p, C.c_int, C.c_int]
        L.orc_cv_term.argtypes = [_f32p, C.c_int, _f32p, _f32p, C.c_int, _f32p, C.c_float, _f32p, C.c_uint32]
        L.orc_cv_term.restype = C.c_float
        L.orc_cv_lookup.argtypes = [_f32p, C.c_int, _f32p, _f32p, _f32p]
        L.orc_cv_lookup.restype = C.c_float
        L.orc_cv_frame.argtypes = [_f32p, C.c_float, _f32p, _f32p]
        L.orc_walks_cv.argtypes = [C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_int, _f32p, C.c_int, C.c_uint64,
                                   C.c_uint32, _f32p, C.c_long, C.c_int, _f32p, _i32p, C.c_int, C.c_int, C.c_int,
                                   _f32p, _f32p, _f32p, C.c_float]
        L.orc_welford.argtypes = [_f32p, C.c_long, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_chan_merge.argtypes = [C.c_double] * 6 + [_f64p]
        L.orc_point_stats.argtypes = [_f32p, C.c_long, C.c_int, _f64p, _f64p, _f64p]
        L.orc_bvh4_query.argtypes = [C.c_void_p, C.c_void_p, C.c_float, C.c_int, _f32p, C.c_long, C.c_void_p,
                                     _f32p, _f32p, _i32p, _i32p, C.c_void_p, C.c_void_p]
        L.orc_cache_layout.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p]
        L.orc_cache_forward.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, _f32p, _f32p, _f32p,
                                        _f32p, _f32p, C.c_long, _f64p]
        L.orc_cache_grads.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, _f32p, _f32p, _f32p,
                                      _f32p, _f32p, _f32p, C.c_long, _f64p, _f64p]
        L.orc_cache_grads.restype = C.c_double
        L.orc_round_half.argtypes = [C.c_double]
        L.orc_round_half.restype = C.c_double
        L.orc_cache_forward_fp16.argtypes = L.orc_cache_forward.argtypes
        L.orc_cache_encode.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, _f32p, _f32p,
                                       _f32p, _f32p, C.c_long, C.c_int, _f64p]
        L.orc_cache_grads_fp16.argtypes = L.orc_cache_grads.argtypes + [C.c_double]
        L.orc_cache_grads_fp16.restype = C.c_double
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    """The reference's CPU core (only where it was compiled from source)."""
    global _ref
    if _ref is None:
        R = C.CDLL(REF_PATH)
        ip = C.POINTER(C.c_int)
        R.ref_bumpy_sphere.argtypes = [C.c_int, C.c_int, C.c_float, C.c_int, ip, ip, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p]
        R.ref_normalize.argtypes = [_f32p, C.c_int, _u32p, C.c_int, _f32p, C.POINTER(C.c_float)]
        R.ref_closest_point_on_triangle.argtypes = [_f32p] * 5
        R.ref_bruteforce.argtypes = [_f32p, C.c_int, _u32p, C.c_int, _f32p, C.c_long, _f32p, _f32p, _i32p]
        R.ref_count_degenerate.argtypes = [_f32p, C.c_int, _u32p, C.c_int, C.c_float]
        R.ref_count_degenerate.restype = C.c_long
        R.ref_load_mesh.argtypes = [C.c_char_p, C.c_int, ip, ip, C.c_void_p, C.c_void_p]
        _ref = R
    return _ref


# ---------------------------------------------------------------- helpers
def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def uv_sphere(stacks: int, slices: int, kind: int, params=(0.0,)):
    nv, nt = C.c_int(), C.c_int()
    if lib().orc_uv_sphere_counts(stacks, slices, C.byref(nv), C.byref(nt)) != 0:
        raise ValueError("stacks >= 4 and slices >= 8 required")
    prm = np.zeros(max(81, len(params)), np.float32)
    prm[: len(params)] = params
    v = np.zeros((nv.value, 3), np.float32)
    t = np.zeros((nt.value, 3), np.uint32)
    lib().orc_uv_sphere(stacks, slices, kind, prm, v, t)
    return v, t


def normalize(verts):
    v = _f32(verts).copy()
    c = np.zeros(3, np.float32)
    s = C.c_float()
    if lib().orc_normalize(v, v.shape[0], c, C.byref(s)) != 0:
        raise ValueError("cannot normalise")
    return v, c, s.value


def bruteforce(verts, tris, pts, nthreads: int = 1):
    pts = _f32(pts).reshape(-1, 3)
    n = pts.shape[0]
    d2 = np.zeros(n, np.float32)
    cp = np.zeros((n, 3), np.float32)
    tri = np.zeros(n, np.int32)
    lib().orc_bruteforce(_f32(verts), np.ascontiguousarray(tris, np.uint32), len(tris), pts, n, d2, cp, tri, nthreads)
    return d2, cp, tri


def closest_point_on_triangle(p, a, b, c):
    out = np.zeros(3, np.float32)
    lib().orc_closest_point_on_triangle(_f32(p), _f32(a), _f32(b), _f32(c), out)
    return out


def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(np.ascontiguousarray(ctr, np.uint32), np.ascontiguousarray(key, np.uint32), out)
    return out


def sincos2pi(v: float):
    s, c = C.c_float(), C.c_float()
    lib().orc_sincos2pi(v, C.byref(s), C.byref(c))
    return s.value, c.value


class Bvh2:
    """The oracle's binary BVH (median split) with brute-force-exact queries."""

    def __init__(self, verts, tris, leaf_size: int = 8):
        self.verts = _f32(verts)
        self.tris = np.ascontiguousarray(tris, np.uint32)
        self.h = lib().orc_bvh2_build(self.verts, len(self.verts), self.tris, len(self.tris), leaf_size)
        if not self.h:
            raise ValueError("bvh build failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_bvh2_free(self.h)
            self.h = None

    def info(self):
        n, d = C.c_int(), C.c_int()
        lib().orc_bvh2_info(self.h, C.byref(n), C.byref(d))
        return n.value, d.value

    def query(self, pts, nthreads: int = 1):
        pts = _f32(pts).reshape(-1, 3)
        n = pts.shape[0]
        d2 = np.zeros(n, np.float32)
        cp = np.zeros((n, 3), np.float32)
        tri = np.zeros(n, np.int32)
        lib().orc_bvh2_query(self.h, pts, n, d2, cp, tri, nthreads)
        return d2, cp, tri

    def walks(self, pts, wpp, *, max_steps=256, eps=1e-4, step_scale=0.999, bmode=BC_CONSTANT_ONE,
              bparams=None, poisson=False, seed=0, stream=0, point_ids=None, depth=-1, nthreads=1,
              want_xm=False, antithetic=False):
        pts = _f32(pts).reshape(-1, 3)
        npts = pts.shape[0]
        prm = _f32(bparams if bparams is not None else np.zeros(1))
        total = npts * wpp
        values = np.zeros(total, np.float32)
        steps = np.zeros(total, np.int32)
        status = np.zeros(total, np.int32)
        xm = np.zeros((total, 3), np.float32) if want_xm else None
        pid = None
        if point_ids is not None:
            pid = np.ascontiguousarray(point_ids, np.uint32)
        rc = lib().orc_walks(self.h, max_steps, eps, step_scale, bmode, prm, int(poisson), seed, stream, pts,
                             pid.ctypes.data if pid is not None else None, npts, wpp, depth, values, steps,
                             status.ctypes.data, xm.ctypes.data if xm is not None else None, nthreads,
                             int(antithetic))
        if rc != 0:
            raise ValueError("orc_walks failed")
        return values, steps, status, xm

    def walks_cv(self, pts, wpp, field, *, points=2, beta=1.0, max_steps=256, eps=1e-4, step_scale=0.999,
                 bmode=BC_CONSTANT_ONE, bparams=None, poisson=False, seed=0, stream=0, nthreads=1):
        """Plain walks with the along-walk control variate; field = (grid
        values[(G+1)^3], G, origin[3], inverse cell size[3])."""
        grid, res, origin, inv_h = field
        pts = _f32(pts).reshape(-1, 3)
        npts = pts.shape[0]
        prm = _f32(bparams if bparams is not None else np.zeros(1))
        values = np.zeros(npts * wpp, np.float32)
        steps = np.zeros(npts * wpp, np.int32)
        rc = lib().orc_walks_cv(self.h, max_steps, eps, step_scale, bmode, prm, int(poisson), seed, stream, pts, npts,
                                wpp, values, steps, nthreads, points, res, _f32(grid).ravel(), _f32(origin),
                                _f32(inv_h), beta)
        if rc != 0:
            raise ValueError("orc_walks_cv failed")
        return values, steps


def cv_term(field, points, p, R, d, psi_word=0):
    """One control-variate term C(p + R d) - (1/m) sum_j C(p + R s_j) of the
    walker (m = points: 2 antipodal, 6 turned octahedron); field as walks_cv."""
    grid, res, origin, inv_h = field
    return lib().orc_cv_term(_f32(grid).ravel(), res, _f32(origin), _f32(inv_h), points, _f32(p), R, _f32(d),
                             psi_word)


def cv_lookup(field, q):
    grid, res, origin, inv_h = field
    return lib().orc_cv_lookup(_f32(grid).ravel(), res, _f32(origin), _f32(inv_h), _f32(q))


def cv_frame(d, psi):
    e1 = np.zeros(3, np.float32)
    e2 = np.zeros(3, np.float32)
    lib().orc_cv_frame(_f32(d), psi, e1, e2)
    return e1, e2


def point_stats(values, npts: int, wpp: int):
    n = np.zeros(npts)
    m = np.zeros(npts)
    s = np.zeros(npts)
    lib().orc_point_stats(_f32(values), npts, wpp, n, m, s)
    return n, m, s


def solution_f64(mode: int, bparams, pts):
    prm = _f32(bparams if bparams is not None else np.zeros(1))
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    return np.array([lib().orc_solution_f64(mode, prm, p) for p in pts])


def boundary_f32(mode: int, bparams, pts):
    prm = _f32(bparams if bparams is not None else np.zeros(1))
    pts = _f32(pts).reshape(-1, 3)
    return np.array([lib().orc_boundary_f32(mode, prm, p) for p in pts], np.float32)


def bvh4_query(nodes, recs, delta, pts, cap=16, thr_init=None):
    """Restated product traversal on exported node/record arrays: results and
    per-query node/triangle visit counts."""
    pts = _f32(pts).reshape(-1, 3)
    n = pts.shape[0]
    d2 = np.zeros(n, np.float32)
    cp = np.zeros((n, 3), np.float32)
    tri = np.zeros(n, np.int32)
    ovf = np.zeros(n, np.int32)
    nn = np.zeros(n, np.uint32)
    nt = np.zeros(n, np.uint32)
    thr = _f32(thr_init) if thr_init is not None else None
    rc = lib().orc_bvh4_query(nodes.ctypes.data, recs.ctypes.data, delta, cap, pts, n,
                              thr.ctypes.data if thr is not None else None, d2, cp, tri, ovf,
                              nn.ctypes.data, nt.ctypes.data)
    if rc != 0:
        raise ValueError("bad ring capacity")
    return d2, cp, tri, ovf, nn, nt


def cache_layout(n_levels, log2_t, base_res, growth):
    off = np.zeros(16, np.uint32)
    sz = np.zeros(16, np.uint32)
    res = np.zeros(16, np.int32)
    dense = np.zeros(16, np.int32)
    tot = np.zeros(1, np.uint32)
    lib().orc_cache_layout(n_levels, log2_t, base_res, growth, off.ctypes.data, sz.ctypes.data, res.ctypes.data,
                           dense.ctypes.data, tot.ctypes.data)
    return {"offsets": off.tolist(), "sizes": sz.tolist(), "res": res.tolist(), "dense": dense.tolist(),
            "total_entries": int(tot[0])}


def _cfg_args(cfg):
    return (cfg.n_levels, cfg.log2_table_size, cfg.base_resolution, float(np.float32(cfg.growth)),
            getattr(cfg, "n_neurons", 64), getattr(cfg, "n_hidden_layers", 4), _f32(cfg.origin), _f32(cfg.extent))


def cache_forward(cfg, grid, weights, pts, fp16=False):
    """FP64 forward of the cache at (n, 3) points (cfg: an n2wos_amd CacheConfig).
    fp16=True rounds weights, features and activations to binary16 where the
    product's MFMA kernels do."""
    pts = _f32(pts).reshape(-1, 3)
    out = np.zeros(len(pts))
    fn = lib().orc_cache_forward_fp16 if fp16 else lib().orc_cache_forward
    fn(*_cfg_args(cfg), _f32(grid), _f32(weights), pts, len(pts), out)
    return out


def cache_encode(cfg, grid, pts, f32=False):
    """Hash-grid features (n, 32) at (n, 3) points: column 2 l + f is feature
    f of level l.  f32=True maps and weights in FP32 like the product kernels."""
    pts = _f32(pts).reshape(-1, 3)
    out = np.zeros((len(pts), 32))
    lib().orc_cache_encode(*_cfg_args(cfg), _f32(grid), pts, len(pts), int(f32), out)
    return out


def cache_grads(cfg, grid, weights, pts, targets):
    """FP64 mean-squared-error gradients: (grid_grad, weight_grad, loss)."""
    pts = _f32(pts).reshape(-1, 3)
    gg = np.zeros(len(grid))
    wg = np.zeros(len(weights))
    loss = lib().orc_cache_grads(*_cfg_args(cfg), _f32(grid), _f32(weights), pts, _f32(targets), len(pts), gg, wg)
    return gg, wg, loss


def cache_grads_fp16(cfg, grid, weights, pts, targets, loss_scale):
    """cache_grads with the product's binary16 rounding points (weights,
    features, activations, loss-scaled dL/dy and pre-activation gradients)."""
    pts = _f32(pts).reshape(-1, 3)
    gg = np.zeros(len(grid))
    wg = np.zeros(len(weights))
    loss = lib().orc_cache_grads_fp16(*_cfg_args(cfg), _f32(grid), _f32(weights), pts, _f32(targets), len(pts), gg,
                                      wg, float(loss_scale))
    return gg, wg, loss
