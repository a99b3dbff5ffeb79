"""CPU SIMT emulation of the persistent walker kernel (the kernel's own source,
64 lanes as coroutines): walk values and step counts bit-exact against the
oracle's walks, FP64 unit statistics against the oracle's reduction tree, and
no divergence at any wave-wide operation.  Runs without a GPU."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "emu")
CSRC = os.path.join(os.path.dirname(HERE), "n2wos-cuda_amd", "csrc")
LIB = os.path.join(EMU, "_build", "libwalker_emu.so")


def _build():
    srcs = [os.path.join(EMU, "walker_emu.cpp"), os.path.join(EMU, "hip", "hip_runtime.h")] + \
        [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hpp")]
    if os.path.exists(LIB) and os.path.getmtime(LIB) >= max(os.path.getmtime(s) for s in srcs):
        return
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    # build under a private name and rename, so parallel test processes never
    # load a half-written library
    tmp = f"{LIB}.{os.getpid()}.tmp"
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-I" + EMU, "-I" + CSRC,
                    "-I" + os.path.join(os.path.dirname(HERE), "include"), os.path.join(EMU, "walker_emu.cpp"),
                    "-o", tmp], check=True)
    os.replace(tmp, LIB)


@pytest.fixture(scope="module")
def emu():
    _build()
    E = C.CDLL(LIB)
    vp, i, f = C.c_void_p, C.c_int, C.c_float
    E.emu_walker.argtypes = [vp, vp, f, vp, i, i, i, f, f, C.c_uint64, i, i, vp, i, vp, vp, vp, vp, vp, vp, vp, vp,
                             vp, vp, i, i, i, i, C.c_long]
    E.emu_walker.restype = i
    E.emu_set_tri_count.argtypes = [C.c_long]
    E.emu_set_unit_ranges.argtypes = [i, i]
    E.emu_set_order.argtypes = [vp]
    E.emu_tri_oob.restype = C.c_long
    E.emu_set_ext.argtypes = [vp, i, i, vp, vp, f, vp, vp, i]
    E.emu_set_cap.argtypes = [i]
    return E


def _case(eng, orc, emu, mesh, npts, wpp, poisson, weights=(10, 10, 6), refill=8, eps=1e-3, seed=77,
          build_method="median", leaf=8, pts=None, cv=None):
    from n2wos_amd import problems as P
    v, t = P.make_mesh(mesh) if isinstance(mesh, str) else mesh
    nodes, recs, delta = eng.build_host(v, t, leaf_size=leaf, build_method=build_method)
    if pts is None:
        rng = np.random.default_rng(npts + wpp)
        d = rng.normal(size=(npts, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        pts = (d * rng.uniform(0, 0.8, (npts, 1))).astype(np.float32)
    npts = len(pts)
    fd2, fcp, ftri, fovf, _, _ = orc.bvh4_query(nodes, recs, delta, pts)
    fovf = fovf.astype(np.int32)
    bc = P.poisson_boundary() if poisson else P.charges_boundary(0)
    ch = np.ascontiguousarray(bc.charges if bc.charges is not None else np.zeros((1, 4)), np.float32)
    units = npts * ((wpp + 63) // 64)
    wv = np.zeros(npts * wpp, np.float32)
    ws = np.zeros(npts * wpp, np.int32)
    um, uq = np.zeros(units), np.zeros(units)
    us, uf = np.zeros(units, np.uint32), np.zeros(units, np.uint32)
    vis = np.zeros(2, np.uint64)
    emu.emu_set_tri_count(len(recs))
    rc = emu.emu_walker(nodes.ctypes.data, recs.ctypes.data, delta, pts.ctypes.data, npts, wpp, 256, eps, 0.999,
                        seed, int(poisson), bc.mode, ch.ctypes.data, 0 if bc.charges is None else len(ch),
                        fd2.ctypes.data, fcp.ctypes.data, fovf.ctypes.data, wv.ctypes.data, ws.ctypes.data,
                        um.ctypes.data, uq.ctypes.data, us.ctypes.data, uf.ctypes.data, vis.ctypes.data, refill,
                        weights[0], weights[1], weights[2], 2_000_000)
    assert rc == 0, "emulated wave hung or diverged"
    assert emu.emu_tri_oob() == 0, "triangle-record load outside the array"
    b2 = orc.Bvh2(v, t)
    if cv is None:
        ov, os_, _, _ = b2.walks(pts, wpp, max_steps=256, eps=eps, step_scale=0.999, bmode=bc.mode,
                                 bparams=bc.oracle_params(), poisson=poisson, seed=seed, nthreads=4)
    else:
        ov, os_ = b2.walks_cv(pts, wpp, cv["field"], points=cv["points"], beta=cv["beta"], max_steps=256, eps=eps,
                              step_scale=0.999, bmode=bc.mode, bparams=bc.oracle_params(), poisson=poisson,
                              seed=seed, nthreads=4)
    assert np.array_equal(ws, os_)
    assert np.array_equal(wv.view(np.uint32), ov.view(np.uint32))
    # unit statistics: the kernel's FP64 tree equals the oracle's fixed order
    n, m, q = orc.point_stats(ov, npts, wpp)
    if wpp <= 64:
        assert np.array_equal(um, m) and np.array_equal(uq, q)
    assert us.sum() == os_.sum()
    if cv is None:
        assert vis[0] > 0 and vis[1] > 0
    return pts, wv, ws


def test_emulated_walker_laplace(eng, orc, emu):
    _case(eng, orc, emu, "mesh1", 6, 70, False)


def test_emulated_walker_ring8(eng, orc, emu):
    # the 8-entry ring variant (tuning key "ring"): more spills, same walks
    emu.emu_set_cap(8)
    try:
        _case(eng, orc, emu, "mesh2", 5, 64, False, weights=(10, 24, 30), refill=2, eps=1e-4, build_method="sah",
              leaf=4)
    finally:
        emu.emu_set_cap(16)


def test_emulated_walker_poisson_small_units(eng, orc, emu):
    _case(eng, orc, emu, "mesh1", 5, 40, True)


@pytest.mark.parametrize("weights,refill", [((1, 1, 1000), 1), ((1000, 1, 1), 64), ((1, 1000, 1), 8)])
def test_emulated_walker_extreme_schedules(eng, orc, emu, weights, refill):
    """Any phase choice and refill threshold gives the same walks."""
    _case(eng, orc, emu, "mesh1", 3, 64, False, weights=weights, refill=refill)


@pytest.mark.parametrize("npts,wpp,poisson", [(37, 4, False), (70, 1, True), (5, 32, False), (9, 16, True)])
def test_emulated_walker_packed_units(eng, orc, emu, npts, wpp, poisson):
    """Plain walks with 1 to 32 walks per point run in packed units (64 / wpp
    points per unit, the last one partial): the oracle's walks bit for bit,
    and per-point statistics in the oracle's reduction order."""
    _case(eng, orc, emu, "mesh1", npts, wpp, poisson)


@pytest.mark.parametrize("ranges,xcd,npts,wpp", [(1, 0, 20, 64), (8, 0, 20, 64), (8, 7, 20, 64), (3, 2, 11, 100),
                                                 (8, 5, 3, 64), (8, 6, 37, 4)])
def test_emulated_walker_unit_ranges(eng, orc, emu, ranges, xcd, npts, wpp):
    """Units claimed range by range (the wave's own XCD's range first, then
    the others in turn; ranges may be empty) are each run exactly once: the
    oracle's walks and statistics bit for bit."""
    emu.emu_set_unit_ranges(ranges, xcd)
    try:
        _case(eng, orc, emu, "mesh1", npts, wpp, False)
    finally:
        emu.emu_set_unit_ranges(8, 3)


@pytest.mark.parametrize("npts,wpp,ranges", [(37, 4, 8), (11, 100, 3), (70, 1, 8), (20, 64, 1)])
def test_emulated_walker_claim_order(eng, orc, emu, npts, wpp, ranges):
    """Units claimed through a permutation of the points (the spatial claim
    order) give each point the oracle's walks and statistics bit for bit."""
    order = np.random.default_rng(npts + wpp).permutation(npts).astype(np.int32)
    emu.emu_set_order(order.ctypes.data)
    emu.emu_set_unit_ranges(ranges, 1)
    try:
        _case(eng, orc, emu, "mesh1", npts, wpp, False)
    finally:
        emu.emu_set_order(None)
        emu.emu_set_unit_ranges(8, 3)


def test_emulated_walker_sah_tree(eng, orc, emu):
    """Walks over a surface-area-heuristic tree (leaf 4) equal the oracle's."""
    _case(eng, orc, emu, "mesh2", 5, 64, False, weights=(10, 10, 30), refill=2, eps=1e-4, build_method="sah", leaf=4)


def test_emulated_walker_exterior_start_points(eng, orc, emu):
    """Walks that start outside the closed mesh move away from it until the
    step cap (256); their query bound grows past the FP32 range after about
    180 steps, which must stay an unbounded query (empty child slots
    unvisited, every triangle-record load inside the array) and give the
    oracle's walks.  Far from the mesh every triangle lies inside the pruning
    margin, so each query scans the mesh: a 224-triangle sphere keeps the
    emulation short."""
    v, t = eng.uv_sphere(8, 16, eng.SPHERE_SIN6, [0.1])
    vn, _, _ = eng.normalize_to_unit_radius(v)
    pts = np.array([[1.6, 0.0, 0.0], [0.0, -1.7, 0.3], [2.5, 2.5, -2.5], [0.3, 0.2, 0.1]], np.float32)
    _case(eng, orc, emu, (vn, t), 4, 64, False, eps=1e-3, pts=pts, leaf=2)


def _smooth_field(v, res, seed):
    """A smooth random field on the grid over the mesh bounds padded 5%:
    (values, res, origin, inverse cell size)."""
    lo, hi = v.min(0).astype(np.float64), v.max(0).astype(np.float64)
    pad = 0.05 * (hi - lo)
    o = (lo - pad).astype(np.float32)
    ext = hi - lo + 2 * pad
    inv = (res / ext).astype(np.float32)
    ax = [o[a] + np.arange(res + 1) / inv[a] for a in range(3)]
    X, Y, Z = np.meshgrid(ax[0], ax[1], ax[2], indexing="ij")
    r = np.random.default_rng(seed)
    val = np.zeros(X.shape)
    for _ in range(5):
        k = r.normal(size=3) * 3
        val += r.normal() * np.sin(k[0] * X + k[1] * Y + k[2] * Z + r.uniform(0, 6.3))
    # vertex (x, y, z) at x + (res+1) (y + (res+1) z): x fastest
    return np.ascontiguousarray(val.transpose(2, 1, 0).ravel(), np.float32), res, o, inv


@pytest.mark.parametrize("points,poisson,npts,wpp", [(2, False, 4, 64), (6, False, 4, 64), (6, True, 4, 64),
                                                     (6, False, 37, 4), (2, False, 70, 1), (6, True, 5, 32)])
def test_emulated_walker_control_variate(eng, orc, emu, points, poisson, npts, wpp):
    """The extended walker's along-walk control variate (antipodal pair or
    turned octahedron over a trilinear field) gives the oracle's walk values
    bit for bit, with the plain walk's step counts; its path record starts at
    the start point with q_0 = 0 and writes exactly steps + 1 entries.  With
    1 to 32 walks per point the units are packed (64 / wpp points per unit,
    the last one partial), and the per-point statistics still equal the
    oracle's reduction order."""
    from n2wos_amd import problems as P
    v, t = P.make_mesh("mesh1")
    field = _smooth_field(v, 24, points + int(poisson))
    vals, res, o, inv = field
    plen = 40
    ppos = np.full(npts * wpp * plen * 3, np.nan, np.float32)
    pq = np.full(npts * wpp * plen, np.nan, np.float32)
    emu.emu_set_ext(vals.ctypes.data, res, points, o.ctypes.data, inv.ctypes.data, 0.75, ppos.ctypes.data,
                    pq.ctypes.data, plen)
    try:
        pts, wv, ws = _case(eng, orc, emu, "mesh1", npts, wpp, poisson, eps=1e-3, seed=91,
                            cv={"field": field, "points": points, "beta": 0.75})
    finally:
        emu.emu_set_ext(None, 0, 0, None, None, 0.0, None, None, 0)
    ppos = ppos.reshape(npts * wpp, plen, 3)
    pq = pq.reshape(npts * wpp, plen)
    for i in range(npts * wpp):
        assert np.array_equal(ppos[i, 0], pts[i // wpp]) and pq[i, 0] == 0.0
        n = min(int(ws[i]) + 1, plen)
        assert np.all(np.isfinite(pq[i, :n])) and np.all(np.isnan(pq[i, n:]))
