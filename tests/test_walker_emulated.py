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
    E.emu_tri_oob.restype = C.c_long
    return E


def _case(eng, orc, emu, mesh, npts, wpp, poisson, weights=(10, 10, 6), refill=8, eps=1e-3, seed=77,
          build_method="median", leaf=8, pts=None):
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
    ov, os_, _, _ = b2.walks(pts, wpp, max_steps=256, eps=eps, step_scale=0.999, bmode=bc.mode,
                             bparams=bc.oracle_params(), poisson=poisson, seed=seed, nthreads=4)
    assert np.array_equal(ws, os_)
    assert np.array_equal(wv.view(np.uint32), ov.view(np.uint32))
    # unit statistics: the kernel's FP64 tree equals the oracle's fixed order
    n, m, q = orc.point_stats(ov, npts, wpp)
    if wpp <= 64:
        assert np.array_equal(um, m) and np.array_equal(uq, q)
    assert us.sum() == os_.sum()
    assert vis[0] > 0 and vis[1] > 0


def test_emulated_walker_laplace(eng, orc, emu):
    _case(eng, orc, emu, "mesh1", 6, 70, False)


def test_emulated_walker_poisson_small_units(eng, orc, emu):
    _case(eng, orc, emu, "mesh1", 5, 40, True)


@pytest.mark.parametrize("weights,refill", [((1, 1, 1000), 1), ((1000, 1, 1), 64), ((1, 1000, 1), 8)])
def test_emulated_walker_extreme_schedules(eng, orc, emu, weights, refill):
    """Any phase choice and refill threshold gives the same walks."""
    _case(eng, orc, emu, "mesh1", 3, 64, False, weights=weights, refill=refill)


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
