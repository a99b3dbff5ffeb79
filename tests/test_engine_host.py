"""CPU-side checks of the engine library: exports, host mesh code, host BVH
build (traversed by the oracle's restatement) and argument validation."""
import ctypes as C

import numpy as np
import pytest


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_library_exports_every_header_symbol(eng):
    lib = eng.lib()
    names = eng.header_functions()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.n2wos_abi_version() == 1


@pytest.mark.parametrize("kind,params", [(0, [0.15]), (1, [0.1]), (2, [0.0])])
def test_uv_sphere_equals_oracle(eng, orc, kind, params):
    v, t = eng.uv_sphere(64, 128, kind, params)
    vo, to = orc.uv_sphere(64, 128, kind, params)
    assert np.array_equal(t, to) and np.array_equal(bits(v), bits(vo))
    vn, c, s = eng.normalize_to_unit_radius(v)
    von, co, so = orc.normalize(vo)
    assert np.array_equal(bits(vn), bits(von)) and np.array_equal(bits(c), bits(co)) and np.float32(s) == np.float32(so)


def test_sh_blob_equals_oracle(eng, orc):
    a = np.random.default_rng(0).normal(0, 0.02, 81).astype(np.float32)
    v, t = eng.uv_sphere(32, 64, 3, a)
    vo, to = orc.uv_sphere(32, 64, 3, a)
    assert np.array_equal(bits(v), bits(vo))


def test_mesh_closed_and_counts(eng):
    for stacks, slices in [(64, 128), (128, 256), (5, 9)]:
        v, t = eng.uv_sphere(stacks, slices, 1, [0.1])
        assert len(t) == 2 * slices * (stacks - 1) and len(v) == 2 + (stacks - 1) * slices
        e = np.sort(np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]), axis=1)
        _, cnt = np.unique(e, axis=0, return_counts=True)
        assert np.all(cnt == 2)  # watertight: every edge shared by two faces
    vn, c, s = eng.normalize_to_unit_radius(v)
    r = np.linalg.norm(vn.astype(np.float64), axis=1)
    assert abs(r.max() - 1.0) < 1e-6
    assert np.allclose((vn.min(0) + vn.max(0)) / 2, 0, atol=1e-6)


def test_degenerate_count(eng, orc):
    v, t = eng.uv_sphere(16, 32, 1, [0.1])
    assert eng.count_degenerate_triangles(v, t) == orc.lib().orc_count_degenerate(v, t, len(t), 1e-20)
    t2 = np.concatenate([t, np.array([[0, 0, 1]], np.uint32)])
    assert eng.count_degenerate_triangles(v, t2) == eng.count_degenerate_triangles(v, t) + 1


def test_host_bvh_traversal_equals_bruteforce(eng, orc):
    v, t = eng.uv_sphere(64, 128, 1, [0.1])
    vn, _, _ = eng.normalize_to_unit_radius(v)
    nodes, recs, delta = eng.build_host(vn, t)
    rng = np.random.default_rng(11)
    pts = np.concatenate([rng.uniform(-1.2, 1.2, (6000, 3)), rng.uniform(-0.05, 0.05, (200, 3)),
                          rng.uniform(-30, 30, (200, 3))]).astype(np.float32)
    d2b, cpb, trb = orc.bruteforce(vn, t, pts, nthreads=8)
    near = (cpb[:1000] + rng.normal(0, 1e-4, (1000, 3))).astype(np.float32)
    pts = np.concatenate([pts, near])
    d2b, cpb, trb = orc.bruteforce(vn, t, pts, nthreads=8)
    for cap in (16, 2):
        d2, cp, tri, ovf, nn, nt = orc.bvh4_query(nodes, recs, delta, pts, cap=cap)
        assert np.array_equal(bits(d2), bits(d2b))
        assert np.array_equal(tri, trb)
        assert np.array_equal(bits(cp), bits(cpb))
        assert ovf.sum() == 0


def test_host_bvh_layout(eng):
    v, t = eng.uv_sphere(16, 32, 1, [0.1])
    nodes, recs, delta = eng.build_host(v, t, leaf_size=4)
    child = nodes.reshape(-1, 128)[:, 96:112].copy().view(np.uint32).reshape(-1, 4)
    leaf = (child & 0x80000000) != 0
    counts = child[leaf] & 31
    firsts = (child[leaf] >> 5) & 0x3FFFFFF
    live = counts > 0
    assert counts[live].sum() == len(t) and counts.max() <= 4
    covered = np.zeros(len(t), int)
    for f, c in zip(firsts[live], counts[live]):
        covered[f:f + c] += 1
    assert np.all(covered == 1)  # leaves tile the record array exactly once
    ids = recs.reshape(-1, 48)[:, 36:40].copy().view(np.int32).ravel()
    assert np.array_equal(np.sort(ids), np.arange(len(t)))


def test_argument_errors(eng):
    lib = eng.lib()
    nv, nt = C.c_int(), C.c_int()
    assert lib.n2wos_uv_sphere_counts(3, 8, C.byref(nv), C.byref(nt)) == eng._lib.ERR_INVALID
    assert b"stacks" in lib.n2wos_last_error()
    v, t = eng.uv_sphere(8, 16, 1, [0.1])
    n = C.c_int64()
    rc = lib.n2wos_bvh_build_host(v.ctypes.data, len(v), t.ctypes.data, len(t), 0, None, None, C.byref(n), None, None)
    assert rc == eng._lib.ERR_INVALID
    rc = lib.n2wos_bvh_build_host(v.ctypes.data, len(v), t.ctypes.data, len(t), 32, None, None, C.byref(n), None, None)
    assert rc == eng._lib.ERR_INVALID
    bad = t.copy()
    bad[0, 0] = len(v)
    rc = lib.n2wos_bvh_build_host(v.ctypes.data, len(v), bad.ctypes.data, len(t), 8, None, None, C.byref(n), None, None)
    assert rc == eng._lib.ERR_INVALID and b"out-of-range" in lib.n2wos_last_error()
    rc = lib.n2wos_bvh_build_host(v.ctypes.data, len(v), t.ctypes.data, 0, 8, None, None, C.byref(n), None, None)
    assert rc == eng._lib.ERR_EMPTY
    rc = lib.n2wos_bvh_build_host(v.ctypes.data, len(v), t.ctypes.data, len(t), 8, b"radix", None, C.byref(n), None,
                                  None)
    assert rc == eng._lib.ERR_UNSUPPORTED and b"radix" in lib.n2wos_last_error()
    with pytest.raises(eng.N2wosError):
        eng.uv_sphere(64, 4)
    rc = lib.n2wos_bvh_query_device(None, None, 10, None, None, None, None, 128, None)
    assert rc == eng._lib.ERR_NULL
