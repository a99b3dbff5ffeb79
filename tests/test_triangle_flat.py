"""The kernels' branch-free closest point on a triangle equals the reference-
order branchy form (and the oracle) bit for bit, on random, region-boundary,
degenerate and non-finite inputs (CPU build of the product header)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "emu")
CSRC = os.path.join(os.path.dirname(HERE), "n2wos-cuda_amd", "csrc")
LIB = os.path.join(EMU, "_build", "libtri_flat.so")


@pytest.fixture(scope="module")
def tri():
    srcs = [os.path.join(EMU, "tri_flat.cpp"), os.path.join(CSRC, "geom.hpp")]
    if not (os.path.exists(LIB) and os.path.getmtime(LIB) >= max(os.path.getmtime(s) for s in srcs)):
        os.makedirs(os.path.dirname(LIB), exist_ok=True)
        tmp = f"{LIB}.{os.getpid()}.tmp"  # private name + rename: safe under parallel test processes
        subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-I" + EMU,
                        "-I" + CSRC, os.path.join(EMU, "tri_flat.cpp"), "-o", tmp], check=True)
        os.replace(tmp, LIB)
    L = C.CDLL(LIB)
    f = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    L.tri_closest_both.argtypes = [f, f, f, f, C.c_long, f, f]
    return L


def _run(tri, p, a, b, c):
    p, a, b, c = (np.ascontiguousarray(x, np.float32).reshape(-1, 3) for x in (p, a, b, c))
    n = len(p)
    r0 = np.zeros((n, 3), np.float32)
    r1 = np.zeros((n, 3), np.float32)
    tri.tri_closest_both(p, a, b, c, n, r0, r1)
    return r0, r1


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    a, b, c = (rng.normal(size=(n, 3)) for _ in range(3))
    # points spread over every Voronoi region: barycentric samples pushed out
    w = rng.uniform(-1.5, 2.5, size=(n, 3))
    w /= w.sum(1, keepdims=True)
    p = w[:, :1] * a + w[:, 1:2] * b + w[:, 2:] * c + rng.normal(scale=0.3, size=(n, 3))
    return p, a, b, c


def test_flat_equals_branchy_random(tri, orc):
    p, a, b, c = _inputs(200000, 1)
    r0, r1 = _run(tri, p, a, b, c)
    assert np.array_equal(r0.view(np.uint32), r1.view(np.uint32))
    # and the branchy form is the oracle's (reference-order) closest point
    k = np.arange(0, 200000, 997)
    for i in k:
        q = orc.closest_point_on_triangle(p[i], a[i], b[i], c[i])
        assert np.array_equal(np.asarray(q, np.float32).view(np.uint32), r0[i].view(np.uint32))


def test_flat_equals_branchy_region_boundaries_and_degenerate(tri):
    rng = np.random.default_rng(2)
    n = 100000
    # integer grid coordinates make the region tests hit their equalities
    a, b, c = (rng.integers(-3, 4, size=(n, 3)).astype(np.float32) for _ in range(3))
    p = rng.integers(-5, 6, size=(n, 3)).astype(np.float32) * 0.5
    # degenerate: collapsed edges / vertices and collinear triangles
    b[: n // 10] = a[: n // 10]
    c[n // 10: n // 5] = b[n // 10: n // 5]
    c[n // 5: n // 4] = 2 * b[n // 5: n // 4] - a[n // 5: n // 4]
    # points exactly at vertices
    p[n // 4: n // 3] = a[n // 4: n // 3]
    r0, r1 = _run(tri, p, a, b, c)
    assert np.array_equal(r0.view(np.uint32), r1.view(np.uint32))


def test_flat_equals_branchy_nonfinite_and_tiny(tri):
    p, a, b, c = _inputs(20000, 3)
    p[:1000, 0] = np.inf
    p[1000:2000, 1] = np.nan
    a[2000:3000] *= 1e-30
    b[3000:4000] = a[3000:4000] + 1e-7
    c[4000:5000] = 1e30
    r0, r1 = _run(tri, p, a, b, c)
    assert np.array_equal(r0.view(np.uint32), r1.view(np.uint32))
