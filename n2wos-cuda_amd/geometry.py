"""Meshes and the BVH closest-point API (host orchestration over the C ABI)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import BvhStats, check, lib

SPHERE_REFERENCE_BUMP = 0
SPHERE_SIN6 = 1
SPHERE_MULTI = 2
SPHERE_SH_BLOB = 3

NODE_BYTES = 128
TRI_BYTES = 48


def _ptr(a) -> int:
    return a.ctypes.data


def uv_sphere(stacks: int, slices: int, kind: int = SPHERE_SIN6, params=(0.1,)):
    """UV sphere with the reference's topology (procedural_bumpy_sphere.cpp:32-90)."""
    nv, nt = C.c_int(), C.c_int()
    check(lib().n2wos_uv_sphere_counts(stacks, slices, C.byref(nv), C.byref(nt)))
    prm = np.ascontiguousarray(params, np.float32)
    verts = np.zeros((nv.value, 3), np.float32)
    tris = np.zeros((nt.value, 3), np.uint32)
    check(lib().n2wos_uv_sphere(stacks, slices, kind, _ptr(prm), len(prm), _ptr(verts), _ptr(tris)))
    return verts, tris


def normalize_to_unit_radius(verts):
    """Returns (normalised copy, centre, scale) as normalize.cpp:18-37."""
    v = np.ascontiguousarray(verts, np.float32).copy()
    c = np.zeros(3, np.float32)
    s = np.zeros(1, np.float32)
    check(lib().n2wos_normalize_to_unit_radius(_ptr(v), len(v), _ptr(c), _ptr(s)))
    return v, c, float(s[0])


def count_degenerate_triangles(verts, tris, area2_eps: float = 1e-20) -> int:
    v = np.ascontiguousarray(verts, np.float32)
    t = np.ascontiguousarray(tris, np.uint32)
    return int(lib().n2wos_count_degenerate_triangles(_ptr(v), len(v), _ptr(t), len(t), area2_eps))


def load_mesh(path: str, kind: str | None = None):
    """(vertices (n, 3) float32, triangles (m, 3) uint32) of an OBJ or PLY file."""
    v, t = C.c_void_p(), C.c_void_p()
    nv, nt = C.c_int(), C.c_int()
    check(lib().n2wos_mesh_load(path.encode(), kind.encode() if kind else None, C.byref(v), C.byref(nv),
                                C.byref(t), C.byref(nt)))
    try:
        verts = np.ctypeslib.as_array(C.cast(v, C.POINTER(C.c_float)), (nv.value, 3)).copy()
        tris = np.ctypeslib.as_array(C.cast(t, C.POINTER(C.c_uint32)), (nt.value, 3)).copy()
    finally:
        lib().n2wos_mesh_free(v, t)
    return verts, tris


def closest_point_bruteforce(verts, tris, points, nthreads: int = 0):
    """Host brute force over every triangle: (d2, closest (n, 3), triangle id)."""
    v = np.ascontiguousarray(verts, np.float32)
    t = np.ascontiguousarray(tris, np.uint32)
    p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    d2 = np.empty(len(p), np.float32)
    cp = np.empty((len(p), 3), np.float32)
    tri = np.empty(len(p), np.int32)
    check(lib().n2wos_closest_point_bruteforce(_ptr(v), len(v), _ptr(t), len(t), _ptr(p), len(p), _ptr(d2),
                                               _ptr(cp), _ptr(tri), nthreads))
    return d2, cp, tri


def build_host(verts, tris, leaf_size: int = 8, build_method: str = "median"):
    """Node and triangle-record arrays exactly as uploaded (no device needed)."""
    v = np.ascontiguousarray(verts, np.float32)
    t = np.ascontiguousarray(tris, np.uint32)
    n = C.c_int64()
    delta = C.c_float()
    m = build_method.encode()
    check(lib().n2wos_bvh_build_host(_ptr(v), len(v), _ptr(t), len(t), leaf_size, m, None, C.byref(n), None, None))
    nodes = np.zeros(n.value * NODE_BYTES, np.uint8)
    recs = np.zeros(len(t) * TRI_BYTES, np.uint8)
    check(lib().n2wos_bvh_build_host(_ptr(v), len(v), _ptr(t), len(t), leaf_size, m, _ptr(nodes), C.byref(n),
                                     _ptr(recs), C.byref(delta)))
    return nodes, recs, delta.value


def _stream_ptr(stream):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    return getattr(stream, "cuda_stream", stream)


class Bvh:
    """Device-resident BVH; the drop-in for CudaBvh / CuBqlBvh (cuda_bvh.hpp:39-63)."""

    def __init__(self, verts, tris, leaf_size: int = 8, build_method: str | None = None):
        v = np.ascontiguousarray(verts, np.float32)
        t = np.ascontiguousarray(tris, np.uint32)
        h = C.c_void_p()
        check(lib().n2wos_bvh_create(_ptr(v), len(v), _ptr(t), len(t), leaf_size,
                                     build_method.encode() if build_method else None, C.byref(h)))
        self.handle = h
        self.triangle_count = len(t)

    def close(self):
        if getattr(self, "handle", None):
            lib().n2wos_bvh_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self) -> dict:
        s = BvhStats()
        check(lib().n2wos_bvh_get_stats(self.handle, C.byref(s)))
        d = {k: getattr(s, k) for k, _ in BvhStats._fields_}
        d["bounds_lo"] = list(s.bounds_lo)
        d["bounds_hi"] = list(s.bounds_hi)
        d["build_method"] = s.build_method.decode()
        return d

    def export(self):
        st = self.stats()
        nodes = np.zeros(st["node_count"] * NODE_BYTES, np.uint8)
        recs = np.zeros(st["triangle_count"] * TRI_BYTES, np.uint8)
        check(lib().n2wos_bvh_export(self.handle, _ptr(nodes), _ptr(recs)))
        return nodes, recs, st["prune_margin"]

    # ---------------------------------------------------- device-pointer API
    def query_device(self, d_points, n, d_distance2, d_closest, d_triangle_id, d_overflow, block_size=128,
                     stream=None, d_active=None):
        """Raw device-pointer query (ints are device addresses), as CudaBvh::query_device."""
        s = _stream_ptr(stream)
        if d_active is None:
            check(lib().n2wos_bvh_query_device(self.handle, d_points, n, d_distance2, d_closest, d_triangle_id,
                                               d_overflow, block_size, s))
        else:
            check(lib().n2wos_bvh_query_device_masked(self.handle, d_points, d_active, n, d_distance2, d_closest,
                                                      d_triangle_id, d_overflow, block_size, s))

    def query_tensors(self, points, active=None, block_size=128, counted=False):
        """Closest points for a (n, 3) float32 CUDA tensor; returns CUDA tensors."""
        import torch
        assert points.is_cuda and points.dtype == torch.float32 and points.is_contiguous()
        n = points.shape[0]
        dev = points.device
        d2 = torch.empty(n, dtype=torch.float32, device=dev)
        cp = torch.empty((n, 3), dtype=torch.float32, device=dev)
        tri = torch.empty(n, dtype=torch.int32, device=dev)
        ovf = torch.empty(n, dtype=torch.int32, device=dev)
        s = torch.cuda.current_stream(dev).cuda_stream
        if counted:
            nn = torch.empty(n, dtype=torch.int32, device=dev)
            nt = torch.empty(n, dtype=torch.int32, device=dev)
            check(lib().n2wos_bvh_query_device_counted(self.handle, points.data_ptr(), n, d2.data_ptr(),
                                                       cp.data_ptr(), tri.data_ptr(), ovf.data_ptr(),
                                                       nn.data_ptr(), nt.data_ptr(), block_size, s))
            return d2, cp, tri, ovf, nn, nt
        if active is None:
            check(lib().n2wos_bvh_query_device(self.handle, points.data_ptr(), n, d2.data_ptr(), cp.data_ptr(),
                                               tri.data_ptr(), ovf.data_ptr(), block_size, s))
        else:
            assert active.dtype == torch.uint8 and active.is_cuda
            check(lib().n2wos_bvh_query_device_masked(self.handle, points.data_ptr(), active.data_ptr(), n,
                                                      d2.data_ptr(), cp.data_ptr(), tri.data_ptr(),
                                                      ovf.data_ptr(), block_size, s))
        return d2, cp, tri, ovf

    def query_host(self, points_np, block_size=128):
        """The C ABI's host-memory query (n2wos_bvh_query_host; CudaBvh::query,
        bvh_query.cu:263-335): (d2, closest, tri, overflow_count, kernel_ms)."""
        pts = np.ascontiguousarray(points_np, np.float32).reshape(-1, 3)
        n = len(pts)
        d2 = np.zeros(n, np.float32)
        cp = np.zeros((n, 3), np.float32)
        tri = np.zeros(n, np.int32)
        ovf, ms = C.c_int(0), C.c_float(0.0)
        check(lib().n2wos_bvh_query_host(self.handle, pts.ctypes.data, n, d2.ctypes.data, cp.ctypes.data,
                                         tri.ctypes.data, C.byref(ovf), C.byref(ms), block_size))
        return d2, cp, tri, ovf.value, ms.value

    def query(self, points_np, block_size=128):
        """Host-vector convenience wrapper (CudaBvh::query, bvh_query.cu:263-335)."""
        import torch
        pts = torch.from_numpy(np.ascontiguousarray(points_np, np.float32).reshape(-1, 3)).cuda()
        d2, cp, tri, ovf = self.query_tensors(pts, block_size=block_size)
        torch.cuda.synchronize()
        return d2.cpu().numpy(), cp.cpu().numpy(), tri.cpu().numpy(), ovf.cpu().numpy()
