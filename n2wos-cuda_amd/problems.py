"""The benchmark problems as concrete synthetic inputs.

Mesh 1: UV sphere stacks=64, slices=128, r = 1 + 0.1 sin(6 theta) sin(6 phi)
        (16,128 triangles), normalised to a centred unit radius.
Mesh 2: the same radius function at stacks=128, slices=256 (65,024 triangles).
Mesh 3: stacks=512, slices=1024, r = 1 + 0.06 sin5t sin7p + 0.03 sin17t cos13p
        (1,046,528 triangles).
Blob:   stacks=256, slices=512 (261,120 triangles), r = 1 + sum_{l<=8} a_lm Y_lm,
        a_lm ~ N(0, 0.02^2) drawn with numpy PCG64(seed).
Charges: 8 point charges c_i uniform on the sphere of radius 2, q_i ~ N(0, 1),
        numpy PCG64(seed); g(x) = sum q_i / |x - c_i| is harmonic inside, so the
        FP64 solution is g itself.
Grid points (problem 2): the 512 x 512 cell centres of [-1, 1]^2 on z = 0 that
        lie inside the mesh.
Random interior points: uniform in the mesh bounding box (numpy PCG64(seed)),
        kept when inside (ray-parity test on the GPU).
"""
from __future__ import annotations

import numpy as np

from . import geometry as G
from .wos import BC_CHARGES, BC_POISSON_SINES, Boundary, SOURCE_NONE, SOURCE_POISSON_SINES

MESHES = {
    "mesh1": (64, 128, G.SPHERE_SIN6, (0.1,)),
    "mesh2": (128, 256, G.SPHERE_SIN6, (0.1,)),
    "mesh3": (512, 1024, G.SPHERE_MULTI, (0.0,)),
    "blob": (256, 512, G.SPHERE_SH_BLOB, None),
}


def sh_blob_coefficients(seed: int = 0) -> np.ndarray:
    return np.random.default_rng(seed).normal(0.0, 0.02, 81).astype(np.float32)


def make_mesh(name: str, seed: int = 0):
    """Returns (normalised vertices, triangles)."""
    stacks, slices, kind, params = MESHES[name]
    if params is None:
        params = sh_blob_coefficients(seed)
    v, t = G.uv_sphere(stacks, slices, kind, params)
    vn, _, _ = G.normalize_to_unit_radius(v)
    return vn, t


def charges(seed: int = 0, n: int = 8, radius: float = 2.0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d = radius * d / np.linalg.norm(d, axis=1, keepdims=True)
    q = rng.normal(size=n)
    return np.concatenate([d, q[:, None]], axis=1).astype(np.float32)


def charges_boundary(seed: int = 0) -> Boundary:
    return Boundary(BC_CHARGES, charges(seed))


def poisson_boundary() -> Boundary:
    return Boundary(BC_POISSON_SINES)


def charges_solution_f64(ch: np.ndarray, pts: np.ndarray) -> np.ndarray:
    """FP64 analytic solution sum q / |x - c| at (n, 3) points."""
    p = np.asarray(pts, np.float64)
    out = np.zeros(len(p))
    for cx, cy, cz, q in np.asarray(ch, np.float64):
        out += q / np.sqrt((p[:, 0] - cx) ** 2 + (p[:, 1] - cy) ** 2 + (p[:, 2] - cz) ** 2)
    return out


def poisson_solution_f64(pts: np.ndarray) -> np.ndarray:
    p = np.asarray(pts, np.float64)
    s = np.sin(2 * np.pi * p[:, 0]) * np.sin(2 * np.pi * p[:, 1]) * np.sin(2 * np.pi * p[:, 2])
    return s + 0.5 * (p * p).sum(1)


def uniform_box_queries(n: int = 65536, seed: int = 0, half: float = 1.2) -> np.ndarray:
    return np.random.default_rng(seed).uniform(-half, half, (n, 3)).astype(np.float32)


def inside_mask(bvh, pts: np.ndarray) -> np.ndarray:
    import torch
    from ._lib import check, lib
    tp = torch.from_numpy(np.ascontiguousarray(pts, np.float32)).cuda()
    out = torch.empty(len(pts), dtype=torch.uint8, device="cuda")
    check(lib().n2wos_bvh_inside_device(bvh.handle, tp.data_ptr(), len(pts), out.data_ptr(),
                                        torch.cuda.current_stream().cuda_stream))
    return out.cpu().numpy().astype(bool)


def interior_points(bvh, verts: np.ndarray, n: int, seed: int = 0) -> np.ndarray:
    """n points uniform in the mesh's interior (rejection in the bounding box)."""
    rng = np.random.default_rng(seed)
    lo, hi = verts.min(0), verts.max(0)
    out = []
    have = 0
    while have < n:
        cand = rng.uniform(lo, hi, (max(2 * (n - have), 1024), 3)).astype(np.float32)
        keep = cand[inside_mask(bvh, cand)]
        out.append(keep)
        have += len(keep)
    return np.concatenate(out)[:n]


def grid_slice_points(bvh, res: int = 512, half: float = 1.0, z: float = 0.0):
    """Interior cell centres of a res x res grid on [-half, half]^2 at height z;
    returns (points, flat cell indices)."""
    c = (np.arange(res, dtype=np.float64) + 0.5) * (2 * half / res) - half
    X, Y = np.meshgrid(c, c, indexing="xy")
    pts = np.stack([X.ravel(), Y.ravel(), np.full(X.size, z)], axis=1).astype(np.float32)
    m = inside_mask(bvh, pts)
    return pts[m], np.nonzero(m)[0]


def surface_shell_queries(verts, tris, n: int, seed: int = 0, dmin: float = 1e-5, dmax: float = 0.02):
    """Area-weighted surface points offset along the face normal by +/- a
    log-uniform distance in [dmin, dmax] (near-boundary shell, |d| < dmax)."""
    rng = np.random.default_rng(seed)
    a, b, c = (verts[tris[:, k]].astype(np.float64) for k in range(3))
    cr = np.cross(b - a, c - a)
    area = np.linalg.norm(cr, axis=1)
    idx = rng.choice(len(tris), size=n, p=area / area.sum())
    u, v = rng.uniform(size=(2, n))
    flip = u + v > 1
    u[flip], v[flip] = 1 - u[flip], 1 - v[flip]
    p = a[idx] + u[:, None] * (b - a)[idx] + v[:, None] * (c - a)[idx]
    nrm = cr[idx] / area[idx, None]
    off = np.exp(rng.uniform(np.log(dmin), np.log(dmax), n)) * rng.choice([-1.0, 1.0], n)
    return (p + off[:, None] * nrm).astype(np.float32)


def wos_prefix_queries(bvh, start: np.ndarray, n: int, seed: int = 0, max_steps: int = 20,
                       step_scale: float = 0.999) -> np.ndarray:
    """Positions after k ~ U{1..max_steps} WoS steps from interior start points
    (uniform directions, numpy PCG64(seed))."""
    import torch
    rng = np.random.default_rng(seed)
    p = start[rng.integers(0, len(start), n)].astype(np.float32)
    k = rng.integers(1, max_steps + 1, n)
    for s in range(max_steps):
        d2, _, _, _ = bvh.query_tensors(torch.from_numpy(p).cuda())
        d = np.sqrt(np.maximum(d2.cpu().numpy(), 0.0))
        dirs = rng.normal(size=(n, 3))
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
        move = (s < k)[:, None] * (step_scale * d)[:, None] * dirs
        p = (p + move).astype(np.float32)
    return p
