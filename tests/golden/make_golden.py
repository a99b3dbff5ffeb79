"""Writes the golden closest-point fixtures from the reference's own CPU core.

Run in a container where the reference tree is present and oracle/_ref was
built from it (``make -C oracle``).  The reference code is compiled from its
sources and run here; only the resulting data (inputs and outputs) is
committed, as closest_point_ref.npz beside this script.

Contents (all produced by the reference functions make_procedural_bumpy_sphere,
normalize_to_unit_radius and closest_point_bruteforce):
  mesh_ref_verts_sha256  sha256 of the normalised 64x128 reference bumpy sphere
                         (amplitude 0.15) vertex bytes, float32 little endian
  mesh_ref_tris_sha256   sha256 of its uint32 triangle indices
  center, scale          its normalisation transform
  pts, d2, cp, tri       2048 queries uniform in [-1.2, 1.2]^3 (numpy PCG64
                         seed 0) against the SIN6 benchmark mesh 64x128, plus
                         512 near-surface queries, with the reference brute
                         force's squared distance, closest point and triangle
  tri_cases              hand-picked point/triangle cases (p, a, b, c) and the
                         reference closest_point_on_triangle output
"""
import ctypes as C
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle as O  # noqa: E402


def main():
    R = O.ref()
    nv, nt = C.c_int(), C.c_int()
    R.ref_bumpy_sphere(64, 128, 0.15, 1, C.byref(nv), C.byref(nt), None, None, None, None)
    v = np.zeros((nv.value, 3), np.float32)
    t = np.zeros((nt.value, 3), np.uint32)
    c = np.zeros(3, np.float32)
    s = np.zeros(1, np.float32)
    R.ref_bumpy_sphere(64, 128, 0.15, 1, C.byref(nv), C.byref(nt), v.ctypes.data, t.ctypes.data, c.ctypes.data,
                       s.ctypes.data)
    ref_v_hash = hashlib.sha256(v.astype("<f4").tobytes()).hexdigest()
    ref_t_hash = hashlib.sha256(t.astype("<u4").tobytes()).hexdigest()

    # benchmark mesh 1 geometry: SIN6 sphere from the oracle restatement, then
    # normalised by the reference's own normalize_to_unit_radius
    mv, mt = O.uv_sphere(64, 128, 1, [0.1])
    mc = np.zeros(3, np.float32)
    ms = C.c_float()
    R.ref_normalize(mv, len(mv), mt, len(mt), mc, C.byref(ms))
    rng = np.random.default_rng(0)
    pts = rng.uniform(-1.2, 1.2, (2048, 3)).astype(np.float32)
    d2 = np.zeros(len(pts), np.float32)
    cp = np.zeros((len(pts), 3), np.float32)
    tri = np.zeros(len(pts), np.int32)
    R.ref_bruteforce(mv, len(mv), mt, len(mt), pts, len(pts), d2, cp, tri)
    near = (cp[:512] + rng.normal(0.0, 1e-3, (512, 3))).astype(np.float32)
    pts2 = np.concatenate([pts, near])
    d2 = np.zeros(len(pts2), np.float32)
    cp = np.zeros((len(pts2), 3), np.float32)
    tri = np.zeros(len(pts2), np.int32)
    R.ref_bruteforce(mv, len(mv), mt, len(mt), pts2, len(pts2), d2, cp, tri)

    a, b, cc = np.array([0, 0, 0], np.float32), np.array([1, 0, 0], np.float32), np.array([0, 1, 0], np.float32)
    probes = [[-1, -1, 1], [2, -0.5, 0.3], [-0.5, 2, -0.2], [0.5, -1, 0.1], [-1, 0.5, 0.7], [1, 1, 0.5],
              [0.25, 0.25, 2], [0.3, 0.1, -0.4], [1e-7, 1e-7, 1e-7], [0.5, 0.5, 0.0]]
    cases = []
    for p in probes:
        p = np.array(p, np.float32)
        out = np.zeros(3, np.float32)
        R.ref_closest_point_on_triangle(p, a, b, cc, out)
        cases.append(np.concatenate([p, a, b, cc, out]))
    skew = rng.uniform(-1, 1, (64, 4, 3)).astype(np.float32)
    for p, x, y, z in skew:
        out = np.zeros(3, np.float32)
        R.ref_closest_point_on_triangle(p, x, y, z, out)
        cases.append(np.concatenate([p, x, y, z, out]))
    np.savez_compressed(os.path.join(HERE, "closest_point_ref.npz"), mesh_ref_verts_sha256=ref_v_hash,
                        mesh_ref_tris_sha256=ref_t_hash, center=c, scale=s, sin6_center=mc,
                        sin6_scale=np.float32(ms.value), pts=pts2, d2=d2, cp=cp, tri=tri,
                        tri_cases=np.array(cases, np.float32))
    print("wrote closest_point_ref.npz", ref_v_hash[:12], len(pts2), "queries")


if __name__ == "__main__":
    main()
