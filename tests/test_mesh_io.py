"""Mesh files and the host brute force (CPU): the engine's OBJ / PLY loaders
return exactly what the reference's loaders return (built from the
reference's own sources where its tree is present) on files covering fan
triangulation, negative and slashed OBJ indices, extra PLY properties and
elements, ascii and binary little-endian, and reject what they reject; the
host brute force is bit-identical to the oracle's; the probe CLI's argument
errors exit with code 1."""
import ctypes as C
import os
import struct
import subprocess

import numpy as np
import pytest

BUNNY = "/root/reference/meshes/bunny_zipper_wataertight.ply"


def ref_load(orc, path, is_ply):
    R = orc.ref()
    nv, nt = C.c_int(), C.c_int()
    if R.ref_load_mesh(path.encode(), int(is_ply), C.byref(nv), C.byref(nt), None, None) != 0:
        return None
    v = np.zeros((nv.value, 3), np.float32)
    t = np.zeros((nt.value, 3), np.uint32)
    assert R.ref_load_mesh(path.encode(), int(is_ply), C.byref(nv), C.byref(nt), v.ctypes.data, t.ctypes.data) == 0
    return v, t


def eng_load(eng, path):
    from n2wos_amd import N2wosError
    from n2wos_amd.geometry import load_mesh
    try:
        return load_mesh(path)
    except N2wosError:
        return None


OBJ_GOOD = """# cube-ish with a quad, a pentagon, slashed and negative indices
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
vt 0 0
vn 0 0 1
f 1 2 3 4
v 0.5 0.5 1.25
f 1/1/1 2/1/1 -1
f -5//1 -3 -1
o ignored
f 2 3 5 4 1
"""

PLY_ASCII = """ply
format ascii 1.0
comment extra properties and a leading element
element material 2
property float shininess
property list uchar int tags
element vertex 5
property float nx
property float x
property float y
property float z
property uchar red
element face 3
property uchar flags
property list uchar int vertex_index
end_header
0.5 2 7 8
1.5 0
0 0 0 0 9
0 1 0 0 9
0 1 1 0 9
0 0 1 0 9
0 0.5 0.5 1.25 9
1 4 0 1 2 3
0 3 0 1 4
2 5 1 2 4 3 0
"""


def write_ply_binary(path):
    verts = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0.5, 0.5, 1.25]], np.float64)
    faces = [[0, 1, 2, 3], [0, 1, 4], [1, 2, 4, 3, 0]]
    head = ("ply\nformat binary_little_endian 1.0\nelement vertex 5\nproperty double x\nproperty double y\n"
            "property double z\nproperty list uchar float junk\nelement face 3\n"
            "property list ushort uint vertex_indices\nproperty short extra\nend_header\n")
    body = b""
    for k, v in enumerate(verts):
        body += struct.pack("<ddd", *v) + struct.pack("<B", k % 2) + struct.pack("<f", 1.0) * (k % 2)
    for f in faces:
        body += struct.pack("<H", len(f)) + struct.pack("<%dI" % len(f), *f) + struct.pack("<h", -3)
    with open(path, "wb") as fh:
        fh.write(head.encode() + body)


BAD = {
    "zero_index.obj": "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 0 1 2\n",
    "out_of_range.obj": "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n",
    "two_vertices.obj": "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2\n",
    "no_faces.obj": "v 0 0 0\nv 1 0 0\nv 0 1 0\n",
    "big_endian.ply": "ply\nformat binary_big_endian 1.0\nelement vertex 0\nend_header\n",
    "no_x.ply": ("ply\nformat ascii 1.0\nelement vertex 3\nproperty float y\nproperty float z\n"
                 "element face 1\nproperty list uchar int vertex_indices\nend_header\n0 0\n0 0\n0 0\n3 0 1 2\n"),
    "bad_index.ply": ("ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\n"
                      "property float z\nelement face 1\nproperty list uchar int vertex_indices\nend_header\n"
                      "0 0 0\n1 0 0\n0 1 0\n3 0 1 3\n"),
    "not_ply.ply": "plx\n",
}


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("meshes")
    out = {}
    for name, text in [("good.obj", OBJ_GOOD), ("ascii.ply", PLY_ASCII)] + list(BAD.items()):
        p = str(d / name)
        with open(p, "w") as fh:
            fh.write(text)
        out[name] = p
    out["binary.ply"] = str(d / "binary.ply")
    write_ply_binary(out["binary.ply"])
    return out


def test_loaders_accept_like_the_reference(eng, orc, files):
    for name in ("good.obj", "ascii.ply", "binary.ply"):
        v, t = eng_load(eng, files[name])
        assert len(t) >= 3
        if orc.ref_available():
            rv, rt = ref_load(orc, files[name], name.endswith(".ply"))
            assert np.array_equal(v.view(np.uint32), rv.view(np.uint32)), name
            assert np.array_equal(t, rt), name
    # closed-form expectations (independent of the reference build)
    v, t = eng_load(eng, files["good.obj"])
    assert t.tolist() == [[0, 1, 2], [0, 2, 3], [0, 1, 4], [0, 2, 4], [1, 2, 4], [1, 4, 3], [1, 3, 0]]
    v2, t2 = eng_load(eng, files["binary.ply"])
    v3, t3 = eng_load(eng, files["ascii.ply"])
    assert np.array_equal(v2, v3) and np.array_equal(t2, t3)
    assert t2.tolist() == [[0, 1, 2], [0, 2, 3], [0, 1, 4], [1, 2, 4], [1, 4, 3], [1, 3, 0]]


def test_loaders_reject_like_the_reference(eng, orc, files):
    for name in BAD:
        assert eng_load(eng, files[name]) is None, name
        if orc.ref_available():
            assert ref_load(orc, files[name], name.endswith(".ply")) is None, name


@pytest.mark.skipif(not os.path.exists(BUNNY), reason="reference mesh tree not present")
def test_bunny_matches_reference_loader(eng, orc):
    v, t = eng_load(eng, BUNNY)
    assert len(t) > 60000
    if orc.ref_available():
        rv, rt = ref_load(orc, BUNNY, True)
        assert np.array_equal(v.view(np.uint32), rv.view(np.uint32))
        assert np.array_equal(t, rt)


def test_host_bruteforce_bit_exact_vs_oracle(eng, orc):
    from n2wos_amd import problems as P
    from n2wos_amd.geometry import closest_point_bruteforce
    v, t = P.make_mesh("mesh1")
    q = P.uniform_box_queries(3000, seed=5)
    d2, cp, tri = closest_point_bruteforce(v, t, q, nthreads=8)
    od2, ocp, otri = orc.bruteforce(v, t, q, nthreads=8)
    assert np.array_equal(d2.view(np.uint32), od2.view(np.uint32))
    assert np.array_equal(cp.view(np.uint32), ocp.view(np.uint32))
    assert np.array_equal(tri, otri)


def test_probe_cli_argument_errors(eng):
    exe = os.path.join(os.path.dirname(eng.LIB_PATH), "..", "bin", "n2wos_probe_cuda_bvh")
    assert os.path.exists(exe), "probe CLI not built"
    r = subprocess.run([exe, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "--bumpy-stacks" in r.stdout
    for args in (["--mesh", "cube"], ["--queries", "0"], ["--mesh", "obj"], ["--frobnicate"], ["--repeat"]):
        r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=60)
        assert r.returncode == 1 and r.stderr.startswith("error:"), args
