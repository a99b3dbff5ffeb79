"""The oracle pinned: against the reference's own compiled CPU core, committed
golden vectors, published known-answer tests and closed forms."""
import ctypes as C
import hashlib
import math
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "closest_point_ref.npz")


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN, allow_pickle=False)


@pytest.fixture(scope="module")
def c1_mesh(orc):
    v, t = orc.uv_sphere(64, 128, 1, [0.1])
    vn, c, s = orc.normalize(v)
    return vn, t, c, s


def test_reference_bumpy_sphere_matches_golden_hash(orc, golden):
    v, t = orc.uv_sphere(64, 128, 0, [0.15])
    vn, c, s = orc.normalize(v)
    assert hashlib.sha256(vn.astype("<f4").tobytes()).hexdigest() == str(golden["mesh_ref_verts_sha256"])
    assert hashlib.sha256(t.astype("<u4").tobytes()).hexdigest() == str(golden["mesh_ref_tris_sha256"])
    assert np.array_equal(bits(c), bits(golden["center"]))
    assert np.float32(s) == golden["scale"][0]


def test_sin6_normalisation_matches_golden(orc, golden, c1_mesh):
    vn, t, c, s = c1_mesh
    assert np.array_equal(bits(c), bits(golden["sin6_center"]))
    assert np.float32(s) == golden["sin6_scale"]


def test_bruteforce_matches_reference_golden_bit_exact(orc, golden, c1_mesh):
    vn, t, _, _ = c1_mesh
    d2, cp, tri = orc.bruteforce(vn, t, golden["pts"], nthreads=4)
    assert np.array_equal(bits(d2), bits(golden["d2"]))
    assert np.array_equal(bits(cp), bits(golden["cp"]))
    assert np.array_equal(tri, golden["tri"])


def test_closest_point_on_triangle_golden_cases(orc, golden):
    for row in golden["tri_cases"]:
        p, a, b, c, want = row[0:3], row[3:6], row[6:9], row[9:12], row[12:15]
        assert np.array_equal(bits(orc.closest_point_on_triangle(p, a, b, c)), bits(want))


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(os.path.dirname(__file__)), "oracle", "_ref",
                                                    "libn2wos_ref.so")), reason="reference core not built here")
def test_oracle_against_compiled_reference_core(orc):
    R = orc.ref()
    for stacks, slices, amp in [(4, 8, 0.0), (16, 32, 0.5), (64, 128, 0.15), (33, 77, 0.3)]:
        nv, nt = C.c_int(), C.c_int()
        R.ref_bumpy_sphere(stacks, slices, amp, 1, C.byref(nv), C.byref(nt), None, None, None, None)
        rv = np.zeros((nv.value, 3), np.float32)
        rt = np.zeros((nt.value, 3), np.uint32)
        rc = np.zeros(3, np.float32)
        rs = np.zeros(1, np.float32)
        R.ref_bumpy_sphere(stacks, slices, amp, 1, C.byref(nv), C.byref(nt), rv.ctypes.data, rt.ctypes.data,
                           rc.ctypes.data, rs.ctypes.data)
        v, t = orc.uv_sphere(stacks, slices, 0, [amp])
        vn, c, s = orc.normalize(v)
        assert np.array_equal(t, rt) and np.array_equal(bits(vn), bits(rv)) and np.float32(s) == rs[0]
        rng = np.random.default_rng(stacks)
        pts = rng.uniform(-1.5, 1.5, (300, 3)).astype(np.float32)
        d2r = np.zeros(300, np.float32)
        cpr = np.zeros((300, 3), np.float32)
        trr = np.zeros(300, np.int32)
        R.ref_bruteforce(vn, len(vn), t, len(t), pts, 300, d2r, cpr, trr)
        d2, cp, tri = orc.bruteforce(vn, t, pts)
        assert np.array_equal(bits(d2), bits(d2r)) and np.array_equal(bits(cp), bits(cpr))
        assert np.array_equal(tri, trr)
        assert orc.lib().orc_count_degenerate(vn, t, len(t), 1e-20) == R.ref_count_degenerate(vn, len(vn), t, len(t),
                                                                                               1e-20)


def test_triangle_regions_closed_form(orc):
    a, b, c = np.array([0, 0, 0.]), np.array([2, 0, 0.]), np.array([0, 2, 0.])
    # face interior: projection, distance = |z|
    assert np.allclose(orc.closest_point_on_triangle([0.5, 0.5, 3.0], a, b, c), [0.5, 0.5, 0.0])
    # vertex regions
    assert np.allclose(orc.closest_point_on_triangle([-1, -1, 1], a, b, c), a)
    assert np.allclose(orc.closest_point_on_triangle([3, -1, 0], a, b, c), b)
    assert np.allclose(orc.closest_point_on_triangle([-1, 3, 0], a, b, c), c)
    # edge regions: ab, ac, bc
    assert np.allclose(orc.closest_point_on_triangle([1, -1, 0], a, b, c), [1, 0, 0])
    assert np.allclose(orc.closest_point_on_triangle([-1, 1, 0], a, b, c), [0, 1, 0])
    assert np.allclose(orc.closest_point_on_triangle([2, 2, 0], a, b, c), [1, 1, 0])
    # degenerate (collinear) triangle: result lies on the segment
    q = orc.closest_point_on_triangle([0.5, 1.0, 0.0], [0, 0, 0.], [1, 0, 0.], [2, 0, 0.])
    assert abs(q[1]) < 1e-7 and 0.0 <= q[0] <= 2.0


def test_closest_point_not_farther_than_any_sample(orc):
    """Brute-force invariant on tiny inputs: the returned point is on the
    triangle (barycentric) and no sampled triangle point is closer."""
    rng = np.random.default_rng(7)
    for _ in range(200):
        p, a, b, c = rng.normal(size=(4, 3)).astype(np.float32)
        q = orc.closest_point_on_triangle(p, a, b, c).astype(np.float64)
        u, v = np.meshgrid(np.linspace(0, 1, 41), np.linspace(0, 1, 41))
        m = (u + v) <= 1
        s = a[None] + u[m][:, None] * (b - a)[None] + v[m][:, None] * (c - a)[None]
        dmin = np.sqrt(((s - p) ** 2).sum(1)).min()
        assert np.linalg.norm(q - p) <= dmin + 1e-5


def test_bruteforce_lowest_index_wins_ties(orc):
    # the same triangle twice: the tie goes to index 0; a permuted copy keeps it
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    t = np.array([[0, 1, 2], [0, 1, 2], [1, 2, 0]], np.uint32)
    d2, cp, tri = orc.bruteforce(v, t, [[0.2, 0.2, 1.0]])
    assert tri[0] == 0


def test_philox_known_answers(orc):
    # Random123 kat_vectors for philox4x32_10
    assert [hex(x) for x in orc.philox([0, 0, 0, 0], [0, 0])] == ["0x6627e8d5", "0xe169c58d", "0xbc57ac4c",
                                                                   "0x9b00dbd8"]
    assert [hex(x) for x in orc.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2)] == ["0x408f276d", "0x41c83b0e",
                                                                                "0xa20bc7c6", "0x6d5451fd"]
    assert [hex(x) for x in orc.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                                       [0xA4093822, 0x299F31D0])] == ["0xd16cfe09", "0x94fdcceb", "0x5001e420",
                                                                      "0x24126ea1"]


def test_u01_open_interval_and_exact(orc):
    L = orc.lib()
    assert L.orc_u01(0) == 2.0 ** -24
    assert L.orc_u01(0xFFFFFFFF) == 1.0 - 2.0 ** -24
    assert 0.0 < L.orc_u01(0x12345678) < 1.0


def test_sincos2pi_accuracy(orc):
    vs = np.concatenate([np.linspace(-2, 2, 4001), np.random.default_rng(1).uniform(0, 1, 4000)]).astype(np.float32)
    worst = 0.0
    for v in vs:
        s, c = orc.sincos2pi(float(v))
        ang = 2 * math.pi * float(v)
        worst = max(worst, abs(s - math.sin(ang)), abs(c - math.cos(ang)))
        assert abs(s * s + c * c - 1.0) < 5e-7
    assert worst < 4e-7


def test_sphere_direction_moments(orc):
    """E[d] = 0 and E[d d^T] = I/3 within 5 sigma for the walk's direction map."""
    n = 20000
    dirs = np.zeros((n, 3))
    L = orc.lib()
    for i in range(n):
        r = orc.philox([i, 0, 0, 0], [1, 2])
        z = 1.0 - 2.0 * L.orc_u01(int(r[0]))
        s, c = orc.sincos2pi(L.orc_u01(int(r[1])))
        rr = math.sqrt(max(0.0, 1 - z * z))
        dirs[i] = [rr * c, rr * s, z]
    assert np.all(np.abs(dirs.mean(0)) < 5 * math.sqrt(1 / 3 / n))
    cov = dirs.T @ dirs / n
    assert np.all(np.abs(cov - np.eye(3) / 3) < 5 * 0.3 / math.sqrt(n))


def test_spherical_harmonics_orthonormal(orc):
    L = orc.lib()
    th = (np.arange(60) + 0.5) * math.pi / 60
    ph = (np.arange(120) + 0.5) * 2 * math.pi / 120
    T, P = np.meshgrid(th, ph, indexing="ij")
    w = np.sin(T) * (math.pi / 60) * (2 * math.pi / 120)
    lm = [(0, 0), (1, -1), (1, 0), (2, 1), (3, -2), (5, 4), (8, -8), (8, 3)]
    Y = {k: np.vectorize(lambda t, p, k=k: L.orc_sh_real(k[0], k[1], t, p))(T, P) for k in lm}
    for i, a in enumerate(lm):
        for b in lm[i:]:
            ip = (Y[a] * Y[b] * w).sum()
            assert abs(ip - (1.0 if a == b else 0.0)) < 2e-3
    assert abs(L.orc_sh_real(0, 0, 0.3, 0.2) - 0.5 / math.sqrt(math.pi)) < 1e-12
    assert abs(L.orc_sh_real(1, 0, 0.3, 0.2) - math.sqrt(3 / (4 * math.pi)) * math.cos(0.3)) < 1e-12


def test_bvh2_equals_bruteforce(orc, c1_mesh):
    vn, t, _, _ = c1_mesh
    rng = np.random.default_rng(3)
    pts = rng.uniform(-1.2, 1.2, (3000, 3)).astype(np.float32)
    b = orc.Bvh2(vn, t)
    d2b, cpb, trb = orc.bruteforce(vn, t, pts, nthreads=4)
    d2, cp, tri = b.query(pts, nthreads=4)
    assert np.array_equal(bits(d2), bits(d2b)) and np.array_equal(tri, trb) and np.array_equal(bits(cp), bits(cpb))


def test_welford_and_chan_against_two_pass(orc):
    rng = np.random.default_rng(5)
    x = rng.normal(3.0, 0.5, 1000).astype(np.float32)
    m, s = C.c_double(), C.c_double()
    orc.lib().orc_welford(x, len(x), C.byref(m), C.byref(s))
    xd = x.astype(np.float64)
    assert abs(m.value - xd.mean()) < 1e-12 and abs(s.value - ((xd - xd.mean()) ** 2).sum()) < 1e-9
    out = np.zeros(3)
    a, b = xd[:300], xd[300:]
    orc.lib().orc_chan_merge(300, a.mean(), ((a - a.mean()) ** 2).sum(), 700, b.mean(), ((b - b.mean()) ** 2).sum(),
                             out)
    assert out[0] == 1000 and abs(out[1] - xd.mean()) < 1e-12 and abs(out[2] - s.value) < 1e-9
    n, mean, m2 = orc.point_stats(x, 4, 250)
    for p in range(4):
        seg = xd[250 * p: 250 * (p + 1)]
        assert n[p] == 250 and abs(mean[p] - seg.mean()) < 1e-12 and abs(m2[p] - ((seg - seg.mean()) ** 2).sum()) < 1e-9


def _mlp_weights(rng, h, layers):
    """Random weights of an h-wide, layers-deep cache MLP in the product's
    layout: A_1 [h x 32], A_2 .. A_L [h x h], A_out [16 x h] (row 0 only)."""
    n = 32 * h + (layers - 1) * h * h + 16 * h
    w = np.zeros(n, np.float32)
    w[:32 * h] = rng.normal(0, 0.3, 32 * h)
    w[32 * h:32 * h + (layers - 1) * h * h] = rng.normal(0, 1.6 / np.sqrt(h), (layers - 1) * h * h)
    out = 32 * h + (layers - 1) * h * h
    w[out:out + h] = rng.normal(0, 0.3, h)
    return w


@pytest.mark.parametrize("h,layers", [(64, 4), (16, 1), (32, 2), (64, 3)])
def test_cache_oracle_gradients_match_finite_differences(orc, h, layers):
    """Central differences of the FP64 cache oracle's loss pin its reverse-mode
    gradients (grid entries and weights of every layer), for each MLP shape
    the cache supports; the forward equals a numpy matrix chain over the
    oracle's hash-grid features in the documented weight layout (a transposed
    or misplaced layer fails it)."""
    from types import SimpleNamespace
    cfg = SimpleNamespace(n_levels=3, log2_table_size=8, base_resolution=2, growth=2.0, origin=(-1, -1, -1),
                          extent=(2, 2, 2), n_neurons=h, n_hidden_layers=layers)
    lay = orc.cache_layout(3, 8, 2, 2.0)
    rng = np.random.default_rng(h + layers)
    grid = rng.normal(0, 0.5, 2 * lay["total_entries"]).astype(np.float32)
    w = _mlp_weights(rng, h, layers)
    pts = rng.uniform(-0.9, 0.9, (7, 3)).astype(np.float32)
    tg = rng.normal(size=7).astype(np.float32)
    gg, wg, loss = orc.cache_grads(cfg, grid, w, pts, tg)
    assert wg.shape == w.shape

    def L(g, ww):
        y = orc.cache_forward(cfg, g, ww, pts)
        return np.mean((y - tg.astype(np.float64)) ** 2)

    # forward: y = A_out[0] relu(A_L ... relu(A_1 x))
    x = orc.cache_encode(cfg, grid, pts).astype(np.float64)
    wd = w.astype(np.float64)
    a = np.maximum(x @ wd[:32 * h].reshape(h, 32).T, 0)
    off = 32 * h
    for _ in range(layers - 1):
        a = np.maximum(a @ wd[off:off + h * h].reshape(h, h).T, 0)
        off += h * h
    y = a @ wd[off:off + h]
    assert np.allclose(orc.cache_forward(cfg, grid, w, pts), y, rtol=1e-12, atol=1e-12)
    assert abs(L(grid, w) - loss) < 1e-12
    # steps of 1e-4 (divided by the FP32 step actually taken) keep clear of
    # the ReLU kinks
    eps = 1e-4
    for k in list(rng.choice(np.nonzero(gg)[0], 6, replace=False)):
        gp, gm = grid.copy(), grid.copy()
        gp[k] += eps
        gm[k] -= eps
        fd = (L(gp, w) - L(gm, w)) / (float(gp[k]) - float(gm[k]))
        assert abs(fd - gg[k]) < 1e-4 * max(1.0, abs(fd))
    picks = [5, 32 * h - 3, off + 3] + [32 * h + j * h * h + 7 * j + 11 for j in range(layers - 1)]
    for k in picks:
        wp, wm = w.copy(), w.copy()
        wp[k] += eps
        wm[k] -= eps
        fd = (L(grid, wp) - L(grid, wm)) / (float(wp[k]) - float(wm[k]))
        assert abs(fd - wg[k]) < 1e-4 * max(1.0, abs(fd))
    assert np.all(wg[off + h:] == 0)  # padded rows of the output layer


def test_cache_oracle_layout_closed_form(orc):
    lay = orc.cache_layout(16, 19, 16, 1.5)
    import math
    for l in range(16):
        n = math.floor(16 * 1.5 ** l)
        assert lay["res"][l] == n
        assert lay["dense"][l] == int((n + 1) ** 3 <= 2 ** 19)
    assert lay["dense"][:4] == [1, 1, 1, 1] and lay["dense"][4] == 0
    assert lay["total_entries"] == sum(lay["sizes"])


def test_half_rounding_matches_numpy_float16(orc):
    """The oracle's binary16 rounding (used by its FP16-rounding cache variant)
    equals numpy's correctly rounded float64 -> float16 conversion: normal
    and subnormal ranges, exact ties (to even), overflow and signed zero."""
    L = orc.lib()
    rng = np.random.default_rng(4)
    x = rng.normal(size=60000) * np.exp(rng.uniform(-22, 12, 60000))
    half = np.float16
    # exact ties between neighbouring binary16 values, normal and subnormal
    a = rng.uniform(-60000, 60000, 2000).astype(half).astype(np.float64)
    b = np.nextafter(a.astype(half), half(np.inf)).astype(np.float64)
    s = (rng.integers(0, 1024, 2000) * 2.0 ** -24).astype(half).astype(np.float64)
    ties = np.concatenate([(a + b) / 2, s + 2.0 ** -25, [65519.0, 65520.0, -65520.0, 2.0 ** -25, 3 * 2.0 ** -25]])
    vals = np.concatenate([x, ties, [0.0, -0.0, 1e-30, -1e-30]])
    got = np.array([L.orc_round_half(float(v)) for v in vals])
    want = vals.astype(half).astype(np.float64)
    assert np.array_equal(np.signbit(got[want == 0]), np.signbit(want[want == 0]))
    assert np.array_equal(got, want)


def _sphere_angles(stacks, slices):
    """(theta, phi) of every UV-sphere vertex in the generator's order, FP64:
    north pole, rings i = 1..stacks-1 of slices vertices, south pole."""
    i = np.repeat(np.arange(1, stacks), slices)
    j = np.tile(np.arange(slices), stacks - 1)
    th = np.concatenate([[0.0], np.pi * i / stacks, [np.pi]])
    ph = np.concatenate([[0.0], 2 * np.pi * j / slices, [0.0]])
    return th, ph


@pytest.mark.parametrize("kind,stacks,slices", [(1, 64, 128), (1, 128, 256), (2, 512, 1024)])
def test_radius_formulas_at_vertices(orc, eng, kind, stacks, slices):
    """Before normalisation every vertex lies at the benchmark radius
    r(theta, phi) evaluated in FP64: 1 + 0.1 sin6t sin6p (meshes 1, 2) and
    1 + 0.06 sin5t sin7p + 0.03 sin17t cos13p (mesh 3).  The bound covers the
    FP32 rounding of the angles and of the position; a dropped term, a wrong
    frequency or a swapped angle is off by 1e-2 or more."""
    th, ph = _sphere_angles(stacks, slices)
    if kind == 1:
        want = 1.0 + 0.1 * np.sin(6 * th) * np.sin(6 * ph)
    else:
        want = 1.0 + 0.06 * np.sin(5 * th) * np.sin(7 * ph) + 0.03 * np.sin(17 * th) * np.cos(13 * ph)
    vo, to = orc.uv_sphere(stacks, slices, kind, [0.1])
    ve, te = eng.uv_sphere(stacks, slices, kind, [0.1])
    assert np.array_equal(bits(vo), bits(ve)) and np.array_equal(to, te)
    r = np.linalg.norm(vo.astype(np.float64), axis=1)
    assert np.max(np.abs(r - want)) < 2e-5
    # the direction is (sin t cos p, sin t sin p, cos t)
    dirs = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], axis=1)
    assert np.max(np.abs(vo.astype(np.float64) - want[:, None] * dirs)) < 2e-5


def _unit_cfg(**kw):
    from n2wos_amd.cache import CacheConfig
    return CacheConfig(origin=(0.0, 0.0, 0.0), extent=(1.0, 1.0, 1.0), **kw)


@pytest.mark.parametrize("f32", [False, True])
def test_hash_grid_dense_level_reproduces_linear_field(orc, eng, f32):
    """Trilinear interpolation reproduces a linear function exactly: with the
    dense level's entries set to a + b x/N + c y/N + d z/N at vertex (x, y, z)
    (entry 8 (bx + B by + B^2 bz) + (x & 1 | (y & 1) << 1 | (z & 1) << 2) of
    brick (bx, by, bz) = (x, y, z) // 2, B = N/2 + 1), the feature at any
    point u of the unit cube is a + b u_x + c u_y + d u_z.  A swapped weight
    (f for 1 - f), a transposed index, a wrong corner bit or stride breaks it."""
    cfg = _unit_cfg(n_levels=2, log2_table_size=19, base_resolution=16, growth=1.5)
    lay = orc.cache_layout(2, 19, 16, 1.5)
    grid = np.zeros(2 * lay["total_entries"], np.float32)
    coef = {0: (0.25, 0.5, -0.75, 1.25), 1: (-0.5, 0.125, 0.375, -0.625)}
    for lvl in range(2):
        assert lay["dense"][lvl]
        n = lay["res"][lvl]
        s = n + 1
        v = np.arange(s ** 3)
        x, y, z = v % s, (v // s) % s, v // (s * s)
        nb = n // 2 + 1
        idx = 8 * ((x // 2) + nb * (y // 2) + nb * nb * (z // 2)) + (x % 2) + 2 * (y % 2) + 4 * (z % 2)
        assert idx.max() < lay["sizes"][lvl]
        for f in range(2):
            a, b, c, d = coef[f] if lvl == 0 else coef[1 - f]
            grid[2 * (lay["offsets"][lvl] + idx) + f] = a + b * x / n + c * y / n + d * z / n
    pts = np.random.default_rng(5).uniform(0, 1, (3000, 3)).astype(np.float32)
    enc = orc.cache_encode(cfg, grid, pts, f32=f32)
    u = pts.astype(np.float64)
    tol = 2e-6 if not f32 else 2e-5
    for lvl in range(2):
        for f in range(2):
            a, b, c, d = coef[f] if lvl == 0 else coef[1 - f]
            assert np.max(np.abs(enc[:, 2 * lvl + f] - (a + b * u[:, 0] + c * u[:, 1] + d * u[:, 2]))) < tol
    assert np.all(enc[:, 4:] == 0.0)


@pytest.mark.parametrize("f32", [False, True])
def test_hash_grid_vertex_values_and_spatial_hash(orc, eng, f32):
    """At a grid vertex the feature equals that vertex's table entry.  Vertex
    (x, y, z) is corner (x & 1) | (y & 1) << 1 | (z & 1) << 2 of brick
    (bx, by, bz) = (x, y, z) // 2, entry 8 brick + corner.  On a hashed level
    (16 levels from N_min 16, growth 1.5, T = 2^19: level 4 has N = 81,
    8 (N/2 + 1)^3 > T) brick = (bx xor 2654435761 by xor 805459861 bz) mod T/8,
    checked at hand-picked integer coordinates; on a dense level brick =
    bx + B by + B^2 bz with B = N/2 + 1."""
    cfg = _unit_cfg(n_levels=5, log2_table_size=19, base_resolution=16, growth=1.5)
    lay = orc.cache_layout(5, 19, 16, 1.5)
    assert lay["res"][4] == 81 and not lay["dense"][4] and lay["dense"][3]
    T = 1 << 19
    grid = np.random.default_rng(6).uniform(-1, 1, 2 * lay["total_entries"]).astype(np.float32)
    verts = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (3, 7, 11), (11, 7, 3), (80, 2, 45), (45, 2, 80),
             (17, 64, 33), (81, 81, 81)]
    for lvl in (3, 4):
        n = lay["res"][lvl]
        pts = np.array([[x / n, y / n, z / n] for x, y, z in verts], np.float32)
        enc = orc.cache_encode(cfg, grid, pts, f32=f32)
        for k, (x, y, z) in enumerate(verts):
            if lay["dense"][lvl]:
                x, y, z = (min(c, n) for c in (x, y, z))
            corner = (x & 1) | ((y & 1) << 1) | ((z & 1) << 2)
            bx, by, bz = x >> 1, y >> 1, z >> 1
            if lay["dense"][lvl]:
                nb = n // 2 + 1
                brick = bx + by * nb + bz * nb * nb
            else:
                m32 = 0xFFFFFFFF
                brick = (bx ^ ((by * 2654435761) & m32) ^ ((bz * 805459861) & m32)) & (T // 8 - 1)
            local = 8 * brick + corner
            e = lay["offsets"][lvl] + local
            for f in range(2):
                # the point x / N rounds to FP32, so a neighbour enters with a
                # weight of about 1e-6; a wrong entry is off by O(1)
                assert abs(enc[k, 2 * lvl + f] - grid[2 * e + f]) < 2e-5, (lvl, x, y, z, f)


def _poly_field(res, coef, lo=-1.5, hi=1.5):
    """A trilinear polynomial c0 + a.x + k1 xy + k2 yz + k3 xz + k4 xyz sampled
    at the vertices of a res^3 grid over [lo, hi]^3 (x fastest), as the
    walker's control-variate field (values, res, origin, inverse cell size)."""
    c0, a, k1, k2, k3, k4 = coef
    ax = lo + (hi - lo) * np.arange(res + 1) / res
    Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")
    vals = c0 + a[0] * X + a[1] * Y + a[2] * Z + k1 * X * Y + k2 * Y * Z + k3 * X * Z + k4 * X * Y * Z
    origin = np.full(3, lo, np.float32)
    inv = np.full(3, res / (hi - lo), np.float32)
    return (np.ascontiguousarray(vals.ravel(), np.float32), res, origin, inv)


def test_cv_lookup_reproduces_trilinear_polynomials(orc):
    """Trilinear interpolation reproduces any trilinear polynomial exactly, so
    the oracle's field lookup must return it at arbitrary points (up to FP32
    rounding): a wrong corner, weight or axis stride fails."""
    rng = np.random.default_rng(31)
    coef = (0.3, rng.normal(size=3), 0.7, -0.4, 0.25, 0.5)
    field = _poly_field(24, coef)
    for q in rng.uniform(-1.45, 1.45, (200, 3)).astype(np.float32):
        x, y, z = q.astype(np.float64)
        want = coef[0] + coef[1] @ q.astype(np.float64) + coef[2] * x * y + coef[3] * y * z + coef[4] * x * z + \
            coef[5] * x * y * z
        assert abs(orc.cv_lookup(field, q) - want) < 2e-5, q


def test_cv_frame_orthonormal_and_turned(orc):
    """The frame (d, e1, e2) is orthonormal for every direction (both
    hemispheres, the poles), and psi turns e1 about d by 2 pi psi."""
    rng = np.random.default_rng(32)
    ds = rng.normal(size=(300, 3))
    ds = np.vstack([ds, [[0, 0, 1], [0, 0, -1], [1, 0, 0], [0, -1, 0]]])
    ds = (ds / np.linalg.norm(ds, axis=1, keepdims=True)).astype(np.float32)
    for d in ds:
        psi = float(rng.uniform())
        e1, e2 = orc.cv_frame(d, psi)
        M = np.stack([d, e1, e2]).astype(np.float64)
        assert np.allclose(M @ M.T, np.eye(3), atol=3e-6), d
        f1, _ = orc.cv_frame(d, 0.0)
        assert abs(float(f1.astype(np.float64) @ e1) - np.cos(2 * np.pi * psi)) < 3e-6


@pytest.mark.parametrize("points", [2, 6])
def test_cv_term_closed_forms(orc, points):
    """For a quadratic field C(x) = c0 + g.x + x^T A x with zero-diagonal A
    (bilinear terms, reproduced exactly by the trilinear grid), the term of
    one step is C(p + R d) minus the mean of C over the design:
      antipodal (m = 2):   R grad C(p) . d
      octahedron (m = 6):  R grad C(p) . d + R^2 d^T A d  (the frame's three
                           axes sum d^T A d + e1^T A e1 + e2^T A e2 = tr A = 0).
    A dropped design point, a wrong sign or a mis-scaled mean fails this."""
    rng = np.random.default_rng(33 + points)
    g = rng.normal(size=3)
    k1, k2, k3 = 0.6, -0.8, 0.45
    field = _poly_field(32, (0.2, g, k1, k2, k3, 0.0))
    A = np.array([[0, k1 / 2, k3 / 2], [k1 / 2, 0, k2 / 2], [k3 / 2, k2 / 2, 0]])
    for _ in range(200):
        p = rng.uniform(-0.9, 0.9, 3).astype(np.float32)
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        R = np.float32(rng.uniform(0.05, 0.5))
        psi_word = int(rng.integers(0, 2 ** 32))
        t = orc.cv_term(field, points, p, R, d, psi_word)
        pd, dd, Rd = p.astype(np.float64), d.astype(np.float64), float(R)
        want = Rd * (g + 2 * A @ pd) @ dd
        if points == 6:
            want += Rd * Rd * dd @ A @ dd
        assert abs(t - want) < 3e-5, (p, d, R, t, want)


def test_cv_term_zero_mean_for_any_field(orc):
    """Every design point is uniform on the sphere of radius R when d and psi
    are, so the term has mean zero for ANY field: here a rough random grid,
    at a fixed point, over 20,000 uniform (d, psi) -- the mean is within 4
    standard errors of 0 for both designs."""
    rng = np.random.default_rng(35)
    res = 16
    vals = rng.normal(size=(res + 1) ** 3).astype(np.float32)
    field = (vals, res, np.full(3, -1.0, np.float32), np.full(3, res / 2.0, np.float32))
    p = np.array([0.13, -0.21, 0.07], np.float32)
    for points in (2, 6):
        d = rng.normal(size=(20000, 3))
        d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
        words = rng.integers(0, 2 ** 32, 20000)
        t = np.array([orc.cv_term(field, points, p, 0.6, d[k], int(words[k])) for k in range(20000)])
        assert np.std(t) > 0.05
        assert abs(t.mean()) < 4 * t.std() / np.sqrt(len(t)), points
