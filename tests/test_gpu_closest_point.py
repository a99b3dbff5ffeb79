"""GPU parity of the closest-point query API against the oracle (bit-exact)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def c1(eng, orc, torch_gpu):
    v, t = eng.uv_sphere(64, 128, 1, [0.1])
    vn, _, _ = eng.normalize_to_unit_radius(v)
    bvh = eng.Bvh(vn, t, leaf_size=8)
    return vn, t, bvh


def test_c1_queries_bit_exact_vs_bruteforce(eng, orc, torch_gpu, c1):
    vn, t, bvh = c1
    torch = torch_gpu
    rng = np.random.default_rng(0)
    pts = rng.uniform(-1.2, 1.2, (65536, 3)).astype(np.float32)
    print("case: c1 uniform 65536 queries", flush=True)
    d2, cp, tri, ovf = bvh.query(pts)
    sub = slice(0, 8192)
    d2b, cpb, trb = orc.bruteforce(vn, t, pts[sub], nthreads=16)
    assert np.array_equal(bits(d2[sub]), bits(d2b))
    assert np.array_equal(tri[sub], trb)
    assert np.array_equal(bits(cp[sub]), bits(cpb))
    # the whole set against the oracle's traversal restatement (same results, exact)
    nodes, recs, delta = bvh.export()
    d2o, cpo, trio, ovfo, nno, nto = orc.bvh4_query(nodes, recs, delta, pts)
    assert np.array_equal(bits(d2), bits(d2o)) and np.array_equal(tri, trio) and np.array_equal(bits(cp), bits(cpo))
    assert ovf.sum() == 0


def test_counted_query_matches_restated_visit_counts(eng, orc, torch_gpu, c1):
    vn, t, bvh = c1
    torch = torch_gpu
    rng = np.random.default_rng(1)
    pts = np.concatenate([rng.uniform(-1.2, 1.2, (4096, 3)), rng.uniform(-0.02, 0.02, (256, 3))]).astype(np.float32)
    print("case: counted query 4352", flush=True)
    d2, cp, tri, ovf, nn, nt = bvh.query_tensors(torch.from_numpy(pts).cuda(), counted=True)
    torch.cuda.synchronize()
    nodes, recs, delta = bvh.export()
    d2o, cpo, trio, ovfo, nno, nto = orc.bvh4_query(nodes, recs, delta, pts)
    assert np.array_equal(nn.cpu().numpy().astype(np.uint32), nno)
    assert np.array_equal(nt.cpu().numpy().astype(np.uint32), nto)
    assert np.array_equal(bits(d2.cpu().numpy()), bits(d2o))


def test_near_surface_and_far_points(eng, orc, torch_gpu, c1):
    vn, t, bvh = c1
    rng = np.random.default_rng(2)
    base = rng.uniform(-1.2, 1.2, (2000, 3)).astype(np.float32)
    _, cpb, _ = orc.bruteforce(vn, t, base, nthreads=16)
    pts = np.concatenate([cpb + rng.normal(0, 1e-5, cpb.shape), cpb, rng.uniform(-50, 50, (500, 3)),
                          vn[:300]]).astype(np.float32)
    print("case: near-surface/far/vertex queries", len(pts), flush=True)
    d2, cp, tri, ovf = bvh.query(pts)
    d2b, cpb2, trb = orc.bruteforce(vn, t, pts, nthreads=16)
    assert np.array_equal(bits(d2), bits(d2b)) and np.array_equal(tri, trb) and np.array_equal(bits(cp), bits(cpb2))


def test_masked_and_nonfinite_and_block_sizes(eng, orc, torch_gpu, c1):
    vn, t, bvh = c1
    torch = torch_gpu
    rng = np.random.default_rng(3)
    pts = rng.uniform(-1.2, 1.2, (1000, 3)).astype(np.float32)
    pts[5] = [np.nan, 0, 0]
    pts[6] = [np.inf, 0, 0]
    active = (rng.uniform(size=1000) < 0.5).astype(np.uint8)
    active[5] = active[6] = 1
    print("case: masked query 1000", flush=True)
    tp = torch.from_numpy(pts).cuda()
    ta = torch.from_numpy(active).cuda()
    d2, cp, tri, ovf = [x.cpu().numpy() for x in bvh.query_tensors(tp, active=ta)]
    torch.cuda.synchronize()
    full = [x.cpu().numpy() for x in bvh.query_tensors(tp)]
    on = active == 1
    assert np.array_equal(bits(d2[on]), bits(full[0][on])) and np.array_equal(tri[on], full[2][on])
    off = ~on
    assert np.all(d2[off] == 0) and np.all(tri[off] == -1) and np.array_equal(cp[off], pts[off])
    assert d2[5] == np.finfo(np.float32).max and tri[5] == -1 and d2[6] == np.finfo(np.float32).max
    for bs in (1, 64, 256, 1024):
        print("case: block size", bs, flush=True)
        r = [x.cpu().numpy() for x in bvh.query_tensors(tp, block_size=bs)]
        assert np.array_equal(bits(r[0]), bits(full[0])) and np.array_equal(r[2], full[2])


def test_query_errors(eng, torch_gpu, c1):
    vn, t, bvh = c1
    torch = torch_gpu
    x = torch.zeros((4, 3), device="cuda")
    d2 = torch.zeros(4, device="cuda")
    with pytest.raises(eng.N2wosError) as e:
        bvh.query_device(x.data_ptr(), -1, d2.data_ptr(), x.data_ptr(), d2.data_ptr(), d2.data_ptr())
    assert e.value.code == eng._lib.ERR_INVALID
    bvh.query_device(x.data_ptr(), 0, 0, 0, 0, 0)  # no-op
    with pytest.raises(eng.N2wosError) as e:
        bvh.query_device(x.data_ptr(), 4, d2.data_ptr(), x.data_ptr(), d2.data_ptr(), d2.data_ptr(), block_size=2048)
    assert e.value.code == eng._lib.ERR_BLOCK
    with pytest.raises(eng.N2wosError) as e:
        bvh.query_device(x.data_ptr(), 4, 0, x.data_ptr(), d2.data_ptr(), d2.data_ptr())
    assert e.value.code == eng._lib.ERR_NULL
    st = bvh.stats()
    assert st["triangle_count"] == len(t) and st["leaf_size"] == 8 and st["build_method"] == "median"


@pytest.mark.parametrize("mesh,leaf", [("mesh2", 4), ("mesh2", 8), ("blob", 4)])
def test_sah_build_bit_exact_vs_bruteforce(eng, orc, torch_gpu, mesh, leaf):
    """The surface-area-heuristic tree changes only the visit order: results
    stay equal to the brute force (uniform, near-surface and WoS-prefix
    queries), and the stack never overflows."""
    from n2wos_amd import problems as P
    v, t = P.make_mesh(mesh)
    bvh = eng.Bvh(v, t, leaf_size=leaf, build_method="sah")
    assert bvh.stats()["build_method"] == "sah"
    rng = np.random.default_rng(7)
    interior = P.interior_points(bvh, v, 512, seed=8)
    pts = np.concatenate([rng.uniform(-1.2, 1.2, (1536, 3)).astype(np.float32),
                          P.surface_shell_queries(v, t, 512, seed=9),
                          P.wos_prefix_queries(bvh, interior, 512, seed=10)]).astype(np.float32)
    print("case: sah", mesh, leaf, len(pts), flush=True)
    d2, cp, tri, ovf = bvh.query(pts)
    d2b, cpb, trb = orc.bruteforce(v, t, pts, nthreads=16)
    assert np.array_equal(bits(d2), bits(d2b)) and np.array_equal(tri, trb) and np.array_equal(bits(cp), bits(cpb))
    assert ovf.sum() == 0


def test_quad_and_per_lane_kernels_agree(eng, torch_gpu, c1):
    """The four-lanes-per-query kernel (default) and the one-lane-per-query
    kernel return the same outputs and visit counts for every query, masked
    or not, including non-finite points, far points and a count that is not
    a multiple of the block."""
    import n2wos_amd._lib as L
    vn, t, bvh = c1
    torch = torch_gpu
    rng = np.random.default_rng(7)
    pts = np.concatenate([rng.uniform(-1.2, 1.2, (3000, 3)), rng.uniform(-40, 40, (200, 3)),
                          vn[rng.integers(0, len(vn), 301)] + rng.normal(0, 1e-3, (301, 3))]).astype(np.float32)
    pts[17] = [np.nan, 0, 0]
    pts[1030] = [0, -np.inf, 0]
    active = (rng.uniform(size=len(pts)) < 0.7).astype(np.uint8)
    tp = torch.from_numpy(pts).cuda()
    ta = torch.from_numpy(active).cuda()
    res = {}
    try:
        for quad in (1, 0):
            L.check(L.lib().n2wos_set_tuning(b"query_quad", quad))
            print("case: query_quad=%d, %d points" % (quad, len(pts)), flush=True)
            plain = [x.cpu().numpy() for x in bvh.query_tensors(tp, counted=True)]
            masked = [x.cpu().numpy() for x in bvh.query_tensors(tp, active=ta)]
            torch.cuda.synchronize()
            res[quad] = plain + masked
    finally:
        L.check(L.lib().n2wos_set_tuning(b"query_quad", 1))
    for a, b in zip(res[1], res[0]):
        assert np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))
    d2, cp, tri, ovf, nn, nt = res[1][:6]
    assert tri[17] == -1 and tri[1030] == -1 and ovf.sum() == 0 and nn.sum() > 0


def test_host_vector_query_equals_device_query(eng, torch_gpu, c1):
    """n2wos_bvh_query_host (the host-memory form, CudaBvh::query) returns the
    device query's results bit for bit, counts overflows, times the launch,
    and handles an empty set and a bad block size."""
    vn, t, bvh = c1
    rng = np.random.default_rng(9)
    pts = rng.uniform(-1.3, 1.3, (20000, 3)).astype(np.float32)
    pts[7] = np.nan
    print("case: host-vector query 20000", flush=True)
    d2, cp, tri, novf, ms = bvh.query_host(pts)
    d2d, cpd, trid, ovfd = bvh.query(pts)
    assert np.array_equal(bits(d2), bits(d2d)) and np.array_equal(bits(cp), bits(cpd))
    assert np.array_equal(tri, trid) and novf == int((ovfd != 0).sum()) == 0
    assert tri[7] == -1 and ms > 0
    e = bvh.query_host(np.zeros((0, 3), np.float32))
    assert len(e[0]) == 0 and e[3] == 0
    with pytest.raises(eng.N2wosError) as err:
        bvh.query_host(pts[:4], block_size=0)
    assert err.value.code == eng._lib.ERR_BLOCK


@pytest.mark.parametrize("cap,slots", [(1, 64), (4, 16), (4, 7), (32, 1)])
@pytest.mark.parametrize("method,leaf", [("median", 8), ("sah", 3)])
def test_capped_queries_finished_by_the_wave_pass_are_exact(eng, orc, torch_gpu, cap, slots, method, leaf):
    """The query kernel stops a query after `query_cap` node visits and
    leaves, and a second pass finishes it with a whole wave (each wave
    scanning `slots` consecutive queries; 4,200 points leave a partial last
    run for 16 and 64 slots).  With tiny caps nearly every query goes through that pass: the
    results still equal the brute force bit for bit (lowest index among
    ties), with no overflow, on uniform, near-surface, vertex, far and centre
    points."""
    from n2wos_amd import _lib
    v, t = eng.uv_sphere(64, 128, 1, [0.1])
    vn, _, _ = eng.normalize_to_unit_radius(v)
    bvh = eng.Bvh(vn, t, leaf_size=leaf, build_method=method)
    rng = np.random.default_rng(7 + cap)
    base = rng.uniform(-1.2, 1.2, (3000, 3)).astype(np.float32)
    _, cpb, _ = orc.bruteforce(vn, t, base[:500], nthreads=16)
    pts = np.concatenate([base, cpb + rng.normal(0, 1e-5, cpb.shape), vn[:200], rng.uniform(-0.05, 0.05, (300, 3)),
                          rng.uniform(-20, 20, (100, 3))]).astype(np.float32)
    print("case: capped queries", cap, slots, method, leaf, len(pts), flush=True)
    _lib.lib().n2wos_set_tuning(b"query_cap", cap)
    _lib.lib().n2wos_set_tuning(b"query_finish_slots", slots)
    try:
        d2, cp, tri, ovf = bvh.query(pts)
    finally:
        _lib.lib().n2wos_set_tuning(b"query_cap", 16)
        _lib.lib().n2wos_set_tuning(b"query_finish_slots", 8)
    d2b, cpb2, trb = orc.bruteforce(vn, t, pts, nthreads=16)
    assert np.array_equal(bits(d2), bits(d2b)) and np.array_equal(tri, trb) and np.array_equal(bits(cp), bits(cpb2))
    assert ovf.sum() == 0


def test_two_pass_query_replays_in_a_hip_graph(eng, orc, torch_gpu, c1):
    """The default query (the four-lane pass, then the wave-wide pass for the
    queries it stopped) captured in a HIP graph: three replays, each after
    eager work that overwrites the outputs, equal the eager call and the brute
    force bit for bit.  Uniform and centre points, so most queries reach the
    second pass."""
    from n2wos_amd._lib import check, lib
    torch = torch_gpu
    vn, t, bvh = c1
    rng = np.random.default_rng(41)
    pts = np.concatenate([rng.uniform(-1.2, 1.2, (3000, 3)), rng.uniform(-0.05, 0.05, (1000, 3))]).astype(np.float32)
    n = len(pts)
    pt = torch.from_numpy(pts).cuda()
    d2 = torch.empty(n, dtype=torch.float32, device="cuda")
    cp = torch.empty((n, 3), dtype=torch.float32, device="cuda")
    tri = torch.empty(n, dtype=torch.int32, device="cuda")
    ovf = torch.empty(n, dtype=torch.int32, device="cuda")

    def call():
        check(lib().n2wos_bvh_query_device(bvh.handle, pt.data_ptr(), n, d2.data_ptr(), cp.data_ptr(), tri.data_ptr(),
                                           ovf.data_ptr(), 128, torch.cuda.current_stream().cuda_stream))

    print("case: two-pass query in a HIP graph, %d points" % n, flush=True)
    call()
    torch.cuda.synchronize()
    ref = [x.cpu().numpy().copy() for x in (d2, cp, tri, ovf)]
    d2b, cpb, trb = orc.bruteforce(vn, t, pts, nthreads=16)
    assert np.array_equal(bits(ref[0]), bits(d2b)) and np.array_equal(ref[2], trb)
    assert np.array_equal(bits(ref[1]), bits(cpb)) and ref[3].sum() == 0
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        call()
    for rep in range(3):
        d2.fill_(-1.0)
        cp.fill_(7.0)
        tri.fill_(-7)
        ovf.fill_(5)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        got = [x.cpu().numpy() for x in (d2, cp, tri, ovf)]
        assert np.array_equal(bits(got[0]), bits(ref[0])), rep
        assert np.array_equal(bits(got[1]), bits(ref[1])), rep
        assert np.array_equal(got[2], ref[2]) and np.array_equal(got[3], ref[3]), rep
