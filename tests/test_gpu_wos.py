"""GPU walker parity: per-walk values and step counts bit-exact against the
oracle's CPU walks, FP64 per-point statistics bit-exact against the oracle's
fixed reduction order, and estimates within 3 sigma of the FP64 solution."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def m1(eng, orc, torch_gpu):
    from n2wos_amd import problems as P
    v, t = P.make_mesh("mesh1")
    bvh = eng.Bvh(v, t)
    return v, t, bvh, orc.Bvh2(v, t)


def run(torch, bvh, bc, wp, pts, wpp, per_walk=True, **kw):
    from n2wos_amd import wos as W
    s = W.Solver(bvh, bc, wp)
    st = s.solve(torch.from_numpy(pts).cuda(), wpp, per_walk=per_walk, **kw)
    torch.cuda.synchronize()
    return st


@pytest.mark.parametrize("poisson", [False, True])
def test_walks_bit_exact_vs_oracle(eng, orc, torch_gpu, m1, poisson):
    from n2wos_amd import problems as P, wos as W
    torch = torch_gpu
    v, t, bvh, b2 = m1
    pts = P.interior_points(bvh, v, 96, seed=1)
    bc = P.poisson_boundary() if poisson else P.charges_boundary(0)
    wp = W.WalkParams(max_steps=256, epsilon=1e-3, seed=77, source=W.SOURCE_POISSON_SINES if poisson else 0)
    print("case: per-walk parity poisson=%s 96x70" % poisson, flush=True)
    st = run(torch, bvh, bc, wp, pts, 70)
    gv = st.extra["walk_values"].cpu().numpy()
    gs = st.extra["walk_steps"].cpu().numpy()
    ov, os_, ost, _ = b2.walks(pts, 70, max_steps=256, eps=1e-3, step_scale=0.999, bmode=bc.mode,
                               bparams=bc.oracle_params(), poisson=poisson, seed=77, nthreads=16)
    assert np.array_equal(gs, os_)
    assert np.array_equal(bits(gv), bits(ov))
    n, m, m2 = orc.point_stats(ov, len(pts), 70)
    assert np.array_equal(st.n.cpu().numpy(), n)
    assert np.array_equal(st.mean.cpu().numpy(), m)
    assert np.array_equal(st.m2.cpu().numpy(), m2)
    assert np.array_equal(st.steps.cpu().numpy(), os_.reshape(len(pts), 70).sum(1))


def test_constant_boundary_exact(eng, torch_gpu, m1):
    from n2wos_amd import problems as P, wos as W
    torch = torch_gpu
    v, t, bvh, _ = m1
    pts = P.interior_points(bvh, v, 500, seed=2)
    st = run(torch, bvh, W.Boundary(W.BC_CONSTANT_ONE), W.WalkParams(epsilon=1e-3), pts, 33, per_walk=False)
    assert np.all(st.mean.cpu().numpy() == 1.0) and np.all(st.m2.cpu().numpy() == 0.0)
    assert np.all(st.n.cpu().numpy() == 33)


def test_forced_max_steps_and_accumulate_and_split(eng, torch_gpu, m1):
    from n2wos_amd import problems as P, wos as W
    torch = torch_gpu
    v, t, bvh, _ = m1
    pts = P.interior_points(bvh, v, 300, seed=3)
    st = run(torch, bvh, P.charges_boundary(0), W.WalkParams(max_steps=1, epsilon=1e-4), pts, 10, per_walk=False)
    # a walk from a point farther than eps from the boundary cannot end in one step
    assert np.all(st.forced.cpu().numpy() == 10) and np.all(st.steps.cpu().numpy() == 10)
    wp = W.WalkParams(epsilon=1e-3, seed=5)
    whole = run(torch, bvh, P.charges_boundary(0), wp, pts, 130, per_walk=False)
    s = W.Solver(bvh, P.charges_boundary(0), wp)
    tp = torch.from_numpy(pts).cuda()
    a = s.solve(tp[:100].contiguous(), 130)
    b = s.solve(tp[100:].contiguous(), 130, point_offset=100)
    torch.cuda.synchronize()
    assert np.array_equal(np.concatenate([a.mean.cpu().numpy(), b.mean.cpu().numpy()]), whole.mean.cpu().numpy())
    # accumulate two batches with different seeds == Chan merge of the two
    st1 = run(torch, bvh, P.charges_boundary(0), wp, pts, 64, per_walk=False)
    st2 = s.solve(tp, 64, seed=6)
    acc = s.solve(tp, 64, stats=W.PointStats(st1.n.clone(), st1.mean.clone(), st1.m2.clone(), st1.steps.clone(),
                                                st1.forced.clone()), accumulate=True, seed=6)
    torch.cuda.synchronize()
    n1, m1_, q1 = st1.n.cpu().numpy(), st1.mean.cpu().numpy(), st1.m2.cpu().numpy()
    n2, m2_, q2 = st2.n.cpu().numpy(), st2.mean.cpu().numpy(), st2.m2.cpu().numpy()
    d = m2_ - m1_
    assert np.allclose(acc.mean.cpu().numpy(), m1_ + d * n2 / (n1 + n2), rtol=0, atol=1e-12)
    assert np.allclose(acc.m2.cpu().numpy(), q1 + q2 + d * d * n1 * n2 / (n1 + n2), rtol=1e-12, atol=1e-12)
    assert np.all(acc.n.cpu().numpy() == 128)


def test_mesh1_problem_against_fp64(eng, torch_gpu, m1):
    """Benchmark problem 1: 4,096 interior points x 128 walks, eps 1e-3,
    8-charge boundary, FP64 analytic truth.  Points within eps of the boundary
    end every walk at step 0 (zero variance, O(eps) shell bias) and are left
    out of the z statistics; 128-walk means are still visibly skewed, so the
    bar here is |z| > 4 for < 1% of points and an unbiased mean z."""
    from n2wos_amd import problems as P, wos as W
    torch = torch_gpu
    v, t, bvh, _ = m1
    pts = P.interior_points(bvh, v, 4096, seed=0)
    bc = P.charges_boundary(0)
    print("case: mesh1 4096x128 statistical", flush=True)
    st = run(torch, bvh, bc, W.WalkParams(epsilon=1e-3, seed=0), pts, 128, per_walk=False)
    mean = st.mean.cpu().numpy()
    var = st.variance().cpu().numpy()
    truth = P.charges_solution_f64(bc.charges, pts)
    ok = var > 0
    assert ok.mean() > 0.99
    z = (mean[ok] - truth[ok]) / np.sqrt(var[ok] / 128)
    assert np.mean(np.abs(z) > 4) < 0.01
    assert abs(z.mean()) < 5 * z.std() / np.sqrt(len(z))
    rmse = np.sqrt(np.mean((mean - truth) ** 2))
    print("mesh1 RMSE", rmse, "mean steps", st.steps.cpu().numpy().sum() / (4096 * 128))
    assert rmse < 0.03


def test_estimator_calibrated_at_high_walk_count(eng, torch_gpu, m1):
    """eps 1e-4, 2,048 walks per point: z-scores against the FP64 solution are
    standard normal (3-sigma rate, unit spread, zero mean)."""
    from n2wos_amd import problems as P, wos as W
    torch = torch_gpu
    v, t, bvh, _ = m1
    pts = P.interior_points(bvh, v, 2048, seed=9)
    bc = P.charges_boundary(0)
    print("case: calibration 2048x2048 eps 1e-4", flush=True)
    st = run(torch, bvh, bc, W.WalkParams(epsilon=1e-4, seed=3), pts, 2048, per_walk=False)
    mean = st.mean.cpu().numpy()
    var = st.variance().cpu().numpy()
    truth = P.charges_solution_f64(bc.charges, pts)
    ok = var > 0
    z = (mean[ok] - truth[ok]) / np.sqrt(var[ok] / 2048)
    assert np.mean(np.abs(z) > 3) < 0.01
    assert 0.9 < z.std() < 1.1
    assert abs(z.mean()) < 5 / np.sqrt(len(z))


def test_inside_test_against_radius(eng, torch_gpu, m1):
    from n2wos_amd import problems as P
    v, t, bvh, _ = m1
    rng = np.random.default_rng(4)
    pts = rng.uniform(-1.1, 1.1, (20000, 3)).astype(np.float32)
    ins = P.inside_mask(bvh, pts)
    r = np.linalg.norm(pts, axis=1)
    rmin = np.linalg.norm(v, axis=1).min()
    assert np.all(ins[r < rmin * 0.99])
    assert not np.any(ins[r > 1.0])


def test_partial_grid_equals_full_grid(eng, torch_gpu, m1):
    """A walker limited to part of the device (max_blocks, as the progressive
    estimator runs beside training) gives the full grid's results bit for
    bit: units are claimed dynamically and reduced in a fixed order."""
    from n2wos_amd import problems as P, wos as W
    torch = torch_gpu
    v, t, bvh, _ = m1
    full = W.full_grid_blocks(W.SOURCE_NONE)
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    assert full > 0 and full % cus == 0
    assert W.full_grid_blocks(W.SOURCE_POISSON_SINES) > 0
    pts = P.interior_points(bvh, v, 2000, seed=9)
    print("case: partial grid 2000x96, full grid %d blocks" % full, flush=True)
    a = run(torch, bvh, P.charges_boundary(0), W.WalkParams(epsilon=1e-3, seed=5), pts, 96, per_walk=False)
    for mb in (1, 7, full // 3):
        b = run(torch, bvh, P.charges_boundary(0), W.WalkParams(epsilon=1e-3, seed=5, max_blocks=mb), pts, 96,
                per_walk=False)
        assert np.array_equal(a.mean.cpu().numpy().view(np.uint64), b.mean.cpu().numpy().view(np.uint64))
        assert np.array_equal(a.m2.cpu().numpy().view(np.uint64), b.m2.cpu().numpy().view(np.uint64))
        assert np.array_equal(a.steps.cpu().numpy(), b.steps.cpu().numpy())


@pytest.mark.parametrize("npts,wpp", [(12000, 16), (9000, 70)])
def test_claim_order_and_ranges_change_nothing(eng, orc, torch_gpu, m1, npts, wpp):
    """Solves of at least 8,192 points claim their units in spatial (Morton
    cell) order, split into per-XCD ranges.  Per-walk values, steps and FP64
    statistics equal those of the caller's order with one range bit for bit,
    and a sample of points equals the oracle's walks."""
    from n2wos_amd import _lib as L, problems as P, wos as W
    torch = torch_gpu
    v, t, bvh, b2 = m1
    pts = P.interior_points(bvh, v, npts, seed=21)
    bc = P.charges_boundary(0)
    wp = W.WalkParams(epsilon=1e-3, seed=8)
    print("case: claim order %dx%d" % (npts, wpp), flush=True)
    runs = []
    try:
        for order, ranges in [(1, 8), (0, 1), (1, 1), (0, 8)]:
            L.check(L.lib().n2wos_set_tuning(b"spatial_order", order))
            L.check(L.lib().n2wos_set_tuning(b"unit_ranges", ranges))
            runs.append(run(torch, bvh, bc, wp, pts, wpp))
    finally:
        L.check(L.lib().n2wos_set_tuning(b"spatial_order", 1))
        L.check(L.lib().n2wos_set_tuning(b"unit_ranges", 8))
    a = runs[0]
    for b in runs[1:]:
        assert np.array_equal(bits(a.extra["walk_values"].cpu().numpy()), bits(b.extra["walk_values"].cpu().numpy()))
        assert np.array_equal(a.extra["walk_steps"].cpu().numpy(), b.extra["walk_steps"].cpu().numpy())
        for f in ("n", "mean", "m2"):
            assert np.array_equal(getattr(a, f).cpu().numpy().view(np.uint64), getattr(b, f).cpu().numpy().view(np.uint64))
        assert np.array_equal(a.steps.cpu().numpy(), b.steps.cpu().numpy())
        assert np.array_equal(a.forced.cpu().numpy(), b.forced.cpu().numpy())
    sel = np.arange(0, npts, npts // 37)[:37]
    ov, os_, _, _ = b2.walks(pts[sel], wpp, max_steps=256, eps=1e-3, step_scale=0.999, bmode=bc.mode,
                             bparams=bc.oracle_params(), seed=8, nthreads=16, point_ids=sel)
    gv = a.extra["walk_values"].cpu().numpy().reshape(npts, wpp)[sel].ravel()
    assert np.array_equal(bits(gv), bits(ov))


@pytest.mark.parametrize("mesh", ["sphere8x16", "mesh1"])
def test_exterior_start_points_bit_exact(eng, orc, torch_gpu, mesh):
    """Walks from points outside the closed mesh drift away until the step cap
    (256); past about 1.8e19 their step bound no longer fits in FP32 and the
    query runs unbounded.  Values and step counts equal the oracle's walks, and
    no walk reports an overflow."""
    from n2wos_amd import problems as P, wos as W
    torch = torch_gpu
    if mesh == "mesh1":
        v, t = P.make_mesh("mesh1")
    else:
        v0, t = eng.uv_sphere(8, 16, eng.SPHERE_SIN6, [0.1])
        v, _, _ = eng.normalize_to_unit_radius(v0)
    bvh = eng.Bvh(v, t)
    pts = np.array([[1.6, 0.0, 0.0], [0.0, -1.7, 0.3], [2.5, 2.5, -2.5], [0.3, 0.2, 0.1]], np.float32)
    bc = P.charges_boundary(0)
    wp = W.WalkParams(max_steps=256, epsilon=1e-3, seed=77)
    print("case: exterior start points %s 4x64" % mesh, flush=True)
    st = run(torch, bvh, bc, wp, pts, 64)
    gv = st.extra["walk_values"].cpu().numpy()
    gs = st.extra["walk_steps"].cpu().numpy()
    ov, os_, _, _ = orc.Bvh2(v, t).walks(pts, 64, max_steps=256, eps=1e-3, step_scale=0.999, bmode=bc.mode,
                                         bparams=bc.oracle_params(), seed=77, nthreads=16)
    assert np.array_equal(gs, os_)
    assert np.array_equal(bits(gv), bits(ov))
    assert (gs[:3 * 64] == 256).mean() > 0.5  # most exterior walks run to the cap


def test_ring_capacity_and_spill_lines_change_nothing(eng, orc, torch_gpu):
    """The walker's LDS ring overflows to a 128-byte line per lane in the
    workspace and then to per-lane scratch.  On mesh 2 with the benchmark's
    SAH leaf-4 tree, the 8-entry ring (about two spills per walk step; on the
    CPU emulator about one step in 40,000 spills past the line into scratch)
    and the default 16-entry ring give the same
    per-walk values, steps and FP64 statistics bit for bit, and a sample of
    points equals the oracle's walks."""
    from n2wos_amd import _lib as L, problems as P, wos as W
    torch = torch_gpu
    v, t = P.make_mesh("mesh2")
    bvh = eng.Bvh(v, t, leaf_size=4, build_method="sah")
    pts = P.interior_points(bvh, v, 3000, seed=31)
    bc = P.charges_boundary(0)
    wp = W.WalkParams(epsilon=1e-4, seed=12)
    runs = []
    try:
        for cap in (16, 8):
            print("case: ring %d, mesh2 3000x64" % cap, flush=True)
            L.check(L.lib().n2wos_set_tuning(b"ring", cap))
            runs.append(run(torch, bvh, bc, wp, pts, 64))
    finally:
        L.check(L.lib().n2wos_set_tuning(b"ring", 16))
    a, b = runs
    assert np.array_equal(bits(a.extra["walk_values"].cpu().numpy()), bits(b.extra["walk_values"].cpu().numpy()))
    assert np.array_equal(a.extra["walk_steps"].cpu().numpy(), b.extra["walk_steps"].cpu().numpy())
    for f in ("n", "mean", "m2"):
        assert np.array_equal(getattr(a, f).cpu().numpy().view(np.uint64), getattr(b, f).cpu().numpy().view(np.uint64))
    assert int(a.forced.sum().item()) == int(b.forced.sum().item())
    sel = np.arange(0, 3000, 100)
    ov, os_, _, _ = orc.Bvh2(v, t).walks(pts[sel], 64, max_steps=256, eps=1e-4, step_scale=0.999, bmode=bc.mode,
                                         bparams=bc.oracle_params(), seed=12, nthreads=16, point_ids=sel)
    gv = a.extra["walk_values"].cpu().numpy().reshape(3000, 64)[sel].ravel()
    assert np.array_equal(bits(gv), bits(ov))
