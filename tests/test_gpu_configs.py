"""The benchmark configurations under test, each against an FP64 truth, and
the engine's plain estimator against the reference's own committed run.

- Problem 2 (headline): mesh 2, the interior cells of the 512 x 512 grid on
  z = 0, 256 walks, eps 1e-4, 8 charges: z-scores against the FP64 solution.
- Problem 3 (Poisson, manufactured u*): Green's-function source sampling on
  mesh 1 at 1,024 walks and on the 1,046,528-triangle mesh 3 (SAH leaf 4) at
  512 walks, z-scores against FP64 u*; and mesh-3 walks bit-exact against the
  oracle's.
- Problem 4 (NC + 2LMC): the 16-level 2^19 hash grid with the 4 x 64 MLP,
  trained online on the solver's coupled walks at depth 4, 16 coupled pairs
  per point: the two-level estimate is unbiased against FP64 and the trained
  cache lowers the residual variance below the untrained one's.
- The reference's committed evaluation run (figures/zebra_k8_m2.json, RTX
  3070, its own RNG): on the same mesh, slice and boundary, this engine's
  plain walks agree with its mean step count, mean sample variance, mean bias
  and RMSE within their standard errors."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def zscores(mean, var, n, truth):
    ok = var > 0
    return (mean[ok] - truth[ok]) / np.sqrt(var[ok] / n), ok


def test_problem2_headline_grid_against_fp64(eng, torch_gpu):
    from n2wos_amd import problems as P, wos as W
    torch = torch_gpu
    v, t = P.make_mesh("mesh2")
    bvh = eng.Bvh(v, t, leaf_size=3, build_method="sah")
    pts, _ = P.grid_slice_points(bvh, 512)
    assert len(pts) == 170132
    bc = P.charges_boundary(0)
    print("case: problem 2, 170,132 grid points x 256 walks", flush=True)
    st = W.Solver(bvh, bc, W.WalkParams(epsilon=1e-4, seed=12345)).solve(torch.from_numpy(pts).cuda(), 256)
    torch.cuda.synchronize()
    mean, var = st.mean.cpu().numpy(), st.variance().cpu().numpy()
    truth = P.charges_solution_f64(bc.charges, pts)
    z, ok = zscores(mean, var, 256, truth)
    assert ok.mean() > 0.999
    rmse = float(np.sqrt(np.mean((mean - truth) ** 2)))
    print("problem 2: RMSE %.5f, z mean %.4f sd %.4f, |z|>4 %.5f" % (rmse, z.mean(), z.std(),
                                                                 np.mean(np.abs(z) > 4)))
    assert np.mean(np.abs(z) > 4) < 0.01
    assert abs(z.mean()) < 0.05 and 0.9 < z.std() < 1.1
    assert rmse < 0.015
    assert int(st.forced.cpu().numpy().sum()) == 0


def _poisson_z(torch, bvh, pts, wpp, seed):
    from n2wos_amd import problems as P, wos as W
    wp = W.WalkParams(epsilon=1e-4, seed=seed, source=W.SOURCE_POISSON_SINES)
    st = W.Solver(bvh, P.poisson_boundary(), wp).solve(torch.from_numpy(pts).cuda(), wpp)
    torch.cuda.synchronize()
    mean, var = st.mean.cpu().numpy(), st.variance().cpu().numpy()
    truth = P.poisson_solution_f64(pts)
    z, ok = zscores(mean, var, wpp, truth)
    return z, ok, float(np.sqrt(np.mean((mean - truth) ** 2)))


def test_poisson_mesh1_against_manufactured_solution(eng, torch_gpu):
    """u* = sin 2pi x sin 2pi y sin 2pi z + |x|^2 / 2, f = -lap u*, g = u*:
    the walks' Green's-function source samples (R^2/6 f(y), y at a Beta(2,2)
    radius fraction) and boundary values give calibrated z-scores.  A wrong
    sign or scale of the source term biases every point by many sigma."""
    from n2wos_amd import problems as P
    torch = torch_gpu
    v, t = P.make_mesh("mesh1")
    bvh = eng.Bvh(v, t)
    pts = P.interior_points(bvh, v, 2048, seed=21)
    print("case: Poisson mesh 1, 2048 x 1024", flush=True)
    z, ok, rmse = _poisson_z(torch, bvh, pts, 1024, 5)
    print("poisson mesh1: RMSE %.5f z mean %.4f sd %.4f" % (rmse, z.mean(), z.std()))
    assert ok.mean() > 0.99
    assert np.mean(np.abs(z) > 3) < 0.01
    assert 0.9 < z.std() < 1.1
    assert abs(z.mean()) < 5 / np.sqrt(len(z))


@pytest.fixture(scope="module")
def mesh3(eng, torch_gpu):
    from n2wos_amd import problems as P
    v, t = P.make_mesh("mesh3")
    assert len(t) == 1046528
    return v, t, eng.Bvh(v, t, leaf_size=4, build_method="sah")


def test_poisson_mesh3_against_manufactured_solution(eng, torch_gpu, mesh3):
    from n2wos_amd import problems as P
    torch = torch_gpu
    v, t, bvh = mesh3
    pts = P.interior_points(bvh, v, 2048, seed=22)
    print("case: Poisson mesh 3 (1,046,528 triangles), 2048 x 512", flush=True)
    z, ok, rmse = _poisson_z(torch, bvh, pts, 512, 6)
    print("poisson mesh3: RMSE %.5f z mean %.4f sd %.4f" % (rmse, z.mean(), z.std()))
    assert ok.mean() > 0.99
    assert np.mean(np.abs(z) > 3) < 0.01
    assert 0.9 < z.std() < 1.1
    assert abs(z.mean()) < 5 / np.sqrt(len(z))


def test_poisson_mesh3_walks_bit_exact_vs_oracle(eng, orc, torch_gpu, mesh3):
    """Per-walk values and step counts of Poisson walks on mesh 3 (SAH tree,
    leaf 4) equal the oracle's walks over its own binary tree."""
    from n2wos_amd import problems as P, wos as W
    torch = torch_gpu
    v, t, bvh = mesh3
    pts = P.interior_points(bvh, v, 32, seed=23)
    wp = W.WalkParams(max_steps=256, epsilon=1e-4, seed=31, source=W.SOURCE_POISSON_SINES)
    print("case: Poisson mesh 3 per-walk parity 32 x 16", flush=True)
    st = W.Solver(bvh, P.poisson_boundary(), wp).solve(torch.from_numpy(pts).cuda(), 16, per_walk=True)
    torch.cuda.synchronize()
    gv, gs = st.extra["walk_values"].cpu().numpy(), st.extra["walk_steps"].cpu().numpy()
    bc = P.poisson_boundary()
    ov, os_, _, _ = orc.Bvh2(v, t).walks(pts, 16, max_steps=256, eps=1e-4, step_scale=0.999, bmode=bc.mode,
                                         bparams=bc.oracle_params(), poisson=True, seed=31, nthreads=16)
    assert np.array_equal(gs, os_)
    assert np.array_equal(bits(gv), bits(ov))


def test_problem4_nc_2lmc_unbiased_and_trained_cache_helps(eng, torch_gpu):
    """The learned-cache configuration on the problem-2 mesh and boundary:
    labels (X_4, W - S_4) from 32 coupled walks per point (their own Philox
    stream), 200 Adam steps at batch 2^18, then NC + 2LMC with 64 coarse
    samples and 16 coupled pairs per point at depth 4."""
    from n2wos_amd import cache as K, nc as N, problems as P, wos as W
    torch = torch_gpu
    v, t = P.make_mesh("mesh2")
    bvh = eng.Bvh(v, t, leaf_size=3, build_method="sah")
    pts, _ = P.grid_slice_points(bvh, 512)
    pts = pts[::16].copy()
    tp = torch.from_numpy(pts).cuda()
    bc = P.charges_boundary(0)
    truth = P.charges_solution_f64(bc.charges, pts)
    ns = N.NcSolver(bvh, bc, W.WalkParams(epsilon=1e-4, seed=12345))
    cfg = K.CacheConfig.for_bounds(v.min(0), v.max(0), seed=3)
    assert (cfg.n_levels, cfg.log2_table_size, cfg.base_resolution, cfg.growth) == (16, 19, 16, 1.5)
    print("case: problem 4 NC+2LMC, %d points, depth 4, 16 pairs" % len(pts), flush=True)
    untrained = K.Cache(cfg)
    trained = K.Cache(cfg)
    b = ns.walks(tp, 32, N.NC_COUPLED, 4, stream_id=7, seed=99)
    x, y = ns.training_pairs(b)
    g = torch.Generator(device="cuda").manual_seed(0)
    for _ in range(200):
        idx = torch.randint(0, x.shape[0], (min(1 << 18, x.shape[0]),), device="cuda", generator=g)
        trained.train_step(x[idx].contiguous(), y[idx].contiguous())
    res = {}
    for name, c in (("untrained", untrained), ("trained", trained)):
        est, co, re, _ = ns.two_level(c, tp, 64, 16, 4, seed=1234)
        torch.cuda.synchronize()
        e = est.cpu().numpy()
        vc = (co.m2 / (co.n - 1)).cpu().numpy()
        vr = (re.m2 / (re.n - 1)).cpu().numpy()
        sd = np.sqrt(vc / 64 + vr / 16)
        ok = sd > 0
        z = (e[ok] - truth[ok]) / sd[ok]
        res[name] = (z, float(vr.mean()), float(np.sqrt(np.mean((e - truth) ** 2))))
        print("%s: residual var %.5f coarse var %.5f RMSE %.5f z mean %.4f sd %.4f" %
              (name, vr.mean(), vc.mean(), res[name][2], z.mean(), z.std()))
        # unbiased: the mean error over the points is within 5 standard
        # errors of zero (the mean of per-point z-scores is not the test:
        # with 16 skewed residual samples per point a t-statistic has a
        # nonzero mean even for an unbiased estimate)
        me = float(np.mean(e - truth))
        se = float(np.sqrt(np.sum(sd ** 2)) / len(sd))
        print("%s: mean error %.3g, %.2f standard errors" % (name, me, me / se))
        assert abs(me) < 5 * se
        assert np.mean(np.abs(z) > 4) < 0.01
        # the RMSE is what the variances predict (no hidden bias)
        assert res[name][2] < 1.1 * np.sqrt(np.mean(sd ** 2))
    assert res["trained"][1] < 0.8 * res["untrained"][1]


def test_plain_walks_agree_with_reference_run(eng, torch_gpu):
    """figures/zebra_k8_m2.json (the reference's evaluation tool on an RTX
    3070): bumpy sphere 128 x 256, amplitude 0.15, normalised; zebra k = 8
    boundary Re((x + iy)^8); the 162,215 interior pixels of the 512^2 xy slice
    at z = 0; plain WoS with eps 1e-4, 64 walks per point.  The walk's law is
    the same here (exact closest point, step 0.999 d, uniform direction), the
    random numbers are not, so the comparison is statistical: each statistic
    differs from the reference's by less than 4 standard errors of the
    difference of two independent runs (sqrt 2 x this run's own standard
    error, computed from its per-walk values and step counts)."""
    from n2wos_amd import eval_nc as EV, geometry as G, wos as W
    torch = torch_gpu
    ref = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_runs_zebra_k8.json")))["zebra_k8_m2"]
    opt = EV.parse_args(["--mesh", "procedural_bumpy_sphere", "--bc", "harmonic_zebra_k8", "--eval-mode", "slice",
                         "--slice-view", "xy", "--slice-plane", "0.0", "--slice-width", "512", "--slice-height",
                         "512"])
    v, t, _, _, _ = EV.make_mesh(opt)
    bvh = G.Bvh(v, t, leaf_size=opt.cubql_leaf_size, build_method=opt.cubql_build_method)
    pts, _, _ = EV.slice_points(opt, bvh, v.min(0), v.max(0))
    assert len(pts) == ref["eval_points"]
    bc = EV.boundary(opt)
    wpp = ref["walks_per_point"]
    print("case: reference run zebra_k8_m2, %d points x %d walks" % (len(pts), wpp), flush=True)
    wp = W.WalkParams(max_steps=ref["max_steps"], epsilon=ref["epsilon"], seed=2024)
    st = W.Solver(bvh, bc, wp).solve(torch.from_numpy(pts).cuda(), wpp, per_walk=True)
    torch.cuda.synchronize()
    val = st.extra["walk_values"].cpu().numpy().astype(np.float64).reshape(-1, wpp)
    stp = st.extra["walk_steps"].cpu().numpy().astype(np.float64).reshape(-1, wpp)
    n = len(pts)
    exact = EV.exact_f64(bc, pts)
    mean = val.mean(1)
    var = val.var(1, ddof=1)
    err = mean - exact
    # statistic, its value here, its standard error for this fixed point set
    m4 = ((val - mean[:, None]) ** 4).mean(1)
    var_of_var = np.maximum(m4 - var ** 2 * (wpp - 3) / (wpp - 1), 0.0) / wpp
    mse_se = np.sqrt(np.sum(2 * (var / wpp) ** 2 + 4 * err ** 2 * var / wpp)) / n
    stats = {
        "mean_steps": (stp.mean(), np.sqrt(np.sum(stp.var(1, ddof=1) / wpp)) / n),
        "mean_sample_variance": (var.mean(), np.sqrt(np.sum(var_of_var)) / n),
        "mean_bias": (err.mean(), np.sqrt(np.sum(var / wpp)) / n),
        "mse": (np.mean(err ** 2), mse_se),
    }
    want = dict(ref, mse=ref["rmse"] ** 2)
    for k, (got, se) in stats.items():
        zd = (got - want[k]) / (np.sqrt(2.0) * se)
        print("%s: here %.6g, reference %.6g, difference %.2f standard errors" % (k, got, want[k], zd))
        assert abs(zd) < 4, k
    assert int(st.forced.cpu().numpy().sum()) == ref["forced_max_steps"]
