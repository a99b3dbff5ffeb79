"""Neural-cache walks and the two-level estimator on the GPU: prefix and
coupled walk outputs bit-exact against the oracle's walks; the combine step
against numpy in FP64; 2LMC unbiased for an untrained cache (3 sigma against
the FP64 solution); with the exact solution as the cache the residual has mean
zero (the reference's oracle-2LMC check, scripts/run_wavefront_oracle_sweep.py:112-160)."""
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


@pytest.mark.parametrize("poisson,anti", [(False, False), (True, False), (False, True), (True, True)])
def test_prefix_and_coupled_walks_bit_exact(eng, orc, torch_gpu, m1, poisson, anti):
    from n2wos_amd import nc as N, problems as P, wos as W
    torch = torch_gpu
    v, t, bvh, b2 = m1
    pts = P.interior_points(bvh, v, 64, seed=4)
    bc = P.poisson_boundary() if poisson else P.charges_boundary(0)
    wp = W.WalkParams(epsilon=1e-3, seed=21, source=W.SOURCE_POISSON_SINES if poisson else 0)
    s = N.NcSolver(bvh, bc, wp)
    tp = torch.from_numpy(pts).cuda()
    print("case: nc walks poisson=%s antithetic=%s" % (poisson, anti), flush=True)
    for mode, stream in ((N.NC_PREFIX, 1), (N.NC_COUPLED, 2)):
        b = s.walks(tp, 37, mode, 3, stream, antithetic=anti)
        torch.cuda.synchronize()
        kw = dict(max_steps=256, eps=1e-3, step_scale=0.999, bmode=bc.mode, bparams=bc.oracle_params(),
                  poisson=poisson, seed=21, stream=stream, nthreads=16, antithetic=anti)
        ov, os_, ost, oxm = b2.walks(pts, 37, depth=3, want_xm=True, **kw)
        need = b.need.cpu().numpy()
        assert np.array_equal(need, (ost == 2).astype(np.uint8))
        assert np.array_equal(bits(b.prefix.cpu().numpy()), bits(ov))
        live = need == 1
        assert np.array_equal(bits(b.xm.cpu().numpy()[live]), bits(oxm[live]))
        if mode == N.NC_COUPLED:
            fv, fs, _, _ = b2.walks(pts, 37, **kw)
            assert np.array_equal(bits(b.full.cpu().numpy()), bits(fv))
            assert np.array_equal(b.steps.cpu().numpy(), fs.reshape(64, 37).sum(1))


def test_antithetic_pairs_unbiased_and_lower_variance(eng, torch_gpu, m1):
    """Depth-1 NC with the exact solution as the cache: mirrored first steps
    keep the estimate unbiased (z-scores of pair means against the FP64
    solution) and cut the variance of a pair mean well below that of two
    independent samples (the odd part of u on the first sphere cancels)."""
    from n2wos_amd import nc as N, problems as P, wos as W
    torch = torch_gpu
    v, t, bvh, _ = m1
    pts = P.interior_points(bvh, v, 1024, seed=13)
    bc = P.charges_boundary(0)
    s = N.NcSolver(bvh, bc, W.WalkParams(epsilon=1e-4, seed=17))
    tp = torch.from_numpy(pts).cuda()
    truth = P.charges_solution_f64(bc.charges, pts)
    out = {}
    for anti in (False, True):
        b = s.walks(tp, 128, N.NC_PREFIX, 1, 1, antithetic=anti)
        need = b.need.cpu().numpy().astype(bool)
        xm = b.xm.cpu().numpy()
        val = b.prefix.cpu().numpy().astype(np.float64)
        val[need] += P.charges_solution_f64(bc.charges, xm[need])
        out[anti] = val.reshape(1024, 64, 2).mean(2)  # pair means
    for anti, pm in out.items():
        se = pm.std(1, ddof=1) / np.sqrt(64)
        z = (pm.mean(1) - truth) / np.maximum(se, 1e-12)
        assert abs(z.mean()) < 5 / np.sqrt(len(z)), anti
        assert np.mean(np.abs(z) > 3.5) < 0.02, anti
    v_ind = out[False].var(1, ddof=1).mean()
    v_anti = out[True].var(1, ddof=1).mean()
    assert v_anti < 0.3 * v_ind, (v_anti, v_ind)


def test_combine_matches_numpy(eng, torch_gpu, m1):
    from n2wos_amd import nc as N, problems as P
    torch = torch_gpu
    v, t, bvh, _ = m1
    pts = P.interior_points(bvh, v, 50, seed=5)
    s = N.NcSolver(bvh, P.charges_boundary(0))
    tp = torch.from_numpy(pts).cuda()
    b = s.walks(tp, 19, N.NC_COUPLED, 2, 2)
    cv = torch.randn(50 * 19, device="cuda")
    st = s.combine(b, cv, 50, 19, N.NC_COUPLED)
    c = s.combine(b, cv, 50, 19, N.NC_PREFIX)
    need = b.need.cpu().numpy().astype(bool)
    pre, full, cvn = b.prefix.cpu().numpy(), b.full.cpu().numpy(), cv.cpu().numpy()
    r = np.where(need, full - (pre + np.where(need, cvn, 0)), 0).astype(np.float32).astype(np.float64).reshape(50, 19)
    a = (pre + np.where(need, cvn, 0)).astype(np.float32).astype(np.float64).reshape(50, 19)
    assert np.allclose(st.mean.cpu().numpy(), r.mean(1), atol=1e-12)
    assert np.allclose(st.m2.cpu().numpy(), ((r - r.mean(1, keepdims=True)) ** 2).sum(1), atol=1e-10)
    assert np.allclose(c.mean.cpu().numpy(), a.mean(1), atol=1e-12)


def test_two_level_unbiased_with_untrained_cache(eng, torch_gpu, m1):
    """Whatever the cache, coarse + residual is unbiased: z-scores against the
    FP64 solution stay standard normal (random-init cache)."""
    from n2wos_amd import cache as K, nc as N, problems as P, wos as W
    torch = torch_gpu
    v, t, bvh, _ = m1
    pts = P.interior_points(bvh, v, 1024, seed=6)
    bc = P.charges_boundary(0)
    s = N.NcSolver(bvh, bc, W.WalkParams(epsilon=1e-4, seed=8))
    cache = K.Cache(K.CacheConfig.for_bounds(v.min(0), v.max(0), seed=2))
    print("case: 2lmc unbiased untrained cache", flush=True)
    est, co, re, _ = s.two_level(cache, torch.from_numpy(pts).cuda(), 256, 256, 2)
    var = co.m2 / (co.n - 1) / co.n + re.m2 / (re.n - 1) / re.n
    z = ((est - torch.from_numpy(P.charges_solution_f64(bc.charges, pts)).cuda()) / var.sqrt()).cpu().numpy()
    assert np.mean(np.abs(z) > 3) < 0.015
    assert abs(z.mean()) < 5 / np.sqrt(len(z))


def test_exact_cache_residual_has_zero_mean(eng, torch_gpu, m1):
    """With C = the exact solution the residual W - S_m - u(X_m) has mean zero:
    checks the coupling of prefix and continuation and the stopping rule."""
    from n2wos_amd import nc as N, problems as P, wos as W
    torch = torch_gpu
    v, t, bvh, _ = m1
    pts = P.interior_points(bvh, v, 512, seed=7)
    bc = P.charges_boundary(0)
    s = N.NcSolver(bvh, bc, W.WalkParams(epsilon=1e-4, seed=9))
    b = s.walks(torch.from_numpy(pts).cuda(), 256, N.NC_COUPLED, 3, 2)
    exact = torch.from_numpy(P.charges_solution_f64(bc.charges, b.xm.cpu().numpy()).astype(np.float32)).cuda()
    re = s.combine(b, exact, 512, 256, N.NC_COUPLED)
    z = (re.mean / (re.m2 / 255 / 256).clamp(min=1e-30).sqrt()).cpu().numpy()
    z = z[np.isfinite(z) & (re.m2.cpu().numpy() > 0)]
    assert abs(z.mean()) < 5 / np.sqrt(len(z))
    assert np.mean(np.abs(z) > 3) < 0.02


def test_trained_cache_gain_over_plain_and_untrained(eng, torch_gpu, m1):
    """What the trained network gains: a 4-level cache is trained on the GPU
    from the solver's own coupled walks (depth 1, 32 per point over 16,384
    interior points, 100 Adam steps), then the neural-cache estimate
    S_1 + C(X_1) with 64 samples per point -- at most 64 walk steps per
    point -- is compared with plain WoS given MORE walk steps per point (4
    full walks, ~98 steps) and with the same estimator on the untrained
    network.  RMSE against the FP64 solution (measured: 0.014 vs 0.085 plain,
    1.08 untrained)."""
    from n2wos_amd import cache as K, nc as N, problems as P, wos as W
    torch = torch_gpu
    v, t, bvh, _ = m1
    pts = P.interior_points(bvh, v, 16384, seed=12)
    tp = torch.from_numpy(pts).cuda()
    bc = P.charges_boundary(0)
    truth = P.charges_solution_f64(bc.charges, pts)
    wp = W.WalkParams(epsilon=1e-4, seed=31)
    s = N.NcSolver(bvh, bc, wp)
    cfg = K.CacheConfig.for_bounds(v.min(0), v.max(0), n_levels=4, seed=5)
    trained, untrained = K.Cache(cfg), K.Cache(cfg)
    print("case: cache gain (4 levels, train depth 1, 32 pairs, 100 steps)", flush=True)
    b = s.walks(tp, 32, N.NC_COUPLED, 1, stream_id=7, seed=41)
    x, y = s.training_pairs(b)
    gen = torch.Generator(device="cuda").manual_seed(3)
    for _ in range(100):
        idx = torch.randint(0, x.shape[0], (1 << 16,), device="cuda", generator=gen)
        trained.train_step(x[idx].contiguous(), y[idx].contiguous())

    def rmse(st):
        return float(np.sqrt(np.mean((st.mean.cpu().numpy() - truth) ** 2)))

    nc_t, wb = s.nc(trained, tp, 64, 1, stream_id=1, seed=51)
    nc_u, _ = s.nc(untrained, tp, 64, 1, stream_id=1, seed=51)
    plain = W.Solver(bvh, bc, wp).solve(tp, 4, seed=61)
    torch.cuda.synchronize()
    nc_steps = float(wb.steps.sum().item()) / len(pts)
    plain_steps = float(plain.steps.sum().item()) / len(pts)
    r_t, r_u, r_p = rmse(nc_t), rmse(nc_u), rmse(plain)
    print("rmse trained %.4g untrained %.4g plain %.4g; steps/point nc %.1f plain %.1f" %
          (r_t, r_u, r_p, nc_steps, plain_steps), flush=True)
    assert nc_steps <= 64 < plain_steps
    assert r_t < 0.4 * r_p
    assert r_t < 0.1 * r_u
