"""Learned cache on the GPU against the FP64 oracle: layout; forward and MSE
gradients against the oracle variant that encodes in FP32 and rounds to
binary16 where the kernels do (FP16 MFMA operands, FP32 accumulation:
relative L2 error <= 1e-3 per weight group, 2e-3 for the grid) and against
the exact FP64 oracle (FP16 operands alone move these gradients by a few
percent, because the loss terms of random targets cancel: <= 6e-2); Adam's
first step in closed form; an overfit test on x^2 - y^2."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / max(1e-30, np.linalg.norm(b))


@pytest.fixture(scope="module")
def small(eng, torch_gpu):
    from n2wos_amd import cache as K
    cfg = K.CacheConfig(n_levels=16, log2_table_size=12, base_resolution=4, growth=1.5, seed=3)
    c = K.Cache(cfg)
    lay = c.layout()
    rng = np.random.default_rng(0)
    grid = rng.normal(0, 0.5, 2 * lay["total_entries"]).astype(np.float32)
    w = np.zeros(K.N_WEIGHTS, np.float32)
    w[:2048] = rng.normal(0, 0.25, 2048)
    w[2048:14336] = rng.normal(0, 0.18, 12288)
    w[14336:14400] = rng.normal(0, 0.3, 64)
    c.set_params(grid, w)
    return cfg, c, grid, w


def test_layout_matches_oracle(eng, orc, small):
    cfg, c, _, _ = small
    lay = c.layout()
    ol = orc.cache_layout(cfg.n_levels, cfg.log2_table_size, cfg.base_resolution, cfg.growth)
    for k in ("offsets", "sizes", "res", "dense", "total_entries"):
        assert lay[k] == ol[k], k


def test_forward_matches_fp64_oracle(eng, orc, torch_gpu, small):
    torch = torch_gpu
    cfg, c, grid, w = small
    pts = np.random.default_rng(1).uniform(-1.1, 1.1, (1000, 3)).astype(np.float32)
    print("case: cache forward 1000", flush=True)
    y = c.infer(torch.from_numpy(pts).cuda()).cpu().numpy()
    ref = orc.cache_forward(cfg, grid, w, pts)
    ref16 = orc.cache_forward(cfg, grid, w, pts, fp16=True)
    print("forward rel vs fp16-rounding oracle %.3e, vs exact %.3e" % (rel(y, ref16), rel(y, ref)), flush=True)
    assert np.std(ref) > 0.1
    assert rel(y, ref16) < 1e-3
    assert rel(y, ref) < 2e-2
    assert np.max(np.abs(y - ref)) < 0.1 * np.max(np.abs(ref))


def test_gradients_match_fp64_oracle(eng, orc, torch_gpu, small):
    torch = torch_gpu
    cfg, c, grid, w = small
    rng = np.random.default_rng(2)
    pts = rng.uniform(-1.0, 1.0, (1000, 3)).astype(np.float32)
    tg = rng.normal(size=1000).astype(np.float32)
    print("case: cache gradients 1000", flush=True)
    gg, wg, loss = c.gradients(torch.from_numpy(pts).cuda(), torch.from_numpy(tg).cuda())
    gg2, wg2, _ = c.gradients(torch.from_numpy(pts).cuda(), torch.from_numpy(tg).cuda())
    assert np.array_equal(wg, wg2)  # weight gradients: fixed reduction order
    ogg, owg, oloss = orc.cache_grads(cfg, grid, w, pts, tg)
    hgg, hwg, hloss = orc.cache_grads_fp16(cfg, grid, w, pts, tg, cfg.loss_scale)
    groups = [(0, 2048), (2048, 6144), (6144, 10240), (10240, 14336), (14336, 14400)]
    print("W rel vs fp16 oracle", ["%.2e" % rel(wg[a:b], hwg[a:b]) for a, b in groups],
          "grid %.2e" % rel(gg, hgg), "| vs exact", ["%.2e" % rel(wg[a:b], owg[a:b]) for a, b in groups],
          "grid %.2e" % rel(gg, ogg), flush=True)
    assert abs(loss - oloss) < 2e-2 * oloss
    for lo, hi in groups:
        assert rel(wg[lo:hi], hwg[lo:hi]) < 1e-3, (lo, hi)
        assert rel(wg[lo:hi], owg[lo:hi]) < 6e-2, (lo, hi)
    assert np.all(wg[14400:] == 0)
    assert rel(gg, hgg) < 2e-3
    assert rel(gg, ogg) < 6e-2


def test_adam_first_step_closed_form(eng, orc, torch_gpu):
    """Step 1 of Adam: m_hat = g, v_hat = g^2, so p1 = p0 - lr g / (|g| + eps)."""
    torch = torch_gpu
    from n2wos_amd import cache as K
    cfg = K.CacheConfig(n_levels=4, log2_table_size=10, base_resolution=4, learning_rate=1e-2, seed=5)
    c = K.Cache(cfg)
    g0, w0 = c.get_params()
    rng = np.random.default_rng(3)
    pts = torch.from_numpy(rng.uniform(-1, 1, (512, 3)).astype(np.float32)).cuda()
    tg = torch.from_numpy(rng.normal(size=512).astype(np.float32)).cuda()
    gg, wg, _ = c.gradients(pts, tg)
    c.train_step(pts, tg)
    torch.cuda.synchronize()
    g1, w1 = c.get_params()
    live = np.abs(wg) > 1e-6
    expect = w0[live] - 1e-2 * wg[live] / np.abs(wg[live])
    assert np.allclose(w1[live], expect, atol=2e-4)
    assert np.array_equal(w1[~live & (wg == 0)], w0[~live & (wg == 0)])
    liveg = np.abs(gg) > 1e-7
    assert np.allclose(g1[liveg], g0[liveg] - 1e-2 * np.sign(gg[liveg]), atol=2e-4)


def test_cache_overfits_smooth_harmonic(eng, torch_gpu):
    """Cache probe shape of the reference (fill x^2 - y^2 targets, train,
    infer, compare): RMSE on held-out points falls below 0.02."""
    torch = torch_gpu
    from n2wos_amd import cache as K
    c = K.Cache(K.CacheConfig(seed=7))
    g = torch.Generator(device="cuda").manual_seed(0)
    print("case: cache overfit x^2-y^2", flush=True)
    for step in range(300):
        p = torch.rand((1 << 16, 3), device="cuda", generator=g) * 2 - 1
        t = p[:, 0] ** 2 - p[:, 1] ** 2
        c.train_step(p.contiguous(), t.contiguous())
    q = torch.rand((1 << 14, 3), device="cuda", generator=g) * 2 - 1
    y = c.infer(q.contiguous())
    rmse = torch.sqrt(torch.mean((y - (q[:, 0] ** 2 - q[:, 1] ** 2)) ** 2)).item()
    print("overfit rmse", rmse)
    assert rmse < 0.02


def test_output_offset_in_inference_and_loss(eng, torch_gpu):
    """The output offset adds exactly to inference, a zero offset changes
    nothing bit for bit, and training with offset o on targets t follows
    training without it on t - o (the loss sees (y + o) - t): after three Adam
    steps the two networks' learned changes agree to 1e-2 relative."""
    from n2wos_amd import cache as K
    torch = torch_gpu
    cfg = K.CacheConfig(n_levels=4, log2_table_size=12, base_resolution=4, growth=1.5, seed=9)
    g = torch.Generator(device="cuda").manual_seed(5)
    x = (torch.rand((4096, 3), device="cuda", generator=g) * 2.2 - 1.1).contiguous()
    t = (torch.sin(2 * x[:, 0]) * x[:, 1] + 3.0).contiguous()
    a, b = K.Cache(cfg), K.Cache(cfg)
    y0 = a.infer(x)
    a.set_output_offset(3.0)
    y1 = a.infer(x)
    assert torch.equal(y1, y0 + 3.0)
    a.set_output_offset(0.0)
    assert torch.equal(a.infer(x), y0)
    a.set_output_offset(3.0)
    t_shift = (t - 3.0).contiguous()
    for _ in range(3):
        a.train_step(x, t)
        b.train_step(x, t_shift)
    da = (a.infer(x) - 3.0 - y0).cpu().numpy()
    db = (b.infer(x) - y0).cpu().numpy()
    assert np.linalg.norm(db) > 1e-3
    assert rel(da, db) < 1e-2
    with pytest.raises(Exception):
        a.set_output_offset(float("nan"))


def test_training_batch_limit_is_set_at_creation(eng, torch_gpu):
    """The activation store is sized by max_batch when the cache is created:
    batches up to it train, a larger one is rejected before anything is
    enqueued (no allocation inside a training step)."""
    from n2wos_amd import cache as K
    from n2wos_amd._lib import ERR_INVALID, N2wosError
    torch = torch_gpu
    c = K.Cache(K.CacheConfig(seed=5, n_levels=2, max_batch=1000))
    g = torch.Generator(device="cuda").manual_seed(1)
    x = (torch.rand((1008, 3), device="cuda", generator=g) * 2 - 1).contiguous()
    y = torch.rand(1008, device="cuda", generator=g)
    c.train_step(x[:1000].contiguous(), y[:1000].contiguous())
    c.train_step(x, y)  # 1008 = 1000 rounded up to whole 16-sample tiles
    with pytest.raises(N2wosError) as e:
        c.train_step(torch.zeros((1009, 3), device="cuda"), torch.zeros(1009, device="cuda"))
    assert e.value.code == ERR_INVALID
    torch.cuda.synchronize()


@pytest.mark.parametrize("h,layers", [(16, 1), (32, 2), (64, 3), (16, 4)])
def test_mlp_shapes_match_fp16_oracle(eng, orc, torch_gpu, h, layers):
    """Every MLP shape (width 16/32/64, depth 1-4): forward and MSE gradients
    against the oracle with the kernels' binary16 rounding points (relative L2
    <= 1e-3 per layer, 2e-3 for the grid), reproducible weight gradients, and
    untouched padded output rows."""
    torch = torch_gpu
    from n2wos_amd import cache as K
    cfg = K.CacheConfig(n_levels=8, log2_table_size=11, base_resolution=4, growth=1.5, seed=11, n_neurons=h,
                        n_hidden_layers=layers, max_batch=2048)
    c = K.Cache(cfg)
    rng = np.random.default_rng(h * 10 + layers)
    grid = rng.normal(0, 0.5, 2 * c.layout()["total_entries"]).astype(np.float32)
    nw = K.n_weights(h, layers)
    off = 32 * h + (layers - 1) * h * h
    w = np.zeros(nw, np.float32)
    w[:32 * h] = rng.normal(0, 0.25, 32 * h)
    w[32 * h:off] = rng.normal(0, 1.4 / np.sqrt(h), off - 32 * h)
    w[off:off + h] = rng.normal(0, 0.3, h)
    c.set_params(grid, w)
    g2, w2 = c.get_params()
    assert np.array_equal(g2, grid) and np.array_equal(w2, w)
    pts = rng.uniform(-1.05, 1.05, (1000, 3)).astype(np.float32)
    tg = rng.normal(size=1000).astype(np.float32)
    print("case: cache shape %dx%d" % (h, layers), flush=True)
    y = c.infer(torch.from_numpy(pts).cuda()).cpu().numpy()
    ref16 = orc.cache_forward(cfg, grid, w, pts, fp16=True)
    assert np.std(ref16) > 0.05
    assert rel(y, ref16) < 1e-3
    gg, wg, loss = c.gradients(torch.from_numpy(pts).cuda(), torch.from_numpy(tg).cuda())
    _, wg2, _ = c.gradients(torch.from_numpy(pts).cuda(), torch.from_numpy(tg).cuda())
    assert np.array_equal(wg, wg2)
    hgg, hwg, hloss = orc.cache_grads_fp16(cfg, grid, w, pts, tg, cfg.loss_scale)
    groups = [(0, 32 * h)] + [(32 * h + k * h * h, 32 * h + (k + 1) * h * h) for k in range(layers - 1)] + \
        [(off, off + h)]
    print("W rel", ["%.2e" % rel(wg[a:b], hwg[a:b]) for a, b in groups], "grid %.2e" % rel(gg, hgg), flush=True)
    assert abs(loss - hloss) < 1e-2 * hloss
    for lo, hi in groups:
        assert rel(wg[lo:hi], hwg[lo:hi]) < 1e-3, (lo, hi)
    assert np.all(wg[off + h:] == 0)
    assert rel(gg, hgg) < 2e-3
    # one training step moves the loss down on the same batch
    p_t, t_t = torch.from_numpy(pts).cuda(), torch.from_numpy(tg).cuda()
    for _ in range(20):
        c.train_step(p_t, t_t)
    _, _, loss2 = c.gradients(p_t, t_t)
    assert loss2 < loss


@pytest.mark.parametrize("bad_value", [float("nan"), float("inf")])
def test_non_finite_step_is_skipped(eng, torch_gpu, bad_value):
    """A batch with one non-finite target makes that step's gradients
    non-finite: the step leaves every parameter as it was (the kernels flag
    the terms where they compute them), and the next finite batch trains."""
    from n2wos_amd import cache as K
    torch = torch_gpu
    c = K.Cache(K.CacheConfig(n_levels=4, log2_table_size=12, base_resolution=4, growth=1.5, seed=5, max_batch=4096))
    g = torch.Generator(device="cuda").manual_seed(3)
    x = (torch.rand((4096, 3), device="cuda", generator=g) * 2 - 1).contiguous()
    y = (x[:, 0] * x[:, 1]).contiguous()
    c.train_step(x, y)
    grid0, w0 = c.get_params()
    yb = y.clone()
    yb[1234] = bad_value
    c.train_step(x, yb.contiguous())
    grid1, w1 = c.get_params()
    assert np.array_equal(grid0.view(np.uint32), grid1.view(np.uint32))
    assert np.array_equal(w0.view(np.uint32), w1.view(np.uint32))
    c.train_step(x, y)
    grid2, w2 = c.get_params()
    assert np.all(np.isfinite(grid2)) and np.all(np.isfinite(w2))
    assert not np.array_equal(w1, w2) and not np.array_equal(grid1, grid2)
