"""Progressive online cache training with snapshot overlap: the estimate of
the overlapped two-stream run equals, bit for bit, a one-stream replay of the
estimator blocks with the snapshots it recorded (the overlap changes nothing
the estimator computes), and the merged NC + 2LMC estimate is unbiased
against the FP64 solution (each block is unbiased whatever its snapshot)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_progressive_overlap_replays_exactly_and_is_unbiased(eng, torch_gpu):
    from n2wos_amd import cache as K, problems as P, progressive as PR, wos as W
    torch = torch_gpu
    v, t = P.make_mesh("mesh1")
    bvh = eng.Bvh(v, t)
    pts = P.interior_points(bvh, v, 4096, seed=21)
    bc = P.charges_boundary(0)
    truth = P.charges_solution_f64(bc.charges, pts)
    params = W.WalkParams(epsilon=1e-4, seed=23)
    cfg = PR.ProgressiveConfig(blocks=3, warmup_steps=20, steps_per_block=20, batch=1 << 14, label_pairs=8,
                               label_stride=2, coarse_wpp=16, residual_wpp=8, seed=5)
    tp = torch.from_numpy(pts).cuda()
    kc = K.CacheConfig.for_bounds(v.min(0), v.max(0), n_levels=2, seed=5)
    pr = PR.ProgressiveTwoLevel(bvh, bc, params, kc, cfg)
    print("case: progressive overlapped run with recorded snapshots", flush=True)
    est, co, re, info = pr.run(tp, overlap=True, record_snapshots=True)
    assert len(info["blocks"]) == 3 and len(info["snapshots"]) == 3
    # the snapshots differ from block to block (the cache kept training)
    assert not np.array_equal(info["snapshots"][0][0], info["snapshots"][2][0])
    print("case: progressive replay on one stream", flush=True)
    est2, _, _ = pr.replay(tp, info["snapshots"])
    e1, e2 = est.cpu().numpy(), est2.cpu().numpy()
    assert np.array_equal(e1.view(np.uint64), e2.view(np.uint64))
    var = (co.m2 / (co.n - 1) / co.n + re.m2 / (re.n - 1) / re.n).cpu().numpy()
    # points within epsilon of the boundary end every walk at step 0 (zero
    # variance, error = FP32 boundary data vs FP64 solution): excluded
    live = var > 0
    assert live.mean() > 0.99
    z = (e1 - truth)[live] / np.sqrt(var[live])
    assert abs(z.mean()) < 5 / np.sqrt(len(z))
    assert np.mean(np.abs(z) > 3.5) < 0.02
    assert int(co.n[0].item()) == 3 * 16 and int(re.n[0].item()) == 3 * 8
