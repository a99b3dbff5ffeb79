"""Progressive online cache training with snapshot overlap: the overlapped
two-stream run gives the same estimate bit for bit as the same schedule on
one stream, and the merged NC + 2LMC estimate is unbiased against the FP64
solution (each block is unbiased whatever its snapshot)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_progressive_overlap_bit_identical_and_unbiased(eng, torch_gpu):
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
    out = {}
    for overlap in (False, True):
        print("case: progressive overlap=%s" % overlap, flush=True)
        kc = K.CacheConfig.for_bounds(v.min(0), v.max(0), n_levels=2, seed=5)
        est, co, re, info = PR.ProgressiveTwoLevel(bvh, bc, params, kc, cfg).run(tp, overlap=overlap)
        out[overlap] = (est.cpu().numpy(), co, re, info)
        assert len(info["blocks"]) == 3
    e0, e1 = out[False][0], out[True][0]
    assert np.array_equal(e0.view(np.uint64), e1.view(np.uint64))
    _, co, re, _ = out[True]
    var = (co.m2 / (co.n - 1) / co.n + re.m2 / (re.n - 1) / re.n).cpu().numpy()
    z = (e1 - truth) / np.sqrt(np.maximum(var, 1e-30))
    assert abs(z.mean()) < 5 / np.sqrt(len(z))
    assert np.mean(np.abs(z) > 3.5) < 0.02
    assert int(co.n[0].item()) == 3 * 16 and int(re.n[0].item()) == 3 * 8
