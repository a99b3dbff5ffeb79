"""n2wos_wos_solve_noalloc (walker workspace in a file-scope device array,
no allocation) against n2wos_wos_solve with a caller workspace: per-point
statistics and per-walk outputs bit for bit, on two calls in a row (the second
smaller, so the array still holds the first call's contents), and a job beyond
the array's size rejected before anything is enqueued."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _outputs(torch, n, wpp, per_walk):
    f64 = dict(dtype=torch.float64, device="cuda")
    o = [torch.full((n,), -1.0, **f64), torch.full((n,), -1.0, **f64), torch.full((n,), -1.0, **f64),
         torch.zeros(n, dtype=torch.int64, device="cuda"), torch.zeros(n, dtype=torch.int32, device="cuda")]
    if per_walk:
        o += [torch.zeros(n * wpp, dtype=torch.float32, device="cuda"),
              torch.zeros(n * wpp, dtype=torch.int32, device="cuda")]
    return o


def _call(L, fn, bvh, bc, wp, pts, wpp, out, stream, ws=None):
    n = pts.shape[0]
    pw = (out[5].data_ptr(), out[6].data_ptr()) if len(out) > 5 else (None, None)
    head = (bvh.handle, C.byref(bc), C.byref(wp), pts.data_ptr(), n, wpp, 0, out[0].data_ptr(), out[1].data_ptr(),
            out[2].data_ptr(), out[3].data_ptr(), out[4].data_ptr(), 0)
    if fn == "ws":
        return L.n2wos_wos_solve(*head, ws.data_ptr(), ws.numel(), *pw, stream)
    return L.n2wos_wos_solve_noalloc(*head, *pw, stream)


def _same(a, b):
    for x, y in zip(a, b):
        xa, ya = x.cpu().numpy(), y.cpu().numpy()
        assert np.array_equal(xa.view(np.uint8), ya.view(np.uint8))


def test_noalloc_solve_equals_workspace_solve(eng, torch_gpu):
    from n2wos_amd import problems as P, wos as W
    from n2wos_amd._lib import ERR_INVALID, check, lib
    torch = torch_gpu
    L = lib()
    v, t = P.make_mesh("mesh1")
    bvh = eng.Bvh(v, t)
    bc = P.charges_boundary(0).to_c()
    wp = W.WalkParams(max_steps=256, epsilon=1e-3, seed=11).to_c()
    stream = torch.cuda.current_stream().cuda_stream
    for n, wpp, per_walk in [(3000, 100, False), (1000, 70, True)]:
        print("case: noalloc vs workspace solve %d x %d per_walk=%s" % (n, wpp, per_walk), flush=True)
        pts = torch.from_numpy(P.interior_points(bvh, v, n, seed=n)).cuda()
        ws = torch.empty(W.workspace_bytes(n, wpp), dtype=torch.uint8, device="cuda")
        ref = _outputs(torch, n, wpp, per_walk)
        got = _outputs(torch, n, wpp, per_walk)
        check(_call(L, "ws", bvh, bc, wp, pts, wpp, ref, stream, ws))
        check(_call(L, "noalloc", bvh, bc, wp, pts, wpp, got, stream))
        torch.cuda.synchronize()
        _same(ref, got)
        assert np.all(got[0].cpu().numpy() == wpp)

    # beyond the scratch array (142 MB of workspace > 96 MiB): rejected on the
    # host.  Buffers are sized for the job, so every pointer is valid.
    n, wpp = 700_000, 256
    assert W.workspace_bytes(n, wpp) > (96 << 20)
    print("case: noalloc size rejection %d x %d" % (n, wpp), flush=True)
    pts = torch.zeros((n, 3), dtype=torch.float32, device="cuda")
    out = _outputs(torch, n, wpp, False)
    rc = _call(L, "noalloc", bvh, bc, wp, pts, wpp, out, stream)
    assert rc == ERR_INVALID
    assert b"n2wos_wos_solve_noalloc holds" in L.n2wos_last_error()
    torch.cuda.synchronize()
    assert np.all(out[0].cpu().numpy() == -1.0)  # nothing was written
