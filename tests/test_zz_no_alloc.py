"""lfts_step_fwd_noalloc (the contour step on device arrays with its intermediate pair in
the library's fixed scratch array) against lfts_step_fwd (host arrays, its own buffer):
bit-identical, on two calls in a row (the second on a smaller mesh, reusing the same
scratch), and a mesh beyond the scratch's 128^3 points refused before any launch."""
import numpy as np
import pytest

import lfts_cpu_amd as L

pytestmark = pytest.mark.gpu

# (mesh, box, N, NA, bead index i): the first call on the larger mesh, the second smaller
CALLS = [((48, 40, 32), (3.7, 3.1, 2.6), 12, 5, 7), ((16, 12, 20), (2.2, 1.9, 2.9), 9, 4, 3)]


def _arrays(M, seed):
    rng = np.random.default_rng(seed)
    q = np.exp(rng.normal(0, 0.3, 2 * M))
    h = np.exp(rng.normal(0, 0.2, 2 * M))
    return q, h


def test_noalloc_step_equals_step_fwd_twice(gpu):
    import torch

    lib = L.lib()
    for n, (m, Lb, N, NA, i) in enumerate(CALLS):
        M = int(np.prod(m))
        print(f"call {n + 1}: mesh {m}, N={N}, NA={NA}, bead {i}", flush=True)
        ctx = L.Context(N, NA, m, Lb)
        q, h = _arrays(M, 10 + n)
        ref = ctx.step_fwd(q, h, i)
        qd = torch.from_numpy(q).cuda()
        hd = torch.from_numpy(h).cuda()
        out = torch.full((2 * M,), float("nan"), dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        rc = lib.lfts_step_fwd_noalloc(ctx.h, qd.data_ptr(), out.data_ptr(), hd.data_ptr(), i)
        assert rc == L.LFTS_OK, lib.lfts_last_error(ctx.h).decode()
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        assert np.array_equal(got, ref), np.abs(got - ref).max()
        # and the other end of the chain (the bead index picks the Boltzmann factors)
        ref2 = ctx.step_fwd(q, h, N - 1)
        assert lib.lfts_step_fwd_noalloc(ctx.h, qd.data_ptr(), out.data_ptr(), hd.data_ptr(), N - 1) == L.LFTS_OK
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), ref2)
        ctx.close()


def test_noalloc_step_refuses_what_it_cannot_serve(gpu):
    import torch

    lib = L.lib()
    big = L.Context(3, 1, (144, 128, 128), (4.0, 4.0, 4.0))  # 2359296 points > 128^3
    d = torch.zeros(8, dtype=torch.float64, device="cuda")
    p = d.data_ptr()
    assert lib.lfts_step_fwd_noalloc(big.h, p, p + 8, p + 16, 1) == L.LFTS_EINVAL
    assert "128^3" in lib.lfts_last_error(big.h).decode()
    big.close()
    small = L.Context(6, 2, (16, 16, 16), (3.0, 3.0, 3.0))
    assert lib.lfts_step_fwd_noalloc(small.h, p, p, p + 16, 1) == L.LFTS_EINVAL  # q_out = q_in
    assert lib.lfts_step_fwd_noalloc(small.h, p, p + 8, p + 16, 6) == L.LFTS_EINVAL  # i outside 1..N-1
    assert lib.lfts_step_fwd_noalloc(small.h, None, p + 8, p + 16, 1) == L.LFTS_EINVAL
    small.close()
