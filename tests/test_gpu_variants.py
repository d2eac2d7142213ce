"""The contour-step kernel shapes the autotuner chooses between (twiddle/kernel tables in
registers or LDS; one or two fields per workgroup) compute bit-identical results, for
the plain chain alone (with and without the fused density sums) and in a replica set,
and for the tangent and adjoint chains."""
import numpy as np
import pytest

import lfts_cpu_amd as L

pytestmark = pytest.mark.gpu


def _fields(M, seed):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.normal(0, 2.0, M), rng.normal(0, 1.0, M)])


@pytest.mark.parametrize("n", [16, 24, 32, 40, 48, 56, 64])
def test_every_shape_gives_the_same_bits(gpu, monkeypatch, n):
    m, N, NA = (n, n, n), 12, 5
    M = n ** 3
    w = _fields(M, n)
    rng = np.random.default_rng(1)
    v, ub = rng.normal(size=M), rng.normal(size=2 * M)
    ref = None
    for variant in range(4):
        print(f"edge {n} variant {variant}", flush=True)
        monkeypatch.setenv("LFTS_SLAB_VARIANT", str(variant))
        c = L.Context(N, NA, m, (4.0, 4.0, 4.0))
        c.set_fields(w)
        lnq = c.calc_concs()
        assert f"variant {variant} (fixed)" in c.tuning()
        got = (lnq, c.get_fields(4), c.pair(N), c.jvp(None, v), c.vjp(ub))
        H = c.hamiltonian(17.0)  # density sums fused into the chain
        got = got + (c.get_fields(4),)
        c.close()
        if ref is None:
            ref, H_ref = got, H
        else:
            assert got[0] == ref[0] and H == H_ref
            for a, b in zip(got[1:], ref[1:]):
                assert np.array_equal(a, b)


def test_tuned_set_matches_fixed_shape(gpu, monkeypatch):
    specs = [(32, 10, 4), (48, 9, 3), (64, 8, 5), (32, 12, 6)]
    ws = [_fields(n ** 3, 10 + k) for k, (n, _, _) in enumerate(specs)]

    def run(fixed):
        if fixed is None:
            monkeypatch.delenv("LFTS_SLAB_VARIANT", raising=False)
        else:
            monkeypatch.setenv("LFTS_SLAB_VARIANT", str(fixed))
        ctxs = [L.Context(N, NA, (n,) * 3, (4.2,) * 3) for n, N, NA in specs]
        for c, w in zip(ctxs, ws):
            c.set_fields(w)
        B = L.Batch(ctxs)
        lnq = B.calc_concs()
        rep = B.tuning()
        out = [c.get_fields(4) for c in ctxs]
        B.close()
        for c in ctxs:
            c.close()
        return lnq, out, rep

    l0, o0, _ = run(0)
    l1, o1, rep = run(None)
    print(rep, flush=True)
    assert "edge 64 x 1 " in rep and "edge 48 x 1 " in rep and "edge 32 x 2 " in rep
    assert np.array_equal(l0, l1)
    for a, b in zip(o0, o1):
        assert np.array_equal(a, b)
