"""The contour-step kernel shapes the autotuner chooses between. The variants of one line
factorisation (twiddle/kernel tables in registers or LDS; one or two fields per
workgroup) compute bit-identical results, for the plain chain alone (with and without the
fused density sums) and in a replica set, and for the tangent and adjoint chains. The
other factorisations R1 x R2 of each edge round differently: they agree with the default
to 1e-12 and with the oracle to 1e-10, and their own variants agree bit for bit."""
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
    monkeypatch.setenv("LFTS_SLAB_SHAPES", "default")
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


ALTERNATIVES = [(16, 2, 8), (24, 6, 4), (32, 8, 4), (40, 10, 4), (48, 12, 4), (56, 14, 4), (64, 16, 4), (64, 4, 16)]
VARIANTS = {(64, 16, 4): (1, 3), (64, 4, 16): (0, 1)}  # the others spill registers and are not built


def _alt_run(n, shape, variant, w2, monkeypatch):
    """One context (calc_concs storing every pair, then an Anderson solve whose chains fuse
    the density sums) and a set of two (calc_concs), factorisation and variant fixed."""
    monkeypatch.setenv("LFTS_SLAB_SHAPES", shape)
    if variant is None:
        monkeypatch.delenv("LFTS_SLAB_VARIANT", raising=False)
    else:
        monkeypatch.setenv("LFTS_SLAB_VARIANT", str(variant))
    m = (n, n, n)
    c = L.Context(11, 4, m, (3.9,) * 3)
    c.set_fields(w2[0])
    lnq = c.calc_concs()
    phi = c.get_fields(4)
    k, _ = c.anderson_mix(30, 1e-6)
    mixed = c.get_fields(4)
    rep = c.tuning()
    c.close()
    a, b = L.Context(11, 4, m, (3.9,) * 3), L.Context(9, 6, m, (4.3,) * 3)
    a.set_fields(w2[0])
    b.set_fields(w2[1])
    B = L.Batch([a, b])
    lnq2 = B.calc_concs()
    rep2 = B.tuning()
    f2 = [a.get_fields(4), b.get_fields(4)]
    B.close()
    a.close()
    b.close()
    return lnq, phi, k, mixed, lnq2, f2, rep + rep2


@pytest.mark.parametrize("n,r1,r2", ALTERNATIVES)
def test_alternative_factorisations(gpu, monkeypatch, n, r1, r2):
    import oracle as O

    M = n ** 3
    w2 = [_fields(M, 70 + n), _fields(M, 71 + n)]
    print(f"edge {n}: {r1}x{r2} against the default", flush=True)
    d = _alt_run(n, "default", 1, w2, monkeypatch)
    runs = [_alt_run(n, f"{n}:{r1}x{r2}", v, w2, monkeypatch) for v in VARIANTS.get((n, r1, r2), (0, 1, 2, 3))]
    for v, g in zip(VARIANTS.get((n, r1, r2), (0, 1, 2, 3)), runs):
        assert f"{r1}x{r2} variant {v} (fixed)" in g[6], g[6]
    g = runs[0]
    # the variants of the alternative factorisation: bit for bit
    for o in runs[1:]:
        assert o[0] == g[0] and o[2] == g[2] and np.array_equal(o[4], g[4])
        for x, y in zip((o[1], o[3]) + tuple(o[5]), (g[1], g[3]) + tuple(g[5])):
            assert np.array_equal(x, y)
    # against the default factorisation and the oracle
    rel = lambda x, y: np.abs(x - y).max() / np.abs(y).max()
    assert abs(g[0] - d[0]) <= 1e-12 * max(1.0, abs(d[0]))
    assert rel(g[1][2 * M:], d[1][2 * M:]) <= 1e-12
    assert np.abs(g[4] - d[4]).max() <= 1e-12 * max(1.0, np.abs(d[4]).max())
    assert g[2] == d[2]  # the same Anderson iteration count
    assert rel(g[3][M:2 * M], d[3][M:2 * M]) <= 1e-9
    wo = np.zeros(4 * M)
    wo[:2 * M] = w2[0]
    lnq_o = O.Diblock(11, 4, (n, n, n), (3.9,) * 3).calc_concs(wo)
    assert abs(g[0] - lnq_o) <= 1e-10 * max(1.0, abs(lnq_o))
    assert rel(g[1][2 * M:], wo[2 * M:]) <= 1e-10
