"""The large plain case at full size: 128^3 mesh, N=128, NA=40, chi N = 20, box 6.0^3,
every pair stored (4.3 GB), disordered Gaussian start (sigma 0.1, seed 3). A propagator
evaluation on the slab path against the oracle (1e-10) and against the three-pass line
path (1e-12); the exact invariant mean(phi+) = 1; one Langevin step with the same host
noise on both GPU paths, which must take the same Anderson iteration counts and agree to
1e-8, the trajectory bar (the Anderson solve amplifies the paths' last-bit differences by
the condition number of its least-squares system: w+ agrees to about 2e-9 here)."""
import numpy as np
import pytest

import lfts_cpu_amd as L
import oracle as O

pytestmark = pytest.mark.gpu

N, NA, XN, C, NDT = 128, 40, 20.0, 100.0, 0.25
E, LB = 128, 6.0


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _start():
    M = E ** 3
    rng = np.random.default_rng(3)
    return rng.normal(0.0, 0.1, 2 * M)


def test_large_case_evaluation(gpu, monkeypatch):
    M = E ** 3
    m, Lb = (E, E, E), (LB, LB, LB)
    w = _start()
    print("128^3 N=128: slab path", flush=True)
    ctx = L.Context(N, NA, m, Lb)
    ctx.set_fields(w)
    lnq = ctx.calc_concs()
    assert "edge 128" in ctx.tuning()
    got = ctx.get_fields(4)
    print("128^3 N=128: line path", flush=True)
    monkeypatch.setenv("LFTS_PATH", "generic")
    gen = L.Context(N, NA, m, Lb)
    gen.set_fields(w)
    lnq_gen = gen.calc_concs()
    got_gen = gen.get_fields(4)
    assert abs(lnq - lnq_gen) <= 1e-12 * max(1.0, abs(lnq_gen))
    assert rel(got[2 * M:], got_gen[2 * M:]) < 1e-12
    for i in (2, N // 2, N):
        assert rel(ctx.pair(i), gen.pair(i)) < 1e-12, i
    gen.close()
    print("128^3 N=128: oracle", flush=True)
    d = O.Diblock(N, NA, m, Lb)
    wo = np.concatenate([w, np.zeros(2 * M)])
    lnq_ref = d.calc_concs(wo)
    assert abs(lnq - lnq_ref) <= 1e-10 * max(1.0, abs(lnq_ref))
    assert rel(got[2 * M:3 * M], wo[2 * M:3 * M]) < 1e-10
    assert rel(got[3 * M:], wo[3 * M:]) < 1e-10
    assert rel(ctx.pair(N), d.pair(N)) < 1e-10
    assert abs(got[3 * M:].mean() - 1.0) < 1e-12
    ctx.close()


def test_large_case_langevin_step_both_paths(gpu, monkeypatch):
    M = E ** 3
    m, Lb = (E, E, E), (LB, LB, LB)
    w = _start()
    sigma = (2.0 * NDT / (N * C)) ** 0.5 * (N ** 0.5)  # any sigma: both paths draw the same noise
    out = []
    for path in ("slab", "generic"):
        print(f"128^3 N=128 Langevin step: {path} path", flush=True)
        if path == "generic":
            monkeypatch.setenv("LFTS_PATH", "generic")
        ctx = L.Context(N, NA, m, Lb)
        ctx.set_fields(w)
        rng = L.Rng(123456789, sigma)
        ctx.langevin_init(rng.normals(M))
        k0, err0 = ctx.anderson_mix(200, 1e-4)
        ctx.langevin_step(rng.normals(M), XN, NDT)
        k1, err1 = ctx.anderson_mix(200, 1e-4)
        lnq = ctx.calc_concs()
        out.append((k0, k1, lnq, ctx.get_fields(4)))
        ctx.close()
    (ka0, ka1, qa, wa), (kb0, kb1, qb, wb) = out
    assert (ka0, ka1) == (kb0, kb1)
    assert abs(qa - qb) <= 1e-8 * max(1.0, abs(qb))
    assert rel(wa[:M], wb[:M]) < 1e-8 and rel(wa[M:2 * M], wb[M:2 * M]) < 1e-8
    assert rel(wa[3 * M:], wb[3 * M:]) < 1e-8


def test_set_with_a_128_member_equals_members_alone(gpu):
    """A replica set spanning all thirteen slab launch groups (every cubic slab edge from
    16^3 to 128^3) plus a line-path member gives every member's Q and fields bit for bit as
    it gets alone."""
    specs = [((128,) * 3, 6, 2), ((64,) * 3, 8, 3), ((48,) * 3, 7, 4), ((40,) * 3, 8, 5), ((32,) * 3, 9, 5),
             ((20, 18, 16), 6, 3), ((96,) * 3, 5, 2), ((16,) * 3, 12, 6), ((24,) * 3, 7, 3),
             ((56,) * 3, 6, 2), ((72,) * 3, 5, 3), ((80,) * 3, 5, 2), ((112,) * 3, 4, 2),
             ((120,) * 3, 4, 1)]

    def make():
        cs = []
        for k, (m, n, na) in enumerate(specs):
            c = L.Context(n, na, m, (4.1, 4.0, 3.9))
            M = int(np.prod(m))
            rng = np.random.default_rng(40 + k)
            c.set_fields(np.concatenate([rng.normal(0, 1.0, M), rng.normal(0, 0.5, M)]))
            cs.append(c)
        return cs

    print("set of 14 with a member of every slab edge", flush=True)
    ctxs = make()
    b = L.Batch(ctxs)
    lnq = b.calc_concs()
    k, err = b.anderson_mix(8, 1e-12)
    got = [c.get_fields(4) for c in ctxs]
    b.close()
    print("members alone", flush=True)
    for i, c in enumerate(make()):
        assert c.calc_concs() == lnq[i], i
        ki, erri = c.anderson_mix(8, 1e-12)
        assert ki == k[i] and erri == err[i], i
        assert np.array_equal(c.get_fields(4), got[i]), i
        c.close()
