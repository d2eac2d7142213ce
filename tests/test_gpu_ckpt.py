"""Checkpointed propagator storage (pairs kept at symmetric segment starts, segments
recomputed with their mirrors) against full storage: the same propagator evaluation,
Anderson solve, tangent and adjoint products and box derivative to 1e-12 relative (only
the order of the bead sums differs), on the slab and line paths, with even and odd N, and the chain-end
factors of NA = 0 and NA = N."""
import numpy as np
import pytest

import lfts_cpu_amd as L

pytestmark = pytest.mark.gpu

CASES = [
    ((32, 32, 32), (4.0, 4.0, 4.0), 16, 7, 4),
    ((32, 32, 32), (4.0, 4.0, 4.0), 15, 6, 3),
    ((48, 48, 48), (4.5, 4.5, 4.5), 11, 5, 2),
    ((40, 40, 40), (4.38, 4.38, 4.38), 13, 6, 3),
    ((128, 128, 128), (6.0, 6.0, 6.0), 9, 4, 3),
    ((96, 96, 96), (5.2, 5.2, 5.2), 8, 3, 3),
    ((16, 16, 16), (3.0, 3.0, 3.0), 14, 7, 4),
    ((72, 72, 72), (4.8, 4.8, 4.8), 10, 4, 3),
    ((24, 40, 16), (3.2, 4.1, 2.6), 11, 5, 3),
    ((6, 5, 8), (2.5, 2.0, 3.5), 13, 13, 3),
    ((8, 6, 5), (3.0, 2.5, 2.0), 12, 0, 5),
    ((10, 12, 7), (3.2, 3.6, 2.9), 21, 9, 4),
]


def _fields(M, seed):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.normal(0, 2.0, M), rng.normal(0, 1.0, M)])


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("m,Lb,N,NA,S", CASES)
def test_checkpointed_matches_full_storage(gpu, m, Lb, N, NA, S):
    M = int(np.prod(m))
    print(f"ckpt case m={m} N={N} NA={NA} stride={S}", flush=True)
    w = _fields(M, N)
    rng = np.random.default_rng(3)
    u, v, ub = rng.normal(size=M), rng.normal(size=M), rng.normal(size=2 * M)
    a, b = L.Context(N, NA, m, Lb), L.Context(N, NA, m, Lb)
    b.set_pair_storage(S)
    st, kept = b.pair_storage()
    assert st == S and kept < N and kept % 2 == 1
    for c in (a, b):
        c.set_fields(w)
    la, lb = a.calc_concs(), b.calc_concs()
    assert abs(la - lb) <= 1e-13 * max(1.0, abs(la))
    assert rel(b.get_fields(4)[2 * M:], a.get_fields(4)[2 * M:]) < 1e-12
    assert rel(b.jvp(u, v), a.jvp(u, v)) < 1e-12
    assert rel(b.vjp(ub), a.vjp(ub)) < 1e-12
    if 1 <= NA <= N - 1:
        (da, qa), (db, qb) = a.box_derivative(), b.box_derivative()
        assert qa == qb or abs(qa - qb) <= 1e-13 * max(1.0, abs(qa))
        assert rel(np.array(db), np.array(da)) < 1e-12
        ka, ea = a.anderson_mix(200, 1e-6)
        kb, eb = b.anderson_mix(200, 1e-6)
        assert ka == kb
        assert rel(b.get_fields(2), a.get_fields(2)) < 1e-9


def test_full_size_checkpointed(gpu):
    m, Lb, N, NA = (64, 64, 64), (4.4, 4.4, 4.4), 90, 45
    M = int(np.prod(m))
    print("ckpt full size 64^3 N=90 stride 10", flush=True)
    w = _fields(M, 5)
    w[:M] *= 4.0
    rng = np.random.default_rng(6)
    v, ub = rng.normal(size=M), rng.normal(size=2 * M)
    a, b = L.Context(N, NA, m, Lb), L.Context(N, NA, m, Lb)
    b.set_pair_storage(10)
    assert b.pair_storage() == (10, 9)
    for c in (a, b):
        c.set_fields(w)
    assert abs(a.calc_concs() - b.calc_concs()) < 1e-12
    assert rel(b.get_fields(4)[2 * M:], a.get_fields(4)[2 * M:]) < 1e-12
    assert rel(b.jvp(None, v), a.jvp(None, v)) < 1e-12
    assert rel(b.vjp(ub), a.vjp(ub)) < 1e-12
    assert rel(np.array(b.box_derivative()[0]), np.array(a.box_derivative()[0])) < 1e-12
    ia = a.saddle_nk(tol=1e-9)
    ib = b.saddle_nk(tol=1e-9)
    assert ib["err"] <= 1e-9
    assert np.abs(b.get_fields(2)[M:] - a.get_fields(2)[M:]).max() < 1e-8, (ia, ib)


def test_checkpointed_limits(gpu):
    c = L.Context(12, 6, (8, 8, 8), (3.0, 3.0, 3.0))
    c.set_pair_storage(3)
    c.set_fields(_fields(512, 4))
    c.calc_concs()
    with pytest.raises(L.LftsError):
        c.pair(4)
    with pytest.raises(L.LftsError):
        L.Batch([c])
    d_ck, _ = c.box_derivative()
    c.set_pair_storage(0)  # back to full storage
    assert c.pair_storage() == (0, 12)
    c.calc_concs()
    c.pair(4)
    assert rel(d_ck, c.box_derivative()[0]) < 1e-12
