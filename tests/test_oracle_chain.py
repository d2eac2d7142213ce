"""Pins the oracle's FFT, g table, contour step and propagator evaluation
(src/step.cc, src/diblock.cc) without the reference build: brute-force DFTs,
closed forms, exact invariants and an independent numpy restatement."""
import math

import numpy as np
import pytest

import oracle as O
from tests import np_ref


def brute_dft(x: np.ndarray, sign: int = -1) -> np.ndarray:
    n = x.size
    k = np.arange(n)
    W = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
    return W @ x


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 15, 16, 20, 24, 40, 48, 64, 11, 13, 22, 26, 44, 143, 17, 34, 323])
def test_fft1d_against_brute_force(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign in (-1, +1):
        ref = brute_dft(x, sign)
        got = O.fft1d(x, sign)
        assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("shape", [(4, 4, 4), (6, 5, 8), (3, 7, 2), (8, 1, 6), (5, 5, 5), (11, 13, 2)])
def test_fft3d_against_brute_force_and_numpy(shape):
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    got = O.fft3d(x, -1)
    # brute force: apply the 1D DFT matrix along each axis
    ref = x.copy()
    for ax, n in enumerate(shape):
        k = np.arange(n)
        W = np.exp(-2j * np.pi * np.outer(k, k) / n)
        ref = np.moveaxis(np.tensordot(W, np.moveaxis(ref, ax, 0), axes=(1, 0)), 0, ax)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.abs(got - np.fft.fftn(x)).max() <= 1e-12 * np.abs(ref).max()
    # inverse is unnormalised: back transform gives M * x
    back = O.fft3d(got, +1)
    assert np.abs(back / x.size - x).max() <= 1e-13 * np.abs(x).max() * 10


def test_fft_parseval_and_plane_wave():
    m = (8, 6, 10)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(m) + 0j
    X = O.fft3d(x)
    assert math.isclose((np.abs(X) ** 2).sum(), x.size * (np.abs(x) ** 2).sum(), rel_tol=1e-13)
    # cos(2 pi (2x/8 + 1y/6)) maps to two spikes of M/2 at +-(2,1,0)
    i, j, k = np.meshgrid(*[np.arange(n) for n in m], indexing="ij")
    pw = np.cos(2 * np.pi * (2 * i / m[0] + 1 * j / m[1])) + 0j
    P = O.fft3d(pw)
    M = pw.size
    assert abs(P[2, 1, 0] - M / 2) < 1e-10 and abs(P[6, 5, 0] - M / 2) < 1e-10
    P[2, 1, 0] = 0
    P[6, 5, 0] = 0
    assert np.abs(P).max() < 1e-10


@pytest.mark.parametrize("m,L", [((3, 4, 5), (2.0, 3.0, 4.5)), ((4, 4, 4), (1.0, 1.0, 1.0)), ((5, 2, 6), (3.3, 1.1, 2.2))])
def test_g_table_closed_form(m, L):
    N = 7
    d = O.Diblock(N, 3, m, L)
    g = d.ghalf().reshape(m[0], m[1], m[2] // 2 + 1)
    M = m[0] * m[1] * m[2]
    for K0 in range(m[0]):
        f0 = K0 if 2 * K0 <= m[0] else K0 - m[0]
        for K1 in range(m[1]):
            f1 = K1 if 2 * K1 <= m[1] else K1 - m[1]
            for k2 in range(m[2] // 2 + 1):
                Ksq = (2 * math.pi) ** 2 * (f0 ** 2 / L[0] ** 2 + f1 ** 2 / L[1] ** 2 + k2 ** 2 / L[2] ** 2)
                assert math.isclose(g[K0, K1, k2], math.exp(-Ksq / (6 * N)) / M, rel_tol=1e-14)
    # the Nyquist index of an even axis carries the positive frequency m/2
    assert np.allclose(g, np_ref.g_half(N, m, L), rtol=1e-15, atol=0)


def _fields(M, rng, a=1.0, b=0.5):
    w = np.zeros(4 * M)
    w[:M] = rng.normal(0, a, M)
    w[M:2 * M] = rng.normal(0, b, M)
    return w


@pytest.mark.parametrize("m,L,N,NA", [((4, 6, 5), (2.0, 2.5, 3.0), 9, 4), ((8, 8, 8), (3.0, 3.0, 3.0), 12, 6),
                                      ((6, 4, 8), (1.5, 2.0, 2.5), 10, 7), ((5, 5, 5), (2.2, 2.2, 2.2), 8, 2)])
def test_calc_concs_matches_independent_numpy_restatement(m, L, N, NA):
    M = int(np.prod(m))
    rng = np.random.default_rng(N * 100 + NA)
    w = _fields(M, rng)
    d = O.Diblock(N, NA, m, L)
    wo = w.copy()
    lnq = d.calc_concs(wo)
    lnq_np, phm, php, q, qd = np_ref.calc_concs(w.copy(), N, NA, m, L)
    assert abs(lnq - lnq_np) <= 1e-12 * max(1, abs(lnq_np))
    assert np.abs(wo[2 * M:3 * M] - phm).max() <= 1e-12 * np.abs(php).max()
    assert np.abs(wo[3 * M:] - php).max() <= 1e-12 * np.abs(php).max()
    # stored pairs follow the reference layout [q_i | q+_{N+1-i}]
    for i in (1, 2, N // 2, N):
        p = d.pair(i)
        assert np.abs(p[:M] - q[i]).max() <= 1e-12 * np.abs(q[i]).max()
        assert np.abs(p[M:] - qd[N + 1 - i]).max() <= 1e-12 * np.abs(qd[N + 1 - i]).max()


def test_step_fwd_h_selection():
    """q uses hA for i < NA else hB; q+ uses hB for i < NB else hA (src/step.cc:53-64)."""
    m, L, N, NA = (4, 4, 6), (2.0, 2.0, 3.0), 10, 3
    M = int(np.prod(m))
    d = O.Diblock(N, NA, m, L)
    rng = np.random.default_rng(0)
    qin = rng.uniform(0.5, 1.5, 2 * M)
    hA = rng.uniform(0.5, 1.5, M)
    hB = rng.uniform(0.5, 1.5, M)
    h = np.concatenate([hA, hB])
    g = np_ref.g_half(N, m, L)
    cq = np_ref.conv(qin[:M], g, m)
    cqd = np_ref.conv(qin[M:], g, m)
    NB = N - NA
    for i in range(1, N):
        out = d.step_fwd(qin, h, i)
        assert np.allclose(out[:M], cq * (hA if i < NA else hB), rtol=1e-13, atol=0)
        assert np.allclose(out[M:], cqd * (hB if i < NB else hA), rtol=1e-13, atol=0)


@pytest.mark.parametrize("N,NA", [(10, 4), (9, 8), (8, 1)])
def test_calc_concs_closed_forms(N, NA):
    m, L = (4, 6, 4), (2.0, 3.0, 2.5)
    M = int(np.prod(m))
    NB = N - NA
    d = O.Diblock(N, NA, m, L)
    # w = 0: Q = 1, phi+ = 1, phi- = (NA-NB)/N
    w = np.zeros(4 * M)
    assert abs(d.calc_concs(w)) < 1e-14
    assert np.allclose(w[3 * M:], 1.0, rtol=0, atol=1e-13)
    assert np.allclose(w[2 * M:3 * M], (NA - NB) / N, rtol=0, atol=1e-13)
    # uniform w+ = c: ln Q = -c, phi+ = 1
    c = 0.37
    w = np.zeros(4 * M)
    w[M:2 * M] = c
    assert abs(d.calc_concs(w) + c) < 1e-13
    assert np.allclose(w[3 * M:], 1.0, rtol=0, atol=1e-12)
    # uniform W- = a: A beads weigh exp(-a/N), B beads exp(+a/N):
    # ln Q = -a (NA - NB)/N, phiA = NA/N, phiB = NB/N
    a = 0.8
    w = np.zeros(4 * M)
    w[:M] = a
    assert abs(d.calc_concs(w) + a * (NA - NB) / N) < 1e-13
    assert np.allclose(w[3 * M:], 1.0, rtol=0, atol=1e-12)
    assert np.allclose(w[2 * M:3 * M], (NA - NB) / N, rtol=0, atol=1e-12)


def test_calc_concs_homopolymer_end_quirk():
    """With NB = 0 the reference still starts q+_N from hB (src/diblock.cc:40), so for a
    uniform W- = a every product q_i q+_i = hB hA^N and phi+ = phi- = exp(2a/N), while
    ln Q = -a is unaffected. The oracle keeps this behaviour."""
    m, L, N, a = (4, 4, 4), (2.0, 2.0, 2.0), 9, 0.8
    M = int(np.prod(m))
    d = O.Diblock(N, N, m, L)
    w = np.zeros(4 * M)
    w[:M] = a
    assert abs(d.calc_concs(w) + a) < 1e-13
    assert np.allclose(w[3 * M:], math.exp(2 * a / N), rtol=1e-12, atol=0)
    assert np.allclose(w[2 * M:3 * M], math.exp(2 * a / N), rtol=1e-12, atol=0)


def test_calc_concs_invariants():
    """Q_i = (1/M) sum_r q_i q+_i / h_i is the same for every bead, and mean(phi+) = 1."""
    m, L, N, NA = (8, 6, 10), (3.0, 2.5, 3.5), 14, 5
    M = int(np.prod(m))
    rng = np.random.default_rng(11)
    w = _fields(M, rng, 2.0, 1.0)
    d = O.Diblock(N, NA, m, L)
    lnq = d.calc_concs(w)
    hA = np.exp(-(w[:M] + w[M:2 * M]) / N)
    hB = np.exp(-(-w[:M] + w[M:2 * M]) / N)
    Q = math.exp(lnq)
    for i in range(1, N + 1):
        qi = d.pair(i)[:M]
        qdi = d.pair(N + 1 - i)[M:]
        hi = hA if i <= NA else hB
        assert math.isclose((qi * qdi / hi).sum() / M, Q, rel_tol=1e-12)
    assert math.isclose(w[3 * M:].mean(), 1.0, rel_tol=0, abs_tol=1e-12)
    assert (w[2 * M:3 * M] + w[3 * M:]).min() >= 0.0


def test_calc_concs_weak_field_linear_response():
    """Small single-mode W- = eps cos(k.r): the phi- response is linear in eps and
    sits on the same mode (discrete-chain linear response); doubling eps doubles it."""
    m, L, N, NA = (8, 8, 8), (3.0, 3.0, 3.0), 12, 6
    M = int(np.prod(m))
    i, j, k = np.meshgrid(*[np.arange(n) for n in m], indexing="ij")
    mode = np.cos(2 * np.pi * 1 * i / m[0]).reshape(-1)
    d = O.Diblock(N, NA, m, L)
    resp = []
    for eps in (1e-4, 2e-4):
        w = np.zeros(4 * M)
        w[:M] = eps * mode
        d.calc_concs(w)
        resp.append(w[2 * M:3 * M].copy())
    r1, r2 = resp
    amp1 = (r1 * mode).sum() / (mode * mode).sum()
    amp2 = (r2 * mode).sum() / (mode * mode).sum()
    assert amp1 < 0  # A-rich where W- < 0
    assert math.isclose(amp2, 2 * amp1, rel_tol=1e-3)
    assert np.abs(r1 - amp1 * mode).max() < 1e-3 * abs(amp1)
