"""Pinning the oracle's derivatives of calc_concs (no reference counterpart).

Each check is fixed by mathematics, not by retyping the oracle:
  * the tangent (JVP) against central finite differences of calc_concs;
  * the adjoint (VJP) against the tangent through <u, J v> = <J^T u, v>;
  * d lnQ/dL against central finite differences over the box edges;
  * at w = 0 every bead has h = 1 and Q = 1, and the response of phi+ to a cosine mode
    of w+ is diagonal: d phi+ = lambda(k) v with
        lambda(k) = -(1/N) ((2/N) sum_{j=1..N} sum_{t<j} G^t - 1),  G = exp(-K^2/(6N)),
    the discrete-chain Debye function (lambda -> -1 as K -> 0). The Newton-Krylov
    preconditioner of the library inverts this same function.
"""
import numpy as np
import pytest

import oracle as O

CASES = [
    ((4, 5, 6), (3.0, 3.5, 4.0), 12, 5),
    ((6, 4, 5), (2.5, 3.0, 2.0), 9, 9),   # NA = N: the q+ chain starts on hB anyway
    ((4, 4, 6), (3.0, 3.0, 3.5), 8, 0),   # NA = 0: the q chain starts on hA anyway
    ((8, 6, 4), (4.0, 3.0, 2.5), 16, 7),
]


def _fields(M, seed):
    rng = np.random.default_rng(seed)
    w = np.zeros(4 * M)
    w[:M] = rng.normal(0, 2.0, M)
    w[M:2 * M] = rng.normal(0, 1.0, M)
    return w


def _phi(d, w):
    w = w.copy()
    d.calc_concs(w)
    return w[2 * d.M:].copy()


@pytest.mark.parametrize("m,Lb,N,NA", CASES)
def test_jvp_matches_central_differences(m, Lb, N, NA):
    d = O.Diblock(N, NA, m, Lb)
    M = d.M
    w = _fields(M, 1)
    rng = np.random.default_rng(2)
    u, v = rng.normal(size=M), rng.normal(size=M)
    J = d.jvp(w, u, v)
    h = 1e-6
    wp, wm = w.copy(), w.copy()
    wp[:M] += h * u
    wp[M:2 * M] += h * v
    wm[:M] -= h * u
    wm[M:2 * M] -= h * v
    fd = (_phi(d, wp) - _phi(d, wm)) / (2 * h)
    assert np.abs(J - fd).max() <= 1e-7 * np.abs(fd).max()
    # each direction alone adds up to the joint one (linearity)
    assert np.allclose(d.jvp(w, u, None) + d.jvp(w, None, v), J, rtol=0, atol=1e-12 * np.abs(J).max())


@pytest.mark.parametrize("m,Lb,N,NA", CASES)
def test_vjp_is_the_transpose_of_jvp(m, Lb, N, NA):
    d = O.Diblock(N, NA, m, Lb)
    M = d.M
    w = _fields(M, 3)
    rng = np.random.default_rng(4)
    for _ in range(3):
        u, v = rng.normal(size=M), rng.normal(size=M)
        ub = rng.normal(size=2 * M)
        lhs = ub @ d.jvp(w, u, v)
        g = d.vjp(w, ub)
        rhs = g[:M] @ u + g[M:] @ v
        assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


@pytest.mark.parametrize("m,Lb,N,NA", [c for c in CASES if 1 <= c[3] <= c[2] - 1])
def test_box_derivative_matches_central_differences(m, Lb, N, NA):
    d = O.Diblock(N, NA, m, Lb)
    M = d.M
    w = _fields(M, 5)
    lnQ, g = d.dlnQ_dL(w)
    assert lnQ == pytest.approx(d.calc_concs(w.copy()), abs=1e-14)
    h = 1e-6
    for k in range(3):
        Lp, Lm = list(Lb), list(Lb)
        Lp[k] += h
        Lm[k] -= h
        fd = (O.Diblock(N, NA, m, Lp).calc_concs(w.copy()) - O.Diblock(N, NA, m, Lm).calc_concs(w.copy())) / (2 * h)
        assert abs(g[k] - fd) <= 1e-7 * max(abs(fd), 1e-3), k


def debye_lambda(G, N):
    part = np.zeros_like(G)
    s = np.zeros_like(G)
    gt = np.ones_like(G)
    for _ in range(N):
        part = part + gt
        gt = gt * G
        s = s + part
    return -(1.0 / N) * ((2.0 / N) * s - 1.0)


@pytest.mark.parametrize("NA", [3, 7])
def test_homogeneous_response_is_the_discrete_debye_function(NA):
    m, Lb, N = (8, 6, 10), (3.0, 2.5, 4.0), 12
    d = O.Diblock(N, NA, m, Lb)
    M = d.M
    w = np.zeros(4 * M)
    x, y, z = np.meshgrid(*[np.arange(n) for n in m], indexing="ij")
    for k in [(1, 0, 0), (0, 2, 1), (3, 1, 4), (4, 3, 5)]:  # includes the Nyquist planes
        v = np.cos(2 * np.pi * (k[0] * x / m[0] + k[1] * y / m[1] + k[2] * z / m[2])).ravel()
        dphi = d.jvp(w, None, v)
        K2 = sum((2 * np.pi * min(k[i], m[i] - k[i]) / Lb[i]) ** 2 for i in range(3))
        lam = debye_lambda(np.array([np.exp(-K2 / (6 * N))]), N)[0]
        assert np.abs(dphi[M:] - lam * v).max() <= 1e-12
        # phi- responds too (the blocks differ), but its mean stays zero
        assert abs(dphi[:M].mean()) <= 1e-13
    assert debye_lambda(np.array([1.0]), N)[0] == pytest.approx(-1.0, abs=1e-15)
