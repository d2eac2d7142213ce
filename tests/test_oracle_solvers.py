"""Pins the oracle's Anderson mixing, LU, Langevin update, noise stream, S(k) and
parameter arithmetic (src/anderson.cc, src/langevin.cc, src/strFunc.cc,
src/lfts_params.cc, src/lfts_simulation.cc) without the reference build."""
import math
import os

import numpy as np
import pytest
import scipy.linalg

import oracle as O
from tests import np_ref


# ---- LU (stands in for GSL) -----------------------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 3, 5, 10])
def test_lu_solve_matches_lapack(n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + n * np.eye(n)
    b = rng.standard_normal(n)
    x = O.lu_solve(A, b)
    assert np.allclose(x, scipy.linalg.solve(A, b), rtol=1e-13, atol=1e-14)


def test_lu_solve_needs_pivoting():
    A = np.array([[0.0, 2.0, 1.0], [1.0, 1.0, 0.0], [3.0, 0.0, 1.0]])
    b = np.array([1.0, 2.0, 3.0])
    x = O.lu_solve(A, b)
    assert np.allclose(A @ x, b, rtol=0, atol=1e-14)


# ---- Anderson mixing --------------------------------------------------------------------

@pytest.mark.parametrize("m,L,N,NA,amp", [((4, 4, 6), (2.0, 2.0, 3.0), 8, 3, 3.0), ((6, 5, 4), (2.5, 2.0, 1.8), 10, 5, 6.0)])
def test_anderson_matches_independent_numpy_restatement(m, L, N, NA, amp):
    M = int(np.prod(m))
    rng = np.random.default_rng(5)
    w = np.zeros(4 * M)
    w[:M] = rng.normal(0, amp, M)
    w[M:2 * M] = rng.normal(0, 1.0, M)
    d = O.Diblock(N, NA, m, L)
    am = O.Anderson(M, 10)
    wo = w.copy()
    k = am.mix(d, 200, 1e-4, wo)
    wn = w.copy()
    k_np, err_np = np_ref.anderson_mix(wn, N, NA, m, L)
    assert k == k_np
    assert k > 12  # the history ring wraps at least once
    assert np.abs(wo[M:2 * M] - wn[M:2 * M]).max() <= 1e-9 * np.abs(wn[M:2 * M]).max()
    assert math.isclose(am.last_err(), err_np, rel_tol=1e-6)
    assert am.last_err() <= 1e-4


def test_anderson_converged_state_is_a_fixed_point():
    m, L, N, NA = (4, 6, 4), (2.0, 3.0, 2.5), 10, 4
    M = int(np.prod(m))
    rng = np.random.default_rng(9)
    w = np.zeros(4 * M)
    w[:M] = rng.normal(0, 3.0, M)
    d = O.Diblock(N, NA, m, L)
    am = O.Anderson(M, 10)
    am.mix(d, 200, 1e-10, w)
    # the incompressibility residual of the final w+ is small
    d.calc_concs(w)
    assert math.sqrt(((w[3 * M:] - 1) ** 2).mean()) < 1e-9
    # a second solve stops after one iteration (k = 2) and leaves w+ in place
    w2 = w.copy()
    assert am.mix(d, 200, 1e-4, w2) == 2
    assert np.abs(w2[M:2 * M] - w[M:2 * M]).max() < 1e-8


# ---- Langevin update and the noise stream --------------------------------------------------

def test_langevin_update_closed_form():
    M = 50
    rng = np.random.default_rng(1)
    w = rng.normal(0, 1, 4 * M)
    ep, en = rng.normal(0, 1, M), rng.normal(0, 1, M)
    XbN, dt = 17.5, 0.3
    ref = w[:M] - (w[2 * M:3 * M] + 2 * w[:M] / XbN) * dt + 0.5 * (ep + en)
    O.langevin_update(w, ep, en, XbN, dt)
    assert np.allclose(w[:M], ref, rtol=1e-15, atol=1e-15)


def test_mt19937_64_known_answer():
    # C++ standard [rand.predef]: the 10000th output of a default-constructed mt19937_64
    assert int(O.mt64(5489, 10000)[-1]) == 9981545732273789042
    eng = np_ref.MT19937_64(123456789)
    got = O.mt64(123456789, 700)
    assert [eng() for _ in range(700)] == [int(v) for v in got]


@pytest.mark.parametrize("sigma,n", [(1.0, 11), (7.845, 1001)])
def test_normal_stream_matches_polar_method_restatement(sigma, n):
    eng = np_ref.MT19937_64(123456789)
    nd = np_ref.PolarNormal(0.0, sigma)
    ref = np.array([nd(eng) for _ in range(n)])
    assert np.array_equal(O.normals(123456789, sigma, n), ref)


# ---- structure function -------------------------------------------------------------------

@pytest.mark.parametrize("m,L", [((8, 6, 10), (3.0, 2.5, 4.0)), ((5, 4, 7), (2.0, 2.0, 2.0)), ((4, 4, 4), (4.4, 4.4, 4.4))])
def test_sk_bins(m, L):
    s = O.StrFunc(m, L, 100.0, 20.0)
    S, K, wt, P, ns = s.get()
    k0 = np_ref.freqs(m[0])[:, None, None]
    k1 = np_ref.freqs(m[1])[None, :, None]
    k2 = np.arange(m[2] // 2 + 1)[None, None, :]
    ksq = (k0 * k0) / (L[0] * L[0]) + (k1 * k1) / (L[1] * L[1]) + (k2 * k2) / (L[2] * L[2])
    # 2*M_PI*pow(ksq, 0.5) with the C library's pow (src/strFunc.cc:98); numpy's vector
    # power and sqrt differ from it in the last bit for some entries
    Kref = np.array([2 * math.pi * math.pow(x, 0.5) for x in ksq.reshape(-1)])
    assert np.array_equal(K, Kref)
    wref = np.where((k2 == 0) | (k2 == m[2] // 2), 1, 2) * np.ones_like(ksq, dtype=np.int32)
    assert np.array_equal(wt, wref.reshape(-1))
    assert np.array_equal(P, np.argsort(Kref, kind="stable"))
    assert ns == 0 and not S.any()


def test_sk_single_mode():
    m, L = (8, 6, 10), (3.0, 2.5, 4.0)
    M = int(np.prod(m))
    A = 1.7
    i, j, k = np.meshgrid(*[np.arange(n) for n in m], indexing="ij")
    wm = (A * np.cos(2 * np.pi * (2 * i / m[0] + 3 * k / m[2]))).reshape(-1)
    s = O.StrFunc(m, L, 100.0, 20.0)
    s.sample(wm)
    s.sample(wm)
    S, K, wt, P, ns = s.get()
    mk2 = m[2] // 2 + 1
    S3 = S.reshape(m[0], m[1], mk2)
    assert ns == 2
    # r2c keeps (2,0,3) and the conjugate of (-2,0,-3), stored at (K0=m0-2, 0, 3)
    assert math.isclose(S3[2, 0, 3], 2 * (A * M / 2) ** 2, rel_tol=1e-12)
    assert S3[2, 0, 3] == pytest.approx(S3[m[0] - 2, 0, 3], rel=1e-12) or S3[m[0] - 2, 0, 3] < 1e-12
    S3c = S3.copy()
    S3c[2, 0, 3] = 0
    assert S3c.max() < 1e-16 * (A * M) ** 2


def test_sk_save_format(tmp_path):
    m, L = (4, 4, 4), (2.0, 2.0, 2.0)
    M = 64
    CV, chi = 3.0, 12.0
    s = O.StrFunc(m, L, CV, chi)
    rng = np.random.default_rng(2)
    s.sample(rng.normal(0, 1, M))
    S, K, wt, P, ns = s.get()
    f = tmp_path / "struct"
    s.save(str(f))
    coeff = CV / (chi * chi * M * M)
    rows, S_sum, n_same = [], 0.0, 0
    for k in range(len(P)):
        S_sum += wt[P[k]] * ((coeff / ns) * S[P[k]] - 0.5 / chi)
        n_same += wt[P[k]]
        if k == len(P) - 1 or abs(K[P[k + 1]] - K[P[k]]) > 1e-5:
            rows.append("%.8f\t%.8f" % (K[P[k]], S_sum / n_same))
            S_sum, n_same = 0.0, 0
    assert f.read_text().splitlines() == rows


# ---- parameters ------------------------------------------------------------------------------

def _z_inf_mp(L, m, N, Nbar, tmax=100):
    import mpmath as mp

    mp.mp.dps = 50
    R0 = mp.sqrt(N)
    R0dL = [R0 * mp.mpf(L[i]) / m[i] for i in range(3)]
    ell = (R0dL[0] * R0dL[1] * R0dL[2]) ** (mp.mpf(1) / 3)
    s = mp.mpf("0.5")
    for t in range(1, tmax + 1):
        X = (mp.pi / ell) * mp.sqrt(mp.mpf(t) / 6)
        p = mp.mpf(1)
        for i in range(3):
            p *= mp.erf(X * ell / R0dL[i])
        s += (mp.sqrt(mp.pi) / (2 * X)) ** 3 * p
    s *= 2 * R0 / (ell ** 3 * mp.sqrt(Nbar))
    X = (mp.pi / ell) * mp.sqrt((mp.mpf("0.5") + tmax) / 6)
    return float(1 - s - 3 * R0 / (ell * mp.sqrt(mp.pi * Nbar) * X))


@pytest.mark.parametrize("L,m,N,C", [((4.0,) * 3, (32,) * 3, 64, 50.0), ((4.38,) * 3, (16,) * 3, 90, 100.0),
                                     ((4.4,) * 3, (64,) * 3, 90, 50.0), ((6.0,) * 3, (128,) * 3, 128, 100.0),
                                     ((4.38,) * 3, (40,) * 3, 90, 100.0)])
def test_z_inf_against_mpmath(L, m, N, C):
    assert math.isclose(O.z_inf(L, m, N, C * C), _z_inf_mp(L, m, N, C * C), rel_tol=1e-12, abs_tol=1e-14)


def test_read_params(tmp_path):
    f = tmp_path / "in"
    f.write_text("90 40 13.15 100.000000 1.000000 1\n16 16 16 4.38 4.38 4.38\n7 5 2 1000 3\n")
    p = O.read_params(str(f))
    assert (p["N"], p["NA"], p["NB"], p["M"], p["Mk"]) == (90, 40, 50, 4096, 16 * 16 * 9)
    assert (p["equil_its"], p["sim_its"]) == (6, 4)  # rounded down to multiples of n_smpl
    z = O.z_inf((4.38,) * 3, (16,) * 3, 90, 100.0 ** 2)
    assert p["XeN"] == 13.15 and math.isclose(p["chi_b"], 13.15 / z, rel_tol=1e-15)
    assert math.isclose(p["sigma"], math.sqrt(2.0 * 4096 * 1.0 / (100.0 * 4.38 ** 3)), rel_tol=1e-15)


# ---- whole simulation ------------------------------------------------------------------------

def _write_input(path, header, w):
    with open(path, "w") as f:
        f.write(header)
        for v in w:
            f.write("%.17g\n" % v)


def test_hamiltonian_closed_form(tmp_path):
    """Uniform W- = a, w+ = c: the saddle point holds already (phi+ = 1), and
    H = -lnQ + (a^2/XbN - c) = a (NA-NB)/N + a^2/XbN."""
    m, N, NA, XbN = (4, 4, 4), 10, 4, 15.0
    M = 64
    a, c = 0.6, 0.25
    f = tmp_path / "in"
    _write_input(f, f"{N} {NA} {XbN} 50 0.5 0\n4 4 4 2.0 2.0 2.0\n0 0 1 1 1\n", [a] * M + [c] * M)
    s = O.Sim(str(f))
    H = s.getH()
    assert math.isclose(H, a * (NA - (N - NA)) / N + a * a / XbN, rel_tol=1e-12)


def test_sim_step_consumes_reference_stream(tmp_path):
    """The first Langevin step uses the second block of M normals of the seed-123456789
    stream (the first block is drawn by the langevin constructor, src/langevin.cc:16)."""
    m, N, NA, XbN, C, dt = (4, 4, 4), 8, 4, 12.0, 50.0, 0.5
    M = 64
    rng = np.random.default_rng(4)
    w0 = np.concatenate([rng.normal(0, 1, M), np.zeros(M)])
    f = tmp_path / "in"
    _write_input(f, f"{N} {NA} {XbN} {C} {dt} 0\n4 4 4 2.0 2.0 2.0\n2 0 1 1 1\n", w0)
    s = O.Sim(str(f))
    p = s.params
    before = s.w()
    sigma = p["sigma"]
    eta = O.normals(123456789, sigma, 2 * M)
    expect = before[:M] - (before[2 * M:3 * M] + 2 * before[:M] / p["chi_b"]) * dt + 0.5 * (eta[:M] + eta[M:])
    s.step()
    after = s.w()
    assert np.allclose(after[:M], expect, rtol=1e-14, atol=1e-14)
