"""Pins the oracle's anderson::mix iterate sequence (src/anderson.cc:24-93) and
strFunc::save (src/strFunc.cc:50-76) by what the mathematics fixes, not by a second
coding of the same loops.

Anderson. Iteration k of mix computes the residual D0 = phi+ - 1 at the current w+ (= x0),
and for k >= 2 the coefficients C of the least-squares problem
    min_C || D0 - sum_j C_j (D0 - D_j) ||
over the residuals D_j of the nh = min(k-1, nhMax) previous iterations (the normal
equations U C = V of src/anderson.cc:62-70), then moves w+ to
    (x0 + lam D0) + sum_j C_j ((x_j + lam D_j) - (x0 + lam D0)),  lam = 1 - 0.9^k
(src/anderson.cc:46,73-78); k = 1 is the simple mix x0 + 0.1 D0. The oracle records every
iteration (D0, x0, the new w+, C and the history slot it writes); the checks below solve
each least-squares problem with numpy.linalg.lstsq (QR/SVD, not the normal equations),
replay the recorded slot writes to see which iteration each slot holds, and compare.
Each of six planted misreadings of src/anderson.cc (oracle fault hooks) fails them.

save. The spherical average of src/strFunc.cc:60-74 is, for an even mz, the mean of the
full-spectrum |FFT(W-)_k|^2 (times C V / (chi_b^2 M^2) per sample) over each shell of
equal |k|, minus 0.5/chi_b: the half spectrum with weight 2 off the k2 = 0 and k2 = mz/2
planes counts every full-grid wavevector exactly once. The checks compute that with
numpy.fft.fftn and exact shells (Fractions), and each planted misreading of save fails.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle as O


# ---- Anderson ------------------------------------------------------------------------------

def _system(m, L, N, NA, amp, seed):
    M = int(np.prod(m))
    rng = np.random.default_rng(seed)
    w = np.zeros(4 * M)
    w[:M] = rng.normal(0, amp, M)
    w[M:2 * M] = rng.normal(0, 1.0, M)
    return O.Diblock(N, NA, m, L), w


def _run_traced(m, L, N, NA, amp, seed, nhMax, maxIter, tol, fault=0):
    d, w = _system(m, L, N, NA, amp, seed)
    M = w.size // 4
    am = O.Anderson(M, nhMax)
    am.set_fault(fault)
    am.record(maxIter)
    k_ret = am.mix(d, maxIter, tol, w)
    return k_ret, am.trace(), am.last_err()


def check_anderson_trace(tr, nhMax, maxIter, tol, k_ret, last_err, upd_tol=1e-7):
    """Returns a list of violations (empty = the trace is the reference's iteration)."""
    bad = []
    # stop rule of src/anderson.cc:29: iterate while k < maxIter and err > tol; mix
    # returns the first k that failed the test
    if k_ret != len(tr) + 1:
        bad.append(f"returned k {k_ret} for {len(tr)} iterations")
    for it in tr[:-1]:
        if not it["err"] > tol:
            bad.append(f"iteration {it['k']} continued with err {it['err']} <= tol")
    if tr and not (tr[-1]["err"] <= tol or len(tr) == maxIter - 1):
        bad.append("stopped early")
    if tr and last_err != tr[-1]["err"]:
        bad.append("last_err is not the last iteration's residual")
    held = {}  # slot -> iteration index, replayed from the recorded writes
    for i, it in enumerate(tr):
        k = i + 1
        D0, x0 = it["Dh0"], it["wh0"]
        if not (np.isfinite(D0).all() and np.isfinite(x0).all() and np.isfinite(it["w_new"]).all()):
            bad.append(f"k={k}: non-finite fields")
            break
        if it["k"] != k:
            bad.append(f"iteration {i} recorded as k={it['k']}")
        # residual and error (src/anderson.cc:35-45)
        if not math.isclose(it["err"], math.sqrt(np.mean(D0 * D0)), rel_tol=1e-12):
            bad.append(f"k={k}: err is not the rms of phi+ - 1")
        lam = 1.0 - 0.9 ** k
        nh = min(k - 1, nhMax)
        if not math.isclose(it["lam"], lam, rel_tol=1e-14):
            bad.append(f"k={k}: lambda {it['lam']} != 1-0.9^k")
        if it["nh"] != nh:
            bad.append(f"k={k}: nh {it['nh']} != min(k-1, nhMax)")
        prev = list(range(i - nh, i))  # the nh previous iterations
        if k == 1:
            expect = x0 + lam * D0
        else:
            A = np.stack([D0 - tr[j]["Dh0"] for j in prev], axis=1)
            C, *_ = np.linalg.lstsq(A, D0, rcond=None)
            y0 = x0 + lam * D0
            expect = y0 + sum(c * ((tr[j]["wh0"] + lam * tr[j]["Dh0"]) - y0) for c, j in zip(C, prev))
            # the slots in use (1..nh) hold exactly the previous nh iterations, and the
            # coefficient recorded for a slot is the least-squares coefficient of the
            # iteration it holds
            in_slots = [held.get(s) for s in range(1, it["nh"] + 1)]
            if sorted(x for x in in_slots if x is not None) != prev or None in in_slots:
                bad.append(f"k={k}: slots 1..nh hold iterations {in_slots}, expected {prev}")
            else:
                Cs = np.array([C[prev.index(j)] for j in in_slots])
                scale = max(1.0, np.abs(C).max())
                if np.abs(it["C"] - Cs).max() > 1e-5 * scale:
                    bad.append(f"k={k}: C {it['C']} != lstsq {Cs}")
        step = np.abs(expect - x0).max()
        dev = np.abs(it["w_new"] - expect).max()
        if dev > upd_tol * step:
            bad.append(f"k={k}: update off by {dev:.3g} (step {step:.3g})")
        held[it["slot"]] = i
    return bad


CASES = [
    # mesh, box, N, NA, amplitude of W-, seed, nhMax, maxIter
    ((4, 4, 6), (2.0, 2.0, 3.0), 8, 3, 3.0, 5, 10, 18),
    ((6, 5, 4), (2.5, 2.0, 1.8), 10, 5, 6.0, 5, 10, 18),
    ((4, 6, 4), (2.0, 3.0, 2.5), 10, 4, 3.0, 9, 3, 14),
    ((5, 4, 4), (2.2, 1.9, 2.0), 12, 7, 4.0, 3, 4, 14),
]


@pytest.mark.parametrize("m,L,N,NA,amp,seed,nhMax,maxIter", CASES)
def test_anderson_iterations_solve_their_least_squares_problems(m, L, N, NA, amp, seed, nhMax, maxIter):
    tol = 1e-13
    k_ret, tr, last_err = _run_traced(m, L, N, NA, amp, seed, nhMax, maxIter, tol)
    assert len(tr) == maxIter - 1  # no early stop: the ring wraps at least once
    assert len(tr) > nhMax + 2
    assert check_anderson_trace(tr, nhMax, maxIter, tol, k_ret, last_err) == []


def test_anderson_first_two_iterations_closed_form():
    """k = 1: w+ <- w+ + 0.1 D0. k = 2: one history, C = <D0, D0-D1> / |D0-D1|^2,
    w+ <- (x1 + 0.19 D1) + C ((x0 + 0.19 D0) - (x1 + 0.19 D1)) with x0, D0 from k = 1."""
    k_ret, tr, _ = _run_traced((4, 4, 6), (2.0, 2.0, 3.0), 8, 3, 3.0, 5, 10, 3, 1e-13)
    assert k_ret == 3 and len(tr) == 2
    a, b = tr
    assert np.array_equal(b["wh0"], a["w_new"])  # k = 2 starts where k = 1 moved w+
    assert np.allclose(a["w_new"], a["wh0"] + 0.1 * a["Dh0"], rtol=0, atol=1e-15)
    d = b["Dh0"] - a["Dh0"]
    C = float(b["Dh0"] @ d) / float(d @ d)
    assert math.isclose(b["C"][0], C, rel_tol=1e-10)
    y0, y1 = a["wh0"] + 0.19 * a["Dh0"], b["wh0"] + 0.19 * b["Dh0"]
    assert np.allclose(b["w_new"], y1 + C * (y0 - y1), rtol=0, atol=1e-13 * np.abs(y1).max())


def test_anderson_schedules_across_the_ring_wrap():
    """lambda = 1 - 0.9^k, nh = min(k-1, nhMax), and the history slot of iteration k is
    1 + (k-1) % nhMax: each new residual replaces the oldest one once the ring is full."""
    for nhMax, maxIter in ((10, 18), (3, 12)):
        _, tr, _ = _run_traced((4, 4, 6), (2.0, 2.0, 3.0), 8, 3, 3.0, 5, nhMax, maxIter, 1e-13)
        lam = [it["lam"] for it in tr]
        assert lam == [1.0 - 0.9 ** k for k in range(1, maxIter)]
        assert [it["nh"] for it in tr] == [min(k - 1, nhMax) for k in range(1, maxIter)]
        slots = [it["slot"] for it in tr]
        # the oldest history is replaced: slot of iteration k is that of iteration k - nhMax
        assert slots[:nhMax] == list(range(1, nhMax + 1))
        assert all(slots[i] == slots[i - nhMax] for i in range(nhMax, len(slots)))


@pytest.mark.parametrize("fault,what", [
    (1, "lambda = 1 - 0.9^(k+1)"),
    (2, "ring slot 1 + k % nhMax"),
    (3, "DD entries past the slot not refreshed"),
    (4, "U laid out with stride nhMax"),
    (5, "nh capped at nhMax - 1"),
    (6, "lambda * Dh_n dropped from the history sum"),
])
def test_each_planted_anderson_misreading_fails_the_pins(fault, what):
    bad = []
    for m, L, N, NA, amp, seed, nhMax, maxIter in CASES:
        tol = 1e-13
        k_ret, tr, last_err = _run_traced(m, L, N, NA, amp, seed, nhMax, maxIter, tol, fault)
        bad += check_anderson_trace(tr, nhMax, maxIter, tol, k_ret, last_err)
    assert bad, f"planted fault '{what}' was not caught"


# ---- strFunc::save --------------------------------------------------------------------------

def _freqs(n):
    # the reference's frequency range k in [-(n-1)/2, n/2] (src/strFunc.cc:87-95)
    return list(range(-((n - 1) // 2), n // 2 + 1))


def _shells(m, L):
    """Exact shells of |k|^2 = sum (k_d / L_d)^2 over the full DFT grid: sorted list of
    (ksq, [(k0, k1, k2), ...])."""
    Lf = [Fraction(x) for x in L]
    sh = {}
    for a in _freqs(m[0]):
        for b in _freqs(m[1]):
            for c in _freqs(m[2]):
                q = Fraction(a * a) / (Lf[0] * Lf[0]) + Fraction(b * b) / (Lf[1] * Lf[1]) + Fraction(c * c) / (Lf[2] * Lf[2])
                sh.setdefault(q, []).append((a, b, c))
    return sorted(sh.items())


def _expected_rows(m, L, fields, CV, chi, weight=lambda kv: 1.0):
    M = int(np.prod(m))
    P = sum(np.abs(np.fft.fftn(f.reshape(m))) ** 2 for f in fields)  # numpy's own FFT
    coeff = CV / (chi * chi * M * M) / len(fields)
    rows = []
    for q, members in _shells(m, L):
        wts = [weight(kv) for kv in members]
        s = sum(wt * P[kv[0] % m[0], kv[1] % m[1], kv[2] % m[2]] for wt, kv in zip(wts, members))
        rows.append((2 * math.pi * math.sqrt(float(q)), coeff * s / sum(wts) - 0.5 / chi))
    return rows


def _read_rows(path):
    return [tuple(float(x) for x in line.split("\t")) for line in open(path).read().splitlines()]


def _save_matches(rows, expect, tol=2e-8):
    if len(rows) != len(expect):
        return False
    return all(abs(a[0] - b[0]) <= tol and abs(a[1] - b[1]) <= tol * max(1.0, abs(b[1])) for a, b in zip(rows, expect))


SAVE_CASES = [((8, 8, 8), (3.0, 3.0, 3.0)), ((6, 4, 10), (2.5, 3.5, 4.0)), ((8, 6, 4), (4.4, 4.4, 4.4))]


@pytest.mark.parametrize("m,L", SAVE_CASES)
def test_sk_save_is_the_full_spectrum_shell_average(m, L, tmp_path):
    M = int(np.prod(m))
    CV, chi = 37.0, 2.5
    rng = np.random.default_rng(11)
    fields = [rng.normal(0, 1.0, M) for _ in range(3)]
    s = O.StrFunc(m, L, CV, chi)
    for f in fields:
        s.sample(f)
    s.save(str(tmp_path / "sk"))
    rows = _read_rows(tmp_path / "sk")
    expect = _expected_rows(m, L, fields, CV, chi)
    # shells are at least 1e-3 apart in K here, so dK = 1e-5 groups exactly the shells
    Ks = [e[0] for e in expect]
    assert min(np.diff(Ks)) > 1e-3
    assert _save_matches(rows, expect)
    # the K = 0 shell holds only the mean mode, so it is (CV/(chi^2 M^2)) <sum W->^2 - 0.5/chi
    mean_sq = np.mean([f.sum() ** 2 for f in fields])
    assert math.isclose(rows[0][1], CV / (chi * chi * M * M) * mean_sq - 0.5 / chi, rel_tol=1e-7, abs_tol=1e-8)


def test_sk_save_single_cosine_mode(tmp_path):
    """W- = A cos(2 pi (2x/mx + 3z/mz)) has |W-_k|^2 = (A M / 2)^2 at +-(2, 0, 3) only: its
    shell reads coeff 2 (A M/2)^2 / n_shell - 0.5/chi_b, every other shell -0.5/chi_b."""
    m, L = (8, 6, 10), (3.0, 2.5, 4.0)
    M = int(np.prod(m))
    A, CV, chi = 1.7, 100.0, 20.0
    i, j, k = np.meshgrid(*[np.arange(n) for n in m], indexing="ij")
    wm = (A * np.cos(2 * np.pi * (2 * i / m[0] + 3 * k / m[2]))).reshape(-1)
    s = O.StrFunc(m, L, CV, chi)
    s.sample(wm)
    s.sample(wm)
    s.save(str(tmp_path / "sk"))
    rows = _read_rows(tmp_path / "sk")
    sh = _shells(m, L)
    assert len(rows) == len(sh)
    q_mode = Fraction(4) / Fraction(3.0) ** 2 + Fraction(9) / Fraction(4.0) ** 2
    coeff = CV / (chi * chi * M * M)
    for (q, members), (K, val) in zip(sh, rows):
        assert abs(K - 2 * math.pi * math.sqrt(float(q))) <= 1e-8
        if q == q_mode:
            assert (2, 0, 3) in members and (-2, 0, -3) in members
            assert math.isclose(val, coeff * 2 * (A * M / 2) ** 2 / len(members) - 0.5 / chi, rel_tol=1e-8)
        else:
            assert abs(val + 0.5 / chi) <= 1e-8


def test_sk_save_odd_mz_weight_quirk(tmp_path):
    """For an odd mz the reference still gives weight 1 to the k2 = mz/2 = (mz-1)/2 plane
    (src/strFunc.cc:99), whose mirror -(mz-1)/2 is a distinct grid point: on the full grid
    those wavevectors count half. The save must keep that. (Equal box edges, so that
    shells mix such wavevectors with others; otherwise both weightings agree.)"""
    m, L = (4, 6, 5), (2.0, 2.0, 2.0)
    M = int(np.prod(m))
    CV, chi = 10.0, 3.0
    rng = np.random.default_rng(12)
    fields = [rng.normal(0, 1.0, M) for _ in range(2)]
    s = O.StrFunc(m, L, CV, chi)
    for f in fields:
        s.sample(f)
    s.save(str(tmp_path / "sk"))
    rows = _read_rows(tmp_path / "sk")
    half = (m[2] - 1) // 2
    quirk = _expected_rows(m, L, fields, CV, chi, weight=lambda kv: 0.5 if abs(kv[2]) == half else 1.0)
    plain = _expected_rows(m, L, fields, CV, chi)
    assert _save_matches(rows, quirk)
    assert not _save_matches(rows, plain)


@pytest.mark.parametrize("fault,what", [
    (1, "weight 2 on the k2 = mz/2 plane"),
    (2, "n_same not reset between shells"),
    (3, "-0.5/chi_b dropped"),
    (4, "dK ignored"),
])
def test_each_planted_save_misreading_fails_the_pins(fault, what, tmp_path):
    caught = False
    for m, L in SAVE_CASES:
        M = int(np.prod(m))
        rng = np.random.default_rng(11)
        fields = [rng.normal(0, 1.0, M) for _ in range(2)]
        s = O.StrFunc(m, L, 37.0, 2.5)
        s.set_fault(fault)
        for f in fields:
            s.sample(f)
        s.save(str(tmp_path / "sk"))
        caught |= not _save_matches(_read_rows(tmp_path / "sk"), _expected_rows(m, L, fields, 37.0, 2.5))
    assert caught, f"planted fault '{what}' was not caught"
