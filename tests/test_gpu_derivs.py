"""GPU parity of the derivatives of a propagator evaluation and of the Newton-Krylov
saddle solve, through the C ABI.

Bars (FP64): the tangent and adjoint products and the box derivative agree with the
oracle to 1e-10 relative (max-norm); at the 64^3, N = 90 workload (beyond the oracle's
reach in a test) the adjoint is checked against the tangent through
<u, J v> = <J^T u, v> (1e-11) and the tangent against central finite differences of the
GPU's own evaluation (h = 1e-6, 1e-6 relative: the truncation error of the difference
quotient); the Newton-Krylov w+ matches Anderson's converged w+ to 1e-8."""
import numpy as np
import pytest

import lfts_cpu_amd as L
import oracle as O

pytestmark = pytest.mark.gpu

# line path (odd sizes, radices 2/3/5, both chain-end quirks) and the slab path (16^3, 32^3, 40^3, 48^3, 56^3, 64^3, 80^3, 96^3, 128^3)
CASES = [
    ((4, 5, 6), (3.0, 3.5, 4.0), 12, 5),
    ((6, 4, 5), (2.5, 3.0, 2.0), 9, 9),
    ((4, 4, 6), (3.0, 3.0, 3.5), 8, 0),
    ((10, 12, 7), (3.2, 3.6, 2.9), 14, 6),
    ((32, 32, 32), (4.0, 4.0, 4.0), 16, 8),
    ((48, 48, 48), (4.5, 4.5, 4.5), 7, 3),
    ((64, 64, 64), (4.4, 4.4, 4.4), 6, 3),
    ((40, 40, 40), (4.38, 4.38, 4.38), 7, 4),
    ((128, 128, 128), (6.0, 6.0, 6.0), 4, 2),
    ((96, 96, 96), (5.2, 5.2, 5.2), 5, 2),
    ((16, 16, 16), (3.0, 3.0, 3.0), 11, 4),
    ((56, 56, 56), (4.6, 4.6, 4.6), 7, 3),
    ((80, 80, 80), (5.0, 5.0, 5.0), 5, 2),
    ((32, 24, 16), (3.6, 3.1, 2.5), 9, 4),     # line path with register-FFT axis passes
]


def _fields(M, seed, a=2.0, b=1.0):
    rng = np.random.default_rng(seed)
    w = np.zeros(4 * M)
    w[:M] = rng.normal(0, a, M)
    w[M:2 * M] = rng.normal(0, b, M)
    return w


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("m,Lb,N,NA", CASES)
def test_jvp_matches_oracle(gpu, m, Lb, N, NA):
    M = int(np.prod(m))
    print(f"jvp case m={m} N={N} NA={NA}", flush=True)
    w = _fields(M, 11)
    rng = np.random.default_rng(12)
    u, v = rng.normal(size=M), rng.normal(size=M)
    ctx = L.Context(N, NA, m, Lb)
    ctx.set_fields(w[:2 * M])
    got = ctx.jvp(u, v)
    ref = O.Diblock(N, NA, m, Lb).jvp(w, u, v)
    assert rel(got, ref) < 1e-10
    # phi in the context is that of the current fields
    phi = ctx.get_fields(4)[2 * M:]
    wo = w.copy()
    O.Diblock(N, NA, m, Lb).calc_concs(wo)
    assert rel(phi, wo[2 * M:]) < 1e-10
    # one direction only
    assert rel(ctx.jvp(None, v), O.Diblock(N, NA, m, Lb).jvp(w, None, v)) < 1e-10


@pytest.mark.parametrize("m,Lb,N,NA", CASES)
def test_vjp_matches_oracle(gpu, m, Lb, N, NA):
    M = int(np.prod(m))
    print(f"vjp case m={m} N={N} NA={NA}", flush=True)
    w = _fields(M, 13)
    ub = np.random.default_rng(14).normal(size=2 * M)
    ctx = L.Context(N, NA, m, Lb)
    ctx.set_fields(w[:2 * M])
    got = ctx.vjp(ub)
    ref = O.Diblock(N, NA, m, Lb).vjp(w, ub)
    assert rel(got, ref) < 1e-10


@pytest.mark.parametrize("m,Lb,N,NA", [c for c in CASES if 1 <= c[3] <= c[2] - 1])
def test_box_derivative_matches_oracle(gpu, m, Lb, N, NA):
    M = int(np.prod(m))
    print(f"box case m={m} N={N} NA={NA}", flush=True)
    w = _fields(M, 15)
    ctx = L.Context(N, NA, m, Lb)
    ctx.set_fields(w[:2 * M])
    dH, lnQ = ctx.box_derivative()
    lnQ_ref, g = O.Diblock(N, NA, m, Lb).dlnQ_dL(w)
    assert abs(lnQ - lnQ_ref) <= 1e-12 * max(1.0, abs(lnQ_ref))
    assert np.abs(dH + g).max() <= 1e-10 * np.abs(g).max()


def test_box_derivative_rejects_chain_end_quirk(gpu):
    ctx = L.Context(8, 8, (4, 4, 4), (3.0, 3.0, 3.0))
    ctx.set_fields(np.zeros(2 * 64))
    with pytest.raises(L.LftsError):
        ctx.box_derivative()


def test_slab_and_line_paths_agree(gpu, monkeypatch):
    m, Lb, N, NA = (32, 32, 32), (4.0, 4.0, 4.0), 20, 9
    M = int(np.prod(m))
    w = _fields(M, 16)
    rng = np.random.default_rng(17)
    v, ub = rng.normal(size=M), rng.normal(size=2 * M)
    a = L.Context(N, NA, m, Lb)
    monkeypatch.setenv("LFTS_PATH", "generic")
    b = L.Context(N, NA, m, Lb)
    for c in (a, b):
        c.set_fields(w[:2 * M])
    assert rel(a.jvp(None, v), b.jvp(None, v)) < 1e-12
    assert rel(a.vjp(ub), b.vjp(ub)) < 1e-12


def _lamellar(m, amp, seed):
    M = int(np.prod(m))
    x = np.arange(m[0])[:, None, None] * np.ones(m)
    rng = np.random.default_rng(seed)
    w = np.zeros(2 * M)
    w[:M] = (amp * np.cos(2 * np.pi * 3 * x / m[0])).ravel() + rng.normal(0, 0.5, M)
    w[M:] = rng.normal(0, 0.3, M)
    return w


def test_full_size_adjoint_identity_and_finite_differences(gpu):
    m, Lb, N, NA = (64, 64, 64), (4.4, 4.4, 4.4), 90, 45
    M = int(np.prod(m))
    print("full-size derivative checks 64^3 N=90", flush=True)
    w = _lamellar(m, 17.0, 2)
    ctx = L.Context(N, NA, m, Lb)
    rng = np.random.default_rng(5)
    u, v, ub = rng.normal(size=M), rng.normal(size=M), rng.normal(size=2 * M)
    ctx.set_fields(w)
    J = ctx.jvp(u, v)
    g = ctx.vjp(ub)
    lhs, rhs = ub @ J, g[:M] @ u + g[M:] @ v
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs)
    h = 1e-6
    wp, wm = w.copy(), w.copy()
    wp[:M] += h * u
    wp[M:] += h * v
    wm[:M] -= h * u
    wm[M:] -= h * v
    ctx.set_fields(wp)
    ctx.calc_concs()
    fp = ctx.get_fields(4)[2 * M:]
    ctx.set_fields(wm)
    ctx.calc_concs()
    fm = ctx.get_fields(4)[2 * M:]
    assert rel(J, (fp - fm) / (2 * h)) < 1e-6
    # box derivative against central differences of lnQ over each edge
    ctx.set_fields(w)
    dH, lnQ = ctx.box_derivative()
    for k in range(3):
        vals = []
        for s in (+1, -1):
            Lk = list(Lb)
            Lk[k] += s * h
            c2 = L.Context(N, NA, m, Lk)
            c2.set_fields(w)
            vals.append(c2.calc_concs())
            c2.close()
        fd = -(vals[0] - vals[1]) / (2 * h)
        assert abs(dH[k] - fd) <= 1e-6 * max(abs(fd), 1e-3), (k, dH[k], fd)


def test_newton_krylov_matches_anderson(gpu):
    # residual 1e-11 for both: the w+ error is about the residual times 1/|lambda_min| ~ N
    m, Lb, N, NA = (32, 32, 32), (4.0, 4.0, 4.0), 32, 16
    M = int(np.prod(m))
    print("newton-krylov vs anderson 32^3 N=32", flush=True)
    w0 = _lamellar(m, 10.0, 3)
    a = L.Context(N, NA, m, Lb)
    a.set_fields(w0)
    k, err = a.anderson_mix(maxIter=3000, tol=1e-11)
    wa = a.get_fields(2)
    b = L.Context(N, NA, m, Lb)
    b.set_fields(w0)
    info = b.saddle_nk(tol=1e-11)
    wn = b.get_fields(4)
    print("anderson evaluations", k - 1, "err", err, "newton-krylov", info, flush=True)
    assert err <= 1e-11 and info["err"] <= 1e-11
    assert np.abs(wn[M:2 * M] - wa[M:2 * M]).max() <= 1e-8
    assert np.array_equal(wn[:M], w0[:M])  # W- untouched
    assert abs(wn[M:2 * M].mean() - w0[M:].mean()) <= 1e-12  # mean of w+ kept
    # phi in the context belongs to the returned w+
    c = L.Context(N, NA, m, Lb)
    c.set_fields(wn[:2 * M])
    c.calc_concs()
    assert rel(c.get_fields(4)[2 * M:], wn[2 * M:]) < 1e-12


def test_autograd_function_gradcheck(gpu):
    import torch
    from tests.autograd_harness import PhiPlus

    m, Lb, N, NA = (4, 4, 3), (3.0, 3.0, 2.5), 6, 3
    M = int(np.prod(m))
    ctx = L.Context(N, NA, m, Lb)
    rng = np.random.default_rng(7)
    wp = torch.tensor(rng.normal(0, 0.5, M), dtype=torch.float64, requires_grad=True)
    wm = torch.tensor(rng.normal(0, 1.0, M), dtype=torch.float64, requires_grad=True)
    f = lambda a, b: PhiPlus.apply(a, b, ctx)  # noqa: E731
    assert torch.autograd.gradcheck(f, (wp, wm), eps=1e-6, atol=1e-7, rtol=1e-6, check_forward_ad=True,
                                    check_backward_ad=True, check_undefined_grad=False, check_batched_grad=False)


def test_langevin_step_with_newton_krylov_is_the_composition(gpu):
    import torch

    m, Lb, N, NA = (32, 32, 32), (4.0, 4.0, 4.0), 24, 12
    M = int(np.prod(m))
    w0 = _lamellar(m, 8.0, 9)
    rng = np.random.default_rng(10)
    eta_prev, eta_new = rng.normal(0, 0.5, M), rng.normal(0, 0.5, M)
    e = torch.from_numpy(eta_new).cuda()
    a, b = L.Context(N, NA, m, Lb), L.Context(N, NA, m, Lb)
    for c in (a, b):
        c.set_fields(w0)
        c.langevin_init(eta_prev)
    lnq_a, info_a = a.full_step_nk_dev(e.data_ptr(), 17.0, 0.5, tol=1e-6)
    b.langevin_step_dev(e.data_ptr(), 17.0, 0.5)
    info_b = b.saddle_nk(tol=1e-6)
    lnq_b = b.calc_concs()
    assert info_a == info_b and info_a["err"] <= 1e-6
    assert lnq_a == lnq_b
    assert np.array_equal(a.get_fields(4), b.get_fields(4))
