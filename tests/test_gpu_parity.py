"""GPU parity: the HIP path, called through the C ABI, against the oracle on the same
seeded inputs. Tolerances (FP64): 1e-12 relative for one contour step, 1e-10 for a
propagator evaluation, 1e-9 for a converged Anderson solve (its LU amplifies the
last-bit differences of the Gram sums by the condition number of U)."""
import math
import os
import subprocess

import numpy as np
import pytest

import lfts_cpu_amd as L
import oracle as O

pytestmark = pytest.mark.gpu

MESHES = [
    ((8, 8, 8), (3.0, 3.0, 3.0), 12, 6),
    ((6, 5, 8), (2.5, 2.0, 3.5), 10, 3),     # radices 2, 3, 5 and an odd axis
    ((16, 12, 10), (3.3, 2.9, 2.5), 16, 9),  # 12 = 4*3, 10 = 2*5
    ((40, 40, 40), (4.38, 4.38, 4.38), 8, 4),  # the reference's 40^3 example mesh
    ((7, 9, 14), (2.1, 2.7, 3.9), 9, 5),     # radix 7 and 9 = 3*3
    ((32, 32, 32), (4.0, 4.0, 4.0), 64, 32),
    ((32, 16, 24), (3.5, 2.8, 3.1), 10, 4),    # line path, register-FFT axis passes (edge lengths)
    ((64, 40, 48), (4.1, 3.3, 3.7), 8, 5),
    ((22, 13, 8), (3.1, 2.2, 1.9), 10, 4),     # radices 11 and 13 (generic line kernel)
    ((11, 26, 6), (2.0, 3.3, 1.5), 9, 5),
    ((17, 6, 38), (2.6, 1.8, 3.4), 8, 3),      # direct-DFT stages of primes 17 and 19
    ((23, 4, 5), (2.7, 1.2, 1.5), 7, 4),
    ((2, 131, 3), (0.9, 6.5, 1.2), 6, 2),      # a long prime axis: 131 products per output
]


def _fields(M, seed, a=2.0, b=1.0):
    rng = np.random.default_rng(seed)
    w = np.zeros(4 * M)
    w[:M] = rng.normal(0, a, M)
    w[M:2 * M] = rng.normal(0, b, M)
    return w


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("m,Lb,N,NA", MESHES)
def test_contour_step(gpu, m, Lb, N, NA):
    M = int(np.prod(m))
    rng = np.random.default_rng(M)
    qin = rng.uniform(0.2, 2.0, 2 * M)
    h = rng.uniform(0.5, 1.5, 2 * M)
    ctx = L.Context(N, NA, m, Lb)
    d = O.Diblock(N, NA, m, Lb)
    for i in (1, NA - 1, NA, N - NA, N - 1):
        i = max(1, min(N - 1, i))
        got = ctx.step_fwd(qin, h, i)
        ref = d.step_fwd(qin, h, i)
        assert rel(got[:M], ref[:M]) < 1e-12 and rel(got[M:], ref[M:]) < 1e-12, i


@pytest.mark.parametrize("m,Lb,N,NA", MESHES)
def test_calc_concs(gpu, m, Lb, N, NA):
    M = int(np.prod(m))
    w = _fields(M, 3)
    ctx = L.Context(N, NA, m, Lb)
    ctx.set_fields(w[:2 * M])
    lnq = ctx.calc_concs()
    got = ctx.get_fields(4)
    d = O.Diblock(N, NA, m, Lb)
    wo = w.copy()
    lnq_ref = d.calc_concs(wo)
    assert abs(lnq - lnq_ref) <= 1e-10 * max(1.0, abs(lnq_ref))
    assert rel(got[2 * M:3 * M], wo[2 * M:3 * M]) < 1e-10
    assert rel(got[3 * M:], wo[3 * M:]) < 1e-10
    for i in (1, 2, N // 2, N):
        assert rel(ctx.pair(i), d.pair(i)) < 1e-10, i
    # exact invariant: mean(phi+) = 1
    assert abs(got[3 * M:].mean() - 1.0) < 1e-12


def test_calc_concs_closed_forms(gpu):
    m, Lb, N, NA = (8, 6, 10), (3.0, 2.5, 3.5), 14, 5
    M = int(np.prod(m))
    ctx = L.Context(N, NA, m, Lb)
    a = 0.8
    w = np.zeros(2 * M)
    w[:M] = a
    ctx.set_fields(w)
    assert abs(ctx.calc_concs() + a * (NA - (N - NA)) / N) < 1e-13
    phi = ctx.get_fields(4)
    assert np.allclose(phi[3 * M:], 1.0, rtol=0, atol=1e-12)
    assert np.allclose(phi[2 * M:3 * M], (2 * NA - N) / N, rtol=0, atol=1e-12)


@pytest.mark.parametrize("m,Lb,N,NA", MESHES[:4])
def test_anderson_mix(gpu, m, Lb, N, NA):
    M = int(np.prod(m))
    w = _fields(M, 4, 3.0, 1.0)
    ctx = L.Context(N, NA, m, Lb)
    ctx.set_fields(w[:2 * M])
    k, err = ctx.anderson_mix(200, 1e-4)
    got = ctx.get_fields(4)
    d = O.Diblock(N, NA, m, Lb)
    am = O.Anderson(M, 10)
    wo = w.copy()
    k_ref = am.mix(d, 200, 1e-4, wo)
    assert k == k_ref
    assert rel(got[M:2 * M], wo[M:2 * M]) < 1e-9
    assert rel(got[2 * M:], wo[2 * M:]) < 1e-9
    assert math.isclose(err, am.last_err(), rel_tol=1e-6)


@pytest.mark.parametrize("m,nhMax", [((8, 6, 10), 1), ((8, 6, 10), 3), ((32, 32, 32), 7), ((8, 6, 10), 15)])
def test_anderson_history_depths(gpu, m, nhMax):
    """anderson(M, maxHist) with other history depths (src/anderson.h:30; the reference
    driver uses 10): the ring wraps after nhMax slots, with equal iteration counts and
    w+, phi to 1e-9 against the oracle; a 32^3 case runs the slab path."""
    M = int(np.prod(m))
    Lb, N, NA = (3.1, 2.8, 3.4), 14, 6
    w = _fields(M, 40 + nhMax, 3.0, 1.0)
    ctx = L.Context(N, NA, m, Lb, nhMax=nhMax)
    ctx.set_fields(w[:2 * M])
    k, err = ctx.anderson_mix(200, 1e-7)
    got = ctx.get_fields(4)
    d = O.Diblock(N, NA, m, Lb)
    am = O.Anderson(M, nhMax)
    wo = w.copy()
    k_ref = am.mix(d, 200, 1e-7, wo)
    print(f"nhMax {nhMax}: k = {k}", flush=True)
    assert k == k_ref and k > nhMax + 2  # the history ring wraps at least once
    assert rel(got[M:2 * M], wo[M:2 * M]) < 1e-9
    assert rel(got[2 * M:], wo[2 * M:]) < 1e-9
    assert math.isclose(err, am.last_err(), rel_tol=1e-6)


def test_langevin_step_host_and_device_noise(gpu):
    import torch

    m, Lb, N, NA = (8, 6, 10), (3.0, 2.5, 3.5), 12, 6
    M = int(np.prod(m))
    w = _fields(M, 5)
    w[2 * M:] = np.random.default_rng(6).normal(0, 0.3, 2 * M)
    e0, e1, e2 = (np.random.default_rng(s).normal(0, 1.5, M) for s in (7, 8, 9))
    XbN, dt = 17.0, 0.5
    ref = w.copy()
    O.langevin_update(ref, e0, e1, XbN, dt)
    O.langevin_update(ref, e1, e2, XbN, dt)
    for dev in (False, True):
        ctx = L.Context(N, NA, m, Lb)
        ctx.set_fields(w)
        ctx.langevin_init(e0)
        if dev:
            t1 = torch.from_numpy(e1).cuda()
            t2 = torch.from_numpy(e2).cuda()
            torch.cuda.synchronize()
            ctx.langevin_step_dev(t1.data_ptr(), XbN, dt)
            ctx.langevin_step_dev(t2.data_ptr(), XbN, dt)
        else:
            ctx.langevin_step(e1, XbN, dt)
            ctx.langevin_step(e2, XbN, dt)
        got = ctx.get_fields(4)
        assert rel(got[:M], ref[:M]) < 1e-14
        assert np.array_equal(got[M:], w[M:])


@pytest.mark.parametrize("m,Lb", [((8, 6, 10), (3.0, 2.5, 4.0)), ((40, 40, 40), (4.38, 4.38, 4.38)),
                                  ((7, 9, 5), (2.0, 2.0, 2.0)), ((11, 13, 4), (2.0, 2.3, 1.0)),
                                  ((17, 19, 2), (2.2, 2.4, 0.8))])
def test_structure_function(gpu, m, Lb):
    M = int(np.prod(m))
    ctx = L.Context(10, 5, m, Lb)
    s = O.StrFunc(m, Lb, 100.0, 20.0)
    for seed in (1, 2, 3):
        w = _fields(M, seed)
        ctx.set_fields(w[:2 * M])
        ctx.sk_sample()
        s.sample(w[:M])
    S, ns = ctx.sk_get()
    So, _, _, _, nso = s.get()
    assert ns == nso == 3
    assert rel(S, So) < 1e-12


def test_hamiltonian_closed_form(gpu):
    m, Lb, N, NA, XbN = (4, 6, 8), (2.0, 2.5, 3.0), 10, 4, 15.0
    M = int(np.prod(m))
    a, c = 0.6, 0.25
    ctx = L.Context(N, NA, m, Lb)
    ctx.set_fields(np.concatenate([np.full(M, a), np.full(M, c)]))
    H = ctx.hamiltonian(XbN)
    assert math.isclose(H, a * (2 * NA - N) / N + a * a / XbN, rel_tol=1e-12)


def _write_input(path, header, w):
    with open(path, "w") as f:
        f.write(header)
        for v in w:
            f.write("%.17g\n" % v)


@pytest.mark.parametrize("m,N,NA,XbN,steps", [((16, 16, 16), 16, 8, 14.0, 12), ((12, 10, 8), 20, 7, 18.0, 8)])
def test_langevin_trajectory(gpu, tmp_path, m, N, NA, XbN, steps):
    """Whole Langevin steps (noise from the reference's stream, Anderson to 1e-4, the
    extra calc_concs) against the oracle's restatement of the driver loop."""
    M = int(np.prod(m))
    Lb = (3.2, 3.0, 2.8)
    w0 = _fields(M, 10, 0.1, 0.1)[:2 * M]
    f = tmp_path / "in"
    _write_input(f, f"{N} {NA} {XbN} 50 0.5 0\n{m[0]} {m[1]} {m[2]} {Lb[0]} {Lb[1]} {Lb[2]}\n"
                    f"{steps} 0 1 1000 1\n", w0)
    sim = O.Sim(str(f))
    p = sim.params
    ctx = L.Context(N, NA, m, Lb)
    ctx.set_fields(w0)
    rng = L.Rng(123456789, p["sigma"])
    ctx.langevin_init(rng.normals(M))
    k0, _ = ctx.anderson_mix(200, 1e-4)
    assert k0 == sim.last_nAM()
    worst = 0.0
    for step in range(steps):
        ctx.langevin_step(rng.normals(M), p["chi_b"], p["dt"])
        k, _ = ctx.anderson_mix(200, 1e-4)
        lnq = ctx.calc_concs()
        lnq_ref = sim.step()
        assert k == sim.last_nAM(), step
        wg, wo = ctx.get_fields(4), sim.w()
        e = max(rel(wg[:M], wo[:M]), rel(wg[M:2 * M], wo[M:2 * M]), rel(wg[2 * M:], wo[2 * M:]),
                abs(lnq - lnq_ref) / max(1.0, abs(lnq_ref)))
        worst = max(worst, e)
        assert e < 1e-8, (step, e)
    print(f"worst relative difference over {steps} steps: {worst:.3e}")
    H = ctx.hamiltonian(p["chi_b"])
    assert math.isclose(H, sim.getH(), rel_tol=1e-8)


@pytest.mark.parametrize("n,N,NA", [(64, 12, 5), (32, 20, 13), (64, 9, 9), (48, 10, 4), (40, 11, 6), (128, 6, 2),
                                      (96, 7, 3), (16, 10, 10), (16, 5, 0), (24, 9, 4), (56, 8, 3),
                                      (72, 7, 7), (80, 6, 2), (112, 5, 3), (120, 5, 1)])
def test_fused_slab_path_against_oracle_and_generic_path(gpu, monkeypatch, n, N, NA):
    """The one-launch-per-step slab path (every cubic slab edge from 16^3 to 128^3) against the oracle
    and the three-pass line path on the same input."""
    m, Lb = (n, n, n), (4.4, 4.4, 4.4)
    M = n ** 3
    w = _fields(M, 12, 3.0, 1.0)
    ctx = L.Context(N, NA, m, Lb)
    ctx.set_fields(w[:2 * M])
    lnq = ctx.calc_concs()
    got = ctx.get_fields(4)
    monkeypatch.setenv("LFTS_PATH", "generic")
    gen = L.Context(N, NA, m, Lb)
    gen.set_fields(w[:2 * M])
    lnq_gen = gen.calc_concs()
    got_gen = gen.get_fields(4)
    d = O.Diblock(N, NA, m, Lb)
    wo = w.copy()
    lnq_ref = d.calc_concs(wo)
    for cand, wc in ((lnq, got), (lnq_gen, got_gen)):
        assert abs(cand - lnq_ref) <= 1e-10 * max(1.0, abs(lnq_ref))
        assert rel(wc[2 * M:], wo[2 * M:]) < 1e-10
    for i in (2, N // 2 + 1, N):
        assert rel(ctx.pair(i), d.pair(i)) < 1e-10, i
        assert rel(ctx.pair(i), gen.pair(i)) < 1e-12, i


@pytest.mark.parametrize("m,Lb,N,NA", [((32, 16, 24), (3.5, 2.8, 3.1), 10, 4), ((128, 64, 56), (5.0, 3.9, 3.3), 6, 3),
                                       ((16, 12, 120), (2.5, 2.2, 6.0), 7, 7),
                                       ((20, 36, 60), (2.6, 3.4, 4.4), 8, 3),     # lengths only the axis kernel serves
                                       ((144, 10, 8), (6.5, 2.4, 2.2), 6, 2),
                                       ((8, 160, 6), (2.2, 6.9, 2.0), 7, 4),
                                       ((6, 9, 192), (2.0, 2.5, 7.4), 6, 0),
                                       ((200, 6, 8), (7.6, 2.0, 2.2), 5, 3),
                                       ((8, 6, 256), (2.2, 2.0, 8.4), 6, 3),
                                       ((30, 50, 12), (3.0, 4.2, 2.4), 7, 2),
                                       ((100, 8, 6), (5.5, 2.2, 2.0), 6, 4),
                                       ((48, 80, 12), (3.7, 5.0, 2.4), 6, 2),    # x/y lines of 16 threads
                                       ((96, 112, 8), (5.2, 5.6, 2.2), 5, 3),
                                       ((144, 160, 6), (6.5, 6.9, 2.0), 5, 2),
                                       ((192, 128, 4), (7.4, 6.0, 1.8), 4, 2),
                                       # the remaining even lengths with factors 2, 3, 5, 7 (x/y and z lines)
                                       ((18, 28, 42), (2.3, 2.8, 3.6), 6, 3),
                                       ((54, 70, 12), (3.8, 4.4, 2.4), 5, 2),
                                       ((84, 90, 6), (4.8, 5.0, 2.0), 5, 3),
                                       ((98, 108, 4), (5.2, 5.5, 1.8), 4, 2),
                                       ((126, 140, 4), (5.9, 6.2, 1.8), 4, 1),
                                       ((150, 162, 4), (6.4, 6.7, 1.8), 4, 3),
                                       ((168, 180, 2), (6.8, 7.0, 1.5), 4, 2),
                                       ((196, 210, 2), (7.3, 7.6, 1.5), 4, 2),
                                       ((216, 224, 2), (7.6, 7.8, 1.5), 4, 2),
                                       ((240, 250, 2), (8.0, 8.2, 1.5), 4, 2),
                                       ((8, 6, 210), (2.2, 2.0, 7.6), 5, 2),
                                       ((6, 8, 168), (2.0, 2.2, 6.8), 5, 3),
                                       ((4, 6, 250), (1.8, 2.0, 8.2), 5, 2),
                                       ((4, 8, 216), (1.8, 2.2, 7.6), 5, 4),
                                       ((252, 6, 4), (8.3, 2.0, 1.8), 4, 2),
                                       ((40, 24, 32), (3.4, 2.9, 3.1), 6, 2),    # z lines of 32 and 48 points
                                       ((12, 20, 48), (2.4, 2.6, 3.7), 6, 3)])
def test_axis_kernel_against_generic_line_kernel(gpu, monkeypatch, m, Lb, N, NA):
    """Line-path meshes whose axes have slab-edge lengths run the register-FFT axis kernel;
    LFTS_PATH=generic runs the generic line kernel on the same input. Pairs, lnQ and phi
    agree to 1e-12 relative (different FFT arithmetic), and both to the oracle (the first
    two meshes, 1e-10)."""
    M = int(np.prod(m))
    w = _fields(M, 21, 3.0, 1.0)
    ctx = L.Context(N, NA, m, Lb)
    ctx.set_fields(w[:2 * M])
    lnq = ctx.calc_concs()
    monkeypatch.setenv("LFTS_PATH", "generic")
    gen = L.Context(N, NA, m, Lb)
    gen.set_fields(w[:2 * M])
    lnq_gen = gen.calc_concs()
    assert abs(lnq - lnq_gen) <= 1e-12 * max(1.0, abs(lnq_gen))
    assert rel(ctx.get_fields(4)[2 * M:], gen.get_fields(4)[2 * M:]) < 1e-12
    for i in (2, N // 2 + 1, N):
        assert rel(ctx.pair(i), gen.pair(i)) < 1e-12, i
    if M <= 64 * 40 * 48:
        d = O.Diblock(N, NA, m, Lb)
        wo = w.copy()
        lnq_ref = d.calc_concs(wo)
        assert abs(lnq - lnq_ref) <= 1e-10 * max(1.0, abs(lnq_ref))
        assert rel(ctx.get_fields(4)[2 * M:], wo[2 * M:]) < 1e-10


@pytest.mark.parametrize("mode", ["direct", "graph"])
def test_device_anderson_matches_host_loop(gpu, monkeypatch, mode):
    """The Anderson loop on the device (speculative chunks, optionally replayed as a HIP
    graph) gives the host loop's iteration counts and bit-identical fields."""
    m, Lb, N, NA = (32, 32, 32), (4.0, 4.0, 4.0), 16, 7
    M = int(np.prod(m))
    w = _fields(M, 8)
    monkeypatch.setenv("LFTS_GRAPH", "1" if mode == "graph" else "0")
    res = {}
    for kind in ("host", "device"):
        monkeypatch.setenv("LFTS_ANDERSON", kind)
        ctx = L.Context(N, NA, m, Lb)
        ctx.set_fields(w[:2 * M])
        out = []
        for tol in (1e-3, 1e-5, 1e-5, 1e-7):  # the later calls start near the solution
            k, err = ctx.anderson_mix(200, tol)
            out.append((k, ctx.get_fields(4)))
        res[kind] = out
        ctx.close()
    for (ka, wa), (kb, wb) in zip(res["host"], res["device"]):
        assert ka == kb
        assert np.array_equal(wa, wb)
