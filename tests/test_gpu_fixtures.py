"""GPU runs against the fixtures of tests/golden/make_golden.py (no oracle at run time):

  - bench64: the 64^3, N=90 workload of the throughput figure, ten whole Langevin steps
    from identical fields (the oracle's w+ after the initial mix);
  - example40_fields: the reference's example input (input_files/input, 40^3, loadType 1
    with its own 2M field values) through the C ABI and through the `lfts-gpu` driver;
  - the batched workload at its configured size: the 64-member set (24 x 32^3, 24 x 48^3,
    16 x 64^3, N=90), initial mix and two whole steps; every member bit-identical to the
    member run alone, and the first member of each mesh size against the oracle.

Bars. The per-step bar of Q (as lnQ), phi+-, H and S(k) is 1e-10 relative, and 1e-8 over a
run, unless the oracle run against itself in another rounding order (its 3D transforms
along x, y, z instead of z, y, x, its reductions summed from the last point) already
drifts further on the same run: then the bar is ten times that measured spread
(tests/golden/spread.json). Two correct evaluations that differ only in rounding cannot be
held closer than they drift apart; the GPU's transforms (separable register FFTs) and
reduction trees are one more such rounding. The same rule with 1e-8 sets the bars of W-
and w+, which a saddle-point solve fixes only to its residual tolerance.
"""
import json
import math
import os
import subprocess

import numpy as np
import pytest

import lfts_cpu_amd as L
from tests import fixture_metrics as F

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SPREAD = json.load(open(os.path.join(GOLD, "spread.json"))) if os.path.exists(os.path.join(GOLD, "spread.json")) \
    else {}
FIELDS = ["W-", "w+", "phi-", "phi+"]


def bar(base, *spread):
    """The bar: base, or ten times the largest measured oracle-vs-oracle spread."""
    s = [x for x in spread if x is not None]
    return max([base] + [10.0 * x for x in s])


def field_bars(per_step, spread_fields):
    """Bars for (W-, w+, phi-, phi+): phi+- at the per-step/run bar, W- and w+ at 1e-8."""
    sf = spread_fields or [None] * 4
    return [bar(1e-8, sf[0]), bar(1e-8, sf[1]), bar(per_step, sf[2]), bar(per_step, sf[3])]


# ---- the 64^3 throughput workload ---------------------------------------------------------

def test_bench64_steps_from_identical_fields(gpu):
    import bench

    g = dict(np.load(os.path.join(GOLD, "bench64.npz")))
    sp = SPREAD.get("bench64", {})
    m, N, NA, XbN, dt = (64, 64, 64), 90, 45, 17.0, 0.5
    M = int(g["M"])
    w0 = bench.initial_fields(m, 17.0, 0.05, 2)
    assert math.fsum(w0[:M].tolist()) == float(g["wm0_fsum"])  # the generator gives the fixture's W-
    sigma = float(g["sigma"])
    print(f"bench64: 64^3, N={N}, NA={NA}, XbN={XbN}, 10 steps from the oracle's post-mix w+", flush=True)
    ctx = L.Context(N, NA, m, (4.4, 4.4, 4.4))
    rng = L.Rng(123456789, sigma)
    ctx.langevin_init(rng.normals(M))
    ctx.set_fields(np.concatenate([w0[:M], g["wp0"]]))
    lnq0 = ctx.calc_concs()
    d0 = F.summary_diffs(F.prefixed(g, "start_"), ctx.get_fields(4), M)
    print(f"start: lnQ {abs(lnq0 - g['lnQ0']):.1e}, fields {['%.1e' % x for x in d0]}", flush=True)
    assert abs(lnq0 - float(g["lnQ0"])) <= 1e-12 * max(1.0, abs(float(g["lnQ0"])))
    assert max(d0) <= 1e-12
    steps_sp = sp.get("steps", [{}] * 10)
    worst = {}
    for s in range(10):
        ctx.langevin_step(rng.normals(M), XbN, dt)
        k, _ = ctx.anderson_mix(200, 1e-4)
        lnq = ctx.calc_concs()
        ctx.sk_sample()
        f = ctx.get_fields(4)
        ref = {"sums": g["sums"][s], "sumsq": g["sumsq"][s], "points": g["points"], "values": g["values"][s]}
        d = F.summary_diffs(ref, f, M)
        got = F.field_summary(f, M)
        H = F.hamiltonian(lnq, got["sums"], got["sumsq"], XbN, M)
        H0 = F.hamiltonian(g["lnQ"][s], g["sums"][s], g["sumsq"][s], XbN, M)
        e_lnq = abs(lnq - g["lnQ"][s]) / max(1.0, abs(g["lnQ"][s]))
        e_H = abs(H - H0) / max(1.0, abs(H0))
        print(f"step {s + 1}: n_AM+1 {k} ({int(g['nAM'][s])}), lnQ {e_lnq:.1e}, H {e_H:.1e}, "
              f"fields {['%.1e' % x for x in d]}", flush=True)
        assert k == int(g["nAM"][s]), s
        base = 1e-10 if s == 0 else 1e-8
        ss = steps_sp[s] if s < len(steps_sp) else {}
        assert e_lnq <= bar(base, ss.get("lnQ")), (s, e_lnq)
        assert e_H <= bar(base, ss.get("H")), (s, e_H)
        for q, (e, b) in enumerate(zip(d, field_bars(base, ss.get("fields")))):
            assert e <= b, (s, FIELDS[q], e, b)
            worst[FIELDS[q]] = max(worst.get(FIELDS[q], 0.0), e)
    S, ns = ctx.sk_get()
    assert ns == int(g["ns"]) == 10
    e_sk = max(F.rel(S[g["sk_points"]], g["sk_values"][0]),
               abs(math.fsum(S.tolist()) - g["sk_sums"][0]) / abs(g["sk_sums"][0]))
    print(f"worst fields over 10 steps {worst}; S(k) {e_sk:.1e}", flush=True)
    assert e_sk <= bar(1e-8, sp.get("sk"))
    ctx.close()


# ---- the reference's example input with its own fields --------------------------------------

def _example():
    g = dict(np.load(os.path.join(GOLD, "example40_fields.npz")))
    hdr = [str(x) for x in g["header"]]
    tok = (hdr[0] + " " + hdr[1]).split()
    N, NA, XN, C, dt, isXeN = int(tok[0]), int(tok[1]), float(tok[2]), float(tok[3]), float(tok[4]), int(tok[5])
    m = tuple(int(v) for v in tok[6:9])
    Lb = tuple(float(v) for v in tok[9:12])
    return g, hdr, N, NA, XN, C, dt, isXeN, m, Lb


def test_example_input_fields_through_the_c_abi(gpu):
    """src/lfts_simulation.cc:8-87 on input_files/input: the langevin constructor's M
    normals, the file's fields (loadType 1: no uniforms), the initial mix, ten steps."""
    g, hdr, N, NA, XN, C, dt, isXeN, m, Lb = _example()
    sp = SPREAD.get("example40_fields", {})
    M = int(np.prod(m))
    XbN = XN / L.z_inf(Lb, m, N, C * C) if isXeN else XN
    assert math.isclose(XbN, float(g["XbN"]), rel_tol=1e-15)
    sigma = math.sqrt(2.0 * M * dt / (C * Lb[0] * Lb[1] * Lb[2]))
    print(f"example input: mesh {m}, N={N}, NA={NA}, XbN={XbN}, its own fields, 10 steps", flush=True)
    ctx = L.Context(N, NA, m, Lb)
    rng = L.Rng(123456789, sigma)
    ctx.langevin_init(rng.normals(M))
    ctx.set_fields(g["fields"])
    k0, _ = ctx.anderson_mix(200, 1e-4)
    d_init = F.summary_diffs(F.prefixed(g, "init_"), ctx.get_fields(4), M)
    print(f"initial mix: n_AM+1 {k0} ({int(g['nAM_init'])}), fields {['%.1e' % x for x in d_init]}", flush=True)
    assert k0 == int(g["nAM_init"])
    lnq, nam = [], []
    for it in range(len(g["lnQ"])):
        ctx.langevin_step(rng.normals(M), XbN, dt)
        k, _ = ctx.anderson_mix(200, 1e-4)
        lnq.append(ctx.calc_concs())
        nam.append(k)
    d_end = F.summary_diffs(F.prefixed(g, "end_"), ctx.get_fields(4), M)
    H = ctx.hamiltonian(XbN)
    ctx.close()
    e_lnq = max(abs(a - b) / max(1.0, abs(b)) for a, b in zip(lnq, g["lnQ"]))
    e_H = abs(H - float(g["H"])) / max(1.0, abs(float(g["H"])))
    print(f"10 steps: n_AM+1 {nam}, lnQ {e_lnq:.1e}, H {e_H:.1e}, fields {['%.1e' % x for x in d_end]}", flush=True)
    assert nam == g["nAM"].tolist()
    assert e_lnq <= bar(1e-8, sp.get("lnQ")) and e_H <= bar(1e-8, sp.get("H"))
    for tag, d in (("init", d_init), ("end", d_end)):
        for q, (e, b) in enumerate(zip(d, field_bars(1e-8, sp.get(tag)))):
            assert e <= b, (tag, FIELDS[q], e, b)


def test_example_input_fields_through_the_driver(gpu, tmp_path):
    """`lfts-gpu` on input_files/input's parameters and fields (ten equilibration steps):
    the printed lnQ lines and H equal the fixture's to their 6 printed digits."""
    g, hdr, *_ = _example()
    inp = tmp_path / "input"
    with open(inp, "w") as f:
        f.write(hdr[0] + "\n" + hdr[1] + "\n" + f"{len(g['lnQ'])} 0 10 1000 1\n")
        for v in g["fields"]:
            f.write("%.17g\n" % v)
    r = subprocess.run([L.DRIVER_PATH, str(inp)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Loading input field..." in r.stdout
    printed = [float(ln[6:]) for ln in r.stdout.splitlines() if ln.startswith("lnQ = ")]
    assert len(printed) == len(g["lnQ"])
    for a, b in zip(printed, g["lnQ"]):
        assert abs(a - b) <= 6e-6 * abs(b), (a, b)
    H = float(r.stdout.strip().splitlines()[-1])
    assert abs(H - float(g["H"])) <= 6e-6 * abs(float(g["H"]))


# ---- the batched workload at its configured size ---------------------------------------------

def test_batched_workload_at_size(gpu):
    import torch

    import bench

    members = bench.batched_members()
    assert len(members) == 64
    g = dict(np.load(os.path.join(GOLD, "batch64_members.npz")))
    sp = SPREAD.get("batch64_members", {})
    nsteps = 2
    print("batched workload: 64 members (24 x 32^3, 24 x 48^3, 16 x 64^3), N=90", flush=True)
    ctxs, etas, eta0 = [], [], []
    for mb in members:
        c = L.Context(mb["N"], mb["NA"], (mb["n"],) * 3, (mb["L"],) * 3)
        c.set_fields(mb["w"])
        rng = L.Rng(mb["seed"], mb["sigma"])
        eta0.append(rng.normals(mb["M"]))
        c.langevin_init(eta0[-1])
        etas.append([torch.from_numpy(rng.normals(mb["M"])).cuda() for _ in range(nsteps)])
        ctxs.append(c)
    torch.cuda.synchronize()
    XbN = [mb["XbN"] for mb in members]
    B = L.Batch(ctxs)
    k0, _ = B.anderson_mix(200, 1e-4)
    print(f"set: initial mix, n_AM+1 in [{k0.min()}, {k0.max()}]", flush=True)
    res = []
    for s in range(nsteps):
        lnq, k = B.full_step_dev([etas[r][s].data_ptr() for r in range(64)], XbN, 0.5)
        res.append((lnq.copy(), k.copy()))
        print(f"set: step {s + 1}, n_AM+1 in [{k.min()}, {k.max()}]", flush=True)
    got = [c.get_fields(4) for c in ctxs]
    B.close()
    # against the oracle: the first member of each mesh size
    for r in (0, 24, 48):
        gr = {k[len(f"m{r}_"):]: v for k, v in g.items() if k.startswith(f"m{r}_")}
        spr = sp.get(f"member{r}", {})
        M = members[r]["M"]
        assert int(gr["n"]) == members[r]["n"]
        assert k0[r] == int(gr["nAM_init"]), r
        assert [int(res[s][1][r]) for s in range(nsteps)] == gr["nAM"].tolist(), r
        e_lnq = max(abs(res[s][0][r] - gr["lnQ"][s]) / max(1.0, abs(gr["lnQ"][s])) for s in range(nsteps))
        d = F.summary_diffs(F.prefixed(gr, "end_"), got[r], M)
        print(f"member {r} ({members[r]['n']}^3) vs oracle: lnQ {e_lnq:.1e}, fields {['%.1e' % x for x in d]}",
              flush=True)
        assert e_lnq <= bar(1e-10, spr.get("lnQ")), r
        for q, (e, b) in enumerate(zip(d, field_bars(1e-10, spr.get("end")))):
            assert e <= b, (r, FIELDS[q], e, b)
    # every member alone, from the same start, gives the same bits
    for r, (mb, c) in enumerate(zip(members, ctxs)):
        c.set_fields(mb["w"])
        c.langevin_init(eta0[r])
        ka, _ = c.anderson_mix(200, 1e-4)
        assert ka == k0[r], r
        for s in range(nsteps):
            lnq, k = c.full_step_dev(etas[r][s].data_ptr(), XbN[r], 0.5)
            assert lnq == res[s][0][r] and k == res[s][1][r], (r, s)
        assert np.array_equal(c.get_fields(4), got[r]), r
        c.close()
    print("every member alone equals its run in the set, bit for bit", flush=True)
