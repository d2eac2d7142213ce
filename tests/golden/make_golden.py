"""Writes the golden fixtures in this directory (data only: inputs and expected outputs).

Inputs are three parameter files held by the reference, copied unchanged:
  - tests/valgrind/input_valgrind (the reference's own test input): 16^3 mesh, N=90,
    random start (loadType 32), one equilibration and one statistics step;
  - input_files/input64 and input_files/input128: 64^3 and 128^3 meshes, N=90, random
    start (loadType 3), two equilibration steps;
and one written here, example40: the reference's example parameters (input_files/input,
40^3) with a random start (loadType 3) and ten equilibration steps.
All fields are generated from the seeded stream (src/lfts_simulation.cc:23,40-44,145-157),
so the three lines of each file fix the whole run.

Expected outputs come from the oracle's restatement of the reference's driver
(src/lfts_simulation.cc:8-135), run the way the reference runs it:
  constructor (M normals, M uniforms for W-, initial Anderson mix), then per step the
  Langevin update, Anderson mix to 1e-4, calc_concs (the printed lnQ), S(k) sampling
  every sample_freq statistics steps; H at the end.
For the 16^3 run the fixture keeps the final fields in full; for the larger runs it keeps
per-field sums, sums of squares and values at 64 fixed points per field.

Three more fixtures:
  - bench64: the 64^3 workload of the throughput figure (bench.py), w+ after the initial
    mix in full, then ten whole Langevin steps from those fields, per step lnQ, Anderson
    counts and field summaries, and S(k) after the ten samples;
  - example40_fields: input_files/input itself (40^3, loadType 1, its 2M field values kept
    as doubles), the reference's constructor and ten equilibration steps;
  - batch64_members: the first 32^3, 48^3 and 64^3 members of the 64-member batched
    workload, initial mix and two whole steps each.
`make_golden.py spread [names]` reruns the oracle with another rounding order against the
fixtures and writes the differences to spread.json (the GPU tests' bars come from it).

The noise-stream fixture holds the first outputs of std::mt19937_64 seeded 123456789,
normal draws of std::normal_distribution(0, 1) on that engine, and uniform draws that
follow them, as the host library produces them.

Run from the repository root in a container that holds the reference at /root/reference:
    python tests/golden/make_golden.py [run names]
(the 128^3 run takes the oracle about half an hour on 8 cores)
"""
import json
import math
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from tests import fixture_metrics as F  # noqa: E402

REF = "/root/reference"
RUNS = {
    "valgrind16": os.path.join(REF, "tests/valgrind/input_valgrind"),
    "input64": os.path.join(REF, "input_files/input64"),
    "input128": os.path.join(REF, "input_files/input128"),
    # the parameters of the reference's example input (input_files/input: N=90, NA=40,
    # XbN=17.148912, C=100, Ndt=1, 40^3, box 4.38^3) with the random start of loadType 3 in
    # place of its 1.6 MB field listing, and ten equilibration steps
    "example40": "90 40 17.148912 100.000000 1.000000 0\n40 40 40 4.38 4.38 4.38\n10 0 10 1000 3\n",
}
FULL_FIELDS = {"valgrind16"}
N_POINTS = 64


def sample_points(M: int) -> np.ndarray:
    return np.sort(np.random.default_rng(7).choice(M, N_POINTS, replace=False)).astype(np.int64)


def field_summary(w: np.ndarray, M: int) -> dict:
    idx = sample_points(M)
    f = w.reshape(4, M)
    return {
        "sums": np.array([math.fsum(f[k].tolist()) for k in range(4)]),
        "sumsq": np.array([math.fsum((f[k] * f[k]).tolist()) for k in range(4)]),
        "points": idx,
        "values": f[:, idx].copy(),
    }


def run(name: str, src: str) -> None:
    inp = os.path.join(HERE, name + ".input")
    if src.startswith("/"):
        shutil.copyfile(src, inp)
    else:
        with open(inp, "w") as f:
            f.write(src)
    sim = O.Sim(inp)
    p = sim.params
    M = p["M"]
    out = {
        "M": M, "XbN": p["chi_b"], "sigma": p["sigma"], "dt": p["dt"],
        "nAM_init": sim.last_nAM(),
    }
    w_init = sim.w()
    lnq, nam = [], []
    for it in range(1, p["equil_its"] + 1):
        lnq.append(sim.step(False))
        nam.append(sim.last_nAM())
        print(f"{name}: equilibration step {it}: lnQ = {lnq[-1]!r}, n_AM + 1 = {nam[-1]}", flush=True)
    for it in range(1, p["sim_its"] + 1):
        lnq.append(sim.step(it % p["sample_freq"] == 0))
        nam.append(sim.last_nAM())
        print(f"{name}: statistics step {it}: lnQ = {lnq[-1]!r}, n_AM + 1 = {nam[-1]}", flush=True)
    w_end = sim.w()
    S, ns = sim.sk()
    H = sim.getH()
    out.update({"lnQ": np.array(lnq), "nAM": np.array(nam, dtype=np.int64), "H": H, "ns": ns})
    if name in FULL_FIELDS:
        out["w_end"] = w_end
        out["sk"] = S
        s0 = field_summary(w_init, M)
        out.update({"init_" + k: v for k, v in s0.items()})
    else:
        for tag, w in (("init_", w_init), ("end_", w_end)):
            s = field_summary(w, M)
            out.update({tag + k: v for k, v in s.items()})
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print(f"{name}: n_AM(initial) + 1 = {out['nAM_init']}, H = {H!r}", flush=True)


# 100 whole Langevin steps of the CPU-runnable workload: diblock N=64, NA=32, XeN=14.0
# (isXeN=1), C=50, Ndt=0.5, 32^3 mesh, box 4.0^3, W- and w+ i.i.d. N(0, 0.1^2) from numpy
# seed 1; 50 equilibration steps, then 50 statistics steps sampling S(k) every step.
TRAJ_HEADER = "64 32 14.0 50 0.5 1\n32 32 32 4.0 4.0 4.0\n50 50 1 1000 1\n"
TRAJ_SEED = 1
TRAJ_POINTS = 16


def traj_fields(M: int) -> np.ndarray:
    return np.random.default_rng(TRAJ_SEED).normal(0.0, 0.1, 2 * M)


def traj_run() -> None:
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        inp = os.path.join(td, "input")
        with open(inp, "w") as f:
            f.write(TRAJ_HEADER)
        p = O.read_params(inp)
        M = p["M"]
        w0 = traj_fields(M)
        with open(inp, "a") as f:
            for v in w0:
                f.write("%.17g\n" % v)
        sim = O.Sim(inp)
        k0 = sim.last_nAM()
        idx = np.sort(np.random.default_rng(11).choice(M, TRAJ_POINTS, replace=False)).astype(np.int64)
        nsteps = p["equil_its"] + p["sim_its"]
        lnq = np.zeros(nsteps)
        nam = np.zeros(nsteps, dtype=np.int64)
        sums = np.zeros((nsteps, 4))
        sumsq = np.zeros((nsteps, 4))
        vals = np.zeros((nsteps, 4, TRAJ_POINTS))
        for s in range(nsteps):
            it = s + 1 - p["equil_its"]
            lnq[s] = sim.step(it >= 1 and it % p["sample_freq"] == 0)
            nam[s] = sim.last_nAM()
            f = sim.w().reshape(4, M)
            for k in range(4):
                sums[s, k] = math.fsum(f[k].tolist())
                sumsq[s, k] = math.fsum((f[k] * f[k]).tolist())
            vals[s] = f[:, idx]
            if s % 10 == 9:
                print(f"traj32: step {s + 1}: lnQ = {lnq[s]!r}, n_AM + 1 = {nam[s]}", flush=True)
        S, ns = sim.sk()
        np.savez_compressed(
            os.path.join(HERE, "traj32.npz"),
            header=np.array(TRAJ_HEADER), seed=TRAJ_SEED, M=M, XbN=p["chi_b"], sigma=p["sigma"],
            w0_sums=np.array([math.fsum(w0[:M].tolist()), math.fsum(w0[M:].tolist())]),
            w0_head=w0[:4], nAM_init=k0, lnQ=lnq, nAM=nam, sums=sums, sumsq=sumsq, points=idx,
            values=vals, sk=S, ns=ns, H_end=sim.getH())


# ---- the 64^3 plain workload of the throughput figure ---------------------------------------
# N=90, NA=45, XbN=17.0 (bare), C=50, Ndt=0.5, mesh 64^3, box 4.4^3; W- = 17 cos(2 pi 3x/64)
# + N(0, 0.05^2) from numpy seed 2, w+ = 0 (bench.initial_fields); noise from mt19937_64
# seed 123456789. The fixture keeps w+ after the initial Anderson mix in full, so that a
# GPU run can start from identical fields: both sides then evaluate phi of those fields
# (calc_concs) and run BENCH_STEPS whole Langevin steps (step_wm, mix, calc_concs),
# sampling S(k) after every step.
BENCH = dict(N=90, NA=45, XbN=17.0, C=50.0, dt=0.5, m=(64, 64, 64), L=(4.4, 4.4, 4.4))
BENCH_STEPS = 10


def bench_start():
    import bench
    m = BENCH["m"]
    M = int(np.prod(m))
    w = np.zeros(4 * M)
    w[:2 * M] = bench.initial_fields(m, 17.0, 0.05, 2)
    V = float(np.prod(BENCH["L"]))
    sigma = math.sqrt(2.0 * M * BENCH["dt"] / (BENCH["C"] * V))
    return w, M, sigma


def bench64_run(order: int = 0, wp0=None) -> dict:
    """order != 0: the oracle's rounding-order variant from the fixture's own w+ (wp0)."""
    O.set_order(order)
    w, M, sigma = bench_start()
    p = BENCH
    eta = O.normals(123456789, sigma, (BENCH_STEPS + 1) * M).reshape(BENCH_STEPS + 1, M)
    d = O.Diblock(p["N"], p["NA"], p["m"], p["L"])
    am = O.Anderson(M, 10)
    out = {"M": M, "sigma": sigma, "XbN": p["XbN"], "wm0_fsum": math.fsum(w[:M].tolist())}
    out["nAM_init"] = am.mix(d, 200, 1e-4, w)
    out["mix_wp"] = w[M:2 * M].copy()  # this run's own initial mix
    if wp0 is not None:
        w[M:2 * M] = wp0
    out["wp0"] = w[M:2 * M].copy()
    out["lnQ0"] = d.calc_concs(w)
    out.update({"start_" + k: v for k, v in F.field_summary(w, M).items()})
    sk = O.StrFunc(p["m"], p["L"], 1.0, p["XbN"])
    lnq, nam, summ = [], [], []
    for s in range(BENCH_STEPS):
        O.langevin_update(w, eta[s], eta[s + 1], p["XbN"], p["dt"])
        nam.append(am.mix(d, 200, 1e-4, w))
        lnq.append(d.calc_concs(w))
        sk.sample(w[:M])
        summ.append(F.field_summary(w, M))
        print(f"bench64 (order {order}): step {s + 1}: lnQ = {lnq[-1]!r}, n_AM + 1 = {nam[-1]}", flush=True)
    S, _, _, _, ns = sk.get()
    out.update({"lnQ": np.array(lnq), "nAM": np.array(nam, dtype=np.int64), "ns": ns,
                "points": summ[0]["points"],
                "sums": np.array([x["sums"] for x in summ]), "sumsq": np.array([x["sumsq"] for x in summ]),
                "values": np.array([x["values"] for x in summ])})
    out.update({"sk_" + k: v for k, v in F.field_summary(S, S.size, 1).items()})
    O.set_order(0)
    return out


def bench64_fixture() -> None:
    out = bench64_run()
    out.pop("mix_wp")
    np.savez_compressed(os.path.join(HERE, "bench64.npz"), **out)


# ---- the reference's example input with its own fields ---------------------------------------
# input_files/input: N=90, NA=40, XbN=17.148912 (bare), C=100, Ndt=1, 40^3, box 4.38^3,
# loadType 1 with 2M field values (src/lfts_simulation.cc:45-48, README.md:50-60). The
# fixture keeps those values as doubles (parsed as the reference's `>>` parses them) and
# the oracle's run of the reference's constructor and EXAMPLE_STEPS equilibration steps.
EXAMPLE_STEPS = 10


def example_fields_input(path: str, fields: np.ndarray, header: list) -> None:
    with open(path, "w") as f:
        f.write(header[0] + "\n" + header[1] + "\n" + f"{EXAMPLE_STEPS} 0 10 1000 1\n")
        for v in fields:
            f.write("%.17g\n" % v)


def example_fields_run(order: int = 0, fields=None, header=None) -> dict:
    import tempfile
    if fields is None:
        lines = open(os.path.join(REF, "input_files/input")).read().split("\n")
        header = [lines[0].strip(), lines[1].strip(), lines[2].strip()]
        fields = np.array([float(x) for x in " ".join(lines[3:]).split()])
    O.set_order(order)
    with tempfile.TemporaryDirectory() as td:
        inp = os.path.join(td, "input")
        example_fields_input(inp, fields, header)
        sim = O.Sim(inp)
        p = sim.params
        M = p["M"]
        assert fields.size == 2 * M
        out = {"header": np.array(header), "fields": fields, "M": M, "XbN": p["chi_b"], "sigma": p["sigma"],
               "nAM_init": sim.last_nAM()}
        out.update({"init_" + k: v for k, v in F.field_summary(sim.w(), M).items()})
        lnq, nam = [], []
        for it in range(EXAMPLE_STEPS):
            lnq.append(sim.step(False))
            nam.append(sim.last_nAM())
            print(f"example40 fields (order {order}): step {it + 1}: lnQ = {lnq[-1]!r}, n_AM + 1 = {nam[-1]}",
                  flush=True)
        out.update({"end_" + k: v for k, v in F.field_summary(sim.w(), M).items()})
        out.update({"lnQ": np.array(lnq), "nAM": np.array(nam, dtype=np.int64), "H": sim.getH()})
    O.set_order(0)
    return out


def example_fields_fixture() -> None:
    np.savez_compressed(os.path.join(HERE, "example40_fields.npz"), **example_fields_run())


# ---- members of the batched workload ----------------------------------------------------------
# One member per mesh size of the 64-member set (bench.batched_members: members 0, 24, 48 are
# the first 32^3, 48^3 and 64^3 ones): the initial mix and BATCH_STEPS whole Langevin steps,
# as the set runs them (langevin's first draw, the initial mix, then step_wm, mix,
# calc_concs), with the member's own noise stream (mt19937_64 seeded 1000+r).
BATCH_REPS = (0, 24, 48)
BATCH_STEPS = 2


def batch_member_run(r: int, order: int = 0) -> dict:
    import bench
    mb = bench.batched_member(r, bench.BATCH_MESHES[r])
    O.set_order(order)
    M, n = mb["M"], mb["n"]
    w = np.zeros(4 * M)
    w[:2 * M] = mb["w"]
    eta = O.normals(mb["seed"], mb["sigma"], (BATCH_STEPS + 1) * M).reshape(BATCH_STEPS + 1, M)
    d = O.Diblock(mb["N"], mb["NA"], (n, n, n), (mb["L"],) * 3)
    am = O.Anderson(M, 10)
    out = {"r": r, "n": n, "M": M, "nAM_init": am.mix(d, 200, 1e-4, w)}
    out.update({"init_" + k: v for k, v in F.field_summary(w, M).items()})
    lnq, nam = [], []
    for s in range(BATCH_STEPS):
        O.langevin_update(w, eta[s], eta[s + 1], mb["XbN"], mb["dt"])
        nam.append(am.mix(d, 200, 1e-4, w))
        lnq.append(d.calc_concs(w))
        print(f"batch member {r} ({n}^3, order {order}): step {s + 1}: lnQ = {lnq[-1]!r}, n_AM + 1 = {nam[-1]}",
              flush=True)
    out.update({"end_" + k: v for k, v in F.field_summary(w, M).items()})
    out.update({"lnQ": np.array(lnq), "nAM": np.array(nam, dtype=np.int64)})
    O.set_order(0)
    return out


def batch_fixture() -> None:
    out = {}
    for r in BATCH_REPS:
        for k, v in batch_member_run(r).items():
            out[f"m{r}_{k}"] = v
    np.savez_compressed(os.path.join(HERE, "batch64_members.npz"), **out)


# ---- the rounding-order spread ----------------------------------------------------------------
# The oracle with its 3D transforms along x, y, z and its reductions summed from the last
# point (O.set_order(3)) is the same arithmetic rounded differently. Run against each
# fixture, its differences measure how far two correct evaluations of the same run drift
# apart; the GPU tests' bars are set from them (spread.json).
SPREAD_ORDER = 3


def spread(names) -> None:
    path = os.path.join(HERE, "spread.json")
    doc = json.load(open(path)) if os.path.exists(path) else {}
    for name in names:
        entry = {"order": SPREAD_ORDER}
        if name in RUNS:
            g = dict(np.load(os.path.join(HERE, name + ".npz")))
            O.set_order(SPREAD_ORDER)
            sim = O.Sim(os.path.join(HERE, name + ".input"))
            p = sim.params
            M = p["M"]
            entry["nAM_init_equal"] = sim.last_nAM() == int(g["nAM_init"])
            w0 = sim.w()
            lnq, nam = [], []
            for it in range(1, p["equil_its"] + p["sim_its"] + 1):
                lnq.append(sim.step(it > p["equil_its"] and (it - p["equil_its"]) % p["sample_freq"] == 0))
                nam.append(sim.last_nAM())
            w1 = sim.w()
            H = sim.getH()
            S, ns = sim.sk()
            O.set_order(0)
            entry["nAM_equal"] = nam == g["nAM"].tolist()
            entry["lnQ"] = max(abs(a - b) / max(1.0, abs(b)) for a, b in zip(lnq, g["lnQ"]))
            entry["H"] = abs(H - float(g["H"])) / max(1.0, abs(float(g["H"])))
            entry["init"] = F.summary_diffs(F.prefixed(g, "init_"), w0, M)
            if "w_end" in g:
                entry["end"] = [F.rel(w1[k * M:(k + 1) * M], g["w_end"][k * M:(k + 1) * M]) for k in range(4)]
                entry["sk"] = F.rel(S, g["sk"])
            else:
                entry["end"] = F.summary_diffs(F.prefixed(g, "end_"), w1, M)
        elif name == "bench64":
            g = dict(np.load(os.path.join(HERE, "bench64.npz")))
            out = bench64_run(SPREAD_ORDER, wp0=g["wp0"])
            M = int(g["M"])
            entry["nAM_init_equal"] = int(out["nAM_init"]) == int(g["nAM_init"])
            entry["initial_mix_wp"] = F.rel(out["mix_wp"], g["wp0"])
            entry["lnQ0"] = abs(out["lnQ0"] - float(g["lnQ0"])) / max(1.0, abs(float(g["lnQ0"])))
            entry["nAM_equal"] = out["nAM"].tolist() == g["nAM"].tolist()
            steps = []
            for s in range(BENCH_STEPS):
                ref = {"sums": g["sums"][s], "sumsq": g["sumsq"][s], "points": g["points"], "values": g["values"][s]}
                got = {"sums": out["sums"][s], "sumsq": out["sumsq"][s], "points": out["points"],
                       "values": out["values"][s]}
                d = []
                for k in range(4):
                    ds = abs(got["sums"][k] - ref["sums"][k]) / max(abs(ref["sums"][k]), M ** 0.5)
                    dq = abs(got["sumsq"][k] - ref["sumsq"][k]) / ref["sumsq"][k]
                    d.append(max(ds, dq, F.rel(got["values"][k], ref["values"][k])))
                H1 = F.hamiltonian(out["lnQ"][s], out["sums"][s], out["sumsq"][s], BENCH["XbN"], M)
                H0 = F.hamiltonian(g["lnQ"][s], g["sums"][s], g["sumsq"][s], BENCH["XbN"], M)
                steps.append({"lnQ": abs(out["lnQ"][s] - g["lnQ"][s]) / max(1.0, abs(g["lnQ"][s])),
                              "H": abs(H1 - H0) / max(1.0, abs(H0)), "fields": d})
            entry["steps"] = steps
            entry["sk"] = max(F.rel(out["sk_values"], g["sk_values"]),
                              abs(out["sk_sums"][0] - g["sk_sums"][0]) / abs(g["sk_sums"][0]))
        elif name == "example40_fields":
            g = dict(np.load(os.path.join(HERE, "example40_fields.npz")))
            out = example_fields_run(SPREAD_ORDER, g["fields"], [str(x) for x in g["header"]])
            M = int(g["M"])
            entry["nAM_init_equal"] = int(out["nAM_init"]) == int(g["nAM_init"])
            entry["nAM_equal"] = out["nAM"].tolist() == g["nAM"].tolist()
            entry["lnQ"] = max(abs(a - b) / max(1.0, abs(b)) for a, b in zip(out["lnQ"], g["lnQ"]))
            entry["H"] = abs(out["H"] - float(g["H"])) / max(1.0, abs(float(g["H"])))
            entry["init"] = _summary_pair(g, out, "init_")
            entry["end"] = _summary_pair(g, out, "end_")
        elif name == "batch64_members":
            g = dict(np.load(os.path.join(HERE, "batch64_members.npz")))
            for r in BATCH_REPS:
                gr = {k[len(f"m{r}_"):]: v for k, v in g.items() if k.startswith(f"m{r}_")}
                out = batch_member_run(r, SPREAD_ORDER)
                entry[f"member{r}"] = {
                    "nAM_init_equal": int(out["nAM_init"]) == int(gr["nAM_init"]),
                    "nAM_equal": out["nAM"].tolist() == gr["nAM"].tolist(),
                    "lnQ": max(abs(a - b) / max(1.0, abs(b)) for a, b in zip(out["lnQ"], gr["lnQ"])),
                    "init": _summary_pair(gr, out, "init_"), "end": _summary_pair(gr, out, "end_")}
        else:
            raise ValueError(name)
        doc[name] = entry
        print(f"spread {name}: {json.dumps(entry)}", flush=True)
        with open(path, "w") as f:
            json.dump(doc, f, indent=1, sort_keys=True)
            f.write("\n")


def _summary_pair(g: dict, out: dict, tag: str) -> list:
    """summary_diffs between two summaries (same points) of one stage."""
    a, b = F.prefixed(out, tag), F.prefixed(g, tag)
    M = int(g["M"])
    res = []
    for k in range(4):
        ds = abs(a["sums"][k] - b["sums"][k]) / max(abs(b["sums"][k]), M ** 0.5)
        dq = abs(a["sumsq"][k] - b["sumsq"][k]) / b["sumsq"][k]
        res.append(max(ds, dq, F.rel(a["values"][k], b["values"][k])))
    return res


def rng_fixture() -> None:
    raw = O.mt64(123456789, 8)
    z = O.normals(123456789, 1.0, 16)
    import lfts_cpu_amd as L
    r = L.Rng(123456789, 1.0)
    zl = r.normals(16)
    u = r.uniforms(8)
    assert np.array_equal(z, zl), "oracle and host library disagree on the normal stream"
    doc = {
        "seed": 123456789,
        "mt19937_64_first8": [int(v) for v in raw],
        "normal_sigma1_first16": [float(v).hex() for v in zl],
        "uniform_after_16_normals_first8": [float(v).hex() for v in u],
    }
    with open(os.path.join(HERE, "rng_seed123456789.json"), "w") as f:
        json.dump(doc, f, indent=1)
        f.write("\n")


EXTRA = {"traj32": traj_run, "bench64": bench64_fixture, "example40_fields": example_fields_fixture,
         "batch64_members": batch_fixture}
SPREAD_NAMES = ["valgrind16", "example40", "input64", "bench64", "example40_fields", "batch64_members", "input128"]


if __name__ == "__main__":
    O.build()
    args = sys.argv[1:]
    if args and args[0] == "spread":
        spread(args[1:] or SPREAD_NAMES)
        sys.exit(0)
    rng_fixture()
    names = args or list(RUNS) + list(EXTRA)
    for n in names:
        if n in EXTRA:
            EXTRA[n]()
        else:
            run(n, RUNS[n])
