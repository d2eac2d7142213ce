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


if __name__ == "__main__":
    O.build()
    rng_fixture()
    names = sys.argv[1:] or list(RUNS) + ["traj32"]
    for n in names:
        if n == "traj32":
            traj_run()
        else:
            run(n, RUNS[n])
