"""The committed golden fixtures (tests/golden/) against the oracle, the pure-Python
restatement of the noise stream and the host library.

- Noise stream (seed 123456789, src/lfts_simulation.cc:23, src/langevin.cc:16,28): the
  fixture's engine outputs, normal draws and the uniform draws after them equal the
  pure-Python mt19937_64 / polar-method restatement bit for bit, and the oracle's and the
  host library's streams.
- The reference's own test input (tests/valgrind/input_valgrind: 16^3, N=90, random
  start, one equilibration and one statistics step): the oracle's run reproduces the
  fixture's Anderson counts exactly and its lnQ, final fields, S(k) and H to 1e-12
  relative (the oracle's threaded sums may add in another order on another host).
- The parameters both fixture inputs derive (chi_b from XeN through z_inf, sigma) agree
  with the host library's z_inf and the closed form of sigma (src/lfts_params.cc:96-117)."""
import json
import math
import os

import numpy as np

import lfts_cpu_amd as L
import oracle as O
from tests import np_ref

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return dict(np.load(os.path.join(GOLD, name + ".npz")))


def _header(name):
    tok = open(os.path.join(GOLD, name + ".input")).read().split()
    N, NA = int(tok[0]), int(tok[1])
    XN, C, dt, isXeN = float(tok[2]), float(tok[3]), float(tok[4]), int(tok[5])
    m = tuple(int(v) for v in tok[6:9])
    Lb = tuple(float(v) for v in tok[9:12])
    loop = [int(v) for v in tok[12:17]]
    return N, NA, XN, C, dt, isXeN, m, Lb, loop


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def test_noise_stream_fixture():
    doc = json.load(open(os.path.join(GOLD, "rng_seed123456789.json")))
    eng = np_ref.MT19937_64(doc["seed"])
    assert [eng() for _ in range(8)] == doc["mt19937_64_first8"]
    assert [int(v) for v in O.mt64(doc["seed"], 8)] == doc["mt19937_64_first8"]
    z = [float.fromhex(v) for v in doc["normal_sigma1_first16"]]
    u = [float.fromhex(v) for v in doc["uniform_after_16_normals_first8"]]
    eng = np_ref.MT19937_64(doc["seed"])
    nd = np_ref.PolarNormal(0.0, 1.0)
    assert [nd(eng) for _ in range(16)] == z
    assert [np_ref.PolarNormal.canonical(eng) for _ in range(8)] == u
    assert O.normals(doc["seed"], 1.0, 16).tolist() == z
    r = L.Rng(doc["seed"], 1.0)
    assert r.normals(16).tolist() == z
    assert r.uniforms(8).tolist() == u


def test_fixture_parameters():
    for name in ("valgrind16", "input64"):
        g = _load(name)
        N, NA, XN, C, dt, isXeN, m, Lb, loop = _header(name)
        M = int(np.prod(m))
        assert int(g["M"]) == M and float(g["dt"]) == dt
        z = L.z_inf(Lb, m, N, C * C)
        XbN = XN / z if isXeN else XN
        assert math.isclose(float(g["XbN"]), XbN, rel_tol=1e-14)
        V = Lb[0] * Lb[1] * Lb[2]
        assert math.isclose(float(g["sigma"]), math.sqrt(2.0 * M * dt / (C * V)), rel_tol=1e-15)


def test_oracle_reproduces_reference_test_input_run():
    g = _load("valgrind16")
    sim = O.Sim(os.path.join(GOLD, "valgrind16.input"))
    p = sim.params
    assert sim.last_nAM() == int(g["nAM_init"])
    w0 = sim.w().reshape(4, -1)
    idx = g["init_points"]
    assert rel(w0[:, idx], g["init_values"]) < 1e-12
    lnq, nam = [], []
    for _ in range(p["equil_its"]):
        lnq.append(sim.step(False))
        nam.append(sim.last_nAM())
    for it in range(1, p["sim_its"] + 1):
        lnq.append(sim.step(it % p["sample_freq"] == 0))
        nam.append(sim.last_nAM())
    assert nam == g["nAM"].tolist()
    assert np.allclose(lnq, g["lnQ"], rtol=1e-12, atol=0)
    M = p["M"]
    w = sim.w()
    for k in range(4):
        assert rel(w[k * M:(k + 1) * M], g["w_end"][k * M:(k + 1) * M]) < 1e-12, k
    S, ns = sim.sk()
    assert ns == int(g["ns"]) == 1
    assert rel(S, g["sk"]) < 1e-12
    assert math.isclose(sim.getH(), float(g["H"]), rel_tol=1e-12)


def test_trajectory_fixture_start_and_first_steps(tmp_path):
    """The 100-step trajectory fixture (32^3, N=64): numpy still generates its starting
    fields, and the oracle reproduces its first three steps (Anderson counts exactly; lnQ,
    field sums and sampled values to 1e-12 relative)."""
    g = _load("traj32")
    header = str(g["header"])
    M = int(g["M"])
    w0 = np.random.default_rng(int(g["seed"])).normal(0.0, 0.1, 2 * M)
    assert np.array_equal(w0[:4], g["w0_head"])
    assert math.fsum(w0[:M].tolist()) == g["w0_sums"][0] and math.fsum(w0[M:].tolist()) == g["w0_sums"][1]
    f = tmp_path / "input"
    with open(f, "w") as fh:
        fh.write(header)
        for v in w0:
            fh.write("%.17g\n" % v)
    sim = O.Sim(str(f))
    assert sim.last_nAM() == int(g["nAM_init"])
    idx = g["points"]
    for s in range(3):
        lnq = sim.step(False)
        assert sim.last_nAM() == int(g["nAM"][s])
        assert abs(lnq - g["lnQ"][s]) <= 1e-12 * max(1.0, abs(g["lnQ"][s]))
        w = sim.w().reshape(4, M)
        for k in range(4):
            assert abs(math.fsum(w[k].tolist()) - g["sums"][s, k]) <= 1e-12 * max(abs(g["sums"][s, k]), M ** 0.5)
            assert rel(w[k, idx], g["values"][s, k]) < 1e-12
