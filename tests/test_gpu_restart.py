"""Lossless checkpoint/restart of the `lfts-gpu` driver: a run stopped with a checkpoint
and restarted from it prints the same lnQ lines and writes byte-identical files as the
uninterrupted run (the reference's own resume path re-seeds its noise and reloads
6-digit fields, so it cannot do this). Also the --vtk export."""
import os
import subprocess

import numpy as np
import pytest

import lfts_cpu_amd as L

pytestmark = pytest.mark.gpu

HEADER = "16 8 14.0 50 0.5 1\n12 12 12 3.0 3.0 3.0\n4 4 2 2 1\n"


def _input(tmp_path):
    M = 12 ** 3
    rng = np.random.default_rng(31)
    w0 = np.concatenate([rng.normal(0, 0.1, M), rng.normal(0, 0.1, M)])
    p = tmp_path / "input"
    with open(p, "w") as f:
        f.write(HEADER)
        for v in w0:
            f.write("%.17g\n" % v)
    return p


# Separate processes round alike only with the same contour-step factorisations: the
# library picks one per mesh edge by timing (lfts_slab_shapes), so the runs compared
# here fix them.
ENV = dict(os.environ, LFTS_SLAB_SHAPES="default")


def _run(args, cwd, env=ENV):
    print("driver", args, flush=True)
    r = subprocess.run([L.DRIVER_PATH, *map(str, args)], cwd=cwd, capture_output=True, text=True, timeout=300,
                       env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def _lnq(out):
    return [ln for ln in out.splitlines() if ln.startswith("lnQ = ")]


@pytest.mark.parametrize("stop", [3, 6])  # inside equilibration / inside statistics
def test_stop_and_restart_is_identical_to_one_run(gpu, tmp_path, stop):
    inp = _input(tmp_path)
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    out_a = _run([inp], a)
    out_b1 = _run([inp, "--stop-after", stop, "--checkpoint", "ck"], b)
    assert "Stopped after %d Langevin steps" % stop in out_b1
    out_b2 = _run([inp, "--restart", "ck", "--checkpoint", "ck2"], b)
    assert _lnq(out_b1) + _lnq(out_b2) == _lnq(out_a)
    assert out_b2.strip().splitlines()[-1] == out_a.strip().splitlines()[-1]  # H
    files_a = sorted(os.listdir(a))
    files_b = sorted(f for f in os.listdir(b) if not f.startswith("ck"))
    assert files_b == files_a
    for f in files_a:
        assert (a / f).read_bytes() == (b / f).read_bytes(), f


def test_periodic_checkpoints_and_a_corrupt_one(gpu, tmp_path):
    inp = _input(tmp_path)
    _run([inp, "--checkpoint-every", 2, "--checkpoint", "ck"], tmp_path)
    assert (tmp_path / "ck").exists() and not (tmp_path / "ck.tmp").exists()
    data = bytearray((tmp_path / "ck").read_bytes())
    data[100] ^= 1
    (tmp_path / "bad").write_bytes(bytes(data))
    r = subprocess.run([L.DRIVER_PATH, str(inp), "--restart", "bad"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 1 and "cannot read the checkpoint" in r.stdout


def test_vtk_export_matches_the_converter(gpu, tmp_path):
    inp = _input(tmp_path)
    _run([inp, "--vtk"], tmp_path)
    for f in ("phi_0", "phi_eq_2", "phi_st_4"):
        got = (tmp_path / (f + ".vtk")).read_bytes()
        os.rename(tmp_path / (f + ".vtk"), tmp_path / (f + ".vtk.driver"))
        r = subprocess.run([L.VTK_PATH, f, "12", "12", "12"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
        assert r.returncode == 0
        assert (tmp_path / (f + ".vtk")).read_bytes() == got, f


def test_replica_set_driver_matches_single_runs(gpu, tmp_path):
    """`lfts-gpu-batch` advances several inputs as one replica set; every replica's files
    and lnQ lines equal those of `lfts-gpu` run on that input alone."""
    specs = [("16 8 14.0 50 0.5 1\n12 12 12 3.0 3.0 3.0\n", 12 ** 3, 41),
             ("12 5 12.0 40 0.4 0\n10 8 6 3.1 2.6 2.2\n", 10 * 8 * 6, 42),
             ("10 4 15.0 50 0.5 1\n32 32 32 4.0 4.0 4.0\n", 32 ** 3, 43)]
    tail = "4 4 2 2 1\n"
    inputs, singles = [], []
    for r, (head, M, seed) in enumerate(specs):
        rng = np.random.default_rng(seed)
        w0 = np.concatenate([rng.normal(0, 0.1, M), rng.normal(0, 0.1, M)])
        text = head + tail + "".join("%.17g\n" % v for v in w0)
        d = tmp_path / f"rep{r}"
        d.mkdir()
        (d / "input").write_text(text)
        inputs.append(d / "input")
        s = tmp_path / f"single{r}"
        s.mkdir()
        (s / "input").write_text(text)
        singles.append(_run([s / "input"], s))
    print("batch driver", flush=True)
    rb = subprocess.run([L.BATCH_DRIVER_PATH, *map(str, inputs)], cwd=tmp_path, capture_output=True, text=True,
                        timeout=300, env=ENV)
    assert rb.returncode == 0, rb.stdout + rb.stderr
    for r in range(len(specs)):
        lb = [ln.split("] ", 1)[1] for ln in rb.stdout.splitlines() if ln.startswith(f"[{r}] lnQ = ")]
        assert lb == _lnq(singles[r])
        hb = [ln.split("= ")[1] for ln in rb.stdout.splitlines() if ln.startswith(f"[{r}] H = ")]
        assert hb == [singles[r].strip().splitlines()[-1]]
        da, db = tmp_path / f"single{r}", tmp_path / f"rep{r}"
        files = sorted(os.listdir(da))
        assert sorted(os.listdir(db)) == files
        for f in files:
            assert (da / f).read_bytes() == (db / f).read_bytes(), (r, f)


def test_restart_keeps_the_stopped_runs_factorisation(gpu, tmp_path):
    """A checkpoint carries the contour-step factorisation its run used: the restart, given
    no LFTS_SLAB_SHAPES of its own, continues with it (and writes it into its own
    checkpoints), and prints and writes what the uninterrupted run does."""
    M = 24 ** 3
    rng = np.random.default_rng(33)
    w0 = np.concatenate([rng.normal(0, 0.1, M), rng.normal(0, 0.1, M)])
    text = "12 5 14.0 50 0.5 1\n24 24 24 3.4 3.4 3.4\n4 4 2 2 1\n" + "".join("%.17g\n" % v for v in w0)
    alt = dict(os.environ, LFTS_SLAB_SHAPES="24:6x4")
    free = {k: v for k, v in os.environ.items() if k != "LFTS_SLAB_SHAPES"}
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    (a / "input").write_text(text)
    (b / "input").write_text(text)
    full = _run([a / "input"], a, alt)
    first = _run([b / "input", "--checkpoint", b / "ck", "--stop-after", 3], b, alt)
    assert b"24:6x4" in (b / "ck").read_bytes()
    rest = _run([b / "input", "--restart", b / "ck", "--checkpoint", b / "ck2", "--checkpoint-every", 1], b, free)
    assert b"24:6x4" in (b / "ck2").read_bytes()
    assert _lnq(first) + _lnq(rest) == _lnq(full)
    for f in sorted(os.listdir(a)):
        if f.startswith(("w_", "phi_", "struct_")):
            assert (a / f).read_bytes() == (b / f).read_bytes(), f
