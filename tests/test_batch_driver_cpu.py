"""`lfts-gpu-batch` refuses, before any GPU work, inputs that would write their output files
into one directory (every replica writes w_0, phi_0, w_eq_*, ... into its input's
directory). Runs on the CPU: the check comes before the first device call."""
import os
import subprocess

import pytest

import lfts_cpu_amd as L

HEADER = "16 8 12.0 50 0.5 0\n8 8 8 2.0 2.0 2.0\n1 0 1 1 3\n"


def _write(path):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write(HEADER)


@pytest.mark.parametrize("spell", ["same", "dotdot", "bare"])
def test_batch_driver_refuses_inputs_sharing_a_directory(tmp_path, spell):
    if not os.path.exists(L.BATCH_DRIVER_PATH):
        pytest.skip("driver not built")
    a = tmp_path / "runs" / "a" / "input"
    _write(str(a))
    if spell == "same":
        b, args, cwd = tmp_path / "runs" / "a" / "input2", None, tmp_path
        _write(str(b))
        args = [str(a), str(b)]
    elif spell == "dotdot":
        (tmp_path / "runs" / "b").mkdir()
        args = [str(a), str(tmp_path / "runs" / "b" / ".." / "a" / "input")]
        cwd = tmp_path
    else:  # two bare names in the working directory
        _write(str(tmp_path / "runs" / "a" / "other"))
        args, cwd = ["input", "other"], tmp_path / "runs" / "a"
    r = subprocess.run([L.BATCH_DRIVER_PATH, *args], cwd=cwd, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1, r.stdout + r.stderr
    assert "in the same directory" in r.stdout
    assert "replica 0" not in r.stdout  # refused before any replica was set up
    assert not os.path.exists(os.path.join(os.path.dirname(str(a)), "w_0"))
