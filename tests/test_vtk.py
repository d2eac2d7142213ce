"""`lfts-vtk` against a restatement of the reference's tools/make_vtk.sh:44-64: a fixed
STRUCTURED_POINTS header, then the first tab-separated field of each of the first M
lines of the field file, verbatim (for w_* files this includes the 3 header lines, as
the script does). With --xfast the values are reordered to VTK's x-fastest order."""
import os
import subprocess

import numpy as np
import pytest

import lfts_cpu_amd as L


def script_output(path, m):
    M = m[0] * m[1] * m[2]
    head = ["# vtk DataFile Version 2.0", "CT Density", "ASCII", "", "DATASET STRUCTURED_POINTS",
            f"DIMENSIONS {m[0]} {m[1]} {m[2]}", "ORIGIN 0.000000 0.000000 0.000000", "SPACING 0.1 0.1 0.1", "",
            f"POINT_DATA {M}", "SCALARS scalars float", "LOOKUP_TABLE default", ""]
    lines = open(path).read().split("\n")
    if lines and lines[-1] == "":
        lines = lines[:-1]
    body = [ln.split("\t")[0] for ln in lines[:M]]
    return "\n".join(head + body) + "\n"


def run_tool(path, m, *extra):
    r = subprocess.run([L.VTK_PATH, str(path), *map(str, m), *extra], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout
    return open(str(path) + ".vtk").read()


@pytest.fixture(scope="module")
def tool():
    if not os.path.exists(L.VTK_PATH):
        from lfts_cpu_amd import build
        build.build_vtk_tool()
    return L.VTK_PATH


def test_phi_file_matches_script(tool, tmp_path):
    m = (4, 3, 5)
    M = int(np.prod(m))
    v = np.random.default_rng(1).normal(size=2 * M)
    p = tmp_path / "phi_eq_10"
    p.write_text("".join("%g\n" % x for x in v))
    assert run_tool(p, m) == script_output(p, m)


def test_w_file_keeps_the_scripts_header_lines(tool, tmp_path):
    m = (3, 3, 2)
    M = int(np.prod(m))
    v = np.random.default_rng(2).normal(size=2 * M)
    p = tmp_path / "w_eq_10"
    p.write_text("64 32 14 50 0.5 1\n3 3 2 4 4 4\n200 0 10 1000 1\n" + "".join("%g\n" % x for x in v))
    out = run_tool(p, m)
    assert out == script_output(p, m)
    assert out.split("\n")[13] == "64 32 14 50 0.5 1"


def test_tabbed_lines_keep_the_first_field(tool, tmp_path):
    m = (2, 2, 2)
    p = tmp_path / "struct_like"
    p.write_text("".join(f"{i}.5\t{i * 7}\n" for i in range(10)))
    assert run_tool(p, m) == script_output(p, m)


def test_xfast_is_the_true_geometry(tool, tmp_path):
    m = (4, 3, 5)
    M = int(np.prod(m))
    x, y, z = np.meshgrid(*[np.arange(n) for n in m], indexing="ij")
    f = (100 * x + 10 * y + z).ravel()  # file order: z fastest
    p = tmp_path / "phi_st_2"
    p.write_text("".join("%d\n" % v for v in f) + "".join("0\n" for _ in range(M)))
    out = run_tool(p, m, "--xfast").split("\n")
    vals = np.array([int(t) for t in out[13:13 + M]])
    # VTK point n = x + mx (y + my z)
    n = np.arange(M)
    xs, ys, zs = n % m[0], (n // m[0]) % m[1], n // (m[0] * m[1])
    assert np.array_equal(vals, 100 * xs + 10 * ys + zs)
