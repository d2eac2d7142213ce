"""End to end: the `lfts-gpu` driver against the oracle's command-line restatement of
`lfts-cpu`, on the same input file, comparing stdout and every output file."""
import os
import subprocess

import numpy as np
import pytest

import lfts_cpu_amd as L
import oracle as O

pytestmark = pytest.mark.gpu


def _write_input(path, header, w):
    with open(path, "w") as f:
        f.write(header)
        for v in w:
            f.write("%.17g\n" % v)


def _numbers(path):
    return np.array([float(x) for x in open(path).read().split()])


# a cubic slab-path mesh, and meshes with factors 11 and 13 and with a larger prime (generic
# line kernel)
@pytest.mark.parametrize("m,Lb", [((12, 12, 12), (3.0, 3.0, 3.0)), ((22, 13, 6), (3.6, 2.2, 1.1)),
                                  ((17, 8, 6), (2.8, 1.4, 1.1))])
def test_driver_matches_oracle_outputs(gpu, tmp_path, m, Lb):
    N, NA = 16, 8
    M = int(np.prod(m))
    rng = np.random.default_rng(21)
    w0 = np.concatenate([rng.normal(0, 0.1, M), rng.normal(0, 0.1, M)])
    inp = tmp_path / "input"
    mesh = " ".join(str(x) for x in m) + " " + " ".join(str(x) for x in Lb)
    _write_input(inp, f"{N} {NA} 14.0 50 0.5 1\n{mesh}\n4 4 2 2 1\n", w0)
    gdir, odir = tmp_path / "gpu", tmp_path / "orc"
    gdir.mkdir()
    odir.mkdir()
    rg = subprocess.run([L.DRIVER_PATH, str(inp)], cwd=gdir, capture_output=True, text=True, timeout=300)
    assert rg.returncode == 0, rg.stdout + rg.stderr
    ro = subprocess.run([O.EXE_PATH, str(inp)], cwd=odir, capture_output=True, text=True, timeout=300)
    assert ro.returncode == 0, ro.stdout + ro.stderr
    lg = [ln for ln in rg.stdout.splitlines() if ln.startswith("lnQ = ")]
    lo = [ln for ln in ro.stdout.splitlines() if ln.startswith("lnQ = ")]
    assert len(lg) == len(lo) == 8
    for a, b in zip(lg, lo):
        assert abs(float(a[6:]) - float(b[6:])) <= 2e-6 * max(1.0, abs(float(b[6:]))), (a, b)
    # same set of files, same headers, values equal to the 6 printed digits
    files = sorted(os.listdir(odir))
    assert sorted(os.listdir(gdir)) == files
    assert {"w_0", "phi_0", "w_eq_2", "w_eq_4", "phi_st_4", "struct_st_2", "struct_st_4"} <= set(files)
    for fn in files:
        tg, to = (gdir / fn).read_text().splitlines(), (odir / fn).read_text().splitlines()
        assert len(tg) == len(to), fn
        if fn.startswith("w_"):
            assert tg[:3] == to[:3], fn
        a, b = _numbers(gdir / fn), _numbers(odir / fn)
        assert np.allclose(a, b, rtol=2e-5, atol=1e-5), fn
    # final H line
    assert abs(float(rg.stdout.strip().splitlines()[-1]) - float(ro.stdout.strip().splitlines()[-1])) < 1e-4
    # the parameter dump follows src/lfts_params.cc:46-70 name for name
    names = [ln.split(" = ")[0] for ln in rg.stdout.splitlines() if " = " in ln and not ln.startswith("lnQ")]
    assert names[:23] == ["N", "NA", "NB", "chi_b", "C", "dt", "XeN", "isXeN", "m[0]", "m[1]", "m[2]", "L[0]",
                          "L[1]", "L[2]", "equil_its", "sim_its", "sample_freq", "save_freq_", "loadType_", "M_",
                          "Mk_", "V_", "sigma_"]
