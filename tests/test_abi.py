"""The C-ABI library without a GPU: it loads, exports every entry point that
include/lfts_gpu.h declares, and its host-side parts (noise stream, S(k) bins and file
format, z_inf) agree bit-for-bit with the oracle. No compute call is made here."""
import os
import re
import subprocess

import numpy as np
import pytest

import lfts_cpu_amd as L
import oracle as O
from tests.conftest import ROOT, has_gpu

HEADER = os.path.join(ROOT, "include", "lfts_gpu.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lfts_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_binding():
    names = declared_functions()
    assert len(names) >= 25
    assert set(names) == set(L.SIGNATURES), set(names) ^ set(L.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_library_is_gfx950_code():
    # the embedded offload bundle (built here) names its only target
    blob = open(L.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob
    assert b"sm_" not in blob.split(b"amdgcn-amd-amdhsa")[0][-64:]


@pytest.mark.parametrize("m,Lb", [((8, 6, 10), (3.0, 2.5, 4.0)), ((5, 4, 7), (2.0, 2.0, 2.0)),
                                  ((64, 64, 64), (4.4, 4.4, 4.4)), ((40, 40, 40), (4.38, 4.38, 4.38))])
def test_sk_bins_bit_exact(m, Lb):
    K, wt, P = L.sk_bins(m, Lb)
    s = O.StrFunc(m, Lb, 1.0, 1.0)
    _, Ko, wto, Po, _ = s.get()
    assert np.array_equal(K, Ko) and np.array_equal(wt, wto) and np.array_equal(P, Po)


def test_sk_save_identical_to_oracle(tmp_path):
    m, Lb, CV, chi = (6, 4, 8), (2.0, 3.0, 2.5), 77.0, 14.0
    M = int(np.prod(m))
    s = O.StrFunc(m, Lb, CV, chi)
    rng = np.random.default_rng(1)
    s.sample(rng.normal(0, 1, M))
    s.sample(rng.normal(0, 1, M))
    S, _, _, _, ns = s.get()
    s.save(str(tmp_path / "a"))
    L.sk_save(str(tmp_path / "b"), m, Lb, S, ns, CV, chi)
    assert (tmp_path / "a").read_bytes() == (tmp_path / "b").read_bytes()


def test_sk_save_zero_samples_prints_nan(tmp_path):
    """n_st = 0 still saves struct_st_0 with nsamples = 0: (coeff/0)*0 -> nan (-nan)."""
    m, Lb = (4, 4, 4), (2.0, 2.0, 2.0)
    L.sk_save(str(tmp_path / "s"), m, Lb, np.zeros(4 * 4 * 3), 0, 1.0, 10.0)
    vals = [ln.split("\t")[1] for ln in (tmp_path / "s").read_text().splitlines()]
    assert all("nan" in v for v in vals)


def test_noise_stream_matches_oracle():
    sigma = 7.845
    r = L.Rng(123456789, sigma)
    a = r.normals(1001)
    b = r.normals(999)
    ref = O.normals(123456789, sigma, 2000)
    assert np.array_equal(np.concatenate([a, b]), ref)


def test_uniforms_follow_the_normals_on_one_engine():
    from tests import np_ref

    r = L.Rng(123456789, 2.0)
    n = r.normals(5)  # odd count: one cached polar value stays in the distribution
    u = r.uniforms(4)
    eng = np_ref.MT19937_64(123456789)
    nd = np_ref.PolarNormal(0.0, 2.0)
    assert np.array_equal(n, [nd(eng) for _ in range(5)])
    assert np.array_equal(u, [np_ref.PolarNormal.canonical(eng) for _ in range(4)])


@pytest.mark.parametrize("Lb,m,N,C", [((4.4,) * 3, (64,) * 3, 90, 50.0), ((4.38,) * 3, (40,) * 3, 90, 100.0)])
def test_z_inf_bit_exact(Lb, m, N, C):
    assert L.z_inf(Lb, m, N, C * C) == O.z_inf(Lb, m, N, C * C)


def test_create_rejects_bad_descriptors():
    with pytest.raises(L.LftsError):
        L.Context(1, 0, (4, 4, 4), (1.0, 1.0, 1.0))
    with pytest.raises(L.LftsError):
        L.Context(10, 11, (4, 4, 4), (1.0, 1.0, 1.0))
    with pytest.raises(L.LftsError):
        L.Context(10, 5, (4, 4, 4), (1.0, -1.0, 1.0))
    with pytest.raises(L.LftsError):
        L.Context(10, 5, (4, 4, 4), (1.0, 1.0, 1.0), nhMax=0)
    for m in ((2053, 4, 4), (4, 11 * 256, 4), (4, 4, 0)):  # a prime above 7 on a long axis, an empty axis
        with pytest.raises(L.LftsError) as e:
            L.Context(10, 5, m, (1.0, 1.0, 1.0))
        assert e.value.code == L.LFTS_EINVAL


@pytest.mark.skipif(has_gpu(), reason="checks the no-device error path")
def test_create_without_gpu_fails_loudly():
    with pytest.raises(L.LftsError) as e:
        L.Context(10, 5, (4, 4, 4), (1.0, 1.0, 1.0))
    assert e.value.code == L.LFTS_EHIP


@pytest.mark.skipif(has_gpu(), reason="checks the no-device error path")
def test_create_accepts_any_prime_factor():
    """Mesh sizes with prime factors above 7 (11, 13 and larger) on axes of up to 2047
    points pass the size check (the generic line transform has direct-DFT stages of any
    prime): without a device they fail only there."""
    for m in ((22, 4, 4), (4, 13, 4), (11, 26, 143), (17, 4, 4), (4, 38, 23), (2047, 1, 2)):
        with pytest.raises(L.LftsError) as e:
            L.Context(10, 5, m, (1.0, 1.0, 1.0))
        assert e.value.code == L.LFTS_EHIP


def test_driver_missing_input(tmp_path):
    r = subprocess.run([L.DRIVER_PATH, str(tmp_path / "nope")], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 1
    assert "ERROR => Cannot open one of the input files." in r.stdout


def test_driver_names_an_unsupported_mesh(tmp_path):
    """A mesh size with a prime factor above 7 on an axis of more than 2047 points stops
    the driver with the reason (the size check comes before any device call)."""
    f = tmp_path / "in"
    f.write_text("10 5 14.0 50 0.5 0\n2053 2 2 3.0 3.0 3.0\n2 2 1 1 0\n")
    r = subprocess.run([L.DRIVER_PATH, str(f)], capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 1
    assert "a mesh size with a prime factor above 7 may have at most 2047 points" in r.stdout


def test_driver_usage():
    r = subprocess.run([L.DRIVER_PATH], capture_output=True, text=True)
    assert r.returncode == 1 and "Please supply an input file name" in r.stdout


def test_driver_rejects_zero_sample_period(tmp_path):
    f = tmp_path / "in"
    f.write_text("10 5 12.0 50 0.5 0\n4 4 4 2.0 2.0 2.0\n2 2 0 1 0\n")
    r = subprocess.run([L.DRIVER_PATH, str(f)], capture_output=True, text=True, cwd=tmp_path)
    assert r.returncode == 1 and "ERROR" in r.stdout


def test_rng_state_round_trip_continues_the_stream():
    # an odd number of draws leaves the polar method's second normal pending
    for n0 in (0, 1, 7, 1000):
        r = L.Rng(123456789, 2.5)
        r.normals(n0)
        state = r.save_state()
        a = r.normals(101)
        u = r.uniforms(5)
        s = L.Rng(1, 1.0)
        s.load_state(state)
        assert np.array_equal(s.normals(101), a)
        assert np.array_equal(s.uniforms(5), u)
    with pytest.raises(L.LftsError):
        L.Rng(1, 1.0).load_state(b"not a state")


# ---- the host LU of the Anderson system (anderson::LUdecomp, src/anderson.cc:106-127) ----

@pytest.mark.parametrize("n", [1, 2, 3, 7, 10, 15])
def test_lu_solve_matches_lapack_and_the_oracle(n):
    import scipy.linalg

    rng = np.random.default_rng(20 + n)
    A = rng.standard_normal((n, n)) + 0.5 * n * np.eye(n)
    y = rng.standard_normal(n)
    x = L.lu_solve(A, y)
    assert np.allclose(x, scipy.linalg.solve(A, y), rtol=1e-12, atol=1e-13)
    assert np.array_equal(x, O.lu_solve(A, y))  # same pivots, same operation order


def test_lu_solve_pivots():
    A = np.array([[0.0, 2.0, 1.0], [1.0, 1.0, 0.0], [3.0, 0.0, 1.0]])
    y = np.array([1.0, 2.0, 3.0])
    assert np.allclose(A @ L.lu_solve(A, y), y, rtol=0, atol=1e-14)


@pytest.mark.parametrize("A", [np.zeros((3, 3)), np.array([[1.0, 2.0], [2.0, 4.0]]),
                               np.array([[2.0, 1.0, 1.0], [4.0, 2.0, 2.0], [1.0, 0.0, 3.0]])])
def test_lu_solve_reports_a_singular_system(A):
    """GSL's LU_solve reports a matrix with a zero pivot as singular (and the reference run
    aborts); the library returns LFTS_ESING instead of dividing by zero."""
    with pytest.raises(L.LftsError) as e:
        L.lu_solve(A, np.ones(A.shape[0]))
    assert e.value.code == L.LFTS_ESING


def test_lu_solve_rejects_bad_sizes():
    x = np.zeros(16)
    assert L.lib().lfts_lu_solve(16, L._d(np.eye(16).reshape(-1)), L._d(np.ones(16)), L._d(x)) == L.LFTS_EINVAL
    assert L.lib().lfts_lu_solve(0, L._d(np.eye(1).reshape(-1)), L._d(np.ones(1)), L._d(x)) == L.LFTS_EINVAL


def test_slab_shapes_text_before_any_tuning():
    """lfts_slab_shapes reports the factorisations chosen so far: none before a GPU ran."""
    import ctypes

    n = ctypes.c_longlong(-1)
    assert L.lib().lfts_slab_shapes(None, 0, ctypes.byref(n)) == L.LFTS_OK
    assert n.value >= 0
    assert L.lib().lfts_slab_shapes(None, 0, None) == L.LFTS_EINVAL
    if not has_gpu():
        assert L.slab_shapes() == ""
