"""ORACLE — test infrastructure only.

ctypes binding of the CPU restatement in ``oracle/lfts_oracle.cc``. It is the parity
checker for the HIP path and the timed CPU baseline; it is imported only by ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg. The product package
``lfts_cpu_amd`` never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liblfts_oracle.so")
EXE_PATH = os.path.join(HERE, "build", "lfts-oracle")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


def build(force: bool = False) -> None:
    """Compile the oracle with its Makefile (g++ only)."""
    args = ["make", "-C", HERE]
    if force:
        args.append("-B")
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "orc_fft1d": (None, [ctypes.c_int, _dp, ctypes.c_int]),
            "orc_fft3d": (None, [_ip, _dp, ctypes.c_int]),
            "orc_create": (_vp, [ctypes.c_int, ctypes.c_int, _ip, _dp]),
            "orc_destroy": (None, [_vp]),
            "orc_ghalf": (None, [_vp, _dp]),
            "orc_calc_concs": (ctypes.c_double, [_vp, _dp]),
            "orc_get_pair": (None, [_vp, ctypes.c_int, _dp]),
            "orc_step_fwd": (None, [_vp, _dp, _dp, _dp, ctypes.c_int]),
            "orc_jvp": (None, [_vp, _dp, _dp, _dp, _dp]),
            "orc_vjp": (None, [_vp, _dp, _dp, _dp]),
            "orc_dlnQ_dL": (ctypes.c_double, [_vp, _dp, _dp]),
            "orc_am_create": (_vp, [ctypes.c_int, ctypes.c_int]),
            "orc_am_destroy": (None, [_vp]),
            "orc_am_mix": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_double, _dp]),
            "orc_am_last_err": (ctypes.c_double, [_vp]),
            "orc_am_set_trace": (None, [_vp, _dp, ctypes.c_int]),
            "orc_am_trace_count": (ctypes.c_int, [_vp]),
            "orc_am_set_fault": (None, [_vp, ctypes.c_int]),
            "orc_sk_set_fault": (None, [_vp, ctypes.c_int]),
            "orc_set_order": (None, [ctypes.c_int]),
            "orc_lu_solve": (None, [ctypes.c_int, _dp, _dp, _dp]),
            "orc_langevin_update": (None, [_dp, _dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_double]),
            "orc_normals": (None, [ctypes.c_ulonglong, ctypes.c_double, ctypes.c_int, _dp]),
            "orc_mt64": (None, [ctypes.c_ulonglong, ctypes.c_int, ctypes.POINTER(ctypes.c_ulonglong)]),
            "orc_sk_create": (_vp, [_ip, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
            "orc_sk_destroy": (None, [_vp]),
            "orc_sk_sample": (None, [_vp, _dp]),
            "orc_sk_get": (None, [_vp, _dp, _dp, _ip, _ip, _ip]),
            "orc_sk_save": (None, [_vp, ctypes.c_char_p, ctypes.c_int]),
            "orc_z_inf": (ctypes.c_double, [_dp, _ip, ctypes.c_int, ctypes.c_double, ctypes.c_int]),
            "orc_read_params": (ctypes.c_int, [ctypes.c_char_p, _ip, _dp]),
            "orc_sim_create": (_vp, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]),
            "orc_sim_destroy": (None, [_vp]),
            "orc_sim_get_w": (None, [_vp, _dp]),
            "orc_sim_set_w": (None, [_vp, _dp]),
            "orc_sim_last_nAM": (ctypes.c_int, [_vp]),
            "orc_sim_step": (ctypes.c_double, [_vp, ctypes.c_int]),
            "orc_sim_equilibrate": (None, [_vp]),
            "orc_sim_statistics": (None, [_vp]),
            "orc_sim_getH": (ctypes.c_double, [_vp]),
            "orc_sim_sk_get": (None, [_vp, _dp, _ip]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def fft1d(x: np.ndarray, sign: int = -1) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().orc_fft1d(a.size, a.view(np.float64).ctypes.data_as(_dp), sign)
    return a


def fft3d(x: np.ndarray, sign: int = -1) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    m = np.array(a.shape, dtype=np.int32)
    lib().orc_fft3d(_i(m), a.view(np.float64).ctypes.data_as(_dp), sign)
    return a


class Diblock:
    """diblock + step restatement (src/diblock.cc, src/step.cc)."""

    def __init__(self, N: int, NA: int, m, L):
        self.N, self.NA = int(N), int(NA)
        self.m = np.array(m, dtype=np.int32)
        self.L = np.array(L, dtype=np.float64)
        self.M = int(np.prod(self.m))
        self.Mk = int(self.m[0] * self.m[1] * (self.m[2] // 2 + 1))
        self.h = lib().orc_create(self.N, self.NA, _i(self.m), _d(self.L))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def ghalf(self) -> np.ndarray:
        out = np.zeros(self.Mk)
        lib().orc_ghalf(self.h, _d(out))
        return out

    def calc_concs(self, w: np.ndarray) -> float:
        assert w.shape == (4 * self.M,)
        return lib().orc_calc_concs(self.h, _d(w))

    def pair(self, i: int) -> np.ndarray:
        out = np.zeros(2 * self.M)
        lib().orc_get_pair(self.h, int(i), _d(out))
        return out

    def jvp(self, w: np.ndarray, dwm, dwp) -> np.ndarray:
        """Directional derivative [dphi- | dphi+] of calc_concs at w along (dW-, dw+)."""
        out = np.zeros(2 * self.M)
        w = np.ascontiguousarray(w, dtype=np.float64)
        a = None if dwm is None else _d(np.ascontiguousarray(dwm, dtype=np.float64))
        b = None if dwp is None else _d(np.ascontiguousarray(dwp, dtype=np.float64))
        lib().orc_jvp(self.h, _d(w), a, b, _d(out))
        return out

    def vjp(self, w: np.ndarray, ubar: np.ndarray) -> np.ndarray:
        """Gradient [dL/dW- | dL/dw+] of L = <u-, phi-> + <u+, phi+> at w."""
        out = np.zeros(2 * self.M)
        lib().orc_vjp(self.h, _d(np.ascontiguousarray(w, dtype=np.float64)),
                      _d(np.ascontiguousarray(ubar, dtype=np.float64)), _d(out))
        return out

    def dlnQ_dL(self, w: np.ndarray):
        """(ln Q, d lnQ/dL) at fixed fields."""
        out = np.zeros(3)
        lnQ = lib().orc_dlnQ_dL(self.h, _d(np.ascontiguousarray(w, dtype=np.float64)), _d(out))
        return lnQ, out

    def step_fwd(self, q_in: np.ndarray, h: np.ndarray, i: int) -> np.ndarray:
        out = np.zeros(2 * self.M)
        lib().orc_step_fwd(self.h, _d(np.ascontiguousarray(q_in)), _d(out), _d(np.ascontiguousarray(h)), int(i))
        return out


class Anderson:
    """anderson restatement (src/anderson.cc)."""

    def __init__(self, M: int, nhMax: int = 10):
        self.M, self.nhMax = int(M), int(nhMax)
        self.h = lib().orc_am_create(self.M, self.nhMax)
        self._trace = None

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_am_destroy(self.h)
            self.h = None

    def mix(self, dbc: Diblock, maxIter: int, tol: float, w: np.ndarray) -> int:
        return lib().orc_am_mix(self.h, dbc.h, int(maxIter), float(tol), _d(w))

    def last_err(self) -> float:
        return lib().orc_am_last_err(self.h)

    def set_fault(self, fault: int):
        """Plant a misreading of src/anderson.cc (tests of the pins only; 0 = none)."""
        lib().orc_am_set_fault(self.h, int(fault))

    def record(self, cap: int):
        """Record up to cap iterations of the following mix calls (see trace())."""
        self._trace = np.zeros((cap, 5 + self.nhMax + 3 * self.M))
        lib().orc_am_set_trace(self.h, _d(self._trace), int(cap))

    def trace(self) -> list[dict]:
        """Per iteration: k, nh, lambda, err, the history slot written at its end, the
        coefficients C (slot order), Dh0 = phi+ - 1, wh0 = w+ before and w+ after."""
        n = lib().orc_am_trace_count(self.h)
        M, H = self.M, self.nhMax
        out = []
        for t in self._trace[:n]:
            nh = int(t[1])
            out.append(dict(k=int(t[0]), nh=nh, lam=t[2], err=t[3], slot=int(t[4]), C=t[5:5 + nh].copy(),
                            Dh0=t[5 + H:5 + H + M].copy(), wh0=t[5 + H + M:5 + H + 2 * M].copy(),
                            w_new=t[5 + H + 2 * M:5 + H + 3 * M].copy()))
        return out


def set_order(order: int) -> None:
    """Rounding-order variant of the whole oracle (0 = the reference's order): bit 0 runs
    the 3D transforms along x, y, z; bit 1 sums the reductions from the last point. Used
    only to measure how far two correct evaluations of a run drift apart."""
    lib().orc_set_order(int(order))


def lu_solve(A: np.ndarray, b: np.ndarray) -> np.ndarray:
    n = b.size
    x = np.zeros(n)
    lib().orc_lu_solve(n, _d(np.ascontiguousarray(A, dtype=np.float64)), _d(np.ascontiguousarray(b, dtype=np.float64)), _d(x))
    return x


def langevin_update(w, eta_prev, eta_new, XbN, dt):
    M = eta_prev.size
    lib().orc_langevin_update(_d(w), _d(eta_prev), _d(eta_new), M, float(XbN), float(dt))


def normals(seed: int, sigma: float, n: int) -> np.ndarray:
    out = np.zeros(n)
    lib().orc_normals(seed, sigma, n, _d(out))
    return out


def mt64(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint64)
    lib().orc_mt64(seed, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_ulonglong)))
    return out


class StrFunc:
    """strFunc restatement (src/strFunc.cc)."""

    def __init__(self, m, L, CV: float, chi_b: float, dK: float = 1e-5):
        self.m = np.array(m, dtype=np.int32)
        self.L = np.array(L, dtype=np.float64)
        self.Mk = int(self.m[0] * self.m[1] * (self.m[2] // 2 + 1))
        self.h = lib().orc_sk_create(_i(self.m), _d(self.L), float(CV), float(chi_b), float(dK))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sk_destroy(self.h)
            self.h = None

    def sample(self, wm: np.ndarray):
        lib().orc_sk_sample(self.h, _d(np.ascontiguousarray(wm)))

    def get(self):
        S = np.zeros(self.Mk)
        K = np.zeros(self.Mk)
        wt = np.zeros(self.Mk, dtype=np.int32)
        P = np.zeros(self.Mk, dtype=np.int32)
        ns = np.zeros(1, dtype=np.int32)
        lib().orc_sk_get(self.h, _d(S), _d(K), _i(wt), _i(P), _i(ns))
        return S, K, wt, P, int(ns[0])

    def save(self, fname: str, dp: int = 8):
        lib().orc_sk_save(self.h, fname.encode(), dp)

    def set_fault(self, fault: int):
        """Plant a misreading of src/strFunc.cc's save (tests of the pins only)."""
        lib().orc_sk_set_fault(self.h, int(fault))


def z_inf(L, m, N: int, Nbar: float, tmax: int = 100) -> float:
    Lv = np.array(L, dtype=np.float64)
    mv = np.array(m, dtype=np.int32)
    return lib().orc_z_inf(_d(Lv), _i(mv), int(N), float(Nbar), int(tmax))


PARAM_INTS = ["N", "NA", "NB", "isXeN", "mx", "my", "mz", "equil_its", "sim_its",
              "sample_freq", "save_freq", "loadType", "M", "Mk"]
PARAM_DBLS = ["chi_b", "C", "dt", "Lx", "Ly", "Lz", "V", "sigma", "XeN"]


def read_params(path: str) -> dict:
    iv = np.zeros(14, dtype=np.int32)
    dv = np.zeros(9)
    if lib().orc_read_params(path.encode(), _i(iv), _d(dv)) != 0:
        raise FileNotFoundError(path)
    out = {k: int(v) for k, v in zip(PARAM_INTS, iv)}
    out.update({k: float(v) for k, v in zip(PARAM_DBLS, dv)})
    return out


class Sim:
    """lfts_simulation restatement (src/lfts_simulation.cc)."""

    def __init__(self, input_file: str, outdir: str = "", write_files: bool = False, initial_mix: bool = True):
        self.params = read_params(input_file)
        self.M = self.params["M"]
        flags = int(write_files) | (0 if initial_mix else 2)
        self.h = lib().orc_sim_create(input_file.encode(), outdir.encode(), flags)
        if not self.h:
            raise FileNotFoundError(input_file)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sim_destroy(self.h)
            self.h = None

    def w(self) -> np.ndarray:
        out = np.zeros(4 * self.M)
        lib().orc_sim_get_w(self.h, _d(out))
        return out

    def set_w(self, w: np.ndarray):
        lib().orc_sim_set_w(self.h, _d(np.ascontiguousarray(w)))

    def last_nAM(self) -> int:
        return lib().orc_sim_last_nAM(self.h)

    def step(self, sample: bool = False) -> float:
        return lib().orc_sim_step(self.h, int(sample))

    def equilibrate(self):
        lib().orc_sim_equilibrate(self.h)

    def statistics(self):
        lib().orc_sim_statistics(self.h)

    def getH(self) -> float:
        return lib().orc_sim_getH(self.h)

    def sk(self):
        Mk = self.params["Mk"]
        S = np.zeros(Mk)
        ns = np.zeros(1, dtype=np.int32)
        lib().orc_sim_sk_get(self.h, _d(S), _i(ns))
        return S, int(ns[0])
