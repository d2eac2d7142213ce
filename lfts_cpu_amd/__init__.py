"""lfts_cpu_amd — MI355X (gfx950) L-FTS Langevin step behind a C ABI.

This module is a thin ctypes binding of ``lib/liblfts_gpu.so`` (declared in
``include/lfts_gpu.h``). Every compute call runs the HIP kernels; there is no CPU
fallback: if the library is missing or no GPU is present, calls raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "liblfts_gpu.so")
DRIVER_PATH = os.path.join(PKG, "bin", "lfts-gpu")
VTK_PATH = os.path.join(PKG, "bin", "lfts-vtk")
BATCH_DRIVER_PATH = os.path.join(PKG, "bin", "lfts-gpu-batch")

LFTS_OK, LFTS_EINVAL, LFTS_EHIP, LFTS_ESTATE, LFTS_ENUM, LFTS_ESING = range(6)


class LftsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"lfts error {code}: {msg}")
        self.code = code


class Desc(ctypes.Structure):
    _fields_ = [
        ("N", ctypes.c_int),
        ("NA", ctypes.c_int),
        ("m", ctypes.c_int * 3),
        ("L", ctypes.c_double * 3),
        ("nhMax", ctypes.c_int),
    ]


class NkOpts(ctypes.Structure):
    _fields_ = [
        ("maxNewton", ctypes.c_int),
        ("tol", ctypes.c_double),
        ("restart", ctypes.c_int),
        ("maxInner", ctypes.c_int),
        ("eta", ctypes.c_double),
        ("precondition", ctypes.c_int),
    ]


class NkInfo(ctypes.Structure):
    _fields_ = [
        ("newton", ctypes.c_int),
        ("krylov", ctypes.c_int),
        ("prop_evals", ctypes.c_int),
        ("err", ctypes.c_double),
    ]


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_llp = ctypes.POINTER(ctypes.c_longlong)

# name -> (restype, argtypes); mirrors include/lfts_gpu.h
SIGNATURES = {
    "lfts_create": (ctypes.c_int, [ctypes.POINTER(Desc), ctypes.POINTER(_vp)]),
    "lfts_destroy": (ctypes.c_int, [_vp]),
    "lfts_last_error": (ctypes.c_char_p, [_vp]),
    "lfts_get_stream": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "lfts_set_fields": (ctypes.c_int, [_vp, _dp, ctypes.c_int]),
    "lfts_get_fields": (ctypes.c_int, [_vp, _dp, ctypes.c_int]),
    "lfts_fields_dev": (ctypes.c_int, [_vp, ctypes.POINTER(_dp)]),
    "lfts_calc_concs": (ctypes.c_int, [_vp, _dp]),
    "lfts_anderson_mix": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, _ip, _dp]),
    "lfts_langevin_init": (ctypes.c_int, [_vp, _dp]),
    "lfts_langevin_step": (ctypes.c_int, [_vp, _dp, ctypes.c_double, ctypes.c_double]),
    "lfts_langevin_step_dev": (ctypes.c_int, [_vp, _vp, ctypes.c_double, ctypes.c_double]),
    "lfts_langevin_full_step_dev": (ctypes.c_int, [_vp, _vp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                                   ctypes.c_double, _dp, _ip]),
    "lfts_sk_sample": (ctypes.c_int, [_vp]),
    "lfts_sk_get": (ctypes.c_int, [_vp, _dp, _ip]),
    "lfts_sk_reset": (ctypes.c_int, [_vp]),
    "lfts_sk_bins": (ctypes.c_int, [_ip, _dp, _dp, _ip, _ip]),
    "lfts_lu_solve": (ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp]),
    "lfts_slab_shapes": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_longlong, ctypes.POINTER(ctypes.c_longlong)]),
    "lfts_step_fwd_noalloc": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int]),
    "lfts_sk_save": (ctypes.c_int, [ctypes.c_char_p, _ip, _dp, _dp, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_int]),
    "lfts_hamiltonian": (ctypes.c_int, [_vp, ctypes.c_double, _dp]),
    "lfts_get_pair": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "lfts_step_fwd": (ctypes.c_int, [_vp, _dp, _dp, _dp, ctypes.c_int]),
    "lfts_counters": (ctypes.c_int, [_vp, _llp, _llp]),
    "lfts_profile": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lfts_profile_read": (ctypes.c_int, [_vp, _dp, _llp, _dp, _llp]),
    "lfts_jvp": (ctypes.c_int, [_vp, _dp, _dp, _dp]),
    "lfts_vjp": (ctypes.c_int, [_vp, _dp, _dp]),
    "lfts_jvp_dev": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "lfts_vjp_dev": (ctypes.c_int, [_vp, _vp, _vp]),
    "lfts_box_derivative": (ctypes.c_int, [_vp, _dp, _dp]),
    "lfts_saddle_nk": (ctypes.c_int, [_vp, ctypes.POINTER(NkOpts), ctypes.POINTER(NkInfo)]),
    "lfts_langevin_full_step_nk_dev": (ctypes.c_int, [_vp, _vp, ctypes.c_double, ctypes.c_double,
                                                      ctypes.POINTER(NkOpts), _dp, ctypes.POINTER(NkInfo)]),
    "lfts_set_pair_storage": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lfts_pair_storage": (ctypes.c_int, [_vp, _ip, _ip]),
    "lfts_batch_create": (ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int, ctypes.POINTER(_vp)]),
    "lfts_batch_destroy": (ctypes.c_int, [_vp]),
    "lfts_batch_last_error": (ctypes.c_char_p, [_vp]),
    "lfts_batch_get_stream": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "lfts_batch_calc_concs": (ctypes.c_int, [_vp, _dp]),
    "lfts_batch_anderson_mix": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, _ip, _dp]),
    "lfts_batch_langevin_step_dev": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), _dp, _dp]),
    "lfts_batch_full_step_dev": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), _dp, _dp, ctypes.c_int, ctypes.c_double,
                                                _dp, _ip]),
    "lfts_batch_tuning": (ctypes.c_char_p, [_vp]),
    "lfts_tuning": (ctypes.c_char_p, [_vp]),
    "lfts_batch_full_step": (ctypes.c_int, [_vp, ctypes.POINTER(_dp), _dp, _dp, ctypes.c_int, ctypes.c_double,
                                            _dp, _ip]),
    "lfts_batch_profile": (ctypes.c_int, [_vp, ctypes.c_int]),
    "lfts_batch_profile_read": (ctypes.c_int, [_vp, _dp, _llp, _dp, _llp]),
    "lfts_rng_create": (ctypes.c_int, [ctypes.c_ulonglong, ctypes.c_double, ctypes.POINTER(_vp)]),
    "lfts_rng_normals": (ctypes.c_int, [_vp, _dp, ctypes.c_longlong]),
    "lfts_rng_uniforms": (ctypes.c_int, [_vp, _dp, ctypes.c_longlong]),
    "lfts_rng_destroy": (ctypes.c_int, [_vp]),
    "lfts_rng_save": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_longlong, _llp]),
    "lfts_rng_load": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_longlong]),
    "lfts_sk_set": (ctypes.c_int, [_vp, _dp, ctypes.c_int]),
    "lfts_langevin_get_prev": (ctypes.c_int, [_vp, _dp]),
    "lfts_z_inf": (ctypes.c_double, [_dp, _ip, ctypes.c_int, ctypes.c_double, ctypes.c_int]),
}

_lib = None


def lib():
    """Load the gfx950 library; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m lfts_cpu_amd.build` first")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a: np.ndarray):
    if a.dtype != np.float64 or not a.flags.c_contiguous:
        raise TypeError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    if a.dtype != np.int32 or not a.flags.c_contiguous:
        raise TypeError("expected a C-contiguous int32 array")
    return a.ctypes.data_as(_ip)


class Context:
    """One system on the GPU: the device state behind diblock/anderson/langevin/strFunc."""

    def __init__(self, N: int, NA: int, m, L, nhMax: int = 10):
        d = Desc()
        d.N, d.NA, d.nhMax = int(N), int(NA), int(nhMax)
        for k in range(3):
            d.m[k] = int(m[k])
            d.L[k] = float(L[k])
        self.N, self.NA = int(N), int(NA)
        self.m = tuple(int(x) for x in m)
        self.L = tuple(float(x) for x in L)
        self.M = int(np.prod(self.m))
        self.Mk = self.m[0] * self.m[1] * (self.m[2] // 2 + 1)
        h = _vp()
        rc = lib().lfts_create(ctypes.byref(d), ctypes.byref(h))
        if rc != LFTS_OK:
            raise LftsError(rc, "lfts_create failed (invalid descriptor or no GPU)")
        self.h = h

    def close(self):
        """Release the context; refused (LftsError, LFTS_ESTATE) while a replica set in use
        holds it."""
        if getattr(self, "h", None):
            rc = lib().lfts_destroy(self.h)
            if rc != LFTS_OK:
                raise LftsError(rc, lib().lfts_last_error(self.h).decode())
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc: int):
        if rc != LFTS_OK:
            raise LftsError(rc, lib().lfts_last_error(self.h).decode())

    # fields
    def set_fields(self, w: np.ndarray):
        w = np.ascontiguousarray(w, dtype=np.float64)
        nf = w.size // self.M
        self._chk(lib().lfts_set_fields(self.h, _d(w), nf))

    def get_fields(self, nfields: int = 4) -> np.ndarray:
        out = np.zeros(nfields * self.M)
        self._chk(lib().lfts_get_fields(self.h, _d(out), nfields))
        return out

    def fields_dev(self) -> int:
        p = _dp()
        self._chk(lib().lfts_fields_dev(self.h, ctypes.byref(p)))
        return ctypes.cast(p, ctypes.c_void_p).value

    def stream(self) -> int:
        p = _vp()
        self._chk(lib().lfts_get_stream(self.h, ctypes.byref(p)))
        return p.value or 0

    # physics
    def calc_concs(self) -> float:
        lnQ = ctypes.c_double()
        self._chk(lib().lfts_calc_concs(self.h, ctypes.byref(lnQ)))
        return lnQ.value

    def anderson_mix(self, maxIter: int = 200, tol: float = 1e-4):
        k = ctypes.c_int()
        err = ctypes.c_double()
        self._chk(lib().lfts_anderson_mix(self.h, int(maxIter), float(tol), ctypes.byref(k), ctypes.byref(err)))
        return k.value, err.value

    def langevin_init(self, eta_prev: np.ndarray):
        self._chk(lib().lfts_langevin_init(self.h, _d(np.ascontiguousarray(eta_prev))))

    def langevin_step(self, eta_new: np.ndarray, XbN: float, dt: float):
        self._chk(lib().lfts_langevin_step(self.h, _d(np.ascontiguousarray(eta_new)), float(XbN), float(dt)))

    def langevin_step_dev(self, eta_dev_ptr: int, XbN: float, dt: float):
        self._chk(lib().lfts_langevin_step_dev(self.h, ctypes.c_void_p(eta_dev_ptr), float(XbN), float(dt)))

    def full_step_dev(self, eta_dev_ptr: int, XbN: float, dt: float, maxIter: int = 200, tol: float = 1e-4):
        lnQ = ctypes.c_double()
        k = ctypes.c_int()
        self._chk(lib().lfts_langevin_full_step_dev(self.h, ctypes.c_void_p(eta_dev_ptr), float(XbN), float(dt),
                                                    int(maxIter), float(tol), ctypes.byref(lnQ), ctypes.byref(k)))
        return lnQ.value, k.value

    def sk_sample(self):
        self._chk(lib().lfts_sk_sample(self.h))

    def sk_get(self):
        S = np.zeros(self.Mk)
        ns = ctypes.c_int()
        self._chk(lib().lfts_sk_get(self.h, _d(S), ctypes.byref(ns)))
        return S, ns.value

    def sk_reset(self):
        self._chk(lib().lfts_sk_reset(self.h))

    def sk_set(self, S: np.ndarray, nsamples: int):
        self._chk(lib().lfts_sk_set(self.h, _d(np.ascontiguousarray(S, dtype=np.float64)), int(nsamples)))

    def langevin_get_prev(self) -> np.ndarray:
        out = np.zeros(self.M)
        self._chk(lib().lfts_langevin_get_prev(self.h, _d(out)))
        return out

    def hamiltonian(self, XbN: float) -> float:
        H = ctypes.c_double()
        self._chk(lib().lfts_hamiltonian(self.h, float(XbN), ctypes.byref(H)))
        return H.value

    def pair(self, i: int) -> np.ndarray:
        out = np.zeros(2 * self.M)
        self._chk(lib().lfts_get_pair(self.h, int(i), _d(out)))
        return out

    def step_fwd(self, q_in: np.ndarray, hh: np.ndarray, i: int) -> np.ndarray:
        out = np.zeros(2 * self.M)
        self._chk(lib().lfts_step_fwd(self.h, _d(np.ascontiguousarray(q_in)), _d(out),
                                      _d(np.ascontiguousarray(hh)), int(i)))
        return out

    # derivatives
    def jvp(self, dWm=None, dwp=None) -> np.ndarray:
        """[dphi- | dphi+] along (dW-, dw+) at the current fields."""
        out = np.zeros(2 * self.M)
        a = None if dWm is None else _d(np.ascontiguousarray(dWm, dtype=np.float64))
        b = None if dwp is None else _d(np.ascontiguousarray(dwp, dtype=np.float64))
        self._chk(lib().lfts_jvp(self.h, a, b, _d(out)))
        return out

    def vjp(self, ubar: np.ndarray) -> np.ndarray:
        """[d/dW- | d/dw+] of <u-, phi-> + <u+, phi+> at the current fields."""
        out = np.zeros(2 * self.M)
        self._chk(lib().lfts_vjp(self.h, _d(np.ascontiguousarray(ubar, dtype=np.float64)), _d(out)))
        return out

    def jvp_dev(self, dWm_ptr, dwp_ptr, dphi_ptr):
        self._chk(lib().lfts_jvp_dev(self.h, ctypes.c_void_p(dWm_ptr or None), ctypes.c_void_p(dwp_ptr or None),
                                     ctypes.c_void_p(dphi_ptr)))

    def vjp_dev(self, ubar_ptr, wbar_ptr):
        self._chk(lib().lfts_vjp_dev(self.h, ctypes.c_void_p(ubar_ptr), ctypes.c_void_p(wbar_ptr)))

    def box_derivative(self):
        """(dH/dL as a length-3 array, lnQ) at fixed fields."""
        out = np.zeros(3)
        lnQ = ctypes.c_double()
        self._chk(lib().lfts_box_derivative(self.h, _d(out), ctypes.byref(lnQ)))
        return out, lnQ.value

    def saddle_nk(self, tol: float = 1e-8, maxNewton: int = 30, restart: int = 20, maxInner: int = 200,
                  eta: float = 0.1, precondition: bool = True) -> dict:
        o = NkOpts(int(maxNewton), float(tol), int(restart), int(maxInner), float(eta), int(bool(precondition)))
        info = NkInfo()
        self._chk(lib().lfts_saddle_nk(self.h, ctypes.byref(o), ctypes.byref(info)))
        return {"newton": info.newton, "krylov": info.krylov, "prop_evals": info.prop_evals, "err": info.err}

    def full_step_nk_dev(self, eta_dev_ptr: int, XbN: float, dt: float, tol: float = 1e-4, maxNewton: int = 30,
                         restart: int = 20, maxInner: int = 200, eta: float = 0.1, precondition: bool = True):
        o = NkOpts(int(maxNewton), float(tol), int(restart), int(maxInner), float(eta), int(bool(precondition)))
        info = NkInfo()
        lnQ = ctypes.c_double()
        self._chk(lib().lfts_langevin_full_step_nk_dev(self.h, ctypes.c_void_p(eta_dev_ptr), float(XbN), float(dt),
                                                       ctypes.byref(o), ctypes.byref(lnQ), ctypes.byref(info)))
        return lnQ.value, {"newton": info.newton, "krylov": info.krylov, "prop_evals": info.prop_evals,
                           "err": info.err}

    def profile(self, enable: bool = True, every: int = 1):
        """Live event timing of every `every`-th propagator evaluation (off with enable=False)."""
        self._chk(lib().lfts_profile(self.h, int(every) if enable else 0))

    def profile_read(self):
        """(ms in contour-step chains, contour steps, ms in whole evaluations, evaluations)"""
        a, c = ctypes.c_double(), ctypes.c_double()
        b, d = ctypes.c_longlong(), ctypes.c_longlong()
        self._chk(lib().lfts_profile_read(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d)))
        return a.value, b.value, c.value, d.value

    def tuning(self) -> str:
        return lib().lfts_tuning(self.h).decode()

    def set_pair_storage(self, stride: int):
        self._chk(lib().lfts_set_pair_storage(self.h, int(stride)))

    def pair_storage(self):
        """(stride, pairs kept)"""
        a, b = ctypes.c_int(), ctypes.c_int()
        self._chk(lib().lfts_pair_storage(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def counters(self):
        a = ctypes.c_longlong()
        b = ctypes.c_longlong()
        self._chk(lib().lfts_counters(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value


class Batch:
    """A replica set: many contexts advanced together (one launch per contour step per
    mesh size). Keep the member contexts alive while the set is in use."""

    def __init__(self, ctxs):
        self.ctxs = list(ctxs)
        self.n = len(self.ctxs)
        arr = (_vp * self.n)(*[c.h for c in self.ctxs])
        h = _vp()
        rc = lib().lfts_batch_create(arr, self.n, ctypes.byref(h))
        if rc != LFTS_OK:
            raise LftsError(rc, "lfts_batch_create failed")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().lfts_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc: int, allow_enum: bool = False):
        if rc != LFTS_OK and not (allow_enum and rc == LFTS_ENUM):
            raise LftsError(rc, lib().lfts_batch_last_error(self.h).decode())

    def stream(self) -> int:
        p = _vp()
        self._chk(lib().lfts_batch_get_stream(self.h, ctypes.byref(p)))
        return p.value or 0

    def calc_concs(self) -> np.ndarray:
        out = np.zeros(self.n)
        self._chk(lib().lfts_batch_calc_concs(self.h, _d(out)))
        return out

    def anderson_mix(self, maxIter: int = 200, tol: float = 1e-4):
        k = np.zeros(self.n, dtype=np.int32)
        err = np.zeros(self.n)
        self._chk(lib().lfts_batch_anderson_mix(self.h, int(maxIter), float(tol), _i(k), _d(err)))
        return k, err

    def _ptrs(self, eta_ptrs):
        return (_vp * self.n)(*[ctypes.c_void_p(int(p)) for p in eta_ptrs])

    def langevin_step_dev(self, eta_ptrs, XbN, dt):
        XbN = np.ascontiguousarray(np.broadcast_to(np.asarray(XbN, dtype=np.float64), (self.n,)))
        dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (self.n,)))
        self._chk(lib().lfts_batch_langevin_step_dev(self.h, self._ptrs(eta_ptrs), _d(XbN), _d(dt)))

    def full_step_dev(self, eta_ptrs, XbN, dt, maxIter: int = 200, tol: float = 1e-4):
        XbN = np.ascontiguousarray(np.broadcast_to(np.asarray(XbN, dtype=np.float64), (self.n,)))
        dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (self.n,)))
        lnq = np.zeros(self.n)
        k = np.zeros(self.n, dtype=np.int32)
        self._chk(lib().lfts_batch_full_step_dev(self.h, self._ptrs(eta_ptrs), _d(XbN), _d(dt), int(maxIter),
                                                 float(tol), _d(lnq), _i(k)))
        return lnq, k

    def tuning(self) -> str:
        return lib().lfts_batch_tuning(self.h).decode()

    def profile(self, enable: bool = True, every: int = 1):
        """Live event timing of every `every`-th evaluation of the set (off with enable=False)."""
        self._chk(lib().lfts_batch_profile(self.h, int(every) if enable else 0))

    def profile_read(self):
        a, c = ctypes.c_double(), ctypes.c_double()
        b, d = ctypes.c_longlong(), ctypes.c_longlong()
        self._chk(lib().lfts_batch_profile_read(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                                ctypes.byref(d)))
        return a.value, b.value, c.value, d.value


class Rng:
    """The reference's host noise stream (std::mt19937_64 + std::normal_distribution)."""

    def __init__(self, seed: int = 123456789, sigma: float = 1.0):
        h = _vp()
        rc = lib().lfts_rng_create(ctypes.c_ulonglong(seed), float(sigma), ctypes.byref(h))
        if rc != LFTS_OK:
            raise LftsError(rc, "lfts_rng_create failed")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().lfts_rng_destroy(self.h)
            self.h = None

    def normals(self, n: int) -> np.ndarray:
        out = np.zeros(int(n))
        lib().lfts_rng_normals(self.h, _d(out), int(n))
        return out

    def save_state(self) -> bytes:
        n = ctypes.c_longlong()
        lib().lfts_rng_save(self.h, None, 0, ctypes.byref(n))
        buf = ctypes.create_string_buffer(n.value)
        rc = lib().lfts_rng_save(self.h, buf, n.value, ctypes.byref(n))
        if rc != LFTS_OK:
            raise LftsError(rc, "lfts_rng_save failed")
        return buf.raw[:n.value]

    def load_state(self, state: bytes):
        rc = lib().lfts_rng_load(self.h, state, len(state))
        if rc != LFTS_OK:
            raise LftsError(rc, "lfts_rng_load failed")

    def uniforms(self, n: int) -> np.ndarray:
        out = np.zeros(int(n))
        lib().lfts_rng_uniforms(self.h, _d(out), int(n))
        return out


def slab_shapes() -> str:
    """The contour-step line factorisations chosen in this process so far, e.g. "64:8x8"."""
    n = ctypes.c_longlong()
    lib().lfts_slab_shapes(None, 0, ctypes.byref(n))
    buf = ctypes.create_string_buffer(max(n.value, 1))
    lib().lfts_slab_shapes(buf, n.value, ctypes.byref(n))
    return buf.raw[:n.value].decode()


def lu_solve(A: np.ndarray, y: np.ndarray) -> np.ndarray:
    """anderson::LUdecomp (src/anderson.cc:106-127) on the host: solves A x = y; raises
    LftsError (LFTS_ESING) when U has a zero pivot."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    x = np.zeros(y.size)
    rc = lib().lfts_lu_solve(int(y.size), _d(A), _d(y), _d(x))
    if rc != LFTS_OK:
        raise LftsError(rc, "singular Anderson system" if rc == LFTS_ESING else "bad argument")
    return x


def sk_bins(m, L):
    mv = np.array(m, dtype=np.int32)
    Lv = np.array(L, dtype=np.float64)
    Mk = int(mv[0] * mv[1] * (mv[2] // 2 + 1))
    K = np.zeros(Mk)
    wt = np.zeros(Mk, dtype=np.int32)
    P = np.zeros(Mk, dtype=np.int32)
    rc = lib().lfts_sk_bins(_i(mv), _d(Lv), _d(K), _i(wt), _i(P))
    if rc != LFTS_OK:
        raise LftsError(rc, "lfts_sk_bins failed")
    return K, wt, P


def sk_save(fname: str, m, L, S: np.ndarray, nsamples: int, CV: float, chi_b: float, dK: float = 1e-5, dp: int = 8):
    mv = np.array(m, dtype=np.int32)
    Lv = np.array(L, dtype=np.float64)
    rc = lib().lfts_sk_save(fname.encode(), _i(mv), _d(Lv), _d(np.ascontiguousarray(S)), int(nsamples),
                            float(CV), float(chi_b), float(dK), int(dp))
    if rc != LFTS_OK:
        raise LftsError(rc, "lfts_sk_save failed")


def z_inf(L, m, N: int, Nbar: float, tmax: int = 100) -> float:
    Lv = np.array(L, dtype=np.float64)
    mv = np.array(m, dtype=np.int32)
    return lib().lfts_z_inf(_d(Lv), _i(mv), int(N), float(Nbar), int(tmax))
