"""Independent numpy restatement of the reference arithmetic, for tiny meshes.

It is coded separately from the C++ oracle and uses numpy's real transforms
(rfftn/irfftn: unnormalised forward, like FFTW's r2c; irfftn divides by M, so it is
multiplied back to match FFTW's unnormalised c2r). It therefore checks the oracle's
own FFT, its two-for-one packing of the batched real transforms and its half-spectrum
layout, besides the chain recursion.
"""
from __future__ import annotations

import math

import numpy as np


def freqs(m: int) -> np.ndarray:
    """Frequency of index K on an axis of m points: K if K <= m/2 else K - m
    (k in [-(m-1)/2, m/2], src/step.cc:73-83)."""
    K = np.arange(m)
    return np.where(K <= m // 2, K, K - m)


def g_half(N: int, m, L) -> np.ndarray:
    """g(k) = exp(-K^2/(6N))/M on the r2c half spectrum, shape (m0, m1, m2//2+1)."""
    M = m[0] * m[1] * m[2]
    k0 = freqs(m[0]).astype(np.float64)[:, None, None]
    k1 = freqs(m[1]).astype(np.float64)[None, :, None]
    k2 = np.arange(m[2] // 2 + 1, dtype=np.float64)[None, None, :]
    ksq = (k0 * k0) / (L[0] * L[0]) + (k1 * k1) / (L[1] * L[1]) + (k2 * k2) / (L[2] * L[2])
    K = 2 * np.pi * np.power(ksq, 0.5)
    return np.exp(-K * K / (6.0 * N)) / M


def conv(q: np.ndarray, g: np.ndarray, m) -> np.ndarray:
    M = m[0] * m[1] * m[2]
    return np.fft.irfftn(np.fft.rfftn(q.reshape(m)) * g, s=m, axes=(0, 1, 2)).reshape(-1) * M


def calc_concs(w: np.ndarray, N: int, NA: int, m, L):
    """Returns (lnQ, phi_minus, phi_plus, q, qd) with q[i], qd[i] for beads 1..N."""
    M = m[0] * m[1] * m[2]
    NB = N - NA
    g = g_half(N, m, L)
    wm, wp = w[:M], w[M:2 * M]
    hA = np.exp(-(wm + wp) / N)
    hB = np.exp(-(-wm + wp) / N)
    q = np.zeros((N + 2, M))
    qd = np.zeros((N + 2, M))
    q[1] = hA
    qd[N] = hB
    for i in range(1, N):
        q[i + 1] = conv(q[i], g, m) * (hA if i < NA else hB)
        qd[N - i] = conv(qd[N + 1 - i], g, m) * (hB if i < NB else hA)
    Q = q[N].sum() / M
    phA = np.zeros(M)
    phB = np.zeros(M)
    for i in range(1, NA + 1):
        phA += q[i] * qd[i]
    for i in range(NA + 1, N + 1):
        phB += q[i] * qd[i]
    phA = phA / (N * Q * hA)
    phB = phB / (N * Q * hB)
    return math.log(Q), phA - phB, phA + phB, q, qd


def lu_solve(A: np.ndarray, b: np.ndarray) -> np.ndarray:
    import scipy.linalg

    return scipy.linalg.lu_solve(scipy.linalg.lu_factor(A), b)


def anderson_mix(w: np.ndarray, N: int, NA: int, m, L, maxIter=200, tol=1e-4, nhMax=10):
    """src/anderson.cc:24-93, with numpy arrays for the histories."""
    M = m[0] * m[1] * m[2]
    Dh = np.zeros((nhMax + 1, M))
    wh = np.zeros((nhMax + 1, M))
    DD = np.zeros((nhMax + 1, nhMax + 1))
    err = 1.0
    k = 1
    while k < maxIter and err > tol:
        _, phm, php, _, _ = calc_concs(w, N, NA, m, L)
        w[2 * M:3 * M] = phm
        w[3 * M:] = php
        Dh[0] = php - 1.0
        wh[0] = w[M:2 * M]
        err = math.sqrt(float(np.dot(Dh[0], Dh[0])) / M)
        lam = 1.0 - 0.9 ** k
        nh = k - 1 if k < nhMax + 1 else nhMax
        for mm in range(nh + 1):
            DD[0, mm] = float(np.dot(Dh[0], Dh[mm]))
        if k < 2:
            w[M:2 * M] += lam * Dh[0]
        else:
            U = np.zeros((nh, nh))
            V = np.zeros(nh)
            for a in range(1, nh + 1):
                V[a - 1] = DD[0, 0] - DD[0, a]
                for b in range(1, a + 1):
                    U[a - 1, b - 1] = U[b - 1, a - 1] = DD[0, 0] - DD[0, a] - DD[0, b] + DD[b, a]
            C = lu_solve(U, V)
            w[M:2 * M] += lam * Dh[0]
            for n in range(1, nh + 1):
                w[M:2 * M] += C[n - 1] * ((wh[n] + lam * Dh[n]) - (wh[0] + lam * Dh[0]))
        n = 1 + (k - 1) % nhMax
        Dh[n] = Dh[0]
        wh[n] = wh[0]
        DD[n, n] = DD[0, 0]
        for mm in range(1, n):
            DD[mm, n] = DD[0, mm]
        for mm in range(n + 1, nh + 1):
            DD[n, mm] = DD[0, mm]
        k += 1
    return k, err


# ---- the reference's host noise stream ---------------------------------------------

class MT19937_64:
    """std::mt19937_64 (C++ [rand.eng.mers] with the [rand.predef] parameters)."""

    W, N, Mm, R = 64, 312, 156, 31
    A = 0xB5026F5AA96619E9
    U, D = 29, 0x5555555555555555
    S, B = 17, 0x71D67FFFEDA60000
    T, C = 37, 0xFFF7EEE000000000
    L = 43
    F = 6364136223846793005
    MASK = (1 << 64) - 1

    def __init__(self, seed: int = 5489):
        self.mt = [0] * self.N
        self.mt[0] = seed & self.MASK
        for i in range(1, self.N):
            self.mt[i] = (self.F * (self.mt[i - 1] ^ (self.mt[i - 1] >> (self.W - 2))) + i) & self.MASK
        self.idx = self.N

    def _twist(self):
        upper = (~((1 << self.R) - 1)) & self.MASK
        lower = (1 << self.R) - 1
        for i in range(self.N):
            x = (self.mt[i] & upper) | (self.mt[(i + 1) % self.N] & lower)
            xA = x >> 1
            if x & 1:
                xA ^= self.A
            self.mt[i] = self.mt[(i + self.Mm) % self.N] ^ xA
        self.idx = 0

    def __call__(self) -> int:
        if self.idx >= self.N:
            self._twist()
        y = self.mt[self.idx]
        self.idx += 1
        y ^= (y >> self.U) & self.D
        y ^= (y << self.S) & self.B
        y ^= (y << self.T) & self.C
        y ^= y >> self.L
        return y & self.MASK


class PolarNormal:
    """libstdc++'s std::normal_distribution<double>: Marsaglia's polar method on two
    generate_canonical<double,53> draws (one 64-bit engine call each, divided by 2^64),
    returning y*mult and caching x*mult for the next call."""

    def __init__(self, mean: float = 0.0, stddev: float = 1.0):
        self.mean, self.stddev = mean, stddev
        self.saved = None

    @staticmethod
    def canonical(eng) -> float:
        r = float(eng()) / 18446744073709551616.0
        return r if r < 1.0 else math.nextafter(1.0, 0.0)

    def __call__(self, eng) -> float:
        if self.saved is not None:
            ret, self.saved = self.saved, None
        else:
            while True:
                x = 2.0 * self.canonical(eng) - 1.0
                y = 2.0 * self.canonical(eng) - 1.0
                r2 = x * x + y * y
                if not (r2 > 1.0 or r2 == 0.0):
                    break
            mult = math.sqrt(-2 * math.log(r2) / r2)
            self.saved = x * mult
            ret = y * mult
        return ret * self.stddev + self.mean
