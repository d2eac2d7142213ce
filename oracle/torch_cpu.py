"""ORACLE — test infrastructure only: a second CPU restatement of the reference's Langevin
step, used as the stronger CPU baseline of bench.py (its `cpu_baseline` leg) and checked
against oracle/lfts_oracle.cc by tests/test_torch_cpu_baseline.py. The product package
never imports it.

It follows the reference's arithmetic (src/step.cc:35-89, src/diblock.cc:30-95,
src/anderson.cc:24-127, src/langevin.cc:23-37) with the transforms done by torch's
threaded CPU FFT (batched-2 r2c / c2r over the q and q-dagger halves of a pair, as the
reference's FFTW plans with howmany = 2, src/step.cc:15-16) and the element-wise loops as
threaded torch operations, so it stands in for a threaded FFTW build of lfts-cpu on the
same host cores.
"""
from __future__ import annotations

import math
import os

import numpy as np
import torch


def fft_library() -> str:
    mkl = getattr(torch.backends, "mkl", None)
    return "torch.fft on CPU (" + ("MKL" if mkl is not None and mkl.is_available() else "pocketfft") + \
        f", {torch.get_num_threads()} threads)"


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


class TorchLfts:
    """diblock + anderson + langevin of the reference on host memory (float64)."""

    def __init__(self, N: int, NA: int, m, L, nhMax: int = 10):
        self.N, self.NA, self.NB = int(N), int(NA), int(N) - int(NA)
        self.m = tuple(int(x) for x in m)
        self.L = tuple(float(x) for x in L)
        self.M = int(np.prod(self.m))
        self.nhMax = nhMax
        # g(k) on the r2c half spectrum (src/step.cc:69-89): exp(-K^2/(6N))/M,
        # K = 2 pi |(k0/Lx, k1/Ly, k2/Lz)|, k in the reference's range [-(n-1)/2, n/2]
        k0 = np.array([K if K <= self.m[0] // 2 else K - self.m[0] for K in range(self.m[0])], dtype=np.float64)
        k1 = np.array([K if K <= self.m[1] // 2 else K - self.m[1] for K in range(self.m[1])], dtype=np.float64)
        k2 = np.arange(self.m[2] // 2 + 1, dtype=np.float64)
        ksq = (k0[:, None, None] ** 2 / self.L[0] ** 2 + k1[None, :, None] ** 2 / self.L[1] ** 2
               + k2[None, None, :] ** 2 / self.L[2] ** 2)
        K = 2.0 * math.pi * np.sqrt(ksq)
        self.g = torch.from_numpy(np.exp(-K * K / (6.0 * self.N)) / self.M)
        self.q = torch.zeros((self.N + 1, 2, self.M), dtype=torch.float64)  # pair i = [q_i | q+_{N+1-i}]
        H = nhMax + 1
        self.Dh = torch.zeros((H, self.M), dtype=torch.float64)
        self.wh = torch.zeros((H, self.M), dtype=torch.float64)
        self.DD = np.zeros((H, H))

    def calc_concs(self, w: torch.Tensor) -> float:
        """src/diblock.cc:30-95 on w = [W- | w+ | phi- | phi+] (4M, modified in place)."""
        M, N, NA, NB = self.M, self.N, self.NA, self.NB
        Wm, wp = w[:M], w[M:2 * M]
        hA = torch.exp(-(Wm + wp) / N)
        hB = torch.exp(-(-Wm + wp) / N)
        q = self.q
        q[1, 0] = hA
        q[1, 1] = hB
        shape = (2,) + self.m
        for i in range(1, N):
            z = torch.fft.rfftn(q[i].view(shape), dim=(1, 2, 3))
            z *= self.g
            r = torch.fft.irfftn(z, s=self.m, dim=(1, 2, 3), norm="forward").reshape(2, M)
            torch.mul(r[0], hA if i < NA else hB, out=q[i + 1, 0])
            torch.mul(r[1], hB if i < NB else hA, out=q[i + 1, 1])
        Q = float(q[N, 0].sum()) / M
        a = q[1:N + 1, 0]
        b = torch.flip(q[1:N + 1, 1], dims=(0,))  # q+_i sits in pair N+1-i
        phiA = (a[:NA] * b[:NA]).sum(0) if NA > 0 else torch.zeros(M, dtype=torch.float64)
        phiB = (a[NA:] * b[NA:]).sum(0) if NA < N else torch.zeros(M, dtype=torch.float64)
        pA = phiA / (N * Q * hA)
        pB = phiB / (N * Q * hB)
        w[2 * M:3 * M] = pA - pB
        w[3 * M:4 * M] = pA + pB
        return math.log(Q)

    def anderson_mix(self, maxIter: int, errTol: float, w: torch.Tensor) -> int:
        """src/anderson.cc:24-93; the LU solve by LAPACK (numpy.linalg.solve)."""
        M, Dh, wh, DD = self.M, self.Dh, self.wh, self.DD
        err, k = 1.0, 1
        while k < maxIter and err > errTol:
            self.calc_concs(w)
            Dh[0] = w[3 * M:] - 1.0
            wh[0] = w[M:2 * M]
            err = math.sqrt(float((Dh[0] * Dh[0]).sum()) / M)
            lam = 1.0 - 0.9 ** k
            nh = k - 1 if k < self.nhMax + 1 else self.nhMax
            DD[0, :nh + 1] = (Dh[:nh + 1] @ Dh[0]).numpy()
            if k < 2:
                w[M:2 * M] += lam * Dh[0]
            else:
                U = np.empty((nh, nh))
                V = np.empty(nh)
                for mm in range(1, nh + 1):
                    V[mm - 1] = DD[0, 0] - DD[0, mm]
                    for nn in range(1, mm + 1):
                        U[mm - 1, nn - 1] = U[nn - 1, mm - 1] = DD[0, 0] - DD[0, mm] - DD[0, nn] + DD[nn, mm]
                C = np.linalg.solve(U, V)
                y0 = wh[0] + lam * Dh[0]
                upd = lam * Dh[0]
                for nn in range(1, nh + 1):
                    upd += float(C[nn - 1]) * ((wh[nn] + lam * Dh[nn]) - y0)
                w[M:2 * M] += upd
            n = 1 + (k - 1) % self.nhMax
            Dh[n] = Dh[0]
            wh[n] = wh[0]
            DD[n, n] = DD[0, 0]
            for mm in range(1, n):
                DD[mm, n] = DD[0, mm]
            for mm in range(n + 1, nh + 1):
                DD[n, mm] = DD[0, mm]
            k += 1
        return k

    def langevin_update(self, w: torch.Tensor, eta_prev: torch.Tensor, eta_new: torch.Tensor, XbN: float, dt: float):
        """src/langevin.cc:31."""
        M = self.M
        w[:M] += -(w[2 * M:3 * M] + 2.0 * w[:M] / XbN) * dt + 0.5 * (eta_prev + eta_new)


def threads() -> int:
    return torch.get_num_threads()


def set_threads(n: int | None = None) -> int:
    n = n or int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    torch.set_num_threads(n)
    return torch.get_num_threads()
