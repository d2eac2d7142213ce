"""The torch-FFT CPU restatement (oracle/torch_cpu.py, bench.py's stronger CPU baseline)
computes the reference's step: it agrees with the C++ oracle on one propagator
evaluation, one Anderson solve and one whole Langevin step (within rounding: its
transforms, dot products and LU are different codings)."""
import numpy as np
import pytest
import torch

import oracle as O
from oracle import torch_cpu as T


@pytest.mark.parametrize("m,L,N,NA", [((8, 6, 10), (3.0, 2.5, 4.0), 12, 5), ((16, 16, 16), (3.3, 3.3, 3.3), 20, 20),
                                      ((9, 10, 6), (2.7, 3.1, 2.2), 11, 0)])
def test_calc_concs_matches_oracle(m, L, N, NA):
    M = int(np.prod(m))
    rng = np.random.default_rng(3)
    w = np.zeros(4 * M)
    w[:M] = rng.normal(0, 2.0, M)
    w[M:2 * M] = rng.normal(0, 1.0, M)
    t = T.TorchLfts(N, NA, m, L)
    wt = torch.from_numpy(w.copy())
    lnq = t.calc_concs(wt)
    wo = w.copy()
    lnq_o = O.Diblock(N, NA, m, L).calc_concs(wo)
    assert abs(lnq - lnq_o) <= 1e-12 * max(1.0, abs(lnq_o))
    assert np.abs(wt.numpy()[2 * M:] - wo[2 * M:]).max() <= 1e-12 * np.abs(wo[2 * M:]).max()


def test_langevin_step_matches_oracle():
    m, L, N, NA, XbN, dt = (16, 12, 10), (3.2, 2.8, 2.4), 16, 7, 14.0, 0.5
    M = int(np.prod(m))
    rng = np.random.default_rng(8)
    w = np.zeros(4 * M)
    w[:M] = rng.normal(0, 2.0, M)
    ep, en = rng.normal(0, 1.0, M), rng.normal(0, 1.0, M)
    t = T.TorchLfts(N, NA, m, L)
    wt = torch.from_numpy(w.copy())
    d, am = O.Diblock(N, NA, m, L), O.Anderson(M, 10)
    wo = w.copy()
    assert t.anderson_mix(200, 1e-4, wt) == am.mix(d, 200, 1e-4, wo)
    t.langevin_update(wt, torch.from_numpy(ep), torch.from_numpy(en), XbN, dt)
    O.langevin_update(wo, ep, en, XbN, dt)
    k = t.anderson_mix(200, 1e-4, wt)
    assert k == am.mix(d, 200, 1e-4, wo) and k > 5
    lnq, lnq_o = t.calc_concs(wt), d.calc_concs(wo)
    assert abs(lnq - lnq_o) <= 1e-9
    assert np.abs(wt.numpy()[M:] - wo[M:]).max() <= 1e-7 * np.abs(wo[M:]).max()


def test_reports_library_and_host():
    assert "torch.fft" in T.fft_library()
    assert T.cpu_model()
