"""Live event timing (lfts_profile): with enable = k, every k-th propagator evaluation is
timed, and lfts_profile_read counts exactly those; 0 times nothing; a negative stride is
rejected."""
import numpy as np
import pytest

import lfts_cpu_amd as L

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("every,expect", [(1, 7), (3, 3), (8, 1), (0, 0)])
def test_profile_samples_every_kth_evaluation(gpu, every, expect):
    N, M = 10, 32 ** 3
    c = L.Context(N, 5, (32, 32, 32), (4.0, 4.0, 4.0))
    c.set_fields(np.random.default_rng(2).normal(0, 0.5, 2 * M))
    c.profile(every > 0, max(every, 1))
    for _ in range(7):
        c.calc_concs()
    ms_chain, n_cs, ms_pe, n_pe = c.profile_read()
    assert n_pe == expect and n_cs == expect * (N - 1)
    assert (ms_chain > 0.0) == (expect > 0) and ms_pe >= ms_chain
    c.profile(False)
    with pytest.raises(L.LftsError):
        c._chk(L.lib().lfts_profile(c.h, -1))
    c.close()
