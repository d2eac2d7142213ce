"""Density sums fused into the chain (one replica on the slab path): an evaluation that
adds the bead products during the second half of the chain (the Anderson iterations and
the Hamiltonian use it) gives phi and H bit-identical to the evaluation that stores every
pair and sums them afterwards (lfts_calc_concs). It keeps pairs 1..ceil(N/2) and N, equal
bit for bit to the stored evaluation's; asking for another pair is an error until
lfts_calc_concs runs again."""
import numpy as np
import pytest

import lfts_cpu_amd as L

pytestmark = pytest.mark.gpu

# (edge, N, NA): even and odd N, NA below, at and above ceil(N/2), both ends, tiny N
CASES = [
    (32, 12, 6),
    (32, 11, 3),
    (32, 11, 8),
    (32, 9, 5),
    (48, 10, 1),
    (48, 13, 12),
    (64, 16, 9),
    (40, 13, 7),
    (16, 9, 4),
    (24, 9, 4),
    (56, 11, 6),
    (32, 3, 1),
    (32, 4, 2),
    (32, 7, 0),
    (32, 8, 8),
]


def _fields(M, seed):
    rng = np.random.default_rng(seed)
    w = np.zeros(2 * M)
    w[:M] = rng.normal(0, 1.5, M)
    w[M:] = rng.normal(0, 0.7, M)
    return w


@pytest.mark.parametrize("e,N,NA", CASES)
def test_fused_sums_match_stored_pairs(gpu, e, N, NA):
    m = (e, e, e)
    M = e ** 3
    ctx = L.Context(N, NA, m, (3.7, 3.7, 3.7))
    ctx.set_fields(_fields(M, 100 * N + NA))
    H_fused = ctx.hamiltonian(17.0)                # fused evaluation
    phi_fused = ctx.get_fields(4)[2 * M:].copy()
    i0 = (N + 1) // 2
    kept = sorted(set(range(1, i0 + 1)) | {N})
    pairs_fused = {i: ctx.pair(i) for i in kept}
    dropped = [i for i in range(1, N + 1) if i not in kept]
    for i in dropped:
        with pytest.raises(L.LftsError):
            ctx.pair(i)
    lnQ = ctx.calc_concs()                         # every pair stored, summed afterwards
    phi_stored = ctx.get_fields(4)[2 * M:]
    assert np.isfinite(lnQ)
    assert np.array_equal(phi_fused, phi_stored)
    for i in kept:
        assert np.array_equal(pairs_fused[i], ctx.pair(i)), i
    for i in dropped:                              # available again after a stored evaluation
        assert np.all(np.isfinite(ctx.pair(i)))
    assert ctx.hamiltonian(17.0) == H_fused


def test_fused_anderson_equals_replica_set(gpu):
    """One context (fused sums) and a two-member set of the same system (no fusion) take
    the same Anderson iterations to bit-identical fields."""
    e, N, NA = 32, 15, 7
    M = e ** 3
    w = _fields(M, 5)
    one = L.Context(N, NA, (e, e, e), (3.9, 3.9, 3.9))
    one.set_fields(w)
    k1, err1 = one.anderson_mix(60, 1e-6)
    members = [L.Context(N, NA, (e, e, e), (3.9, 3.9, 3.9)) for _ in range(2)]
    for c in members:
        c.set_fields(w)
    b = L.Batch(members)
    k2, err2 = b.anderson_mix(60, 1e-6)
    assert list(k2) == [k1, k1]
    for c in members:
        assert np.array_equal(c.get_fields(4), one.get_fields(4))
    del b
