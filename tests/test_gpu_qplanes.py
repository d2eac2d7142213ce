"""Q from the chain's own plane sums: on the slab path the last contour-step pass of a
chain adds up each plane of q_N (src/diblock.cc:46-48: Q = sum of q_N / M), so no pass
over q_N follows the chain. Q must equal the sum of the stored pair N, and the separate
pass (LFTS_Q_PASS=1, another summation order) to rounding; every kernel shape rounds the
plane sums alike (one or two fields per workgroup, tables in registers or LDS)."""
import math

import numpy as np
import pytest

import lfts_cpu_amd as L

pytestmark = pytest.mark.gpu

# (edge, N, NA): every slab edge class, even and odd N, the shortest chains
CASES = [
    (16, 9, 4),
    (24, 6, 3),
    (32, 2, 1),
    (32, 3, 2),
    (40, 7, 3),
    (48, 10, 5),
    (56, 5, 2),
    (64, 12, 6),
    (72, 4, 2),
    (80, 5, 3),
    (96, 4, 2),
    (128, 3, 1),
]


def _ctx(e, N, NA, seed):
    M = e ** 3
    c = L.Context(N, NA, (e, e, e), (3.8, 3.8, 3.8))
    rng = np.random.default_rng(seed)
    c.set_fields(np.concatenate([rng.normal(0, 1.2, M), rng.normal(0, 0.6, M)]))
    return c


@pytest.mark.parametrize("e,N,NA", CASES)
def test_q_from_plane_sums(gpu, monkeypatch, e, N, NA):
    M = e ** 3
    monkeypatch.delenv("LFTS_Q_PASS", raising=False)
    c = _ctx(e, N, NA, 7 * e + N)
    lnq = c.calc_concs()
    qN = c.pair(N)[:M]
    phi = c.get_fields(4)[2 * M:]
    # the sum of the stored q_N, in exact (compensated) arithmetic
    assert abs(lnq - math.log(math.fsum(qN) / M)) < 1e-13
    monkeypatch.setenv("LFTS_Q_PASS", "1")
    p = _ctx(e, N, NA, 7 * e + N)
    lnq_p = p.calc_concs()
    assert abs(lnq - lnq_p) < 1e-13
    assert np.array_equal(qN, p.pair(N)[:M])  # the chain itself is untouched
    # phi+ and phi- (which passes through zero: an absolute bar as well)
    np.testing.assert_allclose(phi, p.get_fields(4)[2 * M:], rtol=1e-13, atol=1e-14)
    c.close()
    p.close()


def test_q_plane_sums_alike_in_every_variant(gpu, monkeypatch):
    """All four kernel variants of a shape and the replica-set kernel give the same Q bits
    (the plane sum's order depends only on a thread's place in its field)."""
    e, N, NA = 32, 11, 5
    out = []
    for v in ("0", "1", "2", "3"):
        monkeypatch.setenv("LFTS_SLAB_VARIANT", v)
        c = _ctx(e, N, NA, 3)
        out.append(c.calc_concs())
        c.close()
    monkeypatch.delenv("LFTS_SLAB_VARIANT")
    members = [_ctx(e, N, NA, 3) for _ in range(3)]
    b = L.Batch(members)
    out.extend(b.calc_concs().tolist())
    b.close()
    assert len(set(out)) == 1, out
