"""Error paths of the C ABI that a caller can reach: a singular Anderson system (the
reference's GSL LU_solve reports it and aborts; here LFTS_ESING), and the rules of replica
sets (a context in a set in use cannot change its pair storage, be destroyed or join a
second set). The refusals are decided on the host; no kernel runs for them."""
import numpy as np
import pytest

import lfts_cpu_amd as L

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("loop", ["host", "device"])
def test_singular_anderson_system_is_reported(gpu, monkeypatch, loop):
    """The homogeneous melt (w = 0) with tolerance -1: the first iteration's simple mix moves
    w+ by 0.1 (phi+ - 1), far below the rounding of h = exp(-w+/N), so the second iteration
    sees the very same residual and its 1 x 1 history matrix U = |D0 - D1|^2 is exactly 0."""
    if loop == "device":
        monkeypatch.setenv("LFTS_ANDERSON", "device")
    print(f"singular Anderson system, {loop} loop, 16^3, N=8", flush=True)
    c = L.Context(8, 4, (16, 16, 16), (3.0, 3.0, 3.0))
    c.set_fields(np.zeros(2 * 16 ** 3))
    with pytest.raises(L.LftsError) as e:
        c.anderson_mix(50, -1.0)
    assert e.value.code == L.LFTS_ESING
    assert "singular" in str(e.value)
    c.close()


def test_set_members_cannot_be_changed_or_destroyed(gpu):
    a = L.Context(8, 4, (16, 16, 16), (3.0, 3.0, 3.0))
    b = L.Context(8, 3, (16, 16, 16), (3.2, 3.2, 3.2))
    s = L.Batch([a, b])
    lib = L.lib()
    assert lib.lfts_set_pair_storage(a.h, 3) == L.LFTS_ESTATE
    assert "replica set" in lib.lfts_last_error(a.h).decode()
    with pytest.raises(L.LftsError) as e:
        a.close()
    assert e.value.code == L.LFTS_ESTATE and a.h is not None
    with pytest.raises(L.LftsError):
        L.Batch([b])  # b already belongs to a set in use
    s.close()
    # once the set is gone the same calls succeed
    assert lib.lfts_set_pair_storage(a.h, 3) == L.LFTS_OK
    assert a.pair_storage()[0] == 3
    s2 = L.Batch([b])
    s2.close()
    a.close()
    b.close()
