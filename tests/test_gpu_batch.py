"""Replica sets: a set of contexts with mixed meshes (the slab path at 32^3, 48^3 and
64^3, the line path elsewhere) gives every member exactly the result the member gets
alone, and the 48^3 slab path matches the oracle."""
import numpy as np
import pytest

import lfts_cpu_amd as L
import oracle as O

pytestmark = pytest.mark.gpu

MEMBERS = [
    # (m, L, N, NA, seed)
    ((32, 32, 32), (3.6, 3.6, 3.6), 16, 7, 1),
    ((48, 48, 48), (4.1, 4.1, 4.1), 16, 9, 2),
    ((64, 64, 64), (4.4, 4.4, 4.4), 12, 6, 3),
    ((12, 10, 8), (3.0, 2.5, 2.0), 14, 5, 4),
    ((32, 32, 32), (3.9, 3.9, 3.9), 20, 12, 5),
]


def _fields(M, seed, a=2.0, b=1.0):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.normal(0, a, M), rng.normal(0, b, M)])


def _make(members):
    ctxs, ws = [], []
    for m, Lb, N, NA, seed in members:
        M = int(np.prod(m))
        c = L.Context(N, NA, m, Lb)
        w = _fields(M, seed)
        c.set_fields(w)
        ctxs.append(c)
        ws.append(w)
    return ctxs, ws


def test_batch_calc_concs_equals_members_alone(gpu):
    ctxs, ws = _make(MEMBERS)
    b = L.Batch(ctxs)
    lnq = b.calc_concs()
    got = [c.get_fields(4) for c in ctxs]
    b.close()
    alone, _ = _make(MEMBERS)
    for i, c in enumerate(alone):
        assert c.calc_concs() == lnq[i]
        assert np.array_equal(c.get_fields(4), got[i]), i


def test_batch_anderson_equals_members_alone(gpu):
    ctxs, _ = _make(MEMBERS)
    b = L.Batch(ctxs)
    k, err = b.anderson_mix(200, 1e-4)
    got = [c.get_fields(4) for c in ctxs]
    b.close()
    assert len(set(k.tolist())) > 1  # members converge after different iteration counts
    alone, _ = _make(MEMBERS)
    for i, c in enumerate(alone):
        ki, erri = c.anderson_mix(200, 1e-4)
        assert ki == k[i] and erri == err[i], i
        assert np.array_equal(c.get_fields(4), got[i]), i


def test_batch_full_steps_equal_members_alone(gpu):
    import torch

    ctxs, _ = _make(MEMBERS)
    for c in ctxs:
        c.langevin_init(np.random.default_rng(99).normal(0, 1.0, c.M))
    b = L.Batch(ctxs)
    etas = [[torch.from_numpy(np.random.default_rng(100 * s + i).normal(0, 1.0, c.M)).cuda() for i, c in
             enumerate(ctxs)] for s in range(2)]
    torch.cuda.synchronize()
    XbN = [12.0, 14.0, 16.0, 13.0, 15.0]
    out = []
    for s in range(2):
        out.append(b.full_step_dev([e.data_ptr() for e in etas[s]], XbN, 0.5))
    got = [c.get_fields(4) for c in ctxs]
    b.close()
    alone, _ = _make(MEMBERS)
    for i, c in enumerate(alone):
        c.langevin_init(np.random.default_rng(99).normal(0, 1.0, c.M))
        for s in range(2):
            lnq, k = c.full_step_dev(etas[s][i].data_ptr(), XbN[i], 0.5)
            assert lnq == out[s][0][i] and k == out[s][1][i], (i, s)
        assert np.array_equal(c.get_fields(4), got[i]), i


def test_slab_48_against_oracle(gpu):
    m, Lb, N, NA = (48, 48, 48), (4.1, 4.1, 4.1), 10, 4
    M = 48 ** 3
    w = np.zeros(4 * M)
    w[:2 * M] = _fields(M, 7, 3.0, 1.0)
    ctx = L.Context(N, NA, m, Lb)
    ctx.set_fields(w[:2 * M])
    lnq = ctx.calc_concs()
    got = ctx.get_fields(4)
    d = O.Diblock(N, NA, m, Lb)
    wo = w.copy()
    lnq_ref = d.calc_concs(wo)
    assert abs(lnq - lnq_ref) <= 1e-10 * max(1.0, abs(lnq_ref))
    rel = np.abs(got[2 * M:] - wo[2 * M:]).max() / np.abs(wo[2 * M:]).max()
    assert rel < 1e-10
    for i in (2, N // 2, N):
        p, po = ctx.pair(i), d.pair(i)
        assert np.abs(p - po).max() / np.abs(po).max() < 1e-10, i


# ---- the Anderson iteration replayed as one HIP graph ---------------------------------------

SLAB_MEMBERS = [
    # slab-path meshes only (graph mode): three sizes, mixed N and NA
    ((32, 32, 32), (3.6, 3.6, 3.6), 16, 7, 11),
    ((48, 48, 48), (4.1, 4.1, 4.1), 12, 9, 12),
    ((16, 16, 16), (2.9, 2.9, 2.9), 20, 8, 13),
    ((32, 32, 32), (3.9, 3.9, 3.9), 14, 3, 14),
]


def _steps(members, monkeypatch, graph, nsteps=2):
    import torch

    monkeypatch.setenv("LFTS_GRAPH", "1" if graph else "0")
    ctxs, _ = _make(members)
    for c in ctxs:
        c.langevin_init(np.random.default_rng(98).normal(0, 1.0, c.M))
    etas = [[torch.from_numpy(np.random.default_rng(50 * s + i).normal(0, 1.0, c.M)).cuda() for i, c in
             enumerate(ctxs)] for s in range(nsteps)]
    torch.cuda.synchronize()
    XbN = [12.0 + i for i in range(len(ctxs))]
    b = L.Batch(ctxs)
    k0, _ = b.anderson_mix(200, 1e-4)
    b.profile(True)
    out = [b.full_step_dev([e.data_ptr() for e in etas[s]], XbN, 0.5) for s in range(nsteps)]
    prof = b.profile_read()
    fields = [c.get_fields(4) for c in ctxs]
    counters = [c.counters() for c in ctxs]
    b.close()
    for c in ctxs:
        c.close()
    return k0, out, fields, counters, prof


def test_graph_iterations_equal_direct_launches(gpu, monkeypatch):
    """Every Anderson iteration after the first as a replay of one captured graph (all
    members in its launches, those out of the loop skipped by flags) gives the bits of
    the direct launches over the active members, member by member."""
    g = _steps(SLAB_MEMBERS, monkeypatch, True)
    d = _steps(SLAB_MEMBERS, monkeypatch, False)
    assert np.array_equal(g[0], d[0])
    assert len(set(g[0].tolist())) > 1  # members leave the loop at different iterations
    for (lg, kg), (ld, kd) in zip(g[1], d[1]):
        assert np.array_equal(lg, ld) and np.array_equal(kg, kd)
        assert len(set(kg.tolist())) > 1
    for fg, fd in zip(g[2], d[2]):
        assert np.array_equal(fg, fd)
    assert g[3] == d[3]  # the same evaluation and contour-step counts per member
    ms_chain, n_pass, ms_pe, n_pe = g[4]
    assert ms_chain > 0 and ms_pe >= ms_chain and n_pass > 0 and n_pe == d[4][3]


def test_graph_iterations_with_line_path_members(gpu, monkeypatch):
    """A set with a line-path member (12 x 10 x 8 mesh) beside slab-path members replays
    its iterations as graphs too: the same bits, iteration counts and counters as the
    direct launches."""
    g = _steps(MEMBERS, monkeypatch, True)
    d = _steps(MEMBERS, monkeypatch, False)
    assert np.array_equal(g[0], d[0])
    for (lg, kg), (ld, kd) in zip(g[1], d[1]):
        assert np.array_equal(lg, ld) and np.array_equal(kg, kd)
    for fg, fd in zip(g[2], d[2]):
        assert np.array_equal(fg, fd)
    assert g[3] == d[3]
    assert g[4][3] == d[4][3]  # evaluations counted alike


def test_graph_set_equals_members_alone(gpu, monkeypatch):
    import torch

    k0, out, fields, counters, _ = _steps(SLAB_MEMBERS, monkeypatch, True)
    alone, _ = _make(SLAB_MEMBERS)
    for i, c in enumerate(alone):
        c.langevin_init(np.random.default_rng(98).normal(0, 1.0, c.M))
        ka, _ = c.anderson_mix(200, 1e-4)
        assert ka == k0[i]
        for s in range(2):
            e = torch.from_numpy(np.random.default_rng(50 * s + i).normal(0, 1.0, c.M)).cuda()
            torch.cuda.synchronize()
            lnq, k = c.full_step_dev(e.data_ptr(), 12.0 + i, 0.5)
            assert lnq == out[s][0][i] and k == out[s][1][i], (i, s)
        assert np.array_equal(c.get_fields(4), fields[i]), i
        assert c.counters() == counters[i]
        c.close()


# ---- a 32^3 group riding in the 64^3 group's launches ----------------------------------------

PAIR_MEMBERS = [
    # 64^3 and 32^3 members with different chain lengths in each group (the shorter chains
    # end while the launches go on), and a 48^3 member launched on its own
    ((64, 64, 64), (4.4, 4.4, 4.4), 14, 6, 31),
    ((32, 32, 32), (3.6, 3.6, 3.6), 18, 7, 32),
    ((64, 64, 64), (4.1, 4.1, 4.1), 9, 5, 33),
    ((32, 32, 32), (3.9, 3.9, 3.9), 11, 4, 34),
    ((48, 48, 48), (4.0, 4.0, 4.0), 12, 6, 35),
]


@pytest.mark.parametrize("graph", [True, False])
def test_paired_launch_equals_separate_launches(gpu, monkeypatch, graph):
    """The 64^3 and 32^3 groups of a set in one launch per contour step (default) give
    the bits of one launch per group (LFTS_SET_PAIR=0): initial Anderson solve and two
    Langevin steps, with and without the iteration graphs."""
    monkeypatch.setenv("LFTS_SET_PAIR", "1")
    a = _steps(PAIR_MEMBERS, monkeypatch, graph)
    monkeypatch.setenv("LFTS_SET_PAIR", "0")
    s = _steps(PAIR_MEMBERS, monkeypatch, graph)
    assert np.array_equal(a[0], s[0])
    for (la, ka), (ls, ks) in zip(a[1], s[1]):
        assert np.array_equal(la, ls) and np.array_equal(ka, ks)
    for fa, fs in zip(a[2], s[2]):
        assert np.array_equal(fa, fs)
    assert a[3] == s[3]


# ---- reduction results stored in pinned host memory, or copied ------------------------------

@pytest.mark.parametrize("graph", [True, False])
def test_host_mirrored_scalars_equal_copied_ones_for_a_set(gpu, monkeypatch, graph):
    """The reductions' results stored by the kernels in pinned host memory (default) give
    the bits of a copy from device memory (LFTS_SCAL_COPY=1): initial Anderson solve and
    two Langevin steps of a mixed set."""
    monkeypatch.setenv("LFTS_SCAL_COPY", "0")
    a = _steps(PAIR_MEMBERS, monkeypatch, graph)
    monkeypatch.setenv("LFTS_SCAL_COPY", "1")
    s = _steps(PAIR_MEMBERS, monkeypatch, graph)
    assert len(set(a[0].tolist())) > 1
    assert np.array_equal(a[0], s[0])
    for (la, ka), (ls, ks) in zip(a[1], s[1]):
        assert np.array_equal(la, ls) and np.array_equal(ka, ks)
    for fa, fs in zip(a[2], s[2]):
        assert np.array_equal(fa, fs)
    assert a[3] == s[3]


def test_host_mirrored_scalars_equal_copied_ones_for_one_replica(gpu, monkeypatch):
    """One 64^3 replica (fused density sums, Gram row depth as a kernel argument): the
    same iteration counts, residuals, lnQ, fields and Hamiltonian over an Anderson solve
    and two Langevin steps whether the scalars are read from host memory as the host
    polls for them (default), read from host memory after a stream synchronisation
    (LFTS_SPIN=0), or copied (LFTS_SCAL_COPY=1)."""
    import torch

    m, Lb, N, NA, seed = (64, 64, 64), (4.4, 4.4, 4.4), 20, 9, 41
    out = []
    for copy, spin in (("0", "1"), ("0", "0"), ("1", "1")):
        monkeypatch.setenv("LFTS_SCAL_COPY", copy)
        monkeypatch.setenv("LFTS_SPIN", spin)
        (c,), _ = _make([(m, Lb, N, NA, seed)])
        c.langevin_init(np.random.default_rng(97).normal(0, 1.0, c.M))
        r = [c.anderson_mix(200, 1e-4)]
        for s in range(2):
            e = torch.from_numpy(np.random.default_rng(60 + s).normal(0, 1.0, c.M)).cuda()
            torch.cuda.synchronize()
            r.append(c.full_step_dev(e.data_ptr(), 14.0, 0.5))
        r.append(c.hamiltonian(14.0))
        out.append((r, c.get_fields(4), c.counters()))
        c.close()
    for r, f, n in out[1:]:
        assert r == out[0][0]
        assert np.array_equal(f, out[0][1])
        assert n == out[0][2]


# a member whose axes have prime factors 13 and 17 (direct-DFT stages of the line kernel)
PRIME_MEMBERS = [MEMBERS[0], ((13, 17, 4), (2.6, 3.1, 1.0), 10, 4, 6), MEMBERS[3]]


def test_prime_mesh_member_in_set_and_graphs(gpu, monkeypatch):
    """A member of prime-factor mesh sizes gives in a set the bits it gets alone, and the
    set's replayed iteration graphs (which then hold that kernel's launches) give the bits
    of the direct launches."""
    ctxs, _ = _make(PRIME_MEMBERS)
    b = L.Batch(ctxs)
    lnq = b.calc_concs()
    got = [c.get_fields(4) for c in ctxs]
    b.close()
    alone, _ = _make(PRIME_MEMBERS)
    for i, c in enumerate(alone):
        assert c.calc_concs() == lnq[i]
        assert np.array_equal(c.get_fields(4), got[i]), i
    g = _steps(PRIME_MEMBERS, monkeypatch, True)
    d = _steps(PRIME_MEMBERS, monkeypatch, False)
    assert np.array_equal(g[0], d[0])
    for (lg, kg), (ld, kd) in zip(g[1], d[1]):
        assert np.array_equal(lg, ld) and np.array_equal(kg, kd)
    for fg, fd in zip(g[2], d[2]):
        assert np.array_equal(fg, fd)
