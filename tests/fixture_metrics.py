"""Shared by the fixture generator (tests/golden/make_golden.py) and the GPU tests that
read its fixtures: field summaries and the relative differences the bars apply to."""
import math

import numpy as np

N_POINTS = 64


def sample_points(M: int, n: int = N_POINTS, seed: int = 7) -> np.ndarray:
    return np.sort(np.random.default_rng(seed).choice(M, n, replace=False)).astype(np.int64)


def field_summary(w: np.ndarray, M: int, nfields: int = 4) -> dict:
    """Per field: exact sum (fsum), exact sum of squares, values at 64 fixed points."""
    idx = sample_points(M)
    f = w.reshape(-1, M)[:nfields]
    return {
        "sums": np.array([math.fsum(f[k].tolist()) for k in range(nfields)]),
        "sumsq": np.array([math.fsum((f[k] * f[k]).tolist()) for k in range(nfields)]),
        "points": idx,
        "values": f[:, idx].copy(),
    }


def rel(a, b) -> float:
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


def summary_diffs(ref: dict, w: np.ndarray, M: int, nfields: int = 4) -> list:
    """Largest relative difference per field of w against the summary ref, over its sum
    (relative to max(|sum|, sqrt(M))), its sum of squares and its sampled values."""
    f = w.reshape(-1, M)[:nfields]
    out = []
    for k in range(nfields):
        s, q = math.fsum(f[k].tolist()), math.fsum((f[k] * f[k]).tolist())
        ds = abs(s - ref["sums"][k]) / max(abs(ref["sums"][k]), M ** 0.5)
        dq = abs(q - ref["sumsq"][k]) / ref["sumsq"][k]
        dp = rel(f[k, ref["points"]], ref["values"][k])
        out.append(max(ds, dq, dp))
    return out


def prefixed(g: dict, tag: str) -> dict:
    return {k[len(tag):]: v for k, v in g.items() if k.startswith(tag)}


def hamiltonian(lnq: float, sums, sumsq, XbN: float, M: int) -> float:
    """H = -lnQ + (sum W-^2 / XbN - sum w+) / M (src/lfts_simulation.cc:122-135) from the
    exact field sums of a summary."""
    return -lnq + (sumsq[0] / XbN - sums[1]) / M
