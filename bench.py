#!/usr/bin/env python
"""Langevin-step throughput of the gfx950 path on one MI355X.

Workload (the 64^3 plain FP64 case): diblock N=90, NA=45, XbN=17.0 (bare, isXeN=0),
C=50, Ndt=0.5, mesh 64^3, box 4.4^3 R0; start W- = 17 cos(2 pi 3x/64) + N(0, 0.05^2)
(numpy seed 2), w+ = 0; noise from the reference's stream (mt19937_64 seed 123456789,
normal(0, sigma)); Anderson mixing to 1e-4 (maxIter 200, 10 histories).

A step is one whole Langevin step of the reference's loop (src/lfts_simulation.cc:72-76):
the W- update, the Anderson saddle-point solve, and the extra propagator evaluation
that prints ln Q. The noise of every timed step is drawn on the host beforehand and is
resident in HBM when timing starts.

Prints one JSON line (rank 0). Extra detail goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "Langevin steps/s at 64³,N=90 FP64 (plain vs batched); % HBM roofline"
HBM_PEAK_GBPS = 8000.0
NK_STEPS = 10  # Langevin steps timed with the Newton-Krylov solver in the derivative leg
HOST_STEPS = 5  # Langevin steps with host noise (PCIe-inclusive rate)
PROF_EVERY = 8  # live event timing of every 8th propagator evaluation (a timed one costs ~10 us)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def initial_fields(m, amp, sigma0, seed):
    M = int(np.prod(m))
    x = np.arange(m[0])[:, None, None] * np.ones(m)
    wm = amp * np.cos(2.0 * np.pi * (3 * x / m[0]))
    rng = np.random.default_rng(seed)
    wm = wm.reshape(-1) + rng.normal(0.0, sigma0, M)
    return np.concatenate([wm, np.zeros(M)])


def algorithmic_bytes(M, N, nAM_list, with_langevin=True):
    """Bytes of one Langevin step per the roofline accounting: each counted array is
    touched once, FFT-internal passes excluded."""
    B_PE = 8.0 * M * (8 * N - 2)
    total = 0.0
    for nAM in nAM_list:
        b = (nAM + 1) * B_PE
        for k in range(1, nAM + 1):
            nh = min(k - 1, 10)
            b += 8.0 * M * (3 * nh + 6)
        if with_langevin:
            b += 40.0 * M
        total += b
    return total


def measured_traffic(mesh):
    """HBM bytes per contour-step launch from the committed FETCH_SIZE/WRITE_SIZE
    summary (two rocprofv3 --pmc runs, scaled by a copy of known size); None when no
    summary exists for this mesh."""
    path = os.path.join(ROOT, "profiles", f"k_slab_io_{mesh}cubed_N90_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def cpu_baseline(w_state, params, m, Lb, N, NA, seconds_hint=30.0):
    """The oracle (CPU restatement, OpenMP over FFT lines) timed on this host for one
    Langevin step of the same workload, started from the GPU's state after warm-up."""
    import oracle as O

    M = int(np.prod(m))
    with tempfile.TemporaryDirectory() as td:
        f = os.path.join(td, "input")
        with open(f, "w") as fh:
            fh.write(f"{N} {NA} {params['XbN']} {params['C']} {params['dt']} 0\n")
            fh.write(f"{m[0]} {m[1]} {m[2]} {Lb[0]} {Lb[1]} {Lb[2]}\n")
            fh.write("1 0 1 1000 1\n")
            np.savetxt(fh, w_state[:2 * M], fmt="%.17g")
        sim = O.Sim(f, initial_mix=False)
        sim.set_w(w_state)
        t0 = time.perf_counter()
        sim.step(False)
        dt = time.perf_counter() - t0
        nAM = sim.last_nAM() - 1
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": 1.0 / dt,
        "unit": "Langevin steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"one Langevin step of the same 64^3 N=90 workload from the GPU state after warm-up "
                  f"(n_AM={nAM}, {dt:.1f} s); OpenMP over FFT lines, element-wise loops serial",
    }


BATCH_MESHES = [32] * 24 + [48] * 24 + [64] * 16


def cpu_baseline_torch(w_state, params, m, Lb, N, NA, sigma, steps=3):
    """The stronger CPU baseline: the reference's Langevin step with torch's threaded CPU
    FFT for the batched-2 r2c/c2r transforms and threaded element-wise loops (the CPU
    restatement oracle/torch_cpu.py, checked against the oracle), timed for `steps` whole
    Langevin steps of the same workload from the GPU's state after warm-up, on all host
    threads torch is given (OMP_NUM_THREADS)."""
    import torch

    from oracle import torch_cpu as T

    n_thr = T.set_threads()
    M = int(np.prod(m))
    t = T.TorchLfts(N, NA, m, Lb)
    w = torch.from_numpy(np.array(w_state, dtype=np.float64))
    g = np.random.default_rng(17)
    etas = [torch.from_numpy(g.normal(0.0, sigma, M)) for _ in range(steps + 1)]
    nam = []
    t0 = time.perf_counter()
    for s in range(steps):
        t.langevin_update(w, etas[s], etas[s + 1], params["XbN"], params["dt"])
        nam.append(t.anderson_mix(200, 1e-4, w) - 1)
        t.calc_concs(w)
    dt_s = time.perf_counter() - t0
    return {
        "value": steps / dt_s,
        "unit": "Langevin steps/s",
        "cores": n_thr,
        "kind": "port",
        "fft_library": T.fft_library(),
        "cpu_model": T.cpu_model(),
        "sample": f"{steps} whole Langevin steps of the same 64^3 N=90 workload from the GPU state after warm-up "
                  f"(n_AM={nam}, {dt_s:.1f} s): the reference's arithmetic with torch's threaded CPU FFT for the "
                  f"batched-2 r2c/c2r transforms, threaded element-wise loops, LAPACK for the Anderson LU",
    }


def batched_member(r, n):
    """Member r (mesh n^3) of the batched workload: N=90, NA uniform in [30,60], XbN
    uniform in [12,20], box edge uniform in [3.5,5.0], fields N(0, 0.1^2), all from numpy
    seed 1000+r; noise from mt19937_64 seeded 1000+r with sigma of src/lfts_params.cc:117
    (C=50, Ndt=0.5)."""
    g = np.random.default_rng(1000 + r)
    NA = int(g.integers(30, 61))
    XbN = float(g.uniform(12.0, 20.0))
    Lb = float(g.uniform(3.5, 5.0))
    M = n ** 3
    w = np.concatenate([g.normal(0.0, 0.1, M), g.normal(0.0, 0.1, M)])
    sigma = math.sqrt(2.0 * M * 0.5 / (50.0 * Lb ** 3))
    return dict(n=n, N=90, NA=NA, XbN=XbN, L=Lb, M=M, w=w, sigma=sigma, seed=1000 + r, dt=0.5)


def batched_members(quick=False):
    meshes = [32] * 3 + [48] * 3 + [64] * 2 if quick else BATCH_MESHES
    return [batched_member(r, n) for r, n in enumerate(meshes)]


def batched_leg(steps, warmup, L_mod, torch, quick=False, sequential=True):
    """Replica set of the batched workload: 64 members (24 x 32^3, 24 x 48^3, 16 x 64^3),
    N=90, NA uniform in [30,60], XbN uniform in [12,20], box edge uniform in [3.5,5.0],
    fields N(0, 0.1^2), member r seeded 1000+r (numpy) for its fields and parameters and
    1000+r (mt19937_64) for its noise. Times `steps` whole Langevin steps of the set, then
    the same steps of every member run alone, one after the other, from the same states."""
    members = batched_members(quick)
    meshes = [mb["n"] for mb in members]
    ctxs = []
    for mb in members:
        c = L_mod.Context(90, mb["NA"], (mb["n"],) * 3, (mb["L"],) * 3)
        c.set_fields(mb["w"])
        rng = L_mod.Rng(mb["seed"], mb["sigma"])
        c.langevin_init(rng.normals(mb["M"]))
        mb["rng"] = rng
        ctxs.append(c)
    XbN = [mb["XbN"] for mb in members]
    nsteps = warmup + steps
    etas = [[torch.from_numpy(mb["rng"].normals(mb["M"])).cuda() for mb in members] for _ in range(nsteps)]
    torch.cuda.synchronize()
    B = L_mod.Batch(ctxs)
    t0 = time.perf_counter()
    k0, _ = B.anderson_mix(200, 1e-4)
    log(f"batched set: initial mix {time.perf_counter() - t0:.2f} s, k in [{k0.min()}, {k0.max()}]")
    for s in range(warmup):
        B.full_step_dev([e.data_ptr() for e in etas[s]], XbN, 0.5)
    # snapshot the states so that the sequential leg starts from the same point
    snap = [(c.get_fields(4), None) for c in ctxs]
    # every set evaluation timed: its chain takes milliseconds, so the events' cost is
    # small, and the bytes of a sampled evaluation would depend on which members are active
    B.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kb = []
    for s in range(warmup, nsteps):
        _, k = B.full_step_dev([e.data_ptr() for e in etas[s]], XbN, 0.5)
        kb.append(k.copy())
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    ms_chain, n_pass, ms_pe, n_pe = B.profile_read()
    tuning = B.tuning()
    B.close()
    # bytes moved per the contour-step accounting (48 M per member step)
    cs_bytes = sum(48.0 * mb["M"] * (90 - 1) * (int(k) - 1 + 1) for ks in kb for mb, k in zip(members, ks))
    # the same steps, member by member
    t_seq = float("nan")
    if sequential:
        for c, (w, _) in zip(ctxs, snap):
            c.set_fields(w)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for r, c in enumerate(ctxs):
            for s in range(warmup, nsteps):
                c.full_step_dev(etas[s][r].data_ptr(), XbN[r], 0.5)
        torch.cuda.synchronize()
        t_seq = time.perf_counter() - t0
    n_members = len(members)
    return {
        "members": n_members,
        "meshes": {"32^3": meshes.count(32), "48^3": meshes.count(48), "64^3": meshes.count(64)},
        "steps": steps,
        "member_steps_per_s": n_members * steps / t_batch,
        "sequential_member_steps_per_s": n_members * steps / t_seq,
        "speedup_vs_sequential": t_seq / t_batch,
        "mean_n_AM": float(np.mean([np.mean(k) - 1 for k in kb])),
        "contour_chain_algorithmic_GBps": cs_bytes / (ms_chain * 1e-3) / 1e9 if ms_chain > 0 else None,
        "chain_frac_of_8TBps": cs_bytes / (ms_chain * 1e-3) / 8e12 if ms_chain > 0 else None,
        "slab_tuning": tuning.strip().splitlines(),
    }


def adjoint_leg(ctx, eta_dev, eta_more, XbN, dt, M, torch, tols=(1e-4, 1e-12)):
    """Newton-Krylov against Anderson on identical W-: from the state after the timed
    steps, one more W- update (as at the start of every Langevin step), then both
    solvers from the same fields to the same residual, at the production tolerance and
    at a tight one (where the two w+ must agree). Also times the tangent and adjoint
    products (each includes its forward evaluation) and the box derivative, and whole
    Langevin steps with the Newton-Krylov solve in place of Anderson (a different solver:
    an extension, not the reference's algorithm)."""
    ctx.langevin_step_dev(eta_dev, XbN, dt)
    w_start = ctx.get_fields(2)
    ctx.saddle_nk(tol=1e-2)  # one-time set-up (workspace, preconditioner table) outside the timing
    rows = []
    for tol in tols:
        ctx.set_fields(w_start)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        k, err_a = ctx.anderson_mix(5000, tol)
        t_and = time.perf_counter() - t0
        w_and = ctx.get_fields(2)
        ctx.set_fields(w_start)
        t0 = time.perf_counter()
        info = ctx.saddle_nk(tol=tol)
        t_nk = time.perf_counter() - t0
        w_nk = ctx.get_fields(2)
        rows.append({
            "tol": tol,
            "anderson_prop_evals": int(k) - 1,
            "anderson_err": err_a,
            "anderson_ms": 1e3 * t_and,
            "nk_newton": info["newton"],
            "nk_prop_evals": info["prop_evals"],
            "nk_tangent_evals": info["krylov"],
            "nk_chain_evals": info["prop_evals"] + info["krylov"],
            "nk_err": info["err"],
            "nk_ms": 1e3 * t_nk,
            "max_abs_dwplus": float(np.abs(w_nk[M:] - w_and[M:]).max()),
        })
    # whole Langevin steps with the Newton-Krylov solve (tol 1e-4) in place of Anderson,
    # from the same start, with the next noise vectors of the reference stream
    ctx.set_fields(w_start)
    nk_steps = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e in eta_more:
        _, inf = ctx.full_step_nk_dev(e.data_ptr(), XbN, dt, tol=1e-4)
        nk_steps.append(inf["prop_evals"] + inf["krylov"])
    t_steps = time.perf_counter() - t0
    dev = torch.empty(6 * M, dtype=torch.float64, device="cuda")
    g = torch.Generator(device="cpu").manual_seed(5)
    dev[:3 * M].copy_(torch.randn(3 * M, generator=g, dtype=torch.float64))
    reps = 5
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.jvp_dev(dev.data_ptr(), dev.data_ptr() + 8 * M, dev.data_ptr() + 8 * 3 * M)
    t_jvp = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.vjp_dev(dev.data_ptr() + 8 * M, dev.data_ptr() + 8 * 3 * M)
    t_vjp = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    ctx.box_derivative()
    t_box = time.perf_counter() - t0
    # the same products with checkpointed pair storage (stride ceil(sqrt(N)))
    stride = int(math.ceil(math.sqrt(ctx.N)))
    ctx.set_pair_storage(stride)
    kept = ctx.pair_storage()[1]
    ctx.calc_concs()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.calc_concs()
    t_pe_ck = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.jvp_dev(dev.data_ptr(), dev.data_ptr() + 8 * M, dev.data_ptr() + 8 * 3 * M)
    t_jvp_ck = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.vjp_dev(dev.data_ptr() + 8 * M, dev.data_ptr() + 8 * 3 * M)
    t_vjp_ck = (time.perf_counter() - t0) / reps
    ctx.set_pair_storage(0)
    return {
        "workload": "64^3 N=90 state after the timed steps plus one W- update; Anderson (10 histories) and "
                    "Newton-Krylov (GMRES(20), homogeneous-melt preconditioner) from the same fields",
        "solves": rows,
        "nk_langevin_steps_per_s": len(eta_more) / t_steps if eta_more else None,
        "nk_chain_evals_per_step": float(np.mean(nk_steps)) if nk_steps else None,
        "jvp_with_forward_ms": 1e3 * t_jvp,
        "vjp_with_forward_ms": 1e3 * t_vjp,
        "box_derivative_ms": 1e3 * t_box,
        "checkpointed": {
            "stride": stride,
            "pairs_kept": kept,
            "pairs_full": ctx.N,
            "calc_concs_ms": 1e3 * t_pe_ck,
            "jvp_with_forward_ms": 1e3 * t_jvp_ck,
            "vjp_with_forward_ms": 1e3 * t_vjp_ck,
        },
    }


def driver_leg(w0, m, Lb, N, NA, XbN, C, dt, warmup, steps, lnq_last):
    """The drop-in driver `lfts-gpu` on the timed steps: as host/lfts_simulation.cc runs a
    step, its noise (M normals of the seed-123456789 stream) is drawn on a host thread while
    the GPU runs the step before, and handed over in host memory. The workload's start
    fields go into an input file (loadType 1, 17 significant digits, read back exactly);
    the driver draws the same stream in the same order (src/lfts_simulation.cc:38,45-48),
    so it runs the same steps as the timed region. One run of `warmup + steps` steps,
    timed from the line it prints after step `warmup` to the line after the last step."""
    import subprocess

    exe = os.path.join(ROOT, "lfts_cpu_amd", "bin", "lfts-gpu")
    if not os.path.exists(exe):
        return None
    n_eq = warmup + steps
    stamps, lnqs = [], []
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "input")
        with open(path, "w") as f:
            f.write(f"{N} {NA} {XbN!r} {C!r} {dt!r} 0\n")
            f.write(f"{m[0]} {m[1]} {m[2]} {Lb[0]!r} {Lb[1]!r} {Lb[2]!r}\n")
            f.write(f"{n_eq} 0 1 {10 * n_eq} 1\n")
            np.savetxt(f, w0, fmt="%.17g")
        import threading

        p = subprocess.Popen([exe, path], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        watchdog = threading.Timer(300.0, p.kill)  # a side leg never holds the bench up
        watchdog.start()
        tail = []
        try:
            for line in p.stdout:  # the driver ends every line with std::endl (flushed)
                if line.startswith("lnQ ="):
                    stamps.append(time.perf_counter())
                    lnqs.append(float(line.split("=")[1]))
                tail = (tail + [line])[-20:]
            rc = p.wait()
        finally:
            watchdog.cancel()
        if rc != 0:
            raise RuntimeError(f"lfts-gpu failed ({rc}): {''.join(tail)}")
    if len(lnqs) != n_eq:
        raise RuntimeError(f"lfts-gpu printed {len(lnqs)} lnQ lines for {n_eq} steps")
    # the driver prints 6 significant digits: the same trajectory as the timed steps
    if abs(lnqs[-1] - lnq_last) > 1e-5 * max(1.0, abs(lnq_last)):
        raise RuntimeError(f"lfts-gpu ended at lnQ = {lnqs[-1]}, the timed steps at {lnq_last}")
    t = stamps[-1] - stamps[warmup - 1] if warmup > 0 else stamps[-1] - stamps[0]
    n = steps if warmup > 0 else steps - 1
    return {
        "steps_per_s": n / t,
        "sample": f"lfts-gpu on the workload's start fields for {n_eq} steps, timed between its printed lnQ "
                  f"lines of steps {max(warmup, 1)} and {n_eq} ({t:.3f} s); its last lnQ {lnqs[-1]} (6 digits) "
                  f"matches the timed steps'",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--mesh", type=int, default=64)
    ap.add_argument("--dims", type=int, nargs=3, default=None,
                    help="a non-cubic mesh mx my mz instead of --mesh (same workload otherwise)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--batched-steps", type=int, default=5, help="timed steps of the replica-set leg (0 skips it)")
    ap.add_argument("--batched-quick", action="store_true", help="8 members instead of 64 (smoke runs)")
    ap.add_argument("--only-batched", action="store_true", help="run only the replica-set leg (profiling)")
    ap.add_argument("--batched-no-sequential", action="store_true",
                    help="skip the member-by-member comparison of the replica-set leg (profiling)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the Newton-Krylov / derivative leg")
    ap.add_argument("--no-driver", action="store_true",
                    help="skip the lfts-gpu driver leg (profiling: a profiler would trace the child process too)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(replicas would shard one per GPU with no exchange per step).", file=sys.stderr)
        sys.exit(2)

    import torch

    import lfts_cpu_amd as L

    if args.only_batched:
        print(json.dumps({"batched": batched_leg(args.batched_steps, 1, L, torch, quick=args.batched_quick,
                                                 sequential=not args.batched_no_sequential)}))
        return

    N, NA, XbN, C, dt = 90, 45, 17.0, 50.0, 0.5
    m = tuple(args.dims) if args.dims else (args.mesh,) * 3
    Lb = (4.4, 4.4, 4.4)
    M = int(np.prod(m))
    V = Lb[0] * Lb[1] * Lb[2]
    sigma = math.sqrt(2.0 * M * dt / (C * V))
    params = {"XbN": XbN, "C": C, "dt": dt}

    w0 = initial_fields(m, XbN, 0.05, 2)
    ctx = L.Context(N, NA, m, Lb)
    ctx.set_fields(w0)
    rng = L.Rng(123456789, sigma)
    ctx.langevin_init(rng.normals(M))
    t0 = time.perf_counter()
    k0, err0 = ctx.anderson_mix(200, 1e-4)
    log(f"initial mix: k={k0} err={err0:.3e} ({time.perf_counter() - t0:.2f} s)")

    nsteps = args.warmup + args.steps
    n_extra = 0 if args.no_adjoint else 1 + NK_STEPS
    eta = torch.empty((nsteps + n_extra, M), dtype=torch.float64, device="cuda")
    for s in range(nsteps + n_extra):
        eta[s].copy_(torch.from_numpy(rng.normals(M)))
    torch.cuda.synchronize()

    # live HIP-event timing of the contour chains on the context's stream (the roofline
    # figure) over the timed steps, on every PROF_EVERY-th propagator evaluation. Already on
    # during the warm-up, so that the timed steps reuse pooled events; enabling it again
    # before the timed steps zeroes the sums.
    ctx.profile(True, PROF_EVERY)
    for s in range(args.warmup):
        lnq, k = ctx.full_step_dev(eta[s].data_ptr(), XbN, dt)
        log(f"warmup {s}: lnQ={lnq:.10g} n_AM={k - 1}")
    torch.cuda.synchronize()

    nAM = []
    marks = []
    ctx.profile(True, PROF_EVERY)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.warmup, nsteps):
        lnq, k = ctx.full_step_dev(eta[s].data_ptr(), XbN, dt)  # returns after the step's final read-back
        marks.append(time.perf_counter())
        nAM.append(k - 1)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    step_ms = np.diff(np.array([t0] + marks)) * 1e3
    log(f"timed {args.steps} steps: {elapsed:.3f} s, n_AM={nAM}, lnQ={lnq:.10g}")
    ms_chain, n_cs, ms_pe, n_pe = ctx.profile_read()
    ctx.profile(False)
    tuning_plain = ctx.tuning().strip()

    steps_per_s = args.steps / elapsed
    cs_ms = ms_chain / max(n_cs, 1)
    cs_bytes = 48.0 * M  # one contour step: read the q pair and h, write the q pair
    achieved = cs_bytes / (cs_ms * 1e-3) / 1e9
    step_bytes = algorithmic_bytes(M, N, nAM)
    step_gbps = step_bytes / elapsed / 1e9
    cs_per_s = sum((N - 1) * (n + 1) for n in nAM) * M / elapsed

    cpu = cpu_oracle = None
    if not args.no_cpu:
        state = ctx.get_fields(4)
        try:
            cpu = cpu_baseline_torch(state, params, m, Lb, N, NA, sigma)
            log(f"cpu baseline (torch FFT): {cpu}")
        except Exception as e:  # the baselines are reported, not required
            log(f"cpu baseline (torch FFT) failed: {e!r}")
        try:
            cpu_oracle = cpu_baseline(state, params, m, Lb, N, NA)
            log(f"cpu baseline (oracle): {cpu_oracle}")
        except Exception as e:
            log(f"cpu baseline (oracle) failed: {e!r}")

    adjoint = None
    if not args.no_adjoint:
        e0 = nsteps
        adjoint = adjoint_leg(ctx, eta[e0].data_ptr(), [eta[e0 + 1 + k] for k in range(NK_STEPS)], XbN, dt, M, torch)
        log(f"adjoint leg: {adjoint}")

    # attainable HBM bandwidth: a device-to-device copy of 1 GiB (read + write counted)
    xs = torch.empty(2 ** 27, dtype=torch.float64, device="cuda").fill_(1.0)
    ys = torch.empty_like(xs)
    ys.copy_(xs)
    torch.cuda.synchronize()
    c0 = time.perf_counter()
    for _ in range(5):
        ys.copy_(xs)
    torch.cuda.synchronize()
    copy_gbps = 5 * 2 * xs.numel() * 8 / (time.perf_counter() - c0) / 1e9
    del xs, ys
    log(f"copy bandwidth: {copy_gbps:.0f} GB/s")

    # PCIe-inclusive rate (reported, never `value`): the same step with the noise handed
    # over in host memory, as the lfts-gpu driver does it (drawn beforehand here)
    host_eta = [rng.normals(M) for _ in range(HOST_STEPS)]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for e in host_eta:
        ctx.langevin_step(e, XbN, dt)
        ctx.anderson_mix(200, 1e-4)
        ctx.calc_concs()
    torch.cuda.synchronize()
    host_rate = HOST_STEPS / (time.perf_counter() - t1)
    log(f"host-noise steps: {host_rate:.2f} steps/s")
    t1 = time.perf_counter()
    rng.normals(M)
    draw_ms = (time.perf_counter() - t1) * 1e3

    # the drop-in driver itself on the timed steps (reported, never `value`)
    driver = None
    if not (args.dims or args.no_driver):
        try:
            driver = driver_leg(w0, m, Lb, N, NA, XbN, C, dt, args.warmup, args.steps, lnq)
        except Exception as e:  # a side measurement: report its failure, keep the bench line
            driver = {"steps_per_s": None, "sample": f"lfts-gpu leg failed: {e}"}
    driver_rate = driver["steps_per_s"] if driver else None
    log(f"driver: {driver}; one draw of M normals {draw_ms:.1f} ms")

    batched = None
    if args.batched_steps > 0:
        ctx.close()
        batched = batched_leg(args.batched_steps, 1, L, torch, quick=args.batched_quick)
        if batched.get("contour_chain_algorithmic_GBps"):
            batched["chain_frac_of_attainable"] = batched["contour_chain_algorithmic_GBps"] / copy_gbps

    out = {
        "metric": METRIC,
        "value": steps_per_s,
        "unit": "Langevin steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: lamellar(300) W- + N(0,0.05^2) (numpy seed 2), w+=0; noise from mt19937_64 seed 123456789",
        "config": {
            "workload": f"plain FP64 Langevin step, diblock N={N} NA={NA} XbN={XbN} (isXeN=0) C={C} Ndt={dt}, "
                        f"mesh {m[0]}x{m[1]}x{m[2]}, box 4.4^3, Anderson tol 1e-4 (maxIter 200, 10 histories)",
            "mesh": list(m),
            "N": N,
            "NA": NA,
            "mean_n_AM": float(np.mean(nAM)),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS,
            "traffic": None if args.dims else measured_traffic(args.mesh),
            "kernel": "contour step (48*M algorithmic bytes; plain and fused-density-sum passes, 44:45 per "
                      f"evaluation), HIP events on the context stream over every {PROF_EVERY}th evaluation of the timed steps",
            "contour_step_us": cs_ms * 1e3,
            "attainable_copy_GBps": copy_gbps,
            "frac_of_attainable": achieved / copy_gbps,
            "slab_tuning": tuning_plain,
        },
        "step_ms_median": float(np.median(step_ms)),
        "steps_per_s_median": float(1e3 / np.median(step_ms)),
        "n_AM_histogram": {str(v): int(c) for v, c in zip(*np.unique(np.array(nAM), return_counts=True))},
        "step_algorithmic_GBps": step_gbps,
        "step_roofline_frac": step_gbps / HBM_PEAK_GBPS,
        "contour_steps_M_per_s": cs_per_s,
        "propagator_eval_ms": ms_pe / max(n_pe, 1),
        "pcie_inclusive_steps_per_s": host_rate,
        "driver_style_steps_per_s": driver_rate,
        "driver_style_sample": driver["sample"] if driver else None,
        "host_noise_draw_ms": draw_ms,
        "cpu_baseline": cpu,
        "cpu_baseline_oracle": cpu_oracle,
        "speedup_vs_cpu_baseline": (steps_per_s / cpu["value"]) if cpu else None,
        "batched": batched,
        "newton_krylov": adjoint,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
