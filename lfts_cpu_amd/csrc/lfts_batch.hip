// Element-wise passes and reductions of the Langevin step over a set of replicas.
//
// Grid (nb_max, nact): blockIdx.y picks the active replica, whose first nb workgroups
// stride over its M mesh points (nb = blocks_for(M)). Reductions write one partial per workgroup in a fixed order and a
// second kernel sums them, so results are identical from run to run and identical
// between a replica run alone and the same replica run inside a larger set.
#include <hip/hip_runtime.h>

#include "lfts_internal.h"

namespace lfts {

namespace {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Sum of v over a 256-thread workgroup; valid in thread 0.
__device__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += red[i];
    return s;
}

// block_sum of nc components at once (component c's result in out[c], written by thread
// c): each component takes exactly block_sum's operations, with one barrier for all.
template <int NC>
__device__ __forceinline__ void block_sums(double (&v)[NC], int nc, double* red, double* out) {
#pragma unroll
    for (int c = 0; c < NC; c++)
        if (c < nc) v[c] = wave_sum(v[c]);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int c = 0; c < NC; c++)
            if (c < nc) red[wv * NC + c] = v[c];
    __syncthreads();
    if ((int)threadIdx.x < nc) {
        double s = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += red[i * NC + threadIdx.x];
        out[threadIdx.x] = s;
    }
}

__device__ __forceinline__ const double* pair_ptr(const Rep& R, int i) {
    return (i == 1) ? R.h : R.q + 2L * (i - 2) * R.M;
}

__device__ __forceinline__ double* part_row(double* part, int slot) {
    return part + ((long)slot * MAX_NB + blockIdx.x) * PSTRIDE;
}

constexpr int PHI_UNROLL = 8;  // beads per group in k_phi's pass over the stored pairs

#define FOR_POINTS(r, R) for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < (R).M; r += (long)(R).nb * blockDim.x)
#define REPLICA(R)                              \
    const Rep R = reps[blockIdx.y];             \
    if ((int)blockIdx.x >= R.nb || (R.skip && *R.skip)) return

}  // namespace

// src/diblock.cc:64-72 (prepare_h): hA = exp(-(W-+w+)/N), hB = exp(-(-W-+w+)/N)
__global__ __launch_bounds__(256) void k_prepare_h(const Rep* __restrict__ reps) {
    REPLICA(R);
    FOR_POINTS(r, R) {
        const double wm = R.w[r], wp = R.w[r + R.M];
        R.h[r] = exp(-(wm + wp) / R.N);
        R.h[r + R.M] = exp(-(-wm + wp) / R.N);
    }
}

// partial sums of q_N (src/diblock.cc:46-48) into qpart[id][workgroup], for the members
// whose chain does not leave its plane sums in qpl (line path); k_phi forms Q from them
__global__ __launch_bounds__(256) void k_q_sum(const Rep* __restrict__ reps, double* __restrict__ qpart) {
    __shared__ double red[8];
    REPLICA(R);
    if (R.qpl) return;
    const double* qN = pair_ptr(R, R.N);
    double s = 0.0;
    FOR_POINTS(r, R) s += qN[r];
    s = block_sum(s, red);
    if (threadIdx.x == 0) qpart[(long)R.id * MAX_NB + blockIdx.x] = s;
}

// Q = (sum of the replica's partials) / M, in every workgroup that asks: the chain's n
// plane sums (qpl) or k_q_sum's nb workgroup sums. k_reduce's operations in k_reduce's
// order (thread t adds rows t, t + 256, ... < rows, then the wavefront and LDS sums), so
// every workgroup gets the same bits. All of a thread's rows are loaded first; indexes
// from the row count up are clamped in bounds and never added.
__device__ __forceinline__ double q_from_partials(const Rep& R, const double* qpart, double* red, double* res) {
    constexpr int ROWS = MAX_NB / 256;  // launched with 256 threads
    const double* p = R.qpl ? R.qpl : qpart + (long)R.id * MAX_NB;
    const int nb = R.qpl ? R.n : R.nb;
    double v[ROWS];
#pragma unroll
    for (int k = 0; k < ROWS; k++) {
        const int b = (int)threadIdx.x + 256 * k;
        v[k] = p[(b < nb) ? b : nb - 1];
    }
    double s[1] = {0.0};
#pragma unroll
    for (int k = 0; k < ROWS; k++)
        if ((int)threadIdx.x + 256 * k < nb) s[0] += v[k];
    block_sums<1>(s, 1, red, res);
    __syncthreads();
    return res[0] / R.M;
}

// Sum over one replica's partial rows p[0..nb) (PSTRIDE doubles each), all PSTRIDE
// components together; component c < ncomp lands in res[c], written (and readable) by
// thread c. Thread t adds rows t, t + 256, ... (< nb) in that order. All of its rows are
// loaded before the first add: the partials were written by workgroups on every XCD, so
// each load is a trip to memory, and a load-add loop paid one trip per row. Row indexes
// from nb up are clamped to nb - 1 (in bounds, never added), and every row is PSTRIDE
// doubles wide, so the loads carry no condition; components from ncomp up are summed but
// never stored.
__device__ __forceinline__ void reduce_rows(const double* p, int nb, int ncomp, double* red, double* res) {
    constexpr int ROWS = MAX_NB / 256;  // rows per thread (256 threads)
    static_assert(MAX_NB % 256 == 0, "whole rows per thread");
    double v[ROWS][PSTRIDE];
#pragma unroll
    for (int k = 0; k < ROWS; k++) {
        const int b = (int)threadIdx.x + 256 * k;
        const long row = (b < nb) ? b : nb - 1;
#pragma unroll
        for (int c = 0; c < PSTRIDE; c++) v[k][c] = p[row * PSTRIDE + c];
    }
    double s[PSTRIDE];
#pragma unroll
    for (int c = 0; c < PSTRIDE; c++) s[c] = 0.0;
#pragma unroll
    for (int k = 0; k < ROWS; k++)
        if ((int)threadIdx.x + 256 * k < nb)
#pragma unroll
            for (int c = 0; c < PSTRIDE; c++) s[c] += v[k][c];
    block_sums<PSTRIDE>(s, ncomp, red, res);
}

// scal[rid][off + c] = sum over the replica's partials (divided by M when asked), c <
// ncomp (<= PSTRIDE); hscal (pinned host memory, may be null) receives the same values.
// hseq (pinned host memory, may be null): once the results are stored, hseq[slot] = seq,
// so that the host can wait for them by polling instead of synchronising the stream.
__global__ __launch_bounds__(256) void k_reduce(const Rep* __restrict__ reps, const double* __restrict__ part,
                                                int ncomp, int div_by_M, double* __restrict__ scal, int off,
                                                double* __restrict__ hscal, unsigned* __restrict__ hseq, unsigned seq) {
    __shared__ double red[4 * PSTRIDE];
    __shared__ double res[PSTRIDE];
    const int slot = blockIdx.x;
    const Rep R = reps[slot];
    if (R.skip && *R.skip) return;
    reduce_rows(part + (long)slot * MAX_NB * PSTRIDE, R.nb, ncomp, red, res);
    if ((int)threadIdx.x < ncomp) {
        const double v = res[threadIdx.x];
        const long at = (long)R.id * SSTRIDE + off + threadIdx.x;
        scal[at] = div_by_M ? v / R.M : v;
        if (hscal) hscal[at] = div_by_M ? v / R.M : v;
    }
    if (hseq && threadIdx.x < 64) {  // wave 0 stored the results (ncomp <= 16)
        __threadfence_system();
        if (threadIdx.x == 0) __hip_atomic_store(hseq + slot, seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// src/diblock.cc:51-53,76-95: phiA = sum_{i<=NA} q_i q+_i, phiB = sum_{i>NA} q_i q+_i,
// normalised by N Q h and written as phi- = phiA - phiB, phi+ = phiA + phiB.
// The bead sums run in the order of the fused chain (fused_phi): s[0]/s[1] over beads
// i0+1..N ascending (block A/B), s[2]/s[3] over beads N-i0..1 descending, each started by
// the first product and continued with fma; the middle bead of an odd N comes last. The
// stored-pair evaluation repeats that order, so phi is bit-identical either way.
//
// NH >= 0: the same pass also forms the Anderson Gram row partials <Dh0, Dh_m>, m = 0..NH
// (Dh0 = phi+ - 1, src/anderson.cc:35-55) over the same points in the same order as
// k_am_dots, so the partial rows are identical and the separate pass is saved.
// a stored pair or running sum, read once per evaluation: non-temporal
__device__ __forceinline__ double ld_once(const double* p) { return __builtin_nontemporal_load(p); }

template <int NH>
__device__ __forceinline__ void phi_row(const Rep& R, double Q, int fuse, double* redn, double* out) {
    const long M = R.M;
    const int N = R.N, NA = R.NA, i0 = (N + 1) / 2;
    const bool fused = fused_phi(R, fuse);
    bool u[4];
    phi_sums_used(N, NA, u);
    constexpr int NG = NH < 0 ? 1 : NH + 1;
    double acc[NG];
#pragma unroll
    for (int m = 0; m < NG; m++) acc[m] = 0.0;
    FOR_POINTS(r, R) {
        // history slots of the Gram row: loaded with the sums (finish) or after the bead
        // loops (stored pairs), so they are never live beside a group of bead loads
        double hist[NG];
        auto load_hist = [&]() {
            if constexpr (NH > 0) {
#pragma unroll
                for (int m = 1; m <= NH; m++) hist[m] = R.Dh[(long)m * M + r];
            }
        };
        double s[4] = {0.0, 0.0, 0.0, 0.0};
        if (fused) {  // phacc holds 4 M doubles: load all four sums, keep the used ones
            double v[4];
#pragma unroll
            for (int a = 0; a < 4; a++) v[a] = ld_once(R.phacc + (long)a * M + r);
            load_hist();
#pragma unroll
            for (int a = 0; a < 4; a++) s[a] = u[a] ? v[a] : 0.0;
        } else {
            // Beads in groups of PHI_UNROLL, each group's loads issued before its products:
            // one load pair per bead in flight did not cover the memory latency.
            constexpr int U = PHI_UNROLL;
            bool t[4] = {false, false, false, false};
            auto add = [&](int a, double x, double y) {
                s[a] = t[a] ? fma(x, y, s[a]) : x * y;
                t[a] = true;
            };
            int j = i0 + 1;
            for (; j + U - 1 <= N; j += U) {
                double x[U], y[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    x[u] = ld_once(rep_pair(R, j + u) + r);              // q_j
                    y[u] = ld_once(rep_pair(R, N + 1 - j - u) + M + r);  // q+_j
                }
#pragma unroll
                for (int u = 0; u < U; u++) add((j + u <= NA) ? 0 : 1, x[u], y[u]);
            }
            for (; j <= N; j++) add((j <= NA) ? 0 : 1, ld_once(rep_pair(R, j) + r), ld_once(rep_pair(R, N + 1 - j) + M + r));
            j = N - i0;
            for (; j - U + 1 >= 1; j -= U) {
                double x[U], y[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    x[u] = ld_once(rep_pair(R, N + 1 - j + u) + M + r);  // q+_j
                    y[u] = ld_once(rep_pair(R, j - u) + r);              // q_j
                }
#pragma unroll
                for (int u = 0; u < U; u++) add((j - u <= NA) ? 2 : 3, x[u], y[u]);
            }
            for (; j >= 1; j--) add((j <= NA) ? 2 : 3, ld_once(rep_pair(R, N + 1 - j) + M + r), ld_once(rep_pair(R, j) + r));
            load_hist();
        }
        double pA = s[0] + s[2], pB = s[1] + s[3];
        if (N & 1) {  // middle bead: both fields of pair i0
            const double pm = rep_pair(R, i0)[r] * rep_pair(R, i0)[M + r];
            if (i0 <= NA) pA += pm;
            else pB += pm;
        }
        const double phA = pA / (R.N * Q * R.h[r]);
        const double phB = pB / (R.N * Q * R.h[r + M]);
        const double pp = phA + phB;
        R.w[2 * M + r] = phA - phB;
        R.w[3 * M + r] = pp;
        if constexpr (NH >= 0) {
            const double d0 = pp - 1.0;
            acc[0] += d0 * d0;
#pragma unroll
            for (int m = 1; m <= NH; m++) acc[m] += d0 * hist[m];
        }
    }
    if constexpr (NH >= 0) block_sums<NH + 1>(acc, NH + 1, redn, out);
}

// With coef, also the Gram row partials of each member's next Anderson iteration
// (history depth coef[id].nh_dots) into part, as k_am_dots writes them; nh_one >= 0
// gives that depth directly (one member, no coefficient array).
// Q comes from the member's partials (q_from_partials); workgroup 0 also stores it in
// scal[id][0] and in hscal[id][0] (pinned host memory; hscal may be null).
__global__ __launch_bounds__(256) void k_phi(const Rep* __restrict__ reps, double* __restrict__ scal,
                                             double* __restrict__ hscal, const double* __restrict__ qpart, int fuse,
                                             const Coef* __restrict__ coef, int nh_one, double* __restrict__ part) {
    __shared__ double redn[4 * (MAX_HIST + 1)];
    __shared__ double qred[4], qres[1];
    REPLICA(R);
    const double Q = q_from_partials(R, qpart, qred, qres);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal[(long)R.id * SSTRIDE] = Q;
        if (hscal) hscal[(long)R.id * SSTRIDE] = Q;
    }
    const int nh = nh_one >= 0 ? nh_one : (coef ? coef[R.id].nh_dots : -1);
    if (nh < 0) {
        phi_row<-1>(R, Q, fuse, redn, nullptr);
        return;
    }
    double* out = part_row(part, blockIdx.y);
    static_assert(MAX_HIST == 15, "one case per history depth");
    switch (nh) {  // nh_dots <= nhMax <= MAX_HIST (checked at creation)
#define PHI_CASE(k) \
    case k: phi_row<k>(R, Q, fuse, redn, out); break;
        PHI_CASE(0) PHI_CASE(1) PHI_CASE(2) PHI_CASE(3) PHI_CASE(4) PHI_CASE(5) PHI_CASE(6) PHI_CASE(7)
        PHI_CASE(8) PHI_CASE(9) PHI_CASE(10) PHI_CASE(11) PHI_CASE(12) PHI_CASE(13) PHI_CASE(14) PHI_CASE(15)
#undef PHI_CASE
        default: break;
    }
}


// src/anderson.cc:35-55: DD[0][m] = <Dh0, Dh_m>, Dh0 = phi+ - 1 (m = 0 uses Dh0 itself).
// One instantiation per history depth NH: a point's NH history loads are unconditional and
// all in flight before the first product (under a run-time `m <= nh` test each load was
// followed by its own wait).
template <int NH>
__device__ __forceinline__ void am_dots_row(const Rep& R, double* redn, double* out) {
    const long M = R.M;
    double acc[NH + 1];
#pragma unroll
    for (int m = 0; m <= NH; m++) acc[m] = 0.0;
    FOR_POINTS(r, R) {
        const double d0 = R.w[3 * M + r] - 1.0;
        double v[NH + 1];
#pragma unroll
        for (int m = 1; m <= NH; m++) v[m] = R.Dh[(long)m * M + r];
        acc[0] += d0 * d0;
#pragma unroll
        for (int m = 1; m <= NH; m++) acc[m] += d0 * v[m];
    }
    block_sums<NH + 1>(acc, NH + 1, redn, out);
}

__global__ __launch_bounds__(256) void k_am_dots(const Rep* __restrict__ reps, const Coef* __restrict__ coef,
                                                 double* __restrict__ part) {
    __shared__ double redn[4 * (MAX_HIST + 1)];
    REPLICA(R);
    double* out = part_row(part, blockIdx.y);
    static_assert(MAX_HIST == 15, "one case per history depth");
    switch (coef[R.id].nh_dots) {  // 0 <= nh_dots <= nhMax <= MAX_HIST (checked at creation)
#define AM_DOTS_CASE(k) \
    case k: am_dots_row<k>(R, redn, out); break;
        AM_DOTS_CASE(0) AM_DOTS_CASE(1) AM_DOTS_CASE(2) AM_DOTS_CASE(3) AM_DOTS_CASE(4) AM_DOTS_CASE(5)
        AM_DOTS_CASE(6) AM_DOTS_CASE(7) AM_DOTS_CASE(8) AM_DOTS_CASE(9) AM_DOTS_CASE(10) AM_DOTS_CASE(11)
        AM_DOTS_CASE(12) AM_DOTS_CASE(13) AM_DOTS_CASE(14) AM_DOTS_CASE(15)
#undef AM_DOTS_CASE
        default: break;
    }
}

// src/anderson.cc:57-89: w+ += lambda Dh0 + sum_n C_n((wh_n + lambda Dh_n) - (wh0 + lambda Dh0)),
// then (Dh0, wh0) become history slot `slot`.
// One instantiation per history depth NH, so a point's 2 NH history loads are all issued
// before the first update term; the terms are added in the order n = 1 .. NH. With
// with_h the pass also forms the Boltzmann factors of the new fields, with k_prepare_h's
// expressions (src/diblock.cc:64-72), so the next evaluation can skip that pass.
template <int NH>
__device__ __forceinline__ void am_update_row(const Rep& R, const Coef& c, int with_h) {
    const double lambda = c.lambda;
    const int slot = c.slot;
    const long M = R.M;
    double C[NH + 1];
#pragma unroll
    for (int n = 1; n <= NH; n++) C[n] = c.C[n - 1];
    FOR_POINTS(r, R) {
        const double d0 = R.w[3 * M + r] - 1.0;
        const double w0 = R.w[M + r];
        const double wm = with_h ? R.w[r] : 0.0;
        double vw[NH + 1], vd[NH + 1];
#pragma unroll
        for (int n = 1; n <= NH; n++) {
            vw[n] = R.wh[(long)n * M + r];
            vd[n] = R.Dh[(long)n * M + r];
        }
        const double base = w0 + lambda * d0;
        double wn = base;
#pragma unroll
        for (int n = 1; n <= NH; n++) wn += C[n] * ((vw[n] + lambda * vd[n]) - base);
        R.w[M + r] = wn;
        R.Dh[(long)slot * M + r] = d0;
        R.wh[(long)slot * M + r] = w0;
        if (with_h) {
            R.h[r] = exp(-(wm + wn) / R.N);
            R.h[r + M] = exp(-(-wm + wn) / R.N);
        }
    }
}

__device__ __forceinline__ void am_update_any(const Rep& R, const Coef& c, int with_h) {
    switch (c.nh) {  // 0 <= nh <= nhMax <= MAX_HIST (checked at creation)
#define AM_UPDATE_CASE(k) \
    case k: am_update_row<k>(R, c, with_h); break;
        AM_UPDATE_CASE(0) AM_UPDATE_CASE(1) AM_UPDATE_CASE(2) AM_UPDATE_CASE(3) AM_UPDATE_CASE(4)
        AM_UPDATE_CASE(5) AM_UPDATE_CASE(6) AM_UPDATE_CASE(7) AM_UPDATE_CASE(8) AM_UPDATE_CASE(9)
        AM_UPDATE_CASE(10) AM_UPDATE_CASE(11) AM_UPDATE_CASE(12) AM_UPDATE_CASE(13) AM_UPDATE_CASE(14)
        AM_UPDATE_CASE(15)
#undef AM_UPDATE_CASE
        default: break;
    }
}

__global__ __launch_bounds__(256) void k_am_update(const Rep* __restrict__ reps, const Coef* __restrict__ coef,
                                                   int with_h) {
    REPLICA(R);
    am_update_any(R, coef[R.id], with_h);
}

// One member, its coefficients passed by value (no coefficient array to upload).
__global__ __launch_bounds__(256) void k_am_update_one(const Rep* __restrict__ reps, const Coef c, int with_h) {
    REPLICA(R);
    am_update_any(R, c, with_h);
}

// src/langevin.cc:31: W- += -(phi- + 2 W-/XbN) dt + 0.5 (eta_prev + eta_new); with keep_new the
// new noise replaces eta_prev in the same pass (the pointer swap of src/langevin.cc:34-36).
__global__ __launch_bounds__(256) void k_langevin(const Rep* __restrict__ reps, const Coef* __restrict__ coef) {
    REPLICA(R);
    const Coef& c = coef[R.id];
    const double XbN = c.XbN, dt = c.dt;
    const double* en = c.eta_new;
    const int keep = c.keep_new;
    const long M = R.M;
    FOR_POINTS(r, R) {
        const double wm = R.w[r];
        const double a = R.eta_prev[r], b = en[r];
        R.w[r] = wm + (-(R.w[2 * M + r] + 2 * wm / XbN) * dt + 0.5 * (a + b));
        if (keep) R.eta_prev[r] = b;
    }
}

// src/lfts_simulation.cc:130-131: partials of sum w+ and sum W-^2
__global__ __launch_bounds__(256) void k_h_sums(const Rep* __restrict__ reps, double* __restrict__ part) {
    __shared__ double red[8];
    REPLICA(R);
    const long M = R.M;
    double sp = 0.0, s2 = 0.0;
    FOR_POINTS(r, R) {
        sp += R.w[M + r];
        s2 += R.w[r] * R.w[r];
    }
    const double a = block_sum(sp, red);
    const double b = block_sum(s2, red);
    if (threadIdx.x == 0) {
        part_row(part, blockIdx.y)[0] = a;
        part_row(part, blockIdx.y)[1] = b;
    }
}

// The body of src/anderson.cc:29-89 after the Gram row: on one thread, in the order of
// the host loop (the LU is the host's lu_solve, same operations in the same order). The
// history Gram matrix, the row and the LU work arrays live in LDS, so the serial part
// touches no global memory until the results are written back.
#pragma clang fp contract(off)
__global__ __launch_bounds__(64) void k_am_solve(AmState* __restrict__ s, const double* __restrict__ scal,
                                                 Coef* __restrict__ coef, const double* __restrict__ pow09, long M,
                                                 int nhMax) {
    constexpr int HM = MAX_HIST + 1;
    __shared__ double DD[HM * HM];
    __shared__ double U[MAX_HIST * MAX_HIST], V[MAX_HIST], y[MAX_HIST], C[MAX_HIST];
    __shared__ int p[MAX_HIST];
    __shared__ int hdr[2];
    const int H = nhMax + 1;
    if (threadIdx.x == 0) {
        hdr[0] = s->stop;
        hdr[1] = s->k;
    }
    __syncthreads();
    if (hdr[0]) return;
    const int k = hdr[1];
    const int nh = (k < nhMax + 1) ? k - 1 : nhMax;
    for (int i = threadIdx.x; i < H * H; i += blockDim.x) DD[i] = s->DD[i];
    __syncthreads();
    for (int m = threadIdx.x; m <= nh; m += blockDim.x) DD[m] = scal[m];  // DD(0, m)
    __syncthreads();
    if (threadIdx.x != 0) goto writeback;
    {
        const double err = sqrt(DD[0] / (double)M);
        const double lambda = 1.0 - pow09[k];
        int nh_upd = 0;
        if (k >= 2) {
            for (int m = 1; m <= nh; m++) {
                V[m - 1] = DD[0] - DD[m];
                for (int n = 1; n <= m; n++)
                    U[(m - 1) * nh + n - 1] = U[(n - 1) * nh + m - 1] = DD[0] - DD[m] - DD[n] + DD[n * H + m];
            }
            // LU with partial pivoting, then the two triangular solves
            for (int i = 0; i < nh; i++) p[i] = i;
            for (int j = 0; j < nh; j++) {
                int ip = j;
                double amax = fabs(U[j * nh + j]);
                for (int i = j + 1; i < nh; i++) {
                    const double v = fabs(U[i * nh + j]);
                    if (v > amax) {
                        amax = v;
                        ip = i;
                    }
                }
                if (ip != j) {
                    for (int kk = 0; kk < nh; kk++) {
                        const double t = U[j * nh + kk];
                        U[j * nh + kk] = U[ip * nh + kk];
                        U[ip * nh + kk] = t;
                    }
                    const int t = p[j];
                    p[j] = p[ip];
                    p[ip] = t;
                }
                const double ajj = U[j * nh + j];
                if (ajj == 0.0) continue;
                for (int i = j + 1; i < nh; i++) {
                    const double l = U[i * nh + j] / ajj;
                    U[i * nh + j] = l;
                    for (int kk = j + 1; kk < nh; kk++) U[i * nh + kk] -= l * U[j * nh + kk];
                }
            }
            bool sing = false;
            for (int i = 0; i < nh; i++) sing |= U[i * nh + i] == 0.0;
            if (sing) {
                // the host LU's zero-pivot rule: no solve, the loop stops, the call fails
                s->singular = 1;
            } else {
                for (int i = 0; i < nh; i++) {
                    double t = V[p[i]];
                    for (int kk = 0; kk < i; kk++) t -= U[i * nh + kk] * y[kk];
                    y[i] = t;
                }
                for (int i = nh - 1; i >= 0; i--) {
                    double t = y[i];
                    for (int kk = i + 1; kk < nh; kk++) t -= U[i * nh + kk] * C[kk];
                    C[i] = t / U[i * nh + i];
                }
                nh_upd = nh;
            }
        }
        const int slot = 1 + (k - 1) % nhMax;
        DD[slot * H + slot] = DD[0];
        for (int m = 1; m < slot; m++) DD[m * H + slot] = DD[m];
        for (int m = slot + 1; m <= nh; m++) DD[slot * H + m] = DD[m];
        Coef& c = coef[0];
        c.lambda = lambda;
        c.nh = nh_upd;
        c.slot = slot;
        for (int i = 0; i < nh_upd; i++) c.C[i] = C[i];
        c.nh_dots = (k + 1 < nhMax + 1) ? k : nhMax;
        s->err = err;
        s->k = k + 1;
        s->stop_next = s->singular || !(k + 1 < s->maxIter && err > s->tol);
    }
writeback:
    __syncthreads();
    for (int i = threadIdx.x; i < H * H; i += blockDim.x) s->DD[i] = DD[i];
}

__global__ void k_am_advance(AmState* __restrict__ s) {
    if (threadIdx.x == 0 && !s->stop) s->stop = s->stop_next;
}

#define LAUNCH(kernel, grid, ...)                                            \
    do {                                                                     \
        hipLaunchKernelGGL(kernel, grid, dim3(256), 0, st, __VA_ARGS__);     \
        return hipGetLastError();                                            \
    } while (0)

hipError_t launch_prepare_h(const Rep* reps, int nact, int nb_max, hipStream_t st) {
    LAUNCH(k_prepare_h, dim3(nb_max, nact), reps);
}
hipError_t launch_q_sum(const Rep* reps, int nact, int nb_max, double* qpart, hipStream_t st) {
    LAUNCH(k_q_sum, dim3(nb_max, nact), reps, qpart);
}
hipError_t launch_reduce(const Rep* reps, int nact, const double* part, int ncomp, int div_by_M, double* scal,
                         int scal_off, hipStream_t st, double* hscal, unsigned* hseq, unsigned seq) {
    if (ncomp < 1 || ncomp > PSTRIDE || scal_off < 0 || scal_off + ncomp > SSTRIDE) return hipErrorInvalidValue;
    LAUNCH(k_reduce, dim3(nact), reps, part, ncomp, div_by_M, scal, scal_off, hscal, hseq, seq);
}
hipError_t launch_phi(const Rep* reps, int nact, int nb_max, double* scal, const double* qpart, int fuse,
                      const Coef* coef, int nh_one, double* part, hipStream_t st, double* hscal) {
    if (nh_one > MAX_HIST || (nh_one >= 0 && nact != 1)) return hipErrorInvalidValue;
    LAUNCH(k_phi, dim3(nb_max, nact), reps, scal, hscal, qpart, fuse, coef, nh_one, part);
}
hipError_t launch_am_dots(const Rep* reps, int nact, int nb_max, const Coef* coef, double* part, hipStream_t st) {
    LAUNCH(k_am_dots, dim3(nb_max, nact), reps, coef, part);
}
hipError_t launch_am_update(const Rep* reps, int nact, int nb_max, const Coef* coef, int with_h, hipStream_t st) {
    LAUNCH(k_am_update, dim3(nb_max, nact), reps, coef, with_h);
}
hipError_t launch_am_update_one(const Rep* reps, int nb_max, const Coef& c, int with_h, hipStream_t st) {
    if (c.nh < 0 || c.nh > MAX_HIST) return hipErrorInvalidValue;
    LAUNCH(k_am_update_one, dim3(nb_max, 1), reps, c, with_h);
}
hipError_t launch_langevin(const Rep* reps, int nact, int nb_max, const Coef* coef, hipStream_t st) {
    LAUNCH(k_langevin, dim3(nb_max, nact), reps, coef);
}
hipError_t launch_h_sums(const Rep* reps, int nact, int nb_max, double* part, hipStream_t st) {
    LAUNCH(k_h_sums, dim3(nb_max, nact), reps, part);
}

hipError_t launch_am_solve(AmState* s, const double* scal, Coef* coef, const double* pow09, long M, int nhMax,
                           hipStream_t st) {
    hipLaunchKernelGGL(k_am_solve, dim3(1), dim3(64), 0, st, s, scal, coef, pow09, M, nhMax);
    return hipGetLastError();
}
hipError_t launch_am_advance(AmState* s, hipStream_t st) {
    hipLaunchKernelGGL(k_am_advance, dim3(1), dim3(64), 0, st, s);
    return hipGetLastError();
}

}  // namespace lfts
