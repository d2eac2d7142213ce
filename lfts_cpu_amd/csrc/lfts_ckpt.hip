// Checkpointed propagator storage: for systems whose N pairs of 2M doubles do not fit in
// HBM, only the pairs at segment boundaries are kept (about sqrt(N) of them) and each
// segment is recomputed when its pairs are needed.
//
// The pair layout of the reference (src/diblock.cc:19-24) puts q_j and q+_{N+1-j} in pair
// j, so every sum over beads (phi, its tangent, the adjoint's final sums) multiplies a
// value of pair j with a value of pair N+1-j. The segments are therefore chosen symmetric
// under j -> N+1-j: segment s and segment K-1-s are mirrors, and processing them together
// (both recomputed into buffers) covers the beads of both. One evaluation costs two chains
// (the sweep that leaves the checkpoints, then one recompute of every segment); the
// tangent and adjoint chains run in lockstep with a forward recompute that supplies their
// sources.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "lfts_gpu.h"
#include "lfts_internal.h"
#include "lfts_state.h"

namespace lfts {

namespace {

#define FOR_R(r, n) for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < (n); r += (long)gridDim.x * blockDim.x)

}  // namespace

enum { MIRROR_PHI = 0, MIRROR_TAN = 1, MIRROR_ADJ = 2 };

struct MirrorArgs {
    long M;
    int N, NA, mode;
    int j0, j1;                         // beads (pairs) j0..j1 of X
    const double* X;                    // forward pairs from x0: pair j at X + 2 (j - x0) M
    int x0;
    const double* Y;                    // forward pairs from y0 (the mirror segment)
    int y0;
    const double* DX;                   // derived-chain pairs at the places of X
    const double* DY;                   // derived-chain pairs at the places of Y
    double* acc;
};

// Raw bead sums of one segment against its mirror:
//   PHI: S_X += q_j q+_j
//   TAN: S_X += q_j q+_j, dS_X += dq_j q+_j + q_j dq+_j
//   ADJ: G_X += qbar_j q_j (q chain's factor of bead j), q+bar_m q+_m (q+ chain's factor)
// with q_j from pair j, q+_j from pair N+1-j (src/diblock.cc:51-53).
__global__ __launch_bounds__(256) void k_mirror_acc(const MirrorArgs a) {
    const long M = a.M;
    const int N = a.N, NA = a.NA;
    FOR_R(r, M) {
        double s0 = 0.0, s1 = 0.0, d0 = 0.0, d1 = 0.0;
        for (int j = a.j0; j <= a.j1; j++) {
            const int jm = N + 1 - j;
            const double* P = a.X + 2L * (j - a.x0) * M;
            const double* Pm = a.Y + 2L * (jm - a.y0) * M;
            if (a.mode == MIRROR_PHI) {
                const double v = P[r] * Pm[M + r];
                if (j <= NA) s0 += v; else s1 += v;
            } else if (a.mode == MIRROR_TAN) {
                const double* D = a.DX + 2L * (j - a.x0) * M;
                const double* Dm = a.DY + 2L * (jm - a.y0) * M;
                const double q = P[r], qd = Pm[M + r];
                if (j <= NA) {
                    s0 += q * qd;
                    d0 += D[r] * qd + q * Dm[M + r];
                } else {
                    s1 += q * qd;
                    d1 += D[r] * qd + q * Dm[M + r];
                }
            } else {
                // adjoint pair j = [qbar_{N+1-j} | q+bar_j] against forward pair N+1-j
                const double* A = a.DX + 2L * (j - a.x0) * M;
                const double x = A[r] * Pm[r];          // bead N+1-j of the q chain
                const double y = A[M + r] * Pm[M + r];  // bead j of the q+ chain
                if (jm == 1 || jm <= NA) s0 += x; else s1 += x;
                if (j != N && j <= NA) s0 += y; else s1 += y;
            }
        }
        a.acc[r] += s0;
        a.acc[M + r] += s1;
        if (a.mode == MIRROR_TAN) {
            a.acc[2 * M + r] += d0;
            a.acc[3 * M + r] += d1;
        }
    }
}

// phi from the raw sums (src/diblock.cc:85-95)
__global__ __launch_bounds__(256) void k_ck_phi(long M, int N, const double* __restrict__ acc,
                                                const double* __restrict__ h, const double* __restrict__ Qp,
                                                double* __restrict__ w) {
    const double Q = *Qp;
    FOR_R(r, M) {
        const double pA = acc[r] / (N * Q * h[r]);
        const double pB = acc[M + r] / (N * Q * h[M + r]);
        w[2 * M + r] = pA - pB;
        w[3 * M + r] = pA + pB;
    }
}

// d phi_X = dS_X / (N Q h_X) - phi_X (dQ/Q + e_X)
__global__ __launch_bounds__(256) void k_ck_tan(long M, int N, const double* __restrict__ acc,
                                                const double* __restrict__ h, const double* __restrict__ e,
                                                const double* __restrict__ Qp, const double* __restrict__ dQp,
                                                double* __restrict__ dphi) {
    const double Q = *Qp, dQ = *dQp;
    FOR_R(r, M) {
        const double hA = h[r], hB = h[M + r];
        const double pA = acc[r] / (N * Q * hA), pB = acc[M + r] / (N * Q * hB);
        const double dpA = acc[2 * M + r] / (N * Q * hA) - pA * (dQ / Q + e[r]);
        const double dpB = acc[3 * M + r] / (N * Q * hB) - pB * (dQ / Q + e[M + r]);
        dphi[r] = dpA - dpB;
        dphi[M + r] = dpA + dpB;
    }
}

// dL/dw from the raw sums and the normalisation terms -u_X phi_X
__global__ __launch_bounds__(256) void k_ck_adj(long M, int N, const double* __restrict__ acc,
                                                const double* __restrict__ ubar, const double* __restrict__ w,
                                                double* __restrict__ wbar) {
    FOR_R(r, M) {
        const double uA = ubar[M + r] + ubar[r], uB = ubar[M + r] - ubar[r];
        const double phA = 0.5 * (w[3 * M + r] + w[2 * M + r]);
        const double phB = 0.5 * (w[3 * M + r] - w[2 * M + r]);
        const double GA = acc[r] - uA * phA, GB = acc[M + r] - uB * phB;
        wbar[r] = -(GA - GB) / N;
        wbar[M + r] = -(GA + GB) / N;
    }
}

// out = [f0 a | f1 b]: an adjoint pair times the factors its chain continues with
__global__ __launch_bounds__(256) void k_pair_mul(long M, const double* __restrict__ in, const double* __restrict__ f0,
                                                  const double* __restrict__ f1, double* __restrict__ out) {
    FOR_R(r, M) {
        out[r] = f0[r] * in[r];
        out[M + r] = f1[r] * in[M + r];
    }
}

__global__ __launch_bounds__(256) void k_ck_fill(double* __restrict__ x, long n) {
    FOR_R(r, n) x[r] = 0.0;
}

}  // namespace lfts

namespace lfts_abi {

namespace {

int K(const lfts_ctx* c) { return (int)c->ck_b.size() - 1; }
int bnd(const lfts_ctx* c, int s) { return c->ck_b[s]; }

// Stored forward (derived) pair j, or nullptr: pairs at segment starts b_1..b_{K-1} and
// pair N are kept; pair 1 of the forward chain is h.
double* ck_slot(lfts_ctx* c, double* base, int j) {
    const int k = K(c);
    if (j == c->N) return base + 2L * (k - 1) * c->M;
    for (int s = 1; s < k; s++)
        if (c->ck_b[s] == j) return base + 2L * (s - 1) * c->M;
    return nullptr;
}

double* SEG(lfts_ctx* c, int which) { return c->ck_seg + 2L * which * c->ck_smax * c->M; }
double* RING(lfts_ctx* c, int k) { return c->ck_ring + 2L * k * c->M; }

int fill0(lfts_ctx* c, double* x, long n) {
    hipLaunchKernelGGL(lfts::k_ck_fill, dim3(lfts::blocks_for(n)), dim3(256), 0, c->self->st, x, n);
    HIPCHK(c, hipGetLastError());
    return LFTS_OK;
}

int copy_pair(lfts_ctx* c, double* dst, const double* src) {
    HIPCHK(c, hipMemcpyAsync(dst, src, sizeof(double) * 2 * c->M, hipMemcpyDeviceToDevice, c->self->st));
    return LFTS_OK;
}

const double* fwd_start(lfts_ctx* c, int s) { return s == 0 ? c->h : ck_slot(c, c->ck_f, bnd(c, s)); }

// Forward segment s recomputed into buf (pair j at buf + 2 (j - b_s) M).
int recompute_fwd(lfts_ctx* c, int s, double* buf) {
    if (int rc = copy_pair(c, buf, fwd_start(c, s))) return rc;
    const long M = c->M;
    const int b0 = bnd(c, s), b1 = bnd(c, s + 1) - 1;
    ChainSpec F;
    F.start = buf;
    F.T0 = c->T0;
    F.T1 = c->T1;
    F.out = [buf, b0, M](int j) { return buf + 2L * (j - b0) * M; };
    ChainSpec* cs[1] = {&F};
    return run_lockstep(c, b0, b1, cs, 1);
}

// Forward segment s into fbuf and the derived chain's segment s into dbuf, in lockstep.
// dstart: the derived pair b_s as stored (slot 0 of dbuf); for the adjoint, its chain
// continues from that pair times the factors of forward pass N - b_s.
int recompute_pair(lfts_ctx* c, int mode, int s, double* fbuf, double* dbuf, const double* dstart) {
    const long M = c->M;
    const int b0 = bnd(c, s), b1 = bnd(c, s + 1) - 1;
    if (int rc = copy_pair(c, fbuf, fwd_start(c, s))) return rc;
    if (int rc = copy_pair(c, dbuf, dstart)) return rc;
    ChainSpec F, D;
    F.start = fbuf;
    F.T0 = c->T0;
    F.T1 = c->T1;
    F.out = [fbuf, b0, M](int j) { return fbuf + 2L * (j - b0) * M; };
    D.mode = mode;
    D.T0 = c->ck_T;
    D.T1 = c->ck_T + 2 * M;
    D.aux = c->dbuf + 4 * M;  // e or alpha
    D.out = [dbuf, b0, M](int j) { return dbuf + 2L * (j - b0) * M; };
    D.fwd = F.out;
    if (mode == lfts::CHAIN_ADJOINT) {
        double* st = RING(c, 3);
        const int q = c->N - b0;  // the continuation uses the factors of forward pass N - b_s
        const double* f0 = (q < c->NA) ? c->h : c->h + M;
        const double* f1 = (q < c->NB) ? c->h + M : c->h;
        hipLaunchKernelGGL(lfts::k_pair_mul, dim3(lfts::blocks_for(M)), dim3(256), 0, c->self->st, M, dstart, f0, f1,
                           st);
        HIPCHK(c, hipGetLastError());
        D.start = st;
        D.post = RING(c, 2);
    } else {
        D.start = dbuf;
    }
    ChainSpec* cs[2] = {&F, &D};
    return run_lockstep(c, b0, b1, cs, 2);
}

int mirror_acc(lfts_ctx* c, int mode, int s, int sm, const double* X, const double* Y, const double* DX,
               const double* DY) {
    lfts::MirrorArgs a{};
    a.M = c->M;
    a.N = c->N;
    a.NA = c->NA;
    a.mode = mode;
    a.j0 = bnd(c, s);
    a.j1 = bnd(c, s + 1) - 1;
    a.X = X;
    a.x0 = bnd(c, s);
    a.Y = Y;
    a.y0 = bnd(c, sm);
    a.DX = DX;
    a.DY = DY;
    a.acc = c->ck_acc;
    hipLaunchKernelGGL(lfts::k_mirror_acc, dim3(lfts::blocks_for(c->M)), dim3(256), 0, c->self->st, a);
    HIPCHK(c, hipGetLastError());
    return LFTS_OK;
}

// Symmetric segment starts for stride S: an odd number of segments whose lengths read the
// same forwards and backwards, so segment s and K-1-s are mirrors under j -> N+1-j.
std::vector<int> symmetric_bounds(int N, int S) {
    int Kc = (N + S - 1) / S;
    if (Kc % 2 == 0) Kc++;
    Kc = std::min(Kc, N % 2 ? N : N - 1);
    std::vector<int> len(Kc, N / Kc);
    int rem = N % Kc;
    if (rem % 2) {
        len[Kc / 2]++;
        rem--;
    }
    for (int t = 0; t < rem / 2; t++) {
        len[t]++;
        len[Kc - 1 - t]++;
    }
    std::vector<int> b(Kc + 1, 1);
    for (int s = 0; s < Kc; s++) b[s + 1] = b[s] + len[s];
    return b;
}

void free_ckpt(lfts_ctx* c) {
    for (double** p : {&c->ck_f, &c->ck_d, &c->ck_seg, &c->ck_ring, &c->ck_T, &c->ck_acc}) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    c->ck_b.clear();
    c->ck_stride = 0;
    c->ck_smax = 0;
}

}  // namespace

// Checkpointed propagator evaluation: the sweep that keeps the boundary pairs and pair N,
// Q, then phi from every segment and its mirror (src/diblock.cc:30-60).
int ckpt_calc_concs(lfts_ctx* c) {
    hipStream_t st = c->self->st;
    const long M = c->M;
    const int N = c->N, Kc = K(c);
    lfts_batch* b = c->self;
    BHIPCHK(b, lfts::launch_prepare_h(b->d_ract, 1, b->nb_max, st));
    ChainSpec F;
    F.start = c->h;
    F.T0 = c->T0;
    F.T1 = c->T1;
    F.out = [c](int j) {
        double* s = ck_slot(c, c->ck_f, j);
        return s ? s : RING(c, j % 2);
    };
    ChainSpec* cs[1] = {&F};
    if (int rc = run_lockstep(c, 1, N, cs, 1)) return rc;
    if (int rc = dev_sum(c, ck_slot(c, c->ck_f, N), nullptr, M, 1.0 / (double)M, c->self->d_scal)) return rc;
    if (int rc = fill0(c, c->ck_acc, 2 * M)) return rc;
    for (int s = 0; s <= (Kc - 1) / 2; s++) {
        const int sm = Kc - 1 - s;
        double* A = SEG(c, 0);
        double* B = (sm == s) ? A : SEG(c, 1);
        if (int rc = recompute_fwd(c, s, A)) return rc;
        if (sm != s)
            if (int rc = recompute_fwd(c, sm, B)) return rc;
        if (int rc = mirror_acc(c, lfts::MIRROR_PHI, s, sm, A, B, nullptr, nullptr)) return rc;
        if (sm != s)
            if (int rc = mirror_acc(c, lfts::MIRROR_PHI, sm, s, B, A, nullptr, nullptr)) return rc;
    }
    hipLaunchKernelGGL(lfts::k_ck_phi, dim3(lfts::blocks_for(M)), dim3(256), 0, st, M, N, (const double*)c->ck_acc,
                       (const double*)c->h, (const double*)c->self->d_scal, c->w);
    HIPCHK(c, hipGetLastError());
    c->n_pe++;
    c->n_cs += 2L * (N - 1);
    return LFTS_OK;
}

// Derived chain (tangent or adjoint) over all beads beside a forward recompute, keeping its
// boundary pairs and pair N; then the mirror sums of every segment pair into ck_acc.
// d1: derived pair 1 as stored; dstart1: the chain's start (the adjoint's multiplied pair).
int ckpt_derived(lfts_ctx* c, int mode, const double* d1, const double* dstart1) {
    const long M = c->M;
    const int N = c->N, Kc = K(c);
    ChainSpec F, D;
    F.start = c->h;
    F.T0 = c->T0;
    F.T1 = c->T1;
    F.out = [c](int j) { return RING(c, j % 2); };
    D.mode = mode;
    D.start = dstart1;
    D.T0 = c->ck_T;
    D.T1 = c->ck_T + 2 * M;
    D.post = RING(c, 2);
    D.aux = c->dbuf + 4 * M;
    D.out = [c](int j) {
        double* s = ck_slot(c, c->ck_d, j);
        return s ? s : RING(c, 3);
    };
    D.fwd = F.out;
    ChainSpec* cs[2] = {&F, &D};
    if (int rc = run_lockstep(c, 1, N, cs, 2)) return rc;
    if (int rc = fill0(c, c->ck_acc, (mode == lfts::CHAIN_TANGENT ? 4 : 2) * M)) return rc;
    const int mm = (mode == lfts::CHAIN_TANGENT) ? lfts::MIRROR_TAN : lfts::MIRROR_ADJ;
    auto dstart = [&](int s) -> const double* { return s == 0 ? d1 : ck_slot(c, c->ck_d, bnd(c, s)); };
    for (int s = 0; s <= (Kc - 1) / 2; s++) {
        const int sm = Kc - 1 - s;
        double *FA = SEG(c, 0), *FB = SEG(c, 1), *DA = SEG(c, 2), *DB = SEG(c, 3);
        if (sm == s) {
            FB = FA;
            DB = DA;
        }
        if (int rc = recompute_pair(c, mode, s, FA, DA, dstart(s))) return rc;
        if (sm != s)
            if (int rc = recompute_pair(c, mode, sm, FB, DB, dstart(sm))) return rc;
        if (int rc = mirror_acc(c, mm, s, sm, FA, FB, DA, DB)) return rc;
        if (sm != s)
            if (int rc = mirror_acc(c, mm, sm, s, FB, FA, DB, DA)) return rc;
    }
    c->n_cs += 3L * (N - 1);
    return LFTS_OK;
}

int ckpt_tan_finish(lfts_ctx* c, const double* e, const double* dQ, double* dphi) {
    hipLaunchKernelGGL(lfts::k_ck_tan, dim3(lfts::blocks_for(c->M)), dim3(256), 0, c->self->st, c->M, c->N,
                       (const double*)c->ck_acc, (const double*)c->h, e, (const double*)c->self->d_scal, dQ, dphi);
    HIPCHK(c, hipGetLastError());
    return LFTS_OK;
}

int ckpt_adj_finish(lfts_ctx* c, const double* ubar, double* wbar) {
    hipLaunchKernelGGL(lfts::k_ck_adj, dim3(lfts::blocks_for(c->M)), dim3(256), 0, c->self->st, c->M, c->N,
                       (const double*)c->ck_acc, ubar, (const double*)c->w, wbar);
    HIPCHK(c, hipGetLastError());
    return LFTS_OK;
}

const double* ckpt_pair_N(lfts_ctx* c, bool derived) { return ck_slot(c, derived ? c->ck_d : c->ck_f, c->N); }

// Beads i = 1..N-1 with q_i and q+_{i+1} (the box derivative's products), after a
// checkpointed evaluation. q_i is field 0 of pair i and q+_{i+1} field 1 of pair N - i: one
// place off the mirror N + 1 - i, so segment s against its mirror covers i = b_s ..
// b_{s+1} - 2, and the segment's last bead takes q+ from the next segment's mirror, with
// q_i carried over. Every segment is recomputed twice (as itself and as a mirror).
int ckpt_for_bead_pairs(lfts_ctx* c, const std::function<int(int, const double*, const double*)>& fn) {
    const long M = c->M;
    const int N = c->N, Kc = K(c);
    double* X = SEG(c, 0);
    double* carry = RING(c, 0);  // q_i of the previous segment's last bead
    for (int s = 0; s < Kc; s++) {
        const int sm = Kc - 1 - s;
        double* Y = (sm == s) ? X : SEG(c, 1);
        if (int rc = recompute_fwd(c, s, X)) return rc;
        if (sm != s)
            if (int rc = recompute_fwd(c, sm, Y)) return rc;
        const int b0 = bnd(c, s), b1 = bnd(c, s + 1) - 1, y0 = bnd(c, sm);
        auto qd = [&](int i) { return Y + 2L * ((N - i) - y0) * M + M; };  // q+_{i+1}
        if (s > 0)
            if (int rc = fn(b0 - 1, carry, qd(b0 - 1))) return rc;
        for (int i = b0; i < b1; i++)
            if (int rc = fn(i, X + 2L * (i - b0) * M, qd(i))) return rc;
        if (b1 < N) HIPCHK(c, hipMemcpyAsync(carry, X + 2L * (b1 - b0) * M, sizeof(double) * M,
                                             hipMemcpyDeviceToDevice, c->self->st));
    }
    c->n_cs += 2L * (N - 1);
    return LFTS_OK;
}

// stride 0: every pair stored; stride S >= 2: pairs kept at about N/S symmetric segment
// boundaries (the derivative chains' storage shrinks the same way).
int set_pair_storage(lfts_ctx* c, int stride) {
    const long M = c->M;
    const int N = c->N;
    if (int rc = sync_batch(c->self)) return rc;
    (void)hipFree(c->dq);
    c->dq = nullptr;
    if (stride <= 1 || N < 4) {
        free_ckpt(c);
        if (!c->q) HIPCHK(c, hipMalloc(&c->q, sizeof(double) * 2 * M * (size_t)(N - 1)));
        c->self->reps[0].q = c->q;
        return LFTS_OK;
    }
    (void)hipFree(c->q);
    c->q = nullptr;
    c->self->reps[0].q = nullptr;
    free_ckpt(c);
    c->ck_b = symmetric_bounds(N, stride);
    c->ck_stride = stride;
    const int Kc = K(c);
    for (int s = 0; s < Kc; s++) c->ck_smax = std::max(c->ck_smax, c->ck_b[s + 1] - c->ck_b[s]);
    HIPCHK(c, hipMalloc(&c->ck_f, sizeof(double) * 2 * M * (size_t)Kc));
    HIPCHK(c, hipMalloc(&c->ck_d, sizeof(double) * 2 * M * (size_t)Kc));
    HIPCHK(c, hipMalloc(&c->ck_seg, sizeof(double) * 2 * M * (size_t)(4 * c->ck_smax)));
    HIPCHK(c, hipMalloc(&c->ck_ring, sizeof(double) * 2 * M * 4));
    HIPCHK(c, hipMalloc(&c->ck_T, sizeof(double) * 4 * M));
    HIPCHK(c, hipMalloc(&c->ck_acc, sizeof(double) * 4 * M));
    return LFTS_OK;
}

void release_pair_storage(lfts_ctx* c) { free_ckpt(c); }

}  // namespace lfts_abi

using namespace lfts_abi;

extern "C" {

int lfts_set_pair_storage(lfts_ctx* c, int stride) {
    if (!c || stride < 0) return fail(c, LFTS_EINVAL, "set_pair_storage: bad argument");
    // a replica set in use holds copies of the pair addresses, which this call replaces
    if (c->n_sets > 0) return fail(c, LFTS_ESTATE, "set_pair_storage: destroy the replica set holding this context first");
    // so does the captured device Anderson iteration (descriptor passed by value)
    if (c->am_exec) (void)hipGraphExecDestroy(c->am_exec);
    if (c->am_graph) (void)hipGraphDestroy(c->am_graph);
    c->am_exec = nullptr;
    c->am_graph = nullptr;
    return set_pair_storage(c, stride);
}

int lfts_pair_storage(lfts_ctx* c, int* stride, int* nkept) {
    if (!c) return LFTS_EINVAL;
    if (stride) *stride = c->ck_stride;
    if (nkept) *nkept = c->ck_stride ? (int)c->ck_b.size() - 1 : c->N;
    return LFTS_OK;
}

}  // extern "C"
