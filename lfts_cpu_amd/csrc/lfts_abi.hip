// C-ABI layer of the L-FTS Langevin-step library.
//
// A context (lfts_ctx) owns one system's device state and plans. All work runs on
// replica sets (lfts_batch): a context's own calls use its private set of one, and
// lfts_batch_* drives many contexts together, one launch per contour step for every
// mesh size present. The Anderson least-squares solve (<= 15 x 15 LU), the random
// stream of the reference and the S(k) binning stay on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstddef>
#include <cstring>
#include <fstream>
#include <mutex>
#include <numeric>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "lfts_gpu.h"
#include "lfts_internal.h"
#include "lfts_state.h"

using lfts::Coef;
using lfts::LineArgs;
using lfts::Rep;

namespace lfts_abi {

int fail(lfts_ctx* c, int code, const std::string& msg) {
    if (c) {
        c->err = msg;
        c->err_lit = nullptr;
    }
    return code;
}
int bfail(lfts_batch* b, int code, const std::string& msg) {
    if (b) {
        b->err = msg;
        for (lfts_ctx* c : b->ctx) {
            c->err = msg;
            c->err_lit = nullptr;
        }
    }
    return code;
}

// Stage radices of a line transform of n points: 4s, 2s, 3, 5 and 7, then any larger primes
// (direct-DFT stages of the generic line kernel, for lines of up to lfts::MAX_PRIME_LINE
// points, whose two LDS tiles then stay within 64 KB).
bool factorize(int n, std::vector<int>& r) {
    r.clear();
    if (n < 1) return false;
    const int n0 = n;
    while (n % 4 == 0) { r.push_back(4); n /= 4; }
    while (n % 2 == 0) { r.push_back(2); n /= 2; }
    for (int p : {3, 5, 7})
        while (n % p == 0) { r.push_back(p); n /= p; }
    if (n == 1) return true;
    for (int p = 11; p * p <= n; p += 2)
        while (n % p == 0) { r.push_back(p); n /= p; }
    if (n > 1) r.push_back(n);
    return n0 <= lfts::MAX_PRIME_LINE;
}

// Position of frequency k after the in-place decimation-in-frequency pass: a stage of
// radix r on a block of size B leaves the sub-DFT (k mod r) in sub-block k mod r, so
// pos_B(k) = (B/r)(k mod r) + pos_{B/r}(k div r).
int dif_pos(int k, const std::vector<int>& rad, size_t s, int B) {
    if (s == rad.size()) return 0;
    const int r = rad[s], S = B / r;
    return S * (k % r) + dif_pos(k / r, rad, s + 1, S);
}

int build_axis(lfts_ctx* c, AxisPlan& p, int n, double L, int N) {
    std::vector<int> rad;
    if (!factorize(n, rad) || (int)rad.size() > lfts::MAXST)
        return fail(c, LFTS_EINVAL, "mesh size " + std::to_string(n) + " is not supported");
    p.n = n;
    p.nst = (int)rad.size();
    for (int s = 0; s < p.nst; s++) p.radix[s] = rad[s];
    std::vector<double2> tw(n);
    for (int k = 0; k < n; k++) {
        const double a = -2.0 * M_PI * (double)k / (double)n;
        tw[k] = make_double2(std::cos(a), std::sin(a));
    }
    // Per-axis factor of g(k) = exp(-K^2/(6N))/M, K = 2 pi |k|/L (src/step.cc:84-85),
    // with the reference's frequency convention k in [-(n-1)/2, n/2] (src/step.cc:73-83).
    // kap = (d g/dL)/g = (2 pi f)^2 / (3 N L^3), for the box derivative
    std::vector<double> gperm(n), gnat(n), kap(n);
    std::vector<int> perm(n);
    for (int k = 0; k < n; k++) {
        const int f = (k <= n / 2) ? k : k - n;
        const double K = 2.0 * M_PI * (double)f / L;
        gnat[k] = std::exp(-K * K / (6.0 * N)) / (double)n;
        kap[k] = K * K / (3.0 * N * L);
        perm[k] = dif_pos(k, rad, 0, n);
        gperm[perm[k]] = gnat[k];
    }
    HIPCHK(c, hipMalloc(&p.tw, sizeof(double2) * n));
    HIPCHK(c, hipMalloc(&p.gperm, sizeof(double) * n));
    HIPCHK(c, hipMalloc(&p.perm, sizeof(int) * n));
    HIPCHK(c, hipMalloc(&p.gnat, sizeof(double) * n));
    HIPCHK(c, hipMalloc(&p.kap, sizeof(double) * n));
    HIPCHK(c, hipMemcpy(p.kap, kap.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(p.tw, tw.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(p.gperm, gperm.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(p.perm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(p.gnat, gnat.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    p.pitch = n + 1;
    p.T = std::max(1, std::min(16, 4096 / n));
    bool prime = false;  // a stage of radix 11 or more: the kernel holds two tiles of T lines
    for (int s = 0; s < p.nst; s++) prime = prime || p.radix[s] > 7;
    if (prime) p.T = std::max(1, std::min(16, 2048 / p.pitch));
    return LFTS_OK;
}

Rep make_rep(const lfts_ctx* c) {
    Rep r;
    std::memset(&r, 0, sizeof r);
    r.w = c->w;
    r.h = c->h;
    r.q = c->q;
    r.T0 = c->T0;
    r.T1 = c->T1;
    r.phacc = c->phacc;
    for (int d = 0; d < 3; d++) r.g[d] = c->ax[d].gnat;
    r.Dh = c->Dh;
    r.wh = c->wh;
    r.eta_prev = c->noise;
    r.M = c->M;
    r.N = c->N;
    r.NA = c->NA;
    r.NB = c->NB;
    r.n = c->fused ? c->d.m[0] : 0;
    r.id = 0;
    r.nb = lfts::blocks_for(c->M);
    return r;
}

LineArgs line_args(const lfts_ctx* c, int axis, int mode) {
    const AxisPlan& p = c->ax[axis];
    LineArgs a;
    std::memset(&a, 0, sizeof a);
    a.n = p.n;
    a.nst = p.nst;
    for (int s = 0; s < p.nst; s++) a.radix[s] = p.radix[s];
    a.tw = p.tw;
    a.gperm = p.gperm;
    a.perm = p.perm;
    a.gnat = p.gnat;
    a.fast = c->fast_lines ? 1 : 0;
    a.axis = axis;
    a.m0 = c->d.m[0];
    a.m1 = c->d.m[1];
    a.m2 = c->d.m[2];
    a.nlines = c->M / p.n;
    a.stride = (axis == 2) ? 1 : (axis == 1 ? c->d.m[2] : (long)c->d.m[1] * c->d.m[2]);
    a.T = p.T;
    a.pitch = p.pitch;
    a.mode = mode;
    return a;
}

double* pair_ptr(lfts_ctx* c, int i) { return (i == 1) ? c->h : c->q + 2L * (i - 2) * c->M; }

// Line path, src/step.cc:35-65: out = h_sel * conv_g(in) for the pair (q_i, q+_{N+1-i}),
// as three launches (z, y, then x with the Boltzmann factor); `tmp` holds 2M doubles.
// skip: the device flag that makes the launches return at once (the replica's flag in the
// set that runs it), or null.
int line_contour_step(lfts_ctx* c, const double* in, double* out, const double* h, int i, double* tmp,
                      hipStream_t st, const int* skip) {
    const long M = c->M;
    LineArgs z = line_args(c, 2, lfts::LINE_CONV);
    z.inA = in; z.inB = in + M; z.outA = tmp; z.outB = tmp + M;
    z.skip = skip;
    HIPCHK(c, lfts::launch_line(z, st));
    LineArgs y = line_args(c, 1, lfts::LINE_CONV);
    y.inA = tmp; y.inB = tmp + M; y.outA = tmp; y.outB = tmp + M;
    y.skip = skip;
    HIPCHK(c, lfts::launch_line(y, st));
    LineArgs x = line_args(c, 0, lfts::LINE_CONV);
    x.inA = tmp; x.inB = tmp + M; x.outA = out; x.outB = out + M;
    x.skip = skip;
    x.hA = (i < c->NA) ? h : h + M;       // q_{i+1}   (src/step.cc:53-57)
    x.hB = (i < c->NB) ? h + M : h;       // q+_{N-i} (src/step.cc:60-64)
    HIPCHK(c, lfts::launch_line(x, st));
    return LFTS_OK;
}

// LU with partial pivoting and solve (the GSL LU_decomp/LU_solve of src/anderson.cc:118-126).
// Returns false, leaving x untouched, when U has a zero pivot: GSL's LU_solve then reports
// the matrix as singular (and its default error handler aborts the reference run).
bool lu_solve(int n, const double* A_in, const double* b, double* x) {
    std::vector<double> A(A_in, A_in + (size_t)n * n);
    std::vector<int> p(n);
    std::iota(p.begin(), p.end(), 0);
    for (int j = 0; j < n; j++) {
        int ip = j;
        double amax = std::fabs(A[(size_t)j * n + j]);
        for (int i = j + 1; i < n; i++) {
            const double v = std::fabs(A[(size_t)i * n + j]);
            if (v > amax) { amax = v; ip = i; }
        }
        if (ip != j) {
            for (int k = 0; k < n; k++) std::swap(A[(size_t)j * n + k], A[(size_t)ip * n + k]);
            std::swap(p[j], p[ip]);
        }
        const double ajj = A[(size_t)j * n + j];
        if (ajj == 0.0) continue;
        for (int i = j + 1; i < n; i++) {
            const double l = A[(size_t)i * n + j] / ajj;
            A[(size_t)i * n + j] = l;
            for (int k = j + 1; k < n; k++) A[(size_t)i * n + k] -= l * A[(size_t)j * n + k];
        }
    }
    for (int j = 0; j < n; j++)
        if (A[(size_t)j * n + j] == 0.0) return false;
    std::vector<double> y(n);
    for (int i = 0; i < n; i++) {
        double s = b[p[i]];
        for (int k = 0; k < i; k++) s -= A[(size_t)i * n + k] * y[k];
        y[i] = s;
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = y[i];
        for (int k = i + 1; k < n; k++) s -= A[(size_t)i * n + k] * x[k];
        x[i] = s / A[(size_t)i * n + i];
    }
    return true;
}

// ---- replica sets ----------------------------------------------------------------------

int get_event(lfts_batch* b, hipEvent_t* e) {
    if (!b->ev_pool.empty()) {
        *e = b->ev_pool.back();
        b->ev_pool.pop_back();
        return LFTS_OK;
    }
    BHIPCHK(b, hipEventCreate(e));
    return LFTS_OK;
}

// Fold completed event quads into the profile sums (the stream has been synchronised).
int resolve_events(lfts_batch* b) {
    for (auto& q : b->ev_pending) {
        float a = 0.f, t = 0.f;
        BHIPCHK(b, hipEventElapsedTime(&a, q.ch0, q.ch1));
        BHIPCHK(b, hipEventElapsedTime(&t, q.pe0, q.pe1));
        b->prof_ms_chain += a;
        b->prof_ms_pe += t;
        b->prof_n_pass += q.npasses;
        b->prof_n_pe += q.nrep;
        for (hipEvent_t e : {q.pe0, q.ch0, q.ch1, q.pe1}) b->ev_pool.push_back(e);
    }
    b->ev_pending.clear();
    return LFTS_OK;
}

int sync_batch(lfts_batch* b) {
    BHIPCHK(b, hipStreamSynchronize(b->st));
    if (!b->ev_pending.empty()) return resolve_events(b);
    return LFTS_OK;
}

// Make `ids` (indices into b->ctx) the active set: slab replicas sorted by edge, then
// line-path replicas; upload their descriptors in that order. The pinned staging is
// only rewritten after the stream has drained (callers synchronise before changing the
// set).
int set_active(lfts_batch* b, const std::vector<int>& ids) {
    b->act.clear();
    b->groups.clear();
    b->line_reps.clear();
    for (int edge : lfts::SLAB_EDGES) {
        const int off = (int)b->act.size();
        for (int id : ids)
            if (b->ctx[id]->fused && b->ctx[id]->d.m[0] == edge) b->act.push_back(id);
        const int cnt = (int)b->act.size() - off;
        if (cnt) {
            const lfts_ctx* c0 = b->ctx[b->act[off]];
            b->groups.push_back({edge, off, cnt, c0->ax[0].tw});
        }
    }
    for (int id : ids)
        if (!b->ctx[id]->fused) {
            b->line_reps.push_back(id);
            b->act.push_back(id);
        }
    if (b->act.size() != ids.size())  // a slab-path replica of an edge with no launch group
        return bfail(b, LFTS_EINVAL, "set_active: a replica's mesh has no slab launch group");
    if (b->act.empty()) return LFTS_OK;
    b->nb_max = 1;
    for (size_t s = 0; s < b->act.size(); s++) {
        b->h_ract[s] = b->reps[b->act[s]];
        b->nb_max = std::max(b->nb_max, b->h_ract[s].nb);
    }
    BHIPCHK(b, hipMemcpyAsync(b->d_ract, b->h_ract, sizeof(Rep) * b->act.size(), hipMemcpyHostToDevice, b->st));
    return LFTS_OK;
}

int upload_coef(lfts_batch* b) {
    BHIPCHK(b, hipMemcpyAsync(b->d_coef, b->h_coef, sizeof(Coef) * b->ctx.size(), hipMemcpyHostToDevice, b->st));
    return LFTS_OK;
}

// The reductions store their results in pinned host memory as well as in d_scal, so the
// host reads them without a copy: one launch fewer per Anderson iteration (one 64^3
// replica 29.19-29.21 against 29.11-29.15 Langevin steps/s; replica sets within their
// run-to-run spread). LFTS_SCAL_COPY=1 copies d_scal instead (same values).
// Summing the Gram row in the phi pass's last workgroup instead of a reduction launch
// measured slower (28.3 against 28.8 steps/s, sets 161 against 185): each of its
// workgroups must release its row with a device-scope fence before counting itself in.
double* host_mirror(const lfts_batch* b) {
    const char* e = std::getenv("LFTS_SCAL_COPY");
    return (e && std::strcmp(e, "1") == 0) ? nullptr : b->h_scal_dev;
}

// one checkpointed context: its calc_concs forms Q in its own pass, into d_scal only
bool ckpt_single(const lfts_batch* b) { return b->ctx.size() == 1 && b->ctx[0]->ck_stride; }

int read_scal(lfts_batch* b, bool on_host) {
    if (!on_host || !host_mirror(b))
        BHIPCHK(b, hipMemcpyAsync(b->h_scal, b->d_scal, sizeof(double) * lfts::SSTRIDE * b->ctx.size(),
                                  hipMemcpyDeviceToHost, b->st));
    return sync_batch(b);
}

// the Gram row of every active member into scal and its host mirror; with `flag` each
// slot's sequence number is stored after its results (for wait_row)
int reduce_row(lfts_batch* b, int nact, int ncomp, bool flag = false) {
    unsigned* hseq = (flag && host_mirror(b)) ? b->h_seq_dev : nullptr;
    if (hseq) b->seq++;
    BHIPCHK(b, lfts::launch_reduce(b->d_ract, nact, b->d_part, ncomp, 0, b->d_scal, 0, b->st, host_mirror(b), hseq,
                                   b->seq));
    return LFTS_OK;
}

// One replica's Anderson loop waits for each Gram row by polling the sequence number the
// reduction stores after it, on the calling thread, instead of synchronising the stream:
// 29.19-29.20 against 29.06-29.09 Langevin steps/s at 64^3, N=90 (the wake-up from a
// stream synchronisation costs about 4 us per iteration). The thread then keeps one core
// busy while the GPU evaluates; LFTS_SPIN=0 blocks instead.
bool spin_on() {
    const char* e = std::getenv("LFTS_SPIN");
    return !(e && std::strcmp(e, "0") == 0);
}

// Waits until slots 0..nact-1 carry the last sequence number. Every 1024 polls the stream
// is queried: once it has drained the flags must be set (an error otherwise), and a
// stream error ends the wait. The reduction stores its results, fences at system scope,
// then stores the flag; the acquire fence after the flags are seen keeps the compiler
// from reading the results earlier (x86 keeps loads in order by itself).
int wait_row(lfts_batch* b, int nact) {
    const volatile unsigned* f = b->h_seq;
    auto done = [&]() {
        for (int s = 0; s < nact; s++)
            if (f[s] != b->seq) return false;
        std::atomic_thread_fence(std::memory_order_acquire);
        return true;
    };
    for (long it = 1;; it++) {
        if (done()) return LFTS_OK;
        if ((it & 1023) == 0) {
            const hipError_t q = hipStreamQuery(b->st);
            if (q == hipSuccess) return done() ? LFTS_OK : bfail(b, LFTS_EHIP, "reduction results not flagged");
            if (q != hipErrorNotReady) return bfail(b, LFTS_EHIP, hipGetErrorString(q));
        }
        __builtin_ia32_pause();
    }
}

double scal(const lfts_batch* b, int id, int k) { return b->h_scal[(size_t)id * lfts::SSTRIDE + k]; }

// ---- kernel-shape autotuner ------------------------------------------------------------
// A slab kernel shape is a line factorisation R1 x R2 (lfts::slab_shape_factors) with a
// variant: tables in registers or LDS, one or two fields per workgroup. Variants of one
// factorisation give bit-identical results; factorisations round differently. So the
// factorisation of an edge is chosen once per process, when the edge is first tuned (like
// a plan by timing that every later use of the size shares), and every context and set of
// the process then rounds alike: a member gives the same bits alone and in any set. The
// variant is chosen per (edge, launch-size class, fused density sums or not).
// LFTS_SLAB_SHAPES fixes factorisations ("default", or e.g. "64:16x4,48:6x8"),
// LFTS_SLAB_VARIANT = 0..3 the variant.
namespace {
std::mutex g_shape_mu;
std::map<int, int> g_shape;  // edge -> factorisation chosen by timing in this process
std::map<int, int> g_used;   // edge -> factorisation last used (chosen or fixed)

// the factorisation LFTS_SLAB_SHAPES fixes for this edge, or -1
int env_shape(int edge) {
    const char* e = std::getenv("LFTS_SLAB_SHAPES");
    if (!e || !*e) return -1;
    if (std::strcmp(e, "default") == 0) return 0;
    const std::string spec(e);
    size_t pos = 0;
    while (pos < spec.size()) {
        size_t end = spec.find(',', pos);
        if (end == std::string::npos) end = spec.size();
        int E = 0, r1 = 0, r2 = 0;
        if (std::sscanf(spec.substr(pos, end - pos).c_str(), "%d:%dx%d", &E, &r1, &r2) == 3 && E == edge)
            for (int sh = 0; sh < lfts::slab_shape_count(edge); sh++) {
                int a = 0, b2 = 0;
                if (lfts::slab_shape_factors(edge, sh, &a, &b2) && a == r1 && b2 == r2) return sh;
            }
        pos = end + 1;
    }
    return -1;
}

std::string shape_name(int edge, int shape) {
    int r1 = 0, r2 = 0;
    lfts::slab_shape_factors(edge, shape, &r1, &r2);
    return std::to_string(r1) + "x" + std::to_string(r2);
}
}  // namespace

// The slab kernel shape (*variant = 4 x factorisation + variant) for `cnt` replicas of edge
// `edge` (descriptors d_reps, host copies h_reps) and passes with fused density sums
// (fuse = 1, one replica) or without. Timed on two contour-step passes after a prologue
// (with fused sums: the two passes from the first one that fuses): three rounds over the
// candidates, the fastest round of each counting. An alternative factorisation replaces
// an edge's default only when it is at least 3 % faster, so that near ties keep the
// default in every run. The tuning passes write the pairs of their beads, the partially
// convolved pairs T0/T1 and, fused, the running density sums: the chain that follows
// rewrites all of them before it reads them.
int slab_variant(lfts_batch* b, int edge, int cnt, const Rep* d_reps, const Rep* h_reps, const double2* tw,
                 hipStream_t s, int* variant, int fuse) {
    // launch sizes are tuned in power-of-two classes (the active set shrinks one member
    // at a time as members converge)
    int cls = 1;
    while (cls < cnt) cls *= 2;
    const auto key = std::make_tuple(edge, cls, fuse);
    const auto it = b->tuned.find(key);
    if (it != b->tuned.end()) {
        *variant = it->second;
        return LFTS_OK;
    }
    const int V = lfts::SLAB_VARIANTS;
    const int nsh = lfts::slab_shape_count(edge);
    // LFTS_SLAB_SHAPES first, then the process's earlier choice, else all candidates
    const int env_sh = env_shape(edge);
    int fixed_sh = env_sh;
    if (fixed_sh < 0) {
        std::lock_guard<std::mutex> lk(g_shape_mu);
        const auto g = g_shape.find(edge);
        if (g != g_shape.end()) fixed_sh = g->second;
    }
    const char* env = std::getenv("LFTS_SLAB_VARIANT");
    const int fixed_v = (env && *env >= '0' && *env < '0' + V && env[1] == 0) ? *env - '0' : -1;
    // candidates: (factorisation, variant)
    std::vector<std::pair<int, int>> cand;
    for (int sh = 0; sh < nsh; sh++) {
        if (fixed_sh >= 0 && sh != fixed_sh) continue;
        const int mask = lfts::slab_shape_variants(edge, sh);
        for (int v = 0; v < V; v++)
            if (((mask >> v) & 1) && (fixed_v < 0 || v == fixed_v || !((mask >> fixed_v) & 1))) cand.push_back({sh, v});
    }
    if (cand.empty()) return bfail(b, LFTS_EINVAL, "slab_variant: no kernel shape for this edge");
    const Rep* one = (cnt == 1) ? h_reps : nullptr;
    int i1 = 1;  // first timed pass
    if (fuse && one) i1 = (one->N + 1) / 2;
    std::vector<float> t(cand.size(), 1e30f);
    if (cand.size() > 1) {
        for (hipEvent_t& e : b->tune_ev)
            if (!e) BHIPCHK(b, hipEventCreate(&e));
        BHIPCHK(b, lfts::launch_slab(d_reps, cnt, edge, 0, tw, s, one, 0));
        for (int round = 0; round < 3; round++)
            for (size_t c = 0; c < cand.size(); c++) {
                const int v = cand[c].first * V + cand[c].second;
                BHIPCHK(b, lfts::launch_slab(d_reps, cnt, edge, i1, tw, s, one, v, fuse));  // warm-up
                BHIPCHK(b, lfts::launch_slab(d_reps, cnt, edge, i1 + 1, tw, s, one, v, fuse));
                BHIPCHK(b, hipEventRecord(b->tune_ev[0], s));
                for (int rep = 0; rep < 4; rep++) {
                    BHIPCHK(b, lfts::launch_slab(d_reps, cnt, edge, i1, tw, s, one, v, fuse));
                    BHIPCHK(b, lfts::launch_slab(d_reps, cnt, edge, i1 + 1, tw, s, one, v, fuse));
                }
                BHIPCHK(b, hipEventRecord(b->tune_ev[1], s));
                BHIPCHK(b, hipEventSynchronize(b->tune_ev[1]));
                float ms = 0.f;
                BHIPCHK(b, hipEventElapsedTime(&ms, b->tune_ev[0], b->tune_ev[1]));
                t[c] = std::min(t[c], ms);
            }
    }
    // fastest variant of each factorisation, then the factorisation
    std::vector<int> bestc(nsh, -1);
    for (size_t c = 0; c < cand.size(); c++) {
        int& bc = bestc[cand[c].first];
        if (bc < 0 || t[c] < t[bc]) bc = (int)c;
    }
    int sh = fixed_sh;
    if (env_sh < 0) {
        std::lock_guard<std::mutex> lk(g_shape_mu);
        const auto g = g_shape.find(edge);
        if (g != g_shape.end()) sh = g->second;  // chosen meanwhile by another set
        if (sh < 0) {
            sh = 0;
            for (int k = 1; k < nsh; k++)
                if (bestc[k] >= 0 && (bestc[sh] < 0 || t[bestc[k]] < 0.97f * t[bestc[sh]])) sh = k;
        }
        g_shape[edge] = sh;
    }
    int bc = bestc[sh];
    if (bc < 0) {  // the factorisation another set chose was not among these candidates
        for (size_t c = 0; c < cand.size(); c++)
            if (cand[c].first == sh) bc = (int)c;
        if (bc < 0) return bfail(b, LFTS_EINVAL, "slab_variant: no variant of the chosen factorisation");
    }
    {
        std::lock_guard<std::mutex> lk(g_shape_mu);
        g_used[edge] = sh;
    }
    const int best = cand[bc].first * V + cand[bc].second;
    std::string line = "edge " + std::to_string(edge) + " x " + std::to_string(cls) + (fuse ? " fused" : "") +
                       " (tuned on " + std::to_string(cnt) + "): " + shape_name(edge, sh) + " variant " +
                       std::to_string(cand[bc].second);
    if (cand.size() > 1) {
        line += " (us per pass:";
        char buf[64];
        for (size_t c = 0; c < cand.size(); c++) {
            std::snprintf(buf, sizeof buf, " %s/%d %.2f", shape_name(edge, cand[c].first).c_str(), cand[c].second,
                          125.0 * t[c]);
            line += buf;
        }
        line += ")";
    } else {
        line += (fixed_sh >= 0 || fixed_v >= 0) ? " (fixed)" : " (the only shape)";
    }
    b->tune_report += line + "\n";
    b->tuned[key] = best;
    *variant = best;
    return LFTS_OK;
}
// One propagator evaluation (src/diblock.cc:30-60) of every active replica; Q of
// replica id lands in d_scal[id][0].
int calc_concs_async(lfts_batch* b, bool store_all, bool h_ready, int dots, int part, long long* npasses_out) {
    const bool whole = part == PART_ALL;
    const int nact = (int)b->act.size();
    if (nact == 0) return LFTS_OK;
    if (ckpt_single(b)) {  // checkpointed storage: its own phi pass
        if (int rc = ckpt_calc_concs(b->ctx[0])) return rc;
        if (dots == DOTS_ONE) return bfail(b, LFTS_EINVAL, "by-value Gram row with checkpointed storage");
        if (dots) BHIPCHK(b, lfts::launch_am_dots(b->d_ract, nact, b->nb_max, b->d_coef, b->d_part, b->st));
        return LFTS_OK;
    }
    // One replica alone runs latency-bound passes: it fuses the density sums into the
    // second half of its chain (no separate pass over the stored pairs). A set moves the
    // same bytes either way and keeps the leaner kernel.
    const int fuse = (!store_all && nact == 1) ? 1 : 0;
    lfts_batch::EvQuad ev{};
    // live timing of every prof-th evaluation (its events cost host and queue time)
    const bool rec = whole && b->prof > 0 && (b->prof_seq++ % b->prof) == 0;
    if (rec) {
        for (hipEvent_t* e : {&ev.pe0, &ev.ch0, &ev.ch1, &ev.pe1})
            if (int rc = get_event(b, e)) return rc;
        BHIPCHK(b, hipEventRecord(ev.pe0, b->st));
    }
    if (!h_ready && part != PART_PHI) BHIPCHK(b, lfts::launch_prepare_h(b->d_ract, nact, b->nb_max, b->st));
    if (rec) BHIPCHK(b, hipEventRecord(ev.ch0, b->st));
    long long npasses = 0;
    if (part != PART_PHI) {
    // slab passes: i = 0 prologue, i = 1..N-1 contour steps (src/diblock.cc:43). With
    // more than one mesh size each group's chain runs on its own stream.
    const int ng = (int)b->groups.size();
    const bool fork = ng > 1;
    // kernel shapes, tuned (on first use of a size class) before the groups run concurrently
    int variant[lfts::N_SLAB_EDGES] = {}, variant_fused[lfts::N_SLAB_EDGES] = {};
    for (int g = 0; g < ng; g++) {
        const auto& G = b->groups[g];
        if (int rc = slab_variant(b, G.edge, G.cnt, b->d_ract + G.off, b->h_ract + G.off, G.tw, b->st, &variant[g], 0))
            return rc;
        variant_fused[g] = variant[g];
        if (fuse && G.cnt == 1 && lfts::fused_phi(b->h_ract[G.off], fuse))
            if (int rc = slab_variant(b, G.edge, 1, b->d_ract + G.off, b->h_ract + G.off, G.tw, b->st,
                                      &variant_fused[g], 1))
                return rc;
    }
    if (fork) {
        BHIPCHK(b, hipEventRecord(b->ev_fork, b->st));
        for (int g = 0; g < ng; g++) BHIPCHK(b, hipStreamWaitEvent(b->gst[g], b->ev_fork, 0));
    }
    // A 32^3 group rides in the 64^3 group's launches (one launch per contour step for
    // both): its passes are latency-bound on their own. Both in their default
    // factorisations; LFTS_SET_PAIR=0 keeps separate launches.
    int pa = -1, pb = -1;
    if (!fuse) {
        const char* pe = std::getenv("LFTS_SET_PAIR");
        if (!(pe && std::strcmp(pe, "0") == 0))
            for (int g = 0; g < ng; g++) {
                if (variant[g] >= lfts::SLAB_VARIANTS) continue;
                if (b->groups[g].edge == 64) pa = g;
                if (b->groups[g].edge == 32) pb = g;
            }
        if (pa < 0 || pb < 0 || !lfts::slab_pair_supported(64, 32)) pa = pb = -1;
    }
    if (pa >= 0) {
        const auto& A = b->groups[pa];
        const auto& B = b->groups[pb];
        int Na = 0, Nb = 0;
        for (int k = 0; k < A.cnt; k++) Na = std::max(Na, b->ctx[b->act[A.off + k]]->N);
        for (int k = 0; k < B.cnt; k++) Nb = std::max(Nb, b->ctx[b->act[B.off + k]]->N);
        for (int i = 0; i < std::max(Na, Nb); i++) {
            BHIPCHK(b, lfts::launch_slab_pair(b->d_ract + A.off, A.cnt, A.edge, b->d_ract + B.off, B.cnt, B.edge, i,
                                              A.tw, B.tw, fork ? b->gst[pa] : b->st));
            if (i > 0) npasses += (i < Na) + (i < Nb);
        }
    }
    for (int g = 0; g < ng; g++) {
        if (g == pa || g == pb) continue;
        const auto& G = b->groups[g];
        hipStream_t s = fork ? b->gst[g] : b->st;
        int Ng = 0;
        for (int k = 0; k < G.cnt; k++) Ng = std::max(Ng, b->ctx[b->act[G.off + k]]->N);
        const int i0 = (Ng + 1) / 2;  // the first pass that fuses (one replica: Ng is its N)
        for (int i = 0; i < Ng; i++) {
            const int v = (fuse && i >= i0) ? variant_fused[g] : variant[g];
            BHIPCHK(b, lfts::launch_slab(b->d_ract + G.off, G.cnt, G.edge, i, G.tw, s, b->h_ract + G.off, v, fuse));
            if (i > 0) npasses++;
        }
    }
    // line-path replicas run on the set's stream meanwhile
    for (int id : b->line_reps) {
        lfts_ctx* c = b->ctx[id];
        for (int i = 1; i < c->N; i++) {
            if (int rc = line_contour_step(c, pair_ptr(c, i), pair_ptr(c, i + 1), c->h, i, c->T0, b->st,
                                           b->reps[id].skip))
                return bfail(b, rc, c->err);
            npasses++;
        }
    }
    if (fork)
        for (int g = 0; g < ng; g++) {
            BHIPCHK(b, hipEventRecord(b->ev_join[g], b->gst[g]));
            BHIPCHK(b, hipStreamWaitEvent(b->st, b->ev_join[g], 0));
        }
    }  // chains
    if (npasses_out) *npasses_out = npasses;
    if (part == PART_CHAIN) return LFTS_OK;
    if (rec) BHIPCHK(b, hipEventRecord(ev.ch1, b->st));
    // Q's partials: the slab chains' last passes leave them (qpl); the others need a pass
    bool q_pass = false;
    for (int s = 0; s < nact; s++) q_pass = q_pass || !b->h_ract[s].qpl;
    if (q_pass) BHIPCHK(b, lfts::launch_q_sum(b->d_ract, nact, b->nb_max, b->d_qpart, b->st));
    const int nh_one = (dots == DOTS_ONE) ? b->h_coef[b->act[0]].nh_dots : -1;
    BHIPCHK(b, lfts::launch_phi(b->d_ract, nact, b->nb_max, b->d_scal, b->d_qpart, fuse,
                                dots == DOTS_COEF ? b->d_coef : nullptr, nh_one, b->d_part, b->st, host_mirror(b)));
    if (rec) {
        BHIPCHK(b, hipEventRecord(ev.pe1, b->st));
        ev.npasses = npasses;
        ev.nrep = nact;
        b->ev_pending.push_back(ev);
    }
    for (int id : b->act) {
        lfts_ctx* c = b->ctx[id];
        if (whole) {  // a captured iteration counts its evaluations at each replay
            c->n_pe++;
            c->n_cs += c->N - 1;
        }
        c->pairs_all = !lfts::fused_phi(make_rep(c), fuse);
    }
    return LFTS_OK;
}

// ---- the Anderson iteration of a replica set as one HIP graph ---------------------------

// Graph mode: sets of two or more members with every pair stored (the checkpointed storage
// runs its own chains), unless LFTS_GRAPH=0. Slab-path and line-path members alike take
// their evaluation flags from the set's descriptors.
bool graph_mode(const lfts_batch* b) {
    if (b->ctx.size() < 2 || !b->d_flags) return false;
    for (const lfts_ctx* c : b->ctx)
        if (c->ck_stride) return false;
    const char* ge = std::getenv("LFTS_GRAPH");
    return !(ge && std::strcmp(ge, "0") == 0);
}

// flags: evaluation flag of member id set unless it is in `eval`, update flag unless in `upd`
int put_flags(lfts_batch* b, const std::vector<int>& eval, const std::vector<int>& upd, bool upload) {
    const int n = (int)b->ctx.size();
    for (int i = 0; i < 2 * n; i++) b->h_flags[i] = 1;
    for (int id : eval) b->h_flags[id] = 0;
    for (int id : upd) b->h_flags[n + id] = 0;
    if (upload) BHIPCHK(b, hipMemcpyAsync(b->d_flags, b->h_flags, sizeof(int) * 2 * n, hipMemcpyHostToDevice, b->st));
    return LFTS_OK;
}

// Every flag back to 0 (every member evaluated and updated). A fill on the stream, not an
// upload: the staging buffer may still be read by a copy enqueued just before (a pinned
// source is read when the copy runs, not when it is enqueued).
int clear_flags(lfts_batch* b) {
    const int n = (int)b->ctx.size();
    BHIPCHK(b, hipMemsetAsync(b->d_flags, 0, sizeof(int) * 2 * n, b->st));
    return LFTS_OK;
}

// One Anderson iteration after the first, every member in the launches, in three parts:
// (0) flags and coefficients (pinned staging, read when the graph runs) and the update
// with the Boltzmann factors; (1) the contour-step chains on the groups' streams; (2) Q,
// phi with the Gram row and its reduction (into pinned host memory; a copy with
// LFTS_SCAL_COPY=1).
int enqueue_iteration_part(lfts_batch* b, int part) {
    const int n = (int)b->ctx.size();
    if (part == 0) {
        BHIPCHK(b, hipMemcpyAsync(b->d_flags, b->h_flags, sizeof(int) * 2 * n, hipMemcpyHostToDevice, b->st));
        if (int rc = upload_coef(b)) return rc;
        BHIPCHK(b, lfts::launch_am_update(b->d_rupd, n, b->nb_max, b->d_coef, 1, b->st));
    } else if (part == 1) {
        if (int rc = calc_concs_async(b, false, true, DOTS_COEF, PART_CHAIN, &b->it_npasses)) return rc;
    } else {
        if (int rc = calc_concs_async(b, false, true, DOTS_COEF, PART_PHI)) return rc;
        if (int rc = reduce_row(b, n, b->it_ncomp)) return rc;
        if (!host_mirror(b))
            BHIPCHK(b, hipMemcpyAsync(b->h_scal, b->d_scal, sizeof(double) * lfts::SSTRIDE * n,
                                      hipMemcpyDeviceToHost, b->st));
    }
    return LFTS_OK;
}

// Replays the captured iteration (capturing its parts first) and waits for it. With live
// timing, events recorded between the parts give the chain and evaluation times.
int replay_iteration(lfts_batch* b) {
    for (int part = 0; part < 3; part++) {
        if (b->it_exec[part]) continue;
        BHIPCHK(b, hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_iteration_part(b, part);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->st, &g);
        if (rc || e != hipSuccess || !g) {
            if (g) (void)hipGraphDestroy(g);
            return rc ? rc : bfail(b, LFTS_EHIP, "Anderson iteration graph: capture failed");
        }
        b->it_graph[part] = g;
        BHIPCHK(b, hipGraphInstantiate(&b->it_exec[part], g, nullptr, nullptr, 0));
    }
    lfts_batch::EvQuad ev{};
    if (b->prof)
        for (hipEvent_t* e : {&ev.ch0, &ev.ch1, &ev.pe1})
            if (int rc = get_event(b, e)) return rc;
    BHIPCHK(b, hipGraphLaunch(b->it_exec[0], b->st));
    if (b->prof) BHIPCHK(b, hipEventRecord(ev.ch0, b->st));
    BHIPCHK(b, hipGraphLaunch(b->it_exec[1], b->st));
    if (b->prof) BHIPCHK(b, hipEventRecord(ev.ch1, b->st));
    BHIPCHK(b, hipGraphLaunch(b->it_exec[2], b->st));
    if (b->prof) BHIPCHK(b, hipEventRecord(ev.pe1, b->st));
    if (int rc = sync_batch(b)) return rc;
    if (b->prof) {
        float ch = 0.f, pe = 0.f;
        BHIPCHK(b, hipEventElapsedTime(&ch, ev.ch0, ev.ch1));
        BHIPCHK(b, hipEventElapsedTime(&pe, ev.ch0, ev.pe1));  // the update wrote h: the evaluation starts here
        for (hipEvent_t e : {ev.ch0, ev.ch1, ev.pe1}) b->ev_pool.push_back(e);
        b->prof_ms_chain += ch;
        b->prof_ms_pe += pe;
        b->prof_n_pass += b->it_npasses;
    }
    return LFTS_OK;
}

std::vector<int> all_ids(const lfts_batch* b) {
    std::vector<int> ids(b->ctx.size());
    std::iota(ids.begin(), ids.end(), 0);
    return ids;
}

int batch_calc_concs(lfts_batch* b, double* lnQ) {
    if (int rc = set_active(b, all_ids(b))) return rc;
    if (int rc = calc_concs_async(b, true)) return rc;
    if (int rc = read_scal(b, !ckpt_single(b))) return rc;
    bool finite = true;
    for (size_t id = 0; id < b->ctx.size(); id++) {
        const double Q = scal(b, (int)id, 0);
        if (lnQ) lnQ[id] = std::log(Q);
        if (!std::isfinite(Q) || Q <= 0.0) finite = false;
    }
    return finite ? LFTS_OK : bfail(b, LFTS_ENUM, "partition function Q is not finite and positive");
}

// src/anderson.cc:24-93, every replica with its own iteration count; replicas leave the
// active set as their own loop condition fails, exactly as separate calls would.
int batch_anderson_mix(lfts_batch* b, int maxIter, double errTol, int* k_ret, double* err_ret) {
    const int n = (int)b->ctx.size();
    std::vector<int> k(n, 1);
    std::vector<double> err(n, 1.0);
    std::vector<int> active;
    for (int id = 0; id < n; id++)
        if (1 < maxIter && err[id] > errTol) active.push_back(id);
    // One coefficient upload per iteration: the one after the host LU also carries the
    // history depth of the next iteration's Gram row (the first iteration's is 0). The
    // pinned coefficients are rewritten only after read_scal has synchronised, so the
    // next evaluation is enqueued while the update still runs, and the descriptors are
    // uploaded again only when the active set changes.
    // One member (not checkpointed) passes its coefficients as kernel arguments instead:
    // no upload, and no copy for the update to wait on.
    const bool byval = n == 1 && !b->ctx[0]->ck_stride;
    if (!active.empty()) {
        for (int id : active) b->h_coef[id].nh_dots = 0;
        if (!byval)
            if (int rc = upload_coef(b)) return rc;
    }
    std::vector<int> cur;
    // From the second iteration on, every member still active was in the previous
    // update, which also wrote its Boltzmann factors (with_h).
    bool h_ready = false;
    // Graph mode: the first iteration runs directly with every member in the launches;
    // each later one is a replay of the captured iteration. `upd` are the members whose
    // update the next replay applies (those of the iteration before it).
    const bool graph = graph_mode(b);
    std::vector<int> upd;
    if (graph) {
        if (int rc = set_active(b, all_ids(b))) return rc;
        if (!b->rupd_ready) {
            // the update kernel's descriptors: every member in launch order, update flags
            std::vector<lfts::Rep> ru(b->h_ract, b->h_ract + n);
            for (lfts::Rep& r : ru) r.skip = b->d_flags + n + r.id;
            BHIPCHK(b, hipStreamSynchronize(b->st));
            BHIPCHK(b, hipMemcpy(b->d_rupd, ru.data(), sizeof(lfts::Rep) * n, hipMemcpyHostToDevice));
            b->it_ncomp = 1;
            for (const lfts_ctx* c : b->ctx) b->it_ncomp = std::max(b->it_ncomp, c->nhMax + 1);
            b->rupd_ready = true;
        }
    }
    while (!active.empty()) {
        const int nact = graph ? n : (int)active.size();
        if (graph) {
            if (!h_ready) {
                if (int rc = put_flags(b, active, {}, true)) return rc;
                if (int rc = calc_concs_async(b, false, false, DOTS_COEF)) return rc;
                if (int rc = reduce_row(b, n, b->it_ncomp)) return rc;
                if (int rc = read_scal(b, true)) return rc;
            } else {
                if (int rc = put_flags(b, active, upd, false)) return rc;
                if (int rc = replay_iteration(b)) return rc;
                for (int id : active) {
                    lfts_ctx* c = b->ctx[id];
                    c->n_pe++;
                    c->n_cs += c->N - 1;
                }
                if (b->prof) b->prof_n_pe += (long long)active.size();
            }
        } else {
            if (active != cur) {
                if (int rc = set_active(b, active)) return rc;
                cur = active;
            }
            // the evaluation's phi pass also writes the Gram row partials (dots)
            if (int rc = calc_concs_async(b, false, h_ready, byval ? DOTS_ONE : DOTS_COEF)) return rc;
            int nh_max = 0;
            for (int id : active) {
                const lfts_ctx* c = b->ctx[id];
                nh_max = std::max(nh_max, (k[id] < c->nhMax + 1) ? k[id] - 1 : c->nhMax);
            }
            // one member: wait for the row by polling when asked (and no timing events are due)
            const bool spin = byval && spin_on() && b->ev_pending.empty() && host_mirror(b);
            if (int rc = reduce_row(b, nact, nh_max + 1, spin)) return rc;
            if (int rc = spin ? wait_row(b, nact) : read_scal(b, true)) return rc;
        }
        for (int id : active) {
            lfts_ctx* c = b->ctx[id];
            const int H = c->nhMax + 1;
            auto DD = [&](int i, int j) -> double& { return c->DD[(size_t)i * H + j]; };
            Coef& co = b->h_coef[id];
            const int nh = (k[id] < c->nhMax + 1) ? k[id] - 1 : c->nhMax;
            co.nh = nh;
            // the next iteration's Gram row (k + 1)
            co.nh_dots = (k[id] + 1 < c->nhMax + 1) ? k[id] : c->nhMax;
            const double S1 = scal(b, id, 0);
            err[id] = std::pow(S1 / c->M, 0.5);
            for (int m = 0; m <= nh; m++) DD(0, m) = scal(b, id, m);
            co.lambda = 1.0 - std::pow(0.9, double(k[id]));
            if (k[id] < 2) {
                co.nh = 0;
            } else {
                for (int m = 1; m <= nh; m++) {
                    c->V[m - 1] = DD(0, 0) - DD(0, m);
                    for (int nn = 1; nn <= m; nn++)
                        c->U[(m - 1) * nh + nn - 1] = c->U[(nn - 1) * nh + m - 1] =
                            DD(0, 0) - DD(0, m) - DD(0, nn) + DD(nn, m);
                }
                if (!lu_solve(nh, c->U.data(), c->V.data(), c->C.data())) {
                    if (graph) (void)clear_flags(b);
                    (void)sync_batch(b);
                    return bfail(b, LFTS_ESING, "singular Anderson system: the history matrix U has a zero pivot "
                                                "(src/anderson.cc:124-125; GSL reports it singular and aborts)");
                }
                for (int nn = 0; nn < nh; nn++) co.C[nn] = c->C[nn];
            }
            co.slot = 1 + (k[id] - 1) % c->nhMax;
        }
        if (graph) {
            upd = active;  // applied by the next replay, or after the loop
        } else if (byval) {
            BHIPCHK(b, lfts::launch_am_update_one(b->d_ract, b->nb_max, b->h_coef[active[0]], 1, b->st));
        } else {
            if (int rc = upload_coef(b)) return rc;
            BHIPCHK(b, lfts::launch_am_update(b->d_ract, nact, b->nb_max, b->d_coef, 1, b->st));
        }
        h_ready = true;
        std::vector<int> still;
        for (int id : active) {
            lfts_ctx* c = b->ctx[id];
            const int H = c->nhMax + 1;
            auto DD = [&](int i, int j) -> double& { return c->DD[(size_t)i * H + j]; };
            const int nh = (k[id] < c->nhMax + 1) ? k[id] - 1 : c->nhMax;
            const int slot = 1 + (k[id] - 1) % c->nhMax;
            DD(slot, slot) = DD(0, 0);
            for (int m = 1; m < slot; m++) DD(m, slot) = DD(0, m);
            for (int m = slot + 1; m <= nh; m++) DD(slot, m) = DD(0, m);
            k[id]++;
            if (k[id] < maxIter && err[id] > errTol) still.push_back(id);
        }
        active.swap(still);
    }
    if (graph) {
        // the last update (of the members that left the loop with it), then every flag back
        // to 0 so that the set's other calls run every member
        if (!upd.empty()) {
            if (int rc = put_flags(b, {}, upd, true)) return rc;
            if (int rc = upload_coef(b)) return rc;
            BHIPCHK(b, lfts::launch_am_update(b->d_rupd, n, b->nb_max, b->d_coef, 1, b->st));
        }
        if (int rc = clear_flags(b)) return rc;
    }
    if (int rc = sync_batch(b)) return rc;
    bool nonfinite = false;
    for (int id = 0; id < n; id++) {
        if (k_ret) k_ret[id] = k[id];
        if (err_ret) err_ret[id] = err[id];
        if (!std::isfinite(err[id])) nonfinite = true;
    }
    return nonfinite ? bfail(b, LFTS_ENUM, "Anderson residual is not finite") : LFTS_OK;
}

// ---- Anderson mixing on the device, one context -----------------------------------------

// The device-resident loop is chosen with LFTS_ANDERSON=device. It is not the default:
// at 64^3, N=90 the host loop (one read-back and a host LU per iteration) measured 29.6
// Langevin steps/s against 28.3 for the device loop (its two extra kernels per
// iteration, the flag test in every kernel and the speculative tail cost more than the
// round trip saves); graph replay of an iteration made no measurable difference.
bool device_anderson_enabled(const lfts_ctx* c) {
    const char* e = std::getenv("LFTS_ANDERSON");
    return !c->ck_stride && e && std::strcmp(e, "device") == 0;
}

int ensure_am_device(lfts_ctx* c, int maxIter) {
    if (!c->am) HIPCHK(c, hipMalloc(&c->am, sizeof(lfts::AmState)));
    if (!c->am_h) HIPCHK(c, hipHostMalloc(&c->am_h, sizeof(lfts::AmState), hipHostMallocDefault));
    if (c->pow09_n < maxIter + 2) {
        // lambda = 1 - 0.9^k with the host's pow (src/anderson.cc:41)
        std::vector<double> t(maxIter + 2);
        for (int k = 0; k < maxIter + 2; k++) t[k] = std::pow(0.9, double(k));
        (void)hipFree(c->pow09);
        c->pow09 = nullptr;
        HIPCHK(c, hipMalloc(&c->pow09, sizeof(double) * t.size()));
        HIPCHK(c, hipMemcpy(c->pow09, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice));
        c->pow09_n = maxIter + 2;
    }
    lfts_batch* b = c->self;
    const int* flag = &c->am->stop;
    if (b->reps[0].skip != flag) {
        b->reps[0].skip = flag;  // the set's kernels return at once while the flag is set
        if (c->am_exec) {
            (void)hipGraphExecDestroy(c->am_exec);
            (void)hipGraphDestroy(c->am_graph);
            c->am_exec = nullptr;
            c->am_graph = nullptr;
        }
    }
    return LFTS_OK;
}

// One iteration of src/anderson.cc:29-89 as device work only: propagator evaluation,
// Gram row, solve, update, loop test.
int am_iteration(lfts_ctx* c) {
    lfts_batch* b = c->self;
    if (int rc = calc_concs_async(b)) return rc;
    BHIPCHK(b, lfts::launch_am_dots(b->d_ract, 1, b->nb_max, b->d_coef, b->d_part, b->st));
    BHIPCHK(b, lfts::launch_reduce(b->d_ract, 1, b->d_part, c->nhMax + 1, 0, b->d_scal, 0, b->st));
    BHIPCHK(b, lfts::launch_am_solve(c->am, b->d_scal, b->d_coef, c->pow09, c->M, c->nhMax, b->st));
    BHIPCHK(b, lfts::launch_am_update(b->d_ract, 1, b->nb_max, b->d_coef, 0, b->st));
    BHIPCHK(b, lfts::launch_am_advance(c->am, b->st));
    return LFTS_OK;
}

// anderson::mix with the loop on the device: iterations are enqueued in chunks (the size
// of the previous solve first), and those enqueued past the end of the loop return at
// once. With LFTS_GRAPH unset or 1 (and no live profiling), one iteration is captured as
// a HIP graph on first use and replayed.
int device_anderson_mix(lfts_ctx* c, int maxIter, double errTol, int* k_ret, double* err_ret) {
    lfts_batch* b = c->self;
    if (int rc = ensure_am_device(c, maxIter)) return rc;
    if (int rc = set_active(b, all_ids(b))) return rc;
    lfts::AmState* h = c->am_h;
    std::memset(h, 0, sizeof(lfts::AmState));
    h->k = 1;
    h->maxIter = maxIter;
    h->err = 1.0;
    h->tol = errTol;
    h->stop = !(1 < maxIter && 1.0 > errTol);
    BHIPCHK(b, hipMemcpyAsync(c->am, h, sizeof(lfts::AmState), hipMemcpyHostToDevice, b->st));
    b->h_coef[0].nh_dots = 0;
    if (int rc = upload_coef(b)) return rc;
    const char* ge = std::getenv("LFTS_GRAPH");
    const bool graph = !(ge && std::strcmp(ge, "0") == 0) && !b->prof;
    const long long pe0 = c->n_pe, cs0 = c->n_cs;
    int enq = 0;
    bool first = true;
    while (true) {
        const int n = first ? std::max(1, c->am_last > 2 ? c->am_last - 2 : 6) : 2;
        for (int t = 0; t < n; t++) {
            if (graph && c->am_exec) {
                BHIPCHK(b, hipGraphLaunch(c->am_exec, b->st));
            } else if (graph && enq > 0) {
                // capture after one direct iteration (kernel shapes are tuned by then)
                BHIPCHK(b, hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
                const int rc = am_iteration(c);
                hipGraph_t g = nullptr;
                const hipError_t e = hipStreamEndCapture(b->st, &g);
                if (rc) return rc;
                BHIPCHK(b, e);
                c->am_graph = g;
                BHIPCHK(b, hipGraphInstantiate(&c->am_exec, g, nullptr, nullptr, 0));
                BHIPCHK(b, hipGraphLaunch(c->am_exec, b->st));
            } else {
                if (int rc = am_iteration(c)) return rc;
            }
            enq++;
        }
        first = false;
        BHIPCHK(b, hipMemcpyAsync(h, c->am, offsetof(lfts::AmState, DD), hipMemcpyDeviceToHost, b->st));
        if (int rc = sync_batch(b)) return rc;
        if (h->stop || enq > maxIter + 2) break;
    }
    const int k = h->k;
    const double err = h->err;
    // the flag back to 0 so that later evaluations run
    h->stop = 0;
    BHIPCHK(b, hipMemcpyAsync(&c->am->stop, &h->stop, sizeof(int), hipMemcpyHostToDevice, b->st));
    if (int rc = sync_batch(b)) return rc;
    c->am_last = k - 1;
    c->n_pe = pe0 + (k - 1);
    c->n_cs = cs0 + (long long)(k - 1) * (c->N - 1);
    if (k_ret) *k_ret = k;
    if (err_ret) *err_ret = err;
    if (h->singular)
        return fail(c, LFTS_ESING, "singular Anderson system: the history matrix U has a zero pivot "
                                   "(src/anderson.cc:124-125; GSL reports it singular and aborts)");
    return std::isfinite(err) ? LFTS_OK : fail(c, LFTS_ENUM, "Anderson residual is not finite");
}

// src/langevin.cc:23-37 for every replica; eta_new are device pointers.
int batch_langevin(lfts_batch* b, const double* const* eta_new, const double* XbN, const double* dt) {
    const int n = (int)b->ctx.size();
    for (int id = 0; id < n; id++) {
        if (!eta_new[id]) return bfail(b, LFTS_EINVAL, "langevin: null noise pointer");
        Coef& co = b->h_coef[id];
        co.XbN = XbN[id];
        co.dt = dt[id];
        co.eta_new = eta_new[id];
        co.keep_new = 1;
    }
    if (int rc = set_active(b, all_ids(b))) return rc;
    if (int rc = upload_coef(b)) return rc;
    BHIPCHK(b, lfts::launch_langevin(b->d_ract, n, b->nb_max, b->d_coef, b->st));
    return sync_batch(b);
}

int batch_create(lfts_ctx* const* ctxs, int n, lfts_batch** out) {
    lfts_batch* b = new lfts_batch();
    b->ctx.assign(ctxs, ctxs + n);
    auto cleanup = [&](int rc) {
        lfts_batch_destroy(b);
        return rc;
    };
#define BCREATE(expr)                                        \
    do {                                                     \
        if ((expr) != hipSuccess) return cleanup(LFTS_EHIP); \
    } while (0)
    BCREATE(hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking));
    if (n > 1) {
        for (int g = 0; g < lfts::N_SLAB_EDGES; g++) {
            BCREATE(hipStreamCreateWithFlags(&b->gst[g], hipStreamNonBlocking));
            BCREATE(hipEventCreateWithFlags(&b->ev_join[g], hipEventDisableTiming));
        }
        BCREATE(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
    }
    b->reps.resize(n);
    for (int i = 0; i < n; i++) {
        b->reps[i] = make_rep(ctxs[i]);
        b->reps[i].id = i;
    }
    BCREATE(hipMalloc(&b->d_ract, sizeof(Rep) * n));
    BCREATE(hipHostMalloc(&b->h_ract, sizeof(Rep) * n, hipHostMallocDefault));
    BCREATE(hipMalloc(&b->d_coef, sizeof(Coef) * n));
    BCREATE(hipHostMalloc(&b->h_coef, sizeof(Coef) * n, hipHostMallocDefault));
    std::memset(b->h_coef, 0, sizeof(Coef) * n);
    BCREATE(hipMalloc(&b->d_part, sizeof(double) * lfts::PSTRIDE * lfts::MAX_NB * n));
    BCREATE(hipMalloc(&b->d_scal, sizeof(double) * lfts::SSTRIDE * n));
    BCREATE(hipMalloc(&b->d_qpart, sizeof(double) * lfts::MAX_NB * n));
    // Slab-path chains leave Q's partials (one per plane) themselves, so no pass over q_N
    // follows them. LFTS_Q_PASS=1 (read when the set is made) keeps that pass instead
    // (k_q_sum's workgroup partials: another summation order, the same Q to rounding).
    const char* qp = std::getenv("LFTS_Q_PASS");
    if (!(qp && std::strcmp(qp, "1") == 0))
        for (int i = 0; i < n; i++)
            if (b->reps[i].n && b->reps[i].N >= 2) b->reps[i].qpl = b->d_qpart + (long)i * lfts::MAX_NB;
    BCREATE(hipHostMalloc(&b->h_scal, sizeof(double) * lfts::SSTRIDE * n, hipHostMallocMapped | hipHostMallocCoherent));
    BCREATE(hipHostGetDevicePointer((void**)&b->h_scal_dev, b->h_scal, 0));
    BCREATE(hipHostMalloc(&b->h_seq, sizeof(unsigned) * n, hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(b->h_seq, 0, sizeof(unsigned) * n);
    BCREATE(hipHostGetDevicePointer((void**)&b->h_seq_dev, b->h_seq, 0));
    if (n > 1) {  // flags of the replayed Anderson iteration (all zero: every member runs)
        BCREATE(hipMalloc(&b->d_flags, sizeof(int) * 2 * n));
        BCREATE(hipMemset(b->d_flags, 0, sizeof(int) * 2 * n));
        BCREATE(hipHostMalloc(&b->h_flags, sizeof(int) * 2 * n, hipHostMallocDefault));
        std::memset(b->h_flags, 0, sizeof(int) * 2 * n);
        BCREATE(hipMalloc(&b->d_rupd, sizeof(Rep) * n));
        for (int i = 0; i < n; i++) b->reps[i].skip = b->d_flags + i;
    }
#undef BCREATE
    *out = b;
    return LFTS_OK;
}

}  // namespace lfts_abi

using namespace lfts_abi;

extern "C" {

int lfts_create(const lfts_desc* desc, lfts_ctx** out) {
    if (!desc || !out) return LFTS_EINVAL;
    *out = nullptr;
    if (desc->N < 2 || desc->NA < 0 || desc->NA > desc->N) return LFTS_EINVAL;
    if (desc->nhMax < 1 || desc->nhMax > lfts::MAX_HIST) return LFTS_EINVAL;
    for (int d = 0; d < 3; d++) {
        if (desc->m[d] < 1 || !(desc->L[d] > 0.0)) return LFTS_EINVAL;
        std::vector<int> rad;  // any factors; a prime factor above 7 up to MAX_PRIME_LINE points
        if (!factorize(desc->m[d], rad) || (int)rad.size() > lfts::MAXST) return LFTS_EINVAL;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return LFTS_EHIP;
    lfts_ctx* c = new lfts_ctx();
    if (hipGetDevice(&c->dev) != hipSuccess) {
        delete c;
        return LFTS_EHIP;
    }
    c->d = *desc;
    c->N = desc->N;
    c->NA = desc->NA;
    c->NB = desc->N - desc->NA;
    c->nhMax = desc->nhMax;
    c->M = (long)desc->m[0] * desc->m[1] * desc->m[2];
    c->Mk = (long)desc->m[0] * desc->m[1] * (desc->m[2] / 2 + 1);
    const long M = c->M;
    int rc = LFTS_OK;
#define CREATE_CHK(expr)     \
    do {                     \
        rc = (expr);         \
        if (rc) {            \
            lfts_destroy(c); \
            return rc;       \
        }                    \
    } while (0)
#define CREATE_HIP(expr) CREATE_CHK((expr) == hipSuccess ? LFTS_OK : LFTS_EHIP)
    for (int d = 0; d < 3; d++) CREATE_CHK(build_axis(c, c->ax[d], desc->m[d], desc->L[d], c->N));
    {
        // LFTS_PATH=generic selects the three-pass generic line kernels (runtime radix
        // stages in LDS) even where the fused slab pass or the register-FFT axis kernel
        // applies (used to cross-check the paths)
        const char* path = std::getenv("LFTS_PATH");
        const bool generic = path && std::strcmp(path, "generic") == 0;
        c->fused = !generic && lfts::slab_supported(desc->m[0], desc->m[1], desc->m[2]);
        c->fast_lines = !generic;  // line passes of slab-edge lengths use register FFTs
    }
    CREATE_HIP(hipMalloc(&c->w, sizeof(double) * 4 * M));
    CREATE_HIP(hipMalloc(&c->h, sizeof(double) * 2 * M));
    // the stored pairs: all of them when they fit comfortably in free HBM, else about
    // sqrt(N) of them at segment starts (checkpointed recomputation)
    int auto_stride = 0;
    {
        size_t free_b = 0, total_b = 0;
        const double need = 2.0 * sizeof(double) * 2 * M * (double)(c->N - 1);  // pairs + a derivative chain
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && need > 0.5 * (double)free_b)
            auto_stride = std::max(2, (int)std::ceil(std::sqrt((double)c->N)));
    }
    if (!auto_stride) CREATE_HIP(hipMalloc(&c->q, sizeof(double) * 2 * M * (size_t)(c->N - 1)));
    CREATE_HIP(hipMalloc(&c->T0, sizeof(double) * 2 * M));
    if (c->fused) CREATE_HIP(hipMalloc(&c->T1, sizeof(double) * 2 * M));
    if (c->fused && desc->m[0] <= 64) CREATE_HIP(hipMalloc(&c->phacc, sizeof(double) * 4 * M));  // see fused_phi
    CREATE_HIP(hipMalloc(&c->Dh, sizeof(double) * M * (size_t)(c->nhMax + 1)));
    CREATE_HIP(hipMalloc(&c->wh, sizeof(double) * M * (size_t)(c->nhMax + 1)));
    CREATE_HIP(hipMalloc(&c->noise, sizeof(double) * 2 * M));
    CREATE_HIP(hipMalloc(&c->S, sizeof(double) * c->Mk));
    CREATE_HIP(hipMemset(c->w, 0, sizeof(double) * 4 * M));
    CREATE_HIP(hipMemset(c->noise, 0, sizeof(double) * 2 * M));
    CREATE_HIP(hipMemset(c->S, 0, sizeof(double) * c->Mk));
    const int H = c->nhMax + 1;
    c->DD.assign((size_t)H * H, 0.0);
    c->U.assign((size_t)c->nhMax * c->nhMax, 0.0);
    c->V.assign(c->nhMax, 0.0);
    c->C.assign(c->nhMax, 0.0);
    CREATE_CHK(batch_create(&c, 1, &c->self));
    (void)noalloc_setup(c->dev);  // its result is kept, and reported by lfts_step_fwd_noalloc
    if (auto_stride) CREATE_CHK(set_pair_storage(c, auto_stride));
#undef CREATE_CHK
#undef CREATE_HIP
    *out = c;
    return LFTS_OK;
}

int lfts_destroy(lfts_ctx* c) {
    if (!c) return LFTS_EINVAL;
    if (c->n_sets > 0) return fail(c, LFTS_ESTATE, "destroy: a replica set in use holds this context; destroy it first");
    if (c->self) lfts_batch_destroy(c->self);
    release_pair_storage(c);
    if (c->am_exec) (void)hipGraphExecDestroy(c->am_exec);
    if (c->am_graph) (void)hipGraphDestroy(c->am_graph);
    (void)hipFree(c->am);
    (void)hipFree(c->pow09);
    if (c->am_h) (void)hipHostFree(c->am_h);
    for (auto& p : c->ax) {
        (void)hipFree(p.tw);
        (void)hipFree(p.gperm);
        (void)hipFree(p.perm);
        (void)hipFree(p.gnat);
        (void)hipFree(p.kap);
    }
    for (void* p : {(void*)c->w, (void*)c->h, (void*)c->q, (void*)c->T0, (void*)c->T1, (void*)c->phacc, (void*)c->Dh, (void*)c->wh,
                    (void*)c->noise, (void*)c->S, (void*)c->zk, (void*)c->test, (void*)c->dq, (void*)c->dbuf,
                    (void*)c->acc, (void*)c->dred, (void*)c->kry, (void*)c->pmul})
        (void)hipFree(p);
    if (c->hred) (void)hipHostFree(c->hred);
    delete c;
    return LFTS_OK;
}

const char* lfts_last_error(const lfts_ctx* c) { return c ? (c->err_lit ? c->err_lit : c->err.c_str()) : "null context"; }

int lfts_get_stream(lfts_ctx* c, void** stream) {
    if (!c || !stream) return LFTS_EINVAL;
    *stream = (void*)c->self->st;
    return LFTS_OK;
}

int lfts_set_fields(lfts_ctx* c, const double* w, int nfields) {
    if (!c || !w || (nfields != 2 && nfields != 4)) return fail(c, LFTS_EINVAL, "set_fields: nfields must be 2 or 4");
    HIPCHK(c, hipMemcpyAsync(c->w, w, sizeof(double) * nfields * c->M, hipMemcpyHostToDevice, c->self->st));
    HIPCHK(c, hipStreamSynchronize(c->self->st));
    return LFTS_OK;
}

int lfts_get_fields(lfts_ctx* c, double* w, int nfields) {
    if (!c || !w || (nfields != 2 && nfields != 4)) return fail(c, LFTS_EINVAL, "get_fields: nfields must be 2 or 4");
    HIPCHK(c, hipMemcpyAsync(w, c->w, sizeof(double) * nfields * c->M, hipMemcpyDeviceToHost, c->self->st));
    HIPCHK(c, hipStreamSynchronize(c->self->st));
    return LFTS_OK;
}

int lfts_fields_dev(lfts_ctx* c, double** w_dev) {
    if (!c || !w_dev) return LFTS_EINVAL;
    *w_dev = c->w;
    return LFTS_OK;
}

int lfts_calc_concs(lfts_ctx* c, double* lnQ) {
    if (!c) return LFTS_EINVAL;
    return batch_calc_concs(c->self, lnQ);
}

int lfts_anderson_mix(lfts_ctx* c, int maxIter, double errTol, int* k_ret, double* err_ret) {
    if (!c) return LFTS_EINVAL;
    if (device_anderson_enabled(c)) return device_anderson_mix(c, maxIter, errTol, k_ret, err_ret);
    return batch_anderson_mix(c->self, maxIter, errTol, k_ret, err_ret);
}

int lfts_langevin_init(lfts_ctx* c, const double* eta_prev) {
    if (!c || !eta_prev) return fail(c, LFTS_EINVAL, "langevin_init: null argument");
    HIPCHK(c, hipMemcpyAsync(c->noise, eta_prev, sizeof(double) * c->M, hipMemcpyHostToDevice, c->self->st));
    HIPCHK(c, hipStreamSynchronize(c->self->st));
    return LFTS_OK;
}

int lfts_langevin_step(lfts_ctx* c, const double* eta_new, double XbN, double dt) {
    if (!c || !eta_new) return fail(c, LFTS_EINVAL, "langevin_step: null argument");
    double* staging = c->noise + c->M;
    HIPCHK(c, hipMemcpyAsync(staging, eta_new, sizeof(double) * c->M, hipMemcpyHostToDevice, c->self->st));
    const double* en[1] = {staging};
    return batch_langevin(c->self, en, &XbN, &dt);
}

int lfts_langevin_step_dev(lfts_ctx* c, const double* eta_new_dev, double XbN, double dt) {
    if (!c || !eta_new_dev) return fail(c, LFTS_EINVAL, "langevin_step_dev: null argument");
    const double* en[1] = {eta_new_dev};
    return batch_langevin(c->self, en, &XbN, &dt);
}

int lfts_langevin_full_step_dev(lfts_ctx* c, const double* eta_new_dev, double XbN, double dt, int maxIter,
                                double errTol, double* lnQ, int* k_ret) {
    if (!c) return LFTS_EINVAL;
    if (!device_anderson_enabled(c))
        return lfts_batch_full_step_dev(c->self, &eta_new_dev, &XbN, &dt, maxIter, errTol, lnQ, k_ret);
    if (!eta_new_dev) return fail(c, LFTS_EINVAL, "full_step_dev: null argument");
    int rc = batch_langevin(c->self, &eta_new_dev, &XbN, &dt);
    if (rc) return rc;
    rc = device_anderson_mix(c, maxIter, errTol, k_ret, nullptr);
    if (rc && rc != LFTS_ENUM) return rc;
    const int rc2 = batch_calc_concs(c->self, lnQ);
    return rc ? rc : rc2;
}

// src/strFunc.cc:36-46 (forward transform of W- in three line passes, |X|^2 accumulated
// into the r2c half spectrum by the last pass)
int lfts_sk_sample(lfts_ctx* c) {
    if (!c) return LFTS_EINVAL;
    hipStream_t st = c->self->st;
    if (!c->zk) HIPCHK(c, hipMalloc(&c->zk, sizeof(double2) * c->M));
    LineArgs z = line_args(c, 2, lfts::LINE_FFT);
    z.inR = c->w;
    z.outC = c->zk;
    HIPCHK(c, lfts::launch_line(z, st));
    LineArgs y = line_args(c, 1, lfts::LINE_FFT);
    y.inC = c->zk;
    y.outC = c->zk;
    HIPCHK(c, lfts::launch_line(y, st));
    LineArgs x = line_args(c, 0, lfts::LINE_FFT);
    x.inC = c->zk;
    x.Sacc = c->S;
    HIPCHK(c, lfts::launch_line(x, st));
    HIPCHK(c, hipStreamSynchronize(st));
    c->nsamples++;
    return LFTS_OK;
}

int lfts_sk_set(lfts_ctx* c, const double* S, int nsamples) {
    if (!c || !S || nsamples < 0) return fail(c, LFTS_EINVAL, "sk_set: bad argument");
    HIPCHK(c, hipMemcpyAsync(c->S, S, sizeof(double) * c->Mk, hipMemcpyHostToDevice, c->self->st));
    HIPCHK(c, hipStreamSynchronize(c->self->st));
    c->nsamples = nsamples;
    return LFTS_OK;
}

int lfts_langevin_get_prev(lfts_ctx* c, double* eta_prev) {
    if (!c || !eta_prev) return fail(c, LFTS_EINVAL, "langevin_get_prev: null argument");
    HIPCHK(c, hipMemcpyAsync(eta_prev, c->noise, sizeof(double) * c->M, hipMemcpyDeviceToHost, c->self->st));
    HIPCHK(c, hipStreamSynchronize(c->self->st));
    return LFTS_OK;
}

int lfts_sk_get(lfts_ctx* c, double* S, int* nsamples) {
    if (!c) return LFTS_EINVAL;
    if (S) {
        HIPCHK(c, hipMemcpyAsync(S, c->S, sizeof(double) * c->Mk, hipMemcpyDeviceToHost, c->self->st));
        HIPCHK(c, hipStreamSynchronize(c->self->st));
    }
    if (nsamples) *nsamples = c->nsamples;
    return LFTS_OK;
}

int lfts_sk_reset(lfts_ctx* c) {
    if (!c) return LFTS_EINVAL;
    HIPCHK(c, hipMemsetAsync(c->S, 0, sizeof(double) * c->Mk, c->self->st));
    HIPCHK(c, hipStreamSynchronize(c->self->st));
    c->nsamples = 0;
    return LFTS_OK;
}

// src/anderson.cc:106-127 (the host LU of the Anderson system)
int lfts_lu_solve(int n, const double* A, const double* y, double* x) {
    if (n < 1 || n > lfts::MAX_HIST || !A || !y || !x) return LFTS_EINVAL;
    return lu_solve(n, A, y, x) ? LFTS_OK : LFTS_ESING;
}

// src/strFunc.cc:27-31, 80-103
int lfts_sk_bins(const int* m, const double* L, double* K, int* wt, int* P) {
    if (!m || !L || !K || !wt || !P) return LFTS_EINVAL;
    const int Mk = m[0] * m[1] * (m[2] / 2 + 1);
    for (int k = 0; k < Mk; k++) wt[k] = 2;
    for (int k0 = -(m[0] - 1) / 2; k0 <= m[0] / 2; k0++) {
        const int K0 = (k0 < 0) ? (k0 + m[0]) : k0;
        const double kx_sq = k0 * k0 / (L[0] * L[0]);
        for (int k1 = -(m[1] - 1) / 2; k1 <= m[1] / 2; k1++) {
            const int K1 = (k1 < 0) ? (k1 + m[1]) : k1;
            const double ky_sq = k1 * k1 / (L[1] * L[1]);
            for (int k2 = 0; k2 <= m[2] / 2; k2++) {
                const double kz_sq = k2 * k2 / (L[2] * L[2]);
                const int k = k2 + (m[2] / 2 + 1) * (K1 + m[1] * K0);
                K[k] = 2 * M_PI * pow(kx_sq + ky_sq + kz_sq, 0.5);
                if ((k2 == 0) || (k2 == m[2] / 2)) wt[k] = 1;
            }
        }
    }
    std::iota(P, P + Mk, 0);
    std::stable_sort(P, P + Mk, [K](int i, int j) { return K[i] < K[j]; });
    return LFTS_OK;
}

// src/strFunc.cc:50-76
int lfts_sk_save(const char* fileName, const int* m, const double* L, const double* S, int nsamples, double CV,
                 double chi_b, double dK, int dp) {
    if (!fileName || !m || !L || !S) return LFTS_EINVAL;
    const int M = m[0] * m[1] * m[2];
    const int Mk = m[0] * m[1] * (m[2] / 2 + 1);
    std::vector<double> K(Mk);
    std::vector<int> wt(Mk), P(Mk);
    lfts_sk_bins(m, L, K.data(), wt.data(), P.data());
    const double coeff = CV / (chi_b * chi_b * M * M);
    std::ofstream out(fileName);
    if (!out) return LFTS_EINVAL;
    out.precision(dp);
    out.setf(std::ios::fixed, std::ios::floatfield);
    double S_sum = 0.0;
    int n_same = 0;
    for (int k = 0; k < Mk; k++) {
        S_sum += wt[P[k]] * ((coeff / nsamples) * S[P[k]] - 0.5 / chi_b);
        n_same += wt[P[k]];
        if ((k == Mk - 1) || (std::fabs(K[P[k + 1]] - K[P[k]]) > dK)) {
            out << K[P[k]] << "\t" << S_sum / n_same << std::endl;
            S_sum = 0.0;
            n_same = 0;
        }
    }
    return LFTS_OK;
}

// src/lfts_simulation.cc:122-135
int lfts_hamiltonian(lfts_ctx* c, double XbN, double* H) {
    if (!c) return LFTS_EINVAL;
    lfts_batch* b = c->self;
    if (int rc = set_active(b, all_ids(b))) return rc;
    if (int rc = calc_concs_async(b)) return rc;
    HIPCHK(c, lfts::launch_h_sums(b->d_ract, 1, b->nb_max, b->d_part, b->st));
    HIPCHK(c, lfts::launch_reduce(b->d_ract, 1, b->d_part, 2, 0, b->d_scal, 1, b->st, host_mirror(b)));
    if (int rc = read_scal(b, !ckpt_single(b))) return rc;
    const double lnQ = std::log(scal(b, 0, 0));
    if (H) *H = -lnQ + (scal(b, 0, 2) / XbN - scal(b, 0, 1)) / c->M;
    return std::isfinite(lnQ) ? LFTS_OK : fail(c, LFTS_ENUM, "partition function Q is not finite and positive");
}

int lfts_get_pair(lfts_ctx* c, int i, double* out) {
    if (!c || !out || i < 1 || i > c->N) return fail(c, LFTS_EINVAL, "get_pair: index out of range");
    if (c->ck_stride && i > 1) return fail(c, LFTS_EINVAL, "get_pair: pairs are not stored (checkpointed storage)");
    if (!c->pairs_all && i > (c->N + 1) / 2 && i != c->N)
        return fail(c, LFTS_EINVAL,
                    "get_pair: the last evaluation (an Anderson iteration or the Hamiltonian) kept only pairs "
                    "1..ceil(N/2) and N; call lfts_calc_concs first");
    HIPCHK(c, hipMemcpyAsync(out, pair_ptr(c, i), sizeof(double) * 2 * c->M, hipMemcpyDeviceToHost, c->self->st));
    HIPCHK(c, hipStreamSynchronize(c->self->st));
    return LFTS_OK;
}

// One contour step on host arrays through the line path (the slab path only runs
// whole chains).
int lfts_step_fwd(lfts_ctx* c, const double* q_in, double* q_out, const double* h, int i) {
    if (!c || !q_in || !q_out || !h) return fail(c, LFTS_EINVAL, "step_fwd: null argument");
    const long M = c->M;
    hipStream_t st = c->self->st;
    if (!c->test) HIPCHK(c, hipMalloc(&c->test, sizeof(double) * 8 * M));
    HIPCHK(c, hipMemcpyAsync(c->test, q_in, sizeof(double) * 2 * M, hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(c->test + 2 * M, h, sizeof(double) * 2 * M, hipMemcpyHostToDevice, st));
    if (int rc = line_contour_step(c, c->test, c->test + 4 * M, c->test + 2 * M, i, c->test + 6 * M, st,
                                   c->self ? c->self->reps[0].skip : nullptr))
        return rc;
    HIPCHK(c, hipMemcpyAsync(q_out, c->test + 4 * M, sizeof(double) * 2 * M, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    return LFTS_OK;
}

int lfts_profile(lfts_ctx* c, int enable) {
    if (!c) return LFTS_EINVAL;
    return lfts_batch_profile(c->self, enable);
}

int lfts_profile_read(lfts_ctx* c, double* ms_chain, long long* n_cs, double* ms_pe, long long* n_pe) {
    if (!c) return LFTS_EINVAL;
    return lfts_batch_profile_read(c->self, ms_chain, n_cs, ms_pe, n_pe);
}

int lfts_counters(lfts_ctx* c, long long* n_pe, long long* n_cs) {
    if (!c) return LFTS_EINVAL;
    if (n_pe) *n_pe = c->n_pe;
    if (n_cs) *n_cs = c->n_cs;
    return LFTS_OK;
}

// ---- replica sets ----------------------------------------------------------------------

int lfts_batch_create(lfts_ctx* const* ctxs, int n, lfts_batch** out) {
    if (!ctxs || n < 1 || !out) return LFTS_EINVAL;
    *out = nullptr;
    for (int i = 0; i < n; i++) {
        if (!ctxs[i]) return LFTS_EINVAL;
        for (int j = 0; j < i; j++)
            if (ctxs[j] == ctxs[i]) return LFTS_EINVAL;
        if (ctxs[i]->ck_stride) return LFTS_EINVAL;  // members keep every pair
        if (ctxs[i]->n_sets > 0) return fail(ctxs[i], LFTS_ESTATE, "batch_create: the context belongs to another set in use");
        // work queued on the contexts' own streams finishes first
        if (hipStreamSynchronize(ctxs[i]->self->st) != hipSuccess) return LFTS_EHIP;
    }
    const int rc = batch_create(ctxs, n, out);
    if (rc == LFTS_OK) {
        (*out)->external = true;
        for (int i = 0; i < n; i++) ctxs[i]->n_sets++;
    }
    return rc;
}

int lfts_batch_destroy(lfts_batch* b) {
    if (!b) return LFTS_EINVAL;
    if (b->st) (void)hipStreamSynchronize(b->st);
    if (b->external)
        for (lfts_ctx* c : b->ctx) c->n_sets--;
    for (auto& q : b->ev_pending)
        for (hipEvent_t e : {q.pe0, q.ch0, q.ch1, q.pe1}) b->ev_pool.push_back(e);
    for (hipEvent_t e : b->ev_pool) (void)hipEventDestroy(e);
    for (hipEvent_t e : b->tune_ev)
        if (e) (void)hipEventDestroy(e);
    for (int g = 0; g < lfts::N_SLAB_EDGES; g++) {
        if (b->gst[g]) {
            (void)hipStreamSynchronize(b->gst[g]);
            (void)hipStreamDestroy(b->gst[g]);
        }
        if (b->ev_join[g]) (void)hipEventDestroy(b->ev_join[g]);
    }
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    for (int part = 0; part < 3; part++) {
        if (b->it_exec[part]) (void)hipGraphExecDestroy(b->it_exec[part]);
        if (b->it_graph[part]) (void)hipGraphDestroy(b->it_graph[part]);
    }
    (void)hipFree(b->d_flags);
    (void)hipFree(b->d_rupd);
    if (b->h_flags) (void)hipHostFree(b->h_flags);
    for (void* p : {(void*)b->d_ract, (void*)b->d_coef, (void*)b->d_part, (void*)b->d_scal, (void*)b->d_qpart})
        (void)hipFree(p);
    for (void* p : {(void*)b->h_ract, (void*)b->h_coef, (void*)b->h_scal, (void*)b->h_seq})
        if (p) (void)hipHostFree(p);
    if (b->st) (void)hipStreamDestroy(b->st);
    delete b;
    return LFTS_OK;
}

const char* lfts_batch_last_error(const lfts_batch* b) { return b ? b->err.c_str() : "null batch"; }

int lfts_batch_get_stream(lfts_batch* b, void** stream) {
    if (!b || !stream) return LFTS_EINVAL;
    *stream = (void*)b->st;
    return LFTS_OK;
}

int lfts_batch_calc_concs(lfts_batch* b, double* lnQ) {
    if (!b) return LFTS_EINVAL;
    return batch_calc_concs(b, lnQ);
}

int lfts_batch_anderson_mix(lfts_batch* b, int maxIter, double errTol, int* k_ret, double* err_ret) {
    if (!b) return LFTS_EINVAL;
    return batch_anderson_mix(b, maxIter, errTol, k_ret, err_ret);
}

int lfts_batch_langevin_step_dev(lfts_batch* b, const double* const* eta_new_dev, const double* XbN,
                                 const double* dt) {
    if (!b || !eta_new_dev || !XbN || !dt) return bfail(b, LFTS_EINVAL, "langevin_step_dev: null argument");
    return batch_langevin(b, eta_new_dev, XbN, dt);
}

int lfts_batch_full_step_dev(lfts_batch* b, const double* const* eta_new_dev, const double* XbN, const double* dt,
                             int maxIter, double errTol, double* lnQ, int* k_ret) {
    if (!b || !eta_new_dev || !XbN || !dt) return bfail(b, LFTS_EINVAL, "full_step_dev: null argument");
    int rc = batch_langevin(b, eta_new_dev, XbN, dt);
    if (rc) return rc;
    rc = batch_anderson_mix(b, maxIter, errTol, k_ret, nullptr);
    if (rc && rc != LFTS_ENUM) return rc;
    const int rc2 = batch_calc_concs(b, lnQ);
    return rc ? rc : rc2;
}

const char* lfts_batch_tuning(const lfts_batch* b) { return b ? b->tune_report.c_str() : ""; }

int lfts_slab_shapes(char* buf, long long cap, long long* len) {
    if (!len) return LFTS_EINVAL;
    std::string t;
    {
        std::lock_guard<std::mutex> lk(g_shape_mu);
        for (const auto& kv : g_used) {
            if (!t.empty()) t += ",";
            t += std::to_string(kv.first) + ":" + shape_name(kv.first, kv.second);
        }
    }
    *len = (long long)t.size();
    if (!buf) return LFTS_OK;
    if (cap < (long long)t.size()) return LFTS_EINVAL;
    std::memcpy(buf, t.data(), t.size());
    return LFTS_OK;
}

const char* lfts_tuning(const lfts_ctx* c) { return (c && c->self) ? c->self->tune_report.c_str() : ""; }

int lfts_batch_full_step(lfts_batch* b, const double* const* eta_new, const double* XbN, const double* dt, int maxIter,
                         double errTol, double* lnQ, int* k_ret) {
    if (!b || !eta_new || !XbN || !dt) return bfail(b, LFTS_EINVAL, "full_step: null argument");
    const int n = (int)b->ctx.size();
    std::vector<const double*> dev(n);
    for (int id = 0; id < n; id++) {
        lfts_ctx* c = b->ctx[id];
        if (!eta_new[id]) return bfail(b, LFTS_EINVAL, "full_step: null noise pointer");
        dev[id] = c->noise + c->M;  // the member's staging buffer
        BHIPCHK(b, hipMemcpyAsync(c->noise + c->M, eta_new[id], sizeof(double) * c->M, hipMemcpyHostToDevice, b->st));
    }
    return lfts_batch_full_step_dev(b, dev.data(), XbN, dt, maxIter, errTol, lnQ, k_ret);
}

int lfts_batch_profile(lfts_batch* b, int enable) {
    if (!b) return LFTS_EINVAL;
    if (int rc = sync_batch(b)) return rc;
    if (enable < 0) return bfail(b, LFTS_EINVAL, "profile: enable must be >= 0");
    b->prof = enable;
    b->prof_seq = 0;
    b->prof_ms_chain = b->prof_ms_pe = 0.0;
    b->prof_n_pass = b->prof_n_pe = 0;
    return LFTS_OK;
}

int lfts_batch_profile_read(lfts_batch* b, double* ms_chain, long long* n_passes, double* ms_pe,
                            long long* n_prop_evals) {
    if (!b) return LFTS_EINVAL;
    if (int rc = sync_batch(b)) return rc;
    if (ms_chain) *ms_chain = b->prof_ms_chain;
    if (n_passes) *n_passes = b->prof_n_pass;
    if (ms_pe) *ms_pe = b->prof_ms_pe;
    if (n_prop_evals) *n_prop_evals = b->prof_n_pe;
    return LFTS_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------
// host random stream and parameters
// ------------------------------------------------------------------------------------
struct lfts_rng {
    std::mt19937_64 eng;
    std::normal_distribution<double> nd;
};

extern "C" {

int lfts_rng_create(unsigned long long seed, double sigma, lfts_rng** out) {
    if (!out) return LFTS_EINVAL;
    lfts_rng* r = new lfts_rng();
    r->eng.seed(seed);
    r->nd = std::normal_distribution<double>(0.0, sigma);
    *out = r;
    return LFTS_OK;
}

int lfts_rng_normals(lfts_rng* r, double* out, long long n) {
    if (!r || (!out && n > 0)) return LFTS_EINVAL;
    for (long long i = 0; i < n; i++) out[i] = r->nd(r->eng);
    return LFTS_OK;
}

int lfts_rng_uniforms(lfts_rng* r, double* out, long long n) {
    if (!r || (!out && n > 0)) return LFTS_EINVAL;
    std::uniform_real_distribution<double> u(0.0, 1.0);
    for (long long i = 0; i < n; i++) out[i] = u(r->eng);
    return LFTS_OK;
}

// The engine and the distribution (including the second normal that the polar method
// keeps for the next call) as text, via the standard library's stream operators.
int lfts_rng_save(lfts_rng* r, char* buf, long long cap, long long* len) {
    if (!r || !len) return LFTS_EINVAL;
    std::ostringstream os;
    os << r->eng << ' ' << r->nd;
    const std::string t = os.str();
    *len = (long long)t.size();
    if (!buf) return LFTS_OK;
    if (cap < *len) return LFTS_EINVAL;
    std::memcpy(buf, t.data(), t.size());
    return LFTS_OK;
}

int lfts_rng_load(lfts_rng* r, const char* buf, long long len) {
    if (!r || !buf || len <= 0) return LFTS_EINVAL;
    std::istringstream is(std::string(buf, (size_t)len));
    std::mt19937_64 eng;
    std::normal_distribution<double> nd;
    is >> eng >> nd;
    if (is.fail()) return LFTS_EINVAL;
    r->eng = eng;
    r->nd = nd;
    return LFTS_OK;
}

int lfts_rng_destroy(lfts_rng* r) {
    delete r;
    return LFTS_OK;
}

double lfts_z_inf(const double* L, const int* m, int N, double Nbar, int tmax) {
    double sum, prod = 1.0, X, ell, R0, R0dL[3];
    sum = 0.5;
    R0 = std::sqrt((double)N);
    for (int i = 0; i < 3; i++) {
        R0dL[i] = R0 * L[i] / m[i];
        prod *= R0dL[i];
    }
    ell = std::pow(prod, 1.0 / 3.0);
    for (int t = 1; t <= tmax; t++) {
        X = (M_PI / ell) * std::sqrt(t / 6.0);
        prod = 1.0;
        for (int i = 0; i < 3; i++) prod *= std::erf(X * ell / R0dL[i]);
        sum += std::pow(std::sqrt(M_PI) / (2 * X), 3.0) * prod;
    }
    sum *= 2 * R0 / (std::pow(ell, 3.0) * std::sqrt(Nbar));
    X = (M_PI / ell) * std::sqrt((0.5 + tmax) / 6.0);
    return 1.0 - sum - 3 * R0 / (ell * std::sqrt(M_PI * Nbar) * X);
}

}  // extern "C"
