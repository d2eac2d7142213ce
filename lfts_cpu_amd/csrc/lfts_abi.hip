// C-ABI layer of the L-FTS Langevin-step library: per-system context, plans, device
// memory, the host control of Anderson mixing (its small LU solve stays on the host),
// the reference's host random stream, and the host-side S(k) binning.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "lfts_gpu.h"
#include "lfts_internal.h"

using lfts::AmCoef;
using lfts::LineArgs;

namespace {

struct AxisPlan {
    int n = 1, nst = 0;
    int radix[lfts::MAXST] = {0};
    double2* tw = nullptr;
    double* gperm = nullptr;
    int* perm = nullptr;
    int T = 1, pitch = 2;
    double* gnat = nullptr;  // kernel in natural frequency order (g(k)/n), for the fused path
};

}  // namespace

struct lfts_ctx {
    lfts_desc d{};
    long M = 0, Mk = 0;
    int N = 0, NA = 0, NB = 0, nhMax = 10;
    hipStream_t st = nullptr;
    AxisPlan ax[3];
    double* w = nullptr;      // 4M: W-, w+, phi-, phi+
    double* h = nullptr;      // 2M: hA, hB (= pair 1)
    double* q = nullptr;      // (N-1) pairs for beads 2..N
    double* tmp = nullptr;    // 2M
    double* tmp2 = nullptr;   // 2M (second partial-convolution buffer of the fused path)
    bool fused = false;       // one launch per contour step (slab passes)
    double* Dh = nullptr;     // (nhMax+1) M
    double* wh = nullptr;     // (nhMax+1) M
    double* noise = nullptr;  // 2M: eta_prev | staging for eta_new
    int eta_prev = 0;         // which half holds eta_prev
    double* S = nullptr;      // Mk
    double2* zk = nullptr;    // M, S(k) transform scratch
    double* part = nullptr;   // nred * PSTRIDE
    double* scal = nullptr;   // 64
    double* pinned = nullptr; // host staging, 64 doubles
    double* test = nullptr;   // 4M scratch for lfts_step_fwd
    int nsamples = 0;
    int nblk = 1, nred = 1;
    long long n_pe = 0, n_cs = 0;
    // live event timing (lfts_profile)
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;  // free events
    struct EvQuad { hipEvent_t pe0, ch0, ch1, pe1; int nsteps; };
    std::vector<EvQuad> ev_pending;
    double prof_ms_chain = 0.0, prof_ms_pe = 0.0;
    long long prof_n_cs = 0, prof_n_pe = 0;
    std::vector<double> DD;   // (nhMax+1)^2
    std::vector<double> U, V, C;
    std::string err;
};

static int fail(lfts_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(ctx, expr)                                                                     \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail((ctx), LFTS_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

static bool factorize(int n, std::vector<int>& r) {
    r.clear();
    while (n % 4 == 0) { r.push_back(4); n /= 4; }
    while (n % 2 == 0) { r.push_back(2); n /= 2; }
    for (int p : {3, 5, 7})
        while (n % p == 0) { r.push_back(p); n /= p; }
    return n == 1;
}

// Position of frequency k after the in-place decimation-in-frequency pass: a stage of
// radix r on a block of size B leaves the sub-DFT (k mod r) in sub-block k mod r, so
// pos_B(k) = (B/r)(k mod r) + pos_{B/r}(k div r).
static int dif_pos(int k, const std::vector<int>& rad, size_t s, int B) {
    if (s == rad.size()) return 0;
    const int r = rad[s], S = B / r;
    return S * (k % r) + dif_pos(k / r, rad, s + 1, S);
}

static int build_axis(lfts_ctx* c, AxisPlan& p, int n, double L, int N) {
    std::vector<int> rad;
    if (!factorize(n, rad) || (int)rad.size() > lfts::MAXST)
        return fail(c, LFTS_EINVAL, "mesh size " + std::to_string(n) + " has a prime factor above 7");
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
    std::vector<double> gperm(n);
    std::vector<int> perm(n);
    for (int k = 0; k < n; k++) {
        const int f = (k <= n / 2) ? k : k - n;
        const double K = 2.0 * M_PI * (double)f / L;
        const int pos = dif_pos(k, rad, 0, n);
        perm[k] = pos;
        gperm[pos] = std::exp(-K * K / (6.0 * N)) / (double)n;
    }
    std::vector<double> gnat(n);
    for (int k = 0; k < n; k++) gnat[k] = gperm[perm[k]];
    HIPCHK(c, hipMalloc(&p.tw, sizeof(double2) * n));
    HIPCHK(c, hipMalloc(&p.gperm, sizeof(double) * n));
    HIPCHK(c, hipMalloc(&p.perm, sizeof(int) * n));
    HIPCHK(c, hipMalloc(&p.gnat, sizeof(double) * n));
    HIPCHK(c, hipMemcpy(p.tw, tw.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(p.gperm, gperm.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(p.perm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(p.gnat, gnat.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    p.T = std::max(1, std::min(16, 4096 / n));
    p.pitch = n + 1;
    return LFTS_OK;
}

static LineArgs line_args(const lfts_ctx* c, int axis, int mode) {
    const AxisPlan& p = c->ax[axis];
    LineArgs a;
    std::memset(&a, 0, sizeof a);
    a.n = p.n;
    a.nst = p.nst;
    for (int s = 0; s < p.nst; s++) a.radix[s] = p.radix[s];
    a.tw = p.tw;
    a.gperm = p.gperm;
    a.perm = p.perm;
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

static double* pair_ptr(lfts_ctx* c, int i) { return (i == 1) ? c->h : c->q + 2L * (i - 2) * c->M; }

// src/step.cc:35-65: out = h_sel * conv_g(in) for the pair (q_i, q+_{N+1-i}).
static int contour_step(lfts_ctx* c, const double* in, double* out, const double* h, int i) {
    const long M = c->M;
    LineArgs z = line_args(c, 2, lfts::LINE_CONV);
    z.inA = in; z.inB = in + M; z.outA = c->tmp; z.outB = c->tmp + M;
    HIPCHK(c, lfts::launch_line(z, c->st));
    LineArgs y = line_args(c, 1, lfts::LINE_CONV);
    y.inA = c->tmp; y.inB = c->tmp + M; y.outA = c->tmp; y.outB = c->tmp + M;
    HIPCHK(c, lfts::launch_line(y, c->st));
    LineArgs x = line_args(c, 0, lfts::LINE_CONV);
    x.inA = c->tmp; x.inB = c->tmp + M; x.outA = out; x.outB = out + M;
    x.hA = (i < c->NA) ? h : h + M;       // q_{i+1}   (src/step.cc:53-57)
    x.hB = (i < c->NB) ? h + M : h;       // q+_{N-i} (src/step.cc:60-64)
    HIPCHK(c, lfts::launch_line(x, c->st));
    c->n_cs++;
    return LFTS_OK;
}

static int get_event(lfts_ctx* c, hipEvent_t* e) {
    if (!c->ev_pool.empty()) {
        *e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return LFTS_OK;
    }
    HIPCHK(c, hipEventCreate(e));
    return LFTS_OK;
}

// Fold completed event quads into the profile sums (the stream has been synchronised).
static int resolve_events(lfts_ctx* c) {
    for (auto& q : c->ev_pending) {
        float a = 0.f, b = 0.f;
        HIPCHK(c, hipEventElapsedTime(&a, q.ch0, q.ch1));
        HIPCHK(c, hipEventElapsedTime(&b, q.pe0, q.pe1));
        c->prof_ms_chain += a;
        c->prof_ms_pe += b;
        c->prof_n_cs += q.nsteps;
        c->prof_n_pe += 1;
        for (hipEvent_t e : {q.pe0, q.ch0, q.ch1, q.pe1}) c->ev_pool.push_back(e);
    }
    c->ev_pending.clear();
    return LFTS_OK;
}

// The N-1 contour steps of src/diblock.cc:43 as alternating slab passes, one launch
// per step: a prologue x-slab pass convolves q_1 along y and z, then pass i (i = 1..N-1)
// finishes step i (y-slab for odd i, x-slab for even i), stores q_{i+1} and starts step
// i+1.
static int fused_chain(lfts_ctx* c) {
    const long M = c->M;
    const int n = c->d.m[0];
    const long m1 = c->d.m[1], m2 = c->d.m[2];
    double* T[2] = {c->tmp, c->tmp2};
    auto setup = [&](lfts::SlabArgs& a, bool xslab) {
        std::memset(&a, 0, sizeof a);
        a.M = M;
        a.tw = c->ax[2].tw;
        a.gb = c->ax[2].gnat;
        if (xslab) {  // plane x0: outer axis y
            a.plane_stride = m1 * m2;
            a.Sa = m2;
            a.ga = c->ax[1].gnat;
        } else {      // plane y0: outer axis x
            a.plane_stride = m2;
            a.Sa = m1 * m2;
            a.ga = c->ax[0].gnat;
        }
    };
    lfts::SlabArgs a;
    setup(a, true);
    a.Tin = c->h;  // pair 1 = [q_1 | q+_N] = [hA | hB]
    a.Tout = T[0];
    a.finish = 0;
    a.cont = 1;
    HIPCHK(c, lfts::launch_slab(a, n, c->d.m[0], c->st));
    for (int i = 1; i < c->N; i++) {
        const bool xslab = (i % 2 == 0);
        setup(a, xslab);
        a.Tin = T[(i - 1) & 1];
        a.Tout = T[i & 1];
        a.qout = pair_ptr(c, i + 1);
        a.hq = (i < c->NA) ? c->h : c->h + M;      // q_{i+1}   (src/step.cc:53-57)
        a.hqd = (i < c->NB) ? c->h + M : c->h;     // q+_{N-i} (src/step.cc:60-64)
        a.finish = 1;
        a.cont = (i < c->N - 1) ? 1 : 0;
        HIPCHK(c, lfts::launch_slab(a, n, xslab ? c->d.m[0] : c->d.m[1], c->st));
        c->n_cs++;
    }
    return LFTS_OK;
}

// src/diblock.cc:30-60, without the host read of Q; scal[0] = Q.
static int calc_concs_async(lfts_ctx* c) {
    const long M = c->M;
    lfts_ctx::EvQuad ev{};
    if (c->prof) {
        for (hipEvent_t* e : {&ev.pe0, &ev.ch0, &ev.ch1, &ev.pe1})
            if (int rc = get_event(c, e)) return rc;
        HIPCHK(c, hipEventRecord(ev.pe0, c->st));
    }
    HIPCHK(c, lfts::launch_prepare_h(c->w, c->h, M, c->N, c->nblk, c->st));
    if (c->prof) HIPCHK(c, hipEventRecord(ev.ch0, c->st));
    if (c->fused) {
        if (int rc = fused_chain(c)) return rc;
    } else {
        for (int i = 1; i < c->N; i++) {
            int rc = contour_step(c, pair_ptr(c, i), pair_ptr(c, i + 1), c->h, i);
            if (rc) return rc;
        }
    }
    if (c->prof) HIPCHK(c, hipEventRecord(ev.ch1, c->st));
    HIPCHK(c, lfts::launch_sum_partial(pair_ptr(c, c->N), M, c->part, c->nred, c->st));
    HIPCHK(c, lfts::launch_reduce_final(c->part, c->nred, 1, (double)M, c->scal, c->st));
    HIPCHK(c, lfts::launch_phi(c->h, c->q, c->w, c->scal, M, c->N, c->NA, c->nblk, c->st));
    if (c->prof) {
        HIPCHK(c, hipEventRecord(ev.pe1, c->st));
        ev.nsteps = c->N - 1;
        c->ev_pending.push_back(ev);
    }
    c->n_pe++;
    return LFTS_OK;
}

static int read_scalars(lfts_ctx* c, int n, double* out) {
    HIPCHK(c, hipMemcpyAsync(c->pinned, c->scal, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    std::memcpy(out, c->pinned, sizeof(double) * n);
    if (!c->ev_pending.empty()) return resolve_events(c);
    return LFTS_OK;
}

// LU with partial pivoting and solve (the GSL LU_decomp/LU_solve of src/anderson.cc:118-126).
static bool lu_solve(int n, const double* A_in, const double* b, double* x) {
    std::vector<double> A(A_in, A_in + (size_t)n * n);
    std::vector<int> p(n);
    std::iota(p.begin(), p.end(), 0);
    bool singular = false;
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
        if (ajj == 0.0) { singular = true; continue; }
        for (int i = j + 1; i < n; i++) {
            const double l = A[(size_t)i * n + j] / ajj;
            A[(size_t)i * n + j] = l;
            for (int k = j + 1; k < n; k++) A[(size_t)i * n + k] -= l * A[(size_t)j * n + k];
        }
    }
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
    return !singular;
}

static int check_ctx(lfts_ctx* c) { return c ? LFTS_OK : LFTS_EINVAL; }

extern "C" {

int lfts_create(const lfts_desc* desc, lfts_ctx** out) {
    if (!desc || !out) return LFTS_EINVAL;
    *out = nullptr;
    lfts_ctx* c = new lfts_ctx();
    c->d = *desc;
    c->N = desc->N;
    c->NA = desc->NA;
    c->NB = desc->N - desc->NA;
    c->nhMax = desc->nhMax;
    auto bad = [&](const char*) {
        delete c;
        return LFTS_EINVAL;
    };
    if (c->N < 2 || c->NA < 0 || c->NA > c->N) return bad("need N >= 2 and 0 <= NA <= N");
    if (c->nhMax < 1 || c->nhMax > lfts::MAX_HIST) return bad("nhMax must be in [1, 15]");
    for (int d = 0; d < 3; d++)
        if (desc->m[d] < 1 || !(desc->L[d] > 0.0)) return bad("mesh sizes and box edges must be positive");
    c->M = (long)desc->m[0] * desc->m[1] * desc->m[2];
    c->Mk = (long)desc->m[0] * desc->m[1] * (desc->m[2] / 2 + 1);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        delete c;
        return LFTS_EHIP;
    }
    const long M = c->M;
    int rc = LFTS_OK;
#define CREATE_CHK(expr)                 \
    do {                                 \
        rc = (expr);                     \
        if (rc) { lfts_destroy(c); return rc; } \
    } while (0)
#define CREATE_HIP(expr) CREATE_CHK((expr) == hipSuccess ? LFTS_OK : LFTS_EHIP)
    CREATE_HIP(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    for (int d = 0; d < 3; d++) CREATE_CHK(build_axis(c, c->ax[d], desc->m[d], desc->L[d], c->N));
    CREATE_HIP(hipMalloc(&c->w, sizeof(double) * 4 * M));
    CREATE_HIP(hipMalloc(&c->h, sizeof(double) * 2 * M));
    CREATE_HIP(hipMalloc(&c->q, sizeof(double) * 2 * M * (size_t)(c->N - 1)));
    CREATE_HIP(hipMalloc(&c->tmp, sizeof(double) * 2 * M));
    {
        // LFTS_PATH=generic selects the three-pass line kernels even where the fused
        // slab pass applies (used to cross-check the two paths)
        const char* path = std::getenv("LFTS_PATH");
        const bool generic = path && std::strcmp(path, "generic") == 0;
        c->fused = !generic && lfts::slab_supported(desc->m[0], desc->m[1], desc->m[2]);
        if (c->fused) CREATE_HIP(hipMalloc(&c->tmp2, sizeof(double) * 2 * M));
    }
    CREATE_HIP(hipMalloc(&c->Dh, sizeof(double) * M * (size_t)(c->nhMax + 1)));
    CREATE_HIP(hipMalloc(&c->wh, sizeof(double) * M * (size_t)(c->nhMax + 1)));
    CREATE_HIP(hipMalloc(&c->noise, sizeof(double) * 2 * M));
    CREATE_HIP(hipMalloc(&c->S, sizeof(double) * c->Mk));
    c->nblk = (int)std::min<long>((M + 255) / 256, 2048);
    c->nred = (int)std::min<long>((M + 255) / 256, 1024);
    CREATE_HIP(hipMalloc(&c->part, sizeof(double) * lfts::PSTRIDE * c->nred));
    CREATE_HIP(hipMalloc(&c->scal, sizeof(double) * 64));
    CREATE_HIP(hipHostMalloc(&c->pinned, sizeof(double) * 64, hipHostMallocDefault));
    CREATE_HIP(hipMemset(c->w, 0, sizeof(double) * 4 * M));
    CREATE_HIP(hipMemset(c->noise, 0, sizeof(double) * 2 * M));
    CREATE_HIP(hipMemset(c->S, 0, sizeof(double) * c->Mk));
    const int H = c->nhMax + 1;
    c->DD.assign((size_t)H * H, 0.0);
    c->U.assign((size_t)c->nhMax * c->nhMax, 0.0);
    c->V.assign(c->nhMax, 0.0);
    c->C.assign(c->nhMax, 0.0);
#undef CREATE_CHK
#undef CREATE_HIP
    *out = c;
    return LFTS_OK;
}

int lfts_destroy(lfts_ctx* c) {
    if (!c) return LFTS_EINVAL;
    if (c->st) hipStreamSynchronize(c->st);
    for (auto& q : c->ev_pending)
        for (hipEvent_t e : {q.pe0, q.ch0, q.ch1, q.pe1}) c->ev_pool.push_back(e);
    for (hipEvent_t e : c->ev_pool) hipEventDestroy(e);
    for (auto& p : c->ax) {
        hipFree(p.tw);
        hipFree(p.gperm);
        hipFree(p.perm);
        hipFree(p.gnat);
    }
    hipFree(c->tmp2);
    hipFree(c->w); hipFree(c->h); hipFree(c->q); hipFree(c->tmp); hipFree(c->Dh); hipFree(c->wh);
    hipFree(c->noise); hipFree(c->S); hipFree(c->zk); hipFree(c->part); hipFree(c->scal); hipFree(c->test);
    if (c->pinned) hipHostFree(c->pinned);
    if (c->st) hipStreamDestroy(c->st);
    delete c;
    return LFTS_OK;
}

const char* lfts_last_error(const lfts_ctx* c) { return c ? c->err.c_str() : "null context"; }

int lfts_get_stream(lfts_ctx* c, void** stream) {
    if (!c || !stream) return LFTS_EINVAL;
    *stream = (void*)c->st;
    return LFTS_OK;
}

int lfts_set_fields(lfts_ctx* c, const double* w, int nfields) {
    if (!c || !w || (nfields != 2 && nfields != 4)) return fail(c, LFTS_EINVAL, "set_fields: nfields must be 2 or 4");
    HIPCHK(c, hipMemcpyAsync(c->w, w, sizeof(double) * nfields * c->M, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return LFTS_OK;
}

int lfts_get_fields(lfts_ctx* c, double* w, int nfields) {
    if (!c || !w || (nfields != 2 && nfields != 4)) return fail(c, LFTS_EINVAL, "get_fields: nfields must be 2 or 4");
    HIPCHK(c, hipMemcpyAsync(w, c->w, sizeof(double) * nfields * c->M, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return LFTS_OK;
}

int lfts_fields_dev(lfts_ctx* c, double** w_dev) {
    if (!c || !w_dev) return LFTS_EINVAL;
    *w_dev = c->w;
    return LFTS_OK;
}

int lfts_calc_concs(lfts_ctx* c, double* lnQ) {
    if (check_ctx(c)) return LFTS_EINVAL;
    int rc = calc_concs_async(c);
    if (rc) return rc;
    double Q;
    if ((rc = read_scalars(c, 1, &Q))) return rc;
    if (lnQ) *lnQ = std::log(Q);
    if (!std::isfinite(Q) || Q <= 0.0) return fail(c, LFTS_ENUM, "partition function Q is not finite and positive");
    return LFTS_OK;
}

// src/anderson.cc:24-93
int lfts_anderson_mix(lfts_ctx* c, int maxIter, double errTol, int* k_ret, double* err_ret) {
    if (check_ctx(c)) return LFTS_EINVAL;
    const long M = c->M;
    const int H = c->nhMax + 1;
    auto DD = [&](int i, int j) -> double& { return c->DD[(size_t)i * H + j]; };
    double err = 1.0, lambda;
    int k, nh, rc;
    double dd[lfts::MAX_HIST + 1];
    for (k = 1; k < maxIter && err > errTol; k++) {
        if ((rc = calc_concs_async(c))) return rc;
        nh = (k < c->nhMax + 1) ? k - 1 : c->nhMax;
        HIPCHK(c, lfts::launch_am_dots(c->w, c->Dh, M, nh, c->part, c->nred, c->st));
        HIPCHK(c, lfts::launch_reduce_final(c->part, c->nred, nh + 1, 1.0, c->scal, c->st));
        if ((rc = read_scalars(c, nh + 1, dd))) return rc;
        const double S1 = dd[0];
        err = std::pow(S1 / M, 0.5);
        lambda = 1.0 - std::pow(0.9, double(k));
        for (int m = 0; m <= nh; m++) DD(0, m) = dd[m];
        AmCoef co;
        std::memset(&co, 0, sizeof co);
        co.lambda = lambda;
        co.nh = 0;
        if (k >= 2) {
            for (int m = 1; m <= nh; m++) {
                c->V[m - 1] = DD(0, 0) - DD(0, m);
                for (int n = 1; n <= m; n++)
                    c->U[(m - 1) * nh + n - 1] = c->U[(n - 1) * nh + m - 1] = DD(0, 0) - DD(0, m) - DD(0, n) + DD(n, m);
            }
            lu_solve(nh, c->U.data(), c->V.data(), c->C.data());
            co.nh = nh;
            for (int n = 0; n < nh; n++) co.C[n] = c->C[n];
        }
        const int n = 1 + (k - 1) % c->nhMax;
        co.slot = n;
        HIPCHK(c, lfts::launch_am_update(c->w, c->Dh, c->wh, M, co, c->nblk, c->st));
        DD(n, n) = DD(0, 0);
        for (int m = 1; m < n; m++) DD(m, n) = DD(0, m);
        for (int m = n + 1; m <= nh; m++) DD(n, m) = DD(0, m);
        if (std::isnan(err)) break;
    }
    if (k_ret) *k_ret = k;
    if (err_ret) *err_ret = err;
    if (!std::isfinite(err)) return fail(c, LFTS_ENUM, "Anderson residual is not finite");
    return LFTS_OK;
}

int lfts_langevin_init(lfts_ctx* c, const double* eta_prev) {
    if (!c || !eta_prev) return fail(c, LFTS_EINVAL, "langevin_init: null argument");
    c->eta_prev = 0;
    HIPCHK(c, hipMemcpyAsync(c->noise, eta_prev, sizeof(double) * c->M, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return LFTS_OK;
}

int lfts_langevin_step(lfts_ctx* c, const double* eta_new, double XbN, double dt) {
    if (!c || !eta_new) return fail(c, LFTS_EINVAL, "langevin_step: null argument");
    const long M = c->M;
    double* ep = c->noise + (long)c->eta_prev * M;
    double* en = c->noise + (long)(1 - c->eta_prev) * M;
    HIPCHK(c, hipMemcpyAsync(en, eta_new, sizeof(double) * M, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, lfts::launch_langevin(c->w, ep, en, M, XbN, dt, 0, c->nblk, c->st));
    c->eta_prev = 1 - c->eta_prev;  // src/langevin.cc:34-36
    HIPCHK(c, hipStreamSynchronize(c->st));
    return LFTS_OK;
}

int lfts_langevin_step_dev(lfts_ctx* c, const double* eta_new_dev, double XbN, double dt) {
    if (!c || !eta_new_dev) return fail(c, LFTS_EINVAL, "langevin_step_dev: null argument");
    double* ep = c->noise + (long)c->eta_prev * c->M;
    HIPCHK(c, lfts::launch_langevin(c->w, ep, eta_new_dev, c->M, XbN, dt, 1, c->nblk, c->st));
    return LFTS_OK;
}

int lfts_langevin_full_step_dev(lfts_ctx* c, const double* eta_new_dev, double XbN, double dt, int maxIter,
                                double errTol, double* lnQ, int* k_ret) {
    int rc = lfts_langevin_step_dev(c, eta_new_dev, XbN, dt);
    if (rc) return rc;
    rc = lfts_anderson_mix(c, maxIter, errTol, k_ret, nullptr);
    if (rc && rc != LFTS_ENUM) return rc;
    int rc2 = lfts_calc_concs(c, lnQ);
    return rc ? rc : rc2;
}

// src/strFunc.cc:36-46
int lfts_sk_sample(lfts_ctx* c) {
    if (check_ctx(c)) return LFTS_EINVAL;
    if (!c->zk) HIPCHK(c, hipMalloc(&c->zk, sizeof(double2) * c->M));
    LineArgs z = line_args(c, 2, lfts::LINE_FFT);
    z.inR = c->w;
    z.outC = c->zk;
    HIPCHK(c, lfts::launch_line(z, c->st));
    LineArgs y = line_args(c, 1, lfts::LINE_FFT);
    y.inC = c->zk;
    y.outC = c->zk;
    HIPCHK(c, lfts::launch_line(y, c->st));
    LineArgs x = line_args(c, 0, lfts::LINE_FFT);
    x.inC = c->zk;
    x.Sacc = c->S;
    HIPCHK(c, lfts::launch_line(x, c->st));
    c->nsamples++;
    return LFTS_OK;
}

int lfts_sk_get(lfts_ctx* c, double* S, int* nsamples) {
    if (check_ctx(c)) return LFTS_EINVAL;
    if (S) {
        HIPCHK(c, hipMemcpyAsync(S, c->S, sizeof(double) * c->Mk, hipMemcpyDeviceToHost, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
    }
    if (nsamples) *nsamples = c->nsamples;
    return LFTS_OK;
}

int lfts_sk_reset(lfts_ctx* c) {
    if (check_ctx(c)) return LFTS_EINVAL;
    HIPCHK(c, hipMemsetAsync(c->S, 0, sizeof(double) * c->Mk, c->st));
    c->nsamples = 0;
    return LFTS_OK;
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
    if (check_ctx(c)) return LFTS_EINVAL;
    int rc = calc_concs_async(c);
    if (rc) return rc;
    HIPCHK(c, lfts::launch_h_partial(c->w, c->M, c->part, c->nred, c->st));
    HIPCHK(c, lfts::launch_reduce_final(c->part, c->nred, 2, 1.0, c->scal + 1, c->st));
    double s[3];
    if ((rc = read_scalars(c, 3, s))) return rc;
    const double lnQ = std::log(s[0]);
    if (H) *H = -lnQ + (s[2] / XbN - s[1]) / c->M;
    return std::isfinite(lnQ) ? LFTS_OK : fail(c, LFTS_ENUM, "partition function Q is not finite and positive");
}

int lfts_get_pair(lfts_ctx* c, int i, double* out) {
    if (!c || !out || i < 1 || i > c->N) return fail(c, LFTS_EINVAL, "get_pair: index out of range");
    HIPCHK(c, hipMemcpyAsync(out, pair_ptr(c, i), sizeof(double) * 2 * c->M, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return LFTS_OK;
}

int lfts_step_fwd(lfts_ctx* c, const double* q_in, double* q_out, const double* h, int i) {
    if (!c || !q_in || !q_out || !h) return fail(c, LFTS_EINVAL, "step_fwd: null argument");
    const long M = c->M;
    if (!c->test) HIPCHK(c, hipMalloc(&c->test, sizeof(double) * 6 * M));
    HIPCHK(c, hipMemcpyAsync(c->test, q_in, sizeof(double) * 2 * M, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync(c->test + 2 * M, h, sizeof(double) * 2 * M, hipMemcpyHostToDevice, c->st));
    int rc = contour_step(c, c->test, c->test + 4 * M, c->test + 2 * M, i);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(q_out, c->test + 4 * M, sizeof(double) * 2 * M, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return LFTS_OK;
}

int lfts_profile(lfts_ctx* c, int enable) {
    if (check_ctx(c)) return LFTS_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (int rc = resolve_events(c)) return rc;
    c->prof = enable != 0;
    c->prof_ms_chain = c->prof_ms_pe = 0.0;
    c->prof_n_cs = c->prof_n_pe = 0;
    return LFTS_OK;
}

int lfts_profile_read(lfts_ctx* c, double* ms_chain, long long* n_cs, double* ms_pe, long long* n_pe) {
    if (check_ctx(c)) return LFTS_EINVAL;
    HIPCHK(c, hipStreamSynchronize(c->st));
    if (int rc = resolve_events(c)) return rc;
    if (ms_chain) *ms_chain = c->prof_ms_chain;
    if (n_cs) *n_cs = c->prof_n_cs;
    if (ms_pe) *ms_pe = c->prof_ms_pe;
    if (n_pe) *n_pe = c->prof_n_pe;
    return LFTS_OK;
}

int lfts_counters(lfts_ctx* c, long long* n_pe, long long* n_cs) {
    if (check_ctx(c)) return LFTS_EINVAL;
    if (n_pe) *n_pe = c->n_pe;
    if (n_cs) *n_cs = c->n_cs;
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
