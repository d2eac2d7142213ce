// Host-side state of contexts and replica sets, shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <map>
#include <tuple>
#include <string>
#include <utility>
#include <vector>

#include "lfts_gpu.h"
#include "lfts_internal.h"

namespace lfts_abi {

struct AxisPlan {
    int n = 1, nst = 0;
    int radix[lfts::MAXST] = {0};
    double2* tw = nullptr;    // exp(-2 pi i k/n)
    double* gperm = nullptr;  // kernel at in-place DIF positions (line path)
    int* perm = nullptr;      // DIF position of frequency k
    double* gnat = nullptr;   // kernel in natural frequency order (slab path)
    double* kap = nullptr;    // d g/dL / g in natural frequency order: (2 pi f)^2 / (3 N L^3)
    int T = 1, pitch = 2;
};

}  // namespace lfts_abi

struct lfts_batch;

struct lfts_ctx {
    lfts_desc d{};
    long M = 0, Mk = 0;
    int N = 0, NA = 0, NB = 0, nhMax = 10;
    lfts_abi::AxisPlan ax[3];
    bool fused = false;       // cubic mesh of a slab edge (SLAB_EDGES): one slab launch per contour step
    bool fast_lines = true;   // line-path convolutions of slab-edge lengths: register-FFT axis kernel
    double* w = nullptr;      // 4M
    double* h = nullptr;      // 2M (pair 1)
    double* q = nullptr;      // (N-1) pairs
    double* T0 = nullptr;     // 2M partial pair (slab path) / line-path scratch
    double* T1 = nullptr;     // 2M partial pair (slab path)
    double* phacc = nullptr;  // 4M fused density sums (slab path)
    bool pairs_all = true;    // the last evaluation stored every pair
    double* Dh = nullptr;     // (nhMax+1) M
    double* wh = nullptr;     // (nhMax+1) M
    double* noise = nullptr;  // 2M: eta_prev | staging for a host eta_new
    double* S = nullptr;      // Mk
    double2* zk = nullptr;    // M, transform scratch (allocated on first use)
    double* test = nullptr;   // 8M, lfts_step_fwd scratch (allocated on first use)
    // derivative chains (allocated on first use)
    double* dq = nullptr;     // (N-1) pairs of a tangent or adjoint chain (pairs 2..N)
    double* dbuf = nullptr;   // 10M: d1 (2M) | d2 (2M) | aux (2M) | io (4M)
    double* acc = nullptr;    // M, spectral accumulator of the box derivative
    double* dred = nullptr;   // reduction partials (32 x MAX_NB) and 64 results
    double* hred = nullptr;   // pinned host copy of the results
    // checkpointed pair storage (ck_stride > 0): pairs kept only at segment starts
    int ck_stride = 0;
    int n_sets = 0;  // replica sets in use (lfts_batch_create) that hold this context's addresses
    std::vector<int> ck_b;      // segment starts b_0 = 1 < ... < b_{K-1}, and b_K = N + 1
    int ck_smax = 0;            // longest segment (pairs)
    double* ck_f = nullptr;     // forward pairs at b_1..b_{K-1}, then pair N
    double* ck_d = nullptr;     // derived-chain pairs at the same places, then pair N
    double* ck_seg = nullptr;   // 4 segment buffers of ck_smax pairs
    double* ck_ring = nullptr;  // 4 pairs: forward pairs in flight, adjoint carry, scratch
    double* ck_T = nullptr;     // 4M: slab pipeline of a derived chain run beside the forward one
    double* ck_acc = nullptr;   // 4M: raw bead sums
    // Newton-Krylov workspace (allocated on first use)
    double* kry = nullptr;    // (kry_m + 1) M Krylov basis, then 5M work vectors
    int kry_m = 0;
    double* pmul = nullptr;   // M, preconditioner multiplier in the x-line transform order
    // Anderson mixing on the device (single context): loop state, (0.9)^k table, the
    // captured graph of one iteration
    lfts::AmState* am = nullptr;
    lfts::AmState* am_h = nullptr;  // pinned
    double* pow09 = nullptr;
    int pow09_n = 0;
    int am_last = 0;                // iterations of the previous solve (chunk size)
    hipGraph_t am_graph = nullptr;
    hipGraphExec_t am_exec = nullptr;
    int nsamples = 0;
    long long n_pe = 0, n_cs = 0, n_tan = 0, n_adj = 0;
    std::vector<double> DD, U, V, C;  // Anderson host state (src/anderson.cc:8-21)
    lfts_batch* self = nullptr;       // set of one for the single-context entry points
    std::string err;
    const char* err_lit = nullptr;  // static text of the last failure of a call that may not allocate
    int dev = 0;                    // the HIP device the context was created on
};

struct lfts_batch {
    std::vector<lfts_ctx*> ctx;
    bool external = false;  // made by lfts_batch_create (not a context's own set of one)
    hipStream_t st = nullptr;
    std::vector<lfts::Rep> reps;  // descriptors of all members (id = member index)
    lfts::Rep* d_ract = nullptr;  // descriptors of the active members, in launch order
    lfts::Rep* h_ract = nullptr;  // pinned staging of d_ract
    int nb_max = 1;               // largest workgroup count among the active members
    lfts::Coef* d_coef = nullptr;
    lfts::Coef* h_coef = nullptr;  // pinned
    double* d_part = nullptr;
    double* d_scal = nullptr;
    double* d_qpart = nullptr;  // Q's partials, MAX_NB per member: plane sums (slab) or k_q_sum's
    // pinned and coherent: the Gram row and Q reductions also store their results here
    // (h_scal_dev is its device address), so no copy follows them
    double* h_scal = nullptr;
    double* h_scal_dev = nullptr;
    // pinned and coherent: per launch slot, the sequence number of the last reduction whose
    // results are stored (the host polls it instead of synchronising the stream)
    unsigned* h_seq = nullptr;
    unsigned* h_seq_dev = nullptr;
    unsigned seq = 0;  // wraps around
    // the active set as launched: slab replicas sorted by edge, then line-path ones
    std::vector<int> act;
    struct Group { int edge, off, cnt; const double2* tw; };
    std::vector<Group> groups;
    std::vector<int> line_reps;
    // one stream per slab group, so the groups of a contour step overlap on the GPU;
    // forked from and joined back to `st` around each propagator chain
    hipStream_t gst[lfts::N_SLAB_EDGES] = {};
    hipEvent_t ev_fork = nullptr;
    hipEvent_t ev_join[lfts::N_SLAB_EDGES] = {};
    // slab kernel variant per (mesh edge, members in the launch), picked by timing
    std::map<std::tuple<int, int, int>, int> tuned;  // (edge, launch-size class, fused) -> shape
    std::string tune_report;
    hipEvent_t tune_ev[2] = {nullptr, nullptr};
    // live event timing
    int prof = 0;            // live timing: every prof-th propagator evaluation (0: off)
    long long prof_seq = 0;  // evaluations since the timing was enabled
    std::vector<hipEvent_t> ev_pool;
    struct EvQuad { hipEvent_t pe0, ch0, ch1, pe1; long long npasses; int nrep; };
    std::vector<EvQuad> ev_pending;
    double prof_ms_chain = 0.0, prof_ms_pe = 0.0;
    long long prof_n_pass = 0, prof_n_pe = 0;
    // A set of two or more slab-path members runs each Anderson iteration after the first
    // as one HIP graph (it_exec): coefficient and flag uploads, the update, the propagator
    // evaluation on the groups' streams, the Gram row and its reduction. Every member stays
    // in every launch of the graph; per-member flags (d_flags: n evaluation flags, then n
    // update flags, each member's descriptors pointing at its own) make the kernels of a
    // member that has left the loop return at once.
    int* d_flags = nullptr;       // 2n ints
    int* h_flags = nullptr;       // pinned staging
    lfts::Rep* d_rupd = nullptr;  // all members' descriptors with the update flags (launch order)
    // three graphs: before the chains (uploads, update), the chains, after them (Q, phi and
    // the Gram row), so that live timing can record events between them
    hipGraph_t it_graph[3] = {};
    hipGraphExec_t it_exec[3] = {};
    long long it_npasses = 0;     // contour-step launches of the chain graph
    int it_ncomp = 0;             // Gram-row components the graph's reduction sums
    bool rupd_ready = false;      // d_rupd uploaded
    std::string err;
};

#define HIPCHK(ctx, expr)                                                                                \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess)                                                                            \
            return lfts_abi::fail((ctx), LFTS_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)
#define BHIPCHK(b, expr)                                                                                  \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            return lfts_abi::bfail((b), LFTS_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace lfts_abi {

int fail(lfts_ctx* c, int code, const std::string& msg);
int bfail(lfts_batch* b, int code, const std::string& msg);
lfts::LineArgs line_args(const lfts_ctx* c, int axis, int mode);
double* pair_ptr(lfts_ctx* c, int i);
int line_contour_step(lfts_ctx* c, const double* in, double* out, const double* h, int i, double* tmp,
                      hipStream_t st, const int* skip);
int set_active(lfts_batch* b, const std::vector<int>& ids);
// store_all = false lets slab-path replicas fuse the density sums into the chain, which
// keeps only pairs 1..ceil(N/2) and N; store_all = true keeps every pair.
// h_ready: the active members' Boltzmann factors already match their fields (written by
// the Anderson update just before), so the prepare_h pass is skipped. dots: the phi pass
// also writes the Gram row partials of each member's next Anderson iteration, with the
// history depths of the uploaded coefficients (DOTS_COEF) or, for one member, that of
// the pinned host coefficients passed as a kernel argument (DOTS_ONE).
enum { DOTS_NONE = 0, DOTS_COEF = 1, DOTS_ONE = 2 };
// part: the whole evaluation (PART_ALL), or for a captured Anderson iteration only its
// chains (PART_CHAIN, *npasses_out receives the contour-step launches) or only Q and phi
// (PART_PHI); the parts record no timing events and count no evaluations
enum { PART_ALL = 0, PART_CHAIN = 1, PART_PHI = 2 };
int calc_concs_async(lfts_batch* b, bool store_all = false, bool h_ready = false, int dots = DOTS_NONE,
                     int part = PART_ALL, long long* npasses_out = nullptr);
int slab_variant(lfts_batch* b, int edge, int cnt, const lfts::Rep* d_reps, const lfts::Rep* h_reps,
                 const double2* tw, hipStream_t s, int* variant, int fuse = 0);
// Waits for the set's stream and makes h_scal current: on_host when every value the caller
// reads was stored there by the kernels (launch_reduce or launch_phi with the host
// mirror), else with a copy from d_scal first.
int read_scal(lfts_batch* b, bool on_host = false);
int sync_batch(lfts_batch* b);
std::vector<int> all_ids(const lfts_batch* b);
bool lu_solve(int n, const double* A_in, const double* b, double* x);
// lfts_noalloc.hip: the scratch array's address on device dev, taken at context creation
int noalloc_setup(int dev);
bool factorize(int n, std::vector<int>& r);

// One propagator-type chain for run_lockstep: passes i0 .. i1-1 produce pairs
// i0+1 .. i1 from pair i0 (`start`; for the adjoint, multiplied by the factors of the
// next pass). A derived chain (tangent, adjoint) reads the forward pair of each pass
// through `fwd`, so a forward chain run in the same call must come first.
struct ChainSpec {
    int mode = lfts::CHAIN_PLAIN;
    const double* start = nullptr;
    double* T0 = nullptr;    // slab pipeline pairs; T0 is the line path's scratch
    double* T1 = nullptr;
    double* post = nullptr;  // adjoint, line path: the multiplied pair carried to the next step
    const double* aux = nullptr;
    std::function<double*(int)> out;        // destination of pair j
    std::function<const double*(int)> fwd;  // forward pair j
};
int run_lockstep(lfts_ctx* c, int i0, int i1, ChainSpec* const* chains, int n);

// out[0] = scale * sum_r x[r] y[r] (y null: 1), deterministic two-stage sum
int dev_sum(lfts_ctx* c, const double* x, const double* y, long n, double scale, double* out);

// checkpointed storage (lfts_ckpt.hip)
int ckpt_calc_concs(lfts_ctx* c);
int ckpt_derived(lfts_ctx* c, int mode, const double* d1, const double* dstart1);
int ckpt_tan_finish(lfts_ctx* c, const double* e, const double* dQ, double* dphi);
int ckpt_adj_finish(lfts_ctx* c, const double* ubar, double* wbar);
const double* ckpt_pair_N(lfts_ctx* c, bool derived);
int ckpt_for_bead_pairs(lfts_ctx* c, const std::function<int(int, const double*, const double*)>& fn);
int set_pair_storage(lfts_ctx* c, int stride);
void release_pair_storage(lfts_ctx* c);
int dif_pos(int k, const std::vector<int>& rad, size_t s, int B);

}  // namespace lfts_abi
