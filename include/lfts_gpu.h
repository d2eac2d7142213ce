/*
 * lfts_gpu.h — C ABI of the MI355X (gfx950) L-FTS Langevin-step library.
 *
 * The reference (tmbeardsley/lfts_cpu) has no FFI: its boundary is the in-process C++
 * class API that src/fts_cpu.cc drives. Each entry point below replaces one method of
 * that API (cited). Conventions:
 *   - one opaque context per system (replica); calls on one context are stream-ordered
 *     on the context's own HIP stream and are not thread-safe per context;
 *   - every call returns 0 on success and a non-zero LFTS_E* code on failure; the text
 *     of the last failure is available from lfts_last_error(ctx);
 *   - host buffers are caller-owned and never retained after a call returns;
 *   - "dev" entry points take device pointers (caller-owned HBM), for callers that keep
 *     their inputs resident on the GPU;
 *   - fields are FP64, in the reference's order [W- | w+ | phi- | phi+], each of length
 *     M = m[0]*m[1]*m[2], row-major p = (i*m[1] + j)*m[2] + k (src/field_generator.h:18).
 */
#ifndef LFTS_GPU_H
#define LFTS_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    LFTS_OK = 0,
    LFTS_EINVAL = 1,   /* bad argument or unsupported mesh size */
    LFTS_EHIP = 2,     /* HIP runtime error (device missing, out of memory, launch failure) */
    LFTS_ESTATE = 3,   /* call out of order (e.g. S(k) read before any sample) */
    LFTS_ENUM = 4,     /* non-finite value detected (NaN/Inf guard) */
    LFTS_ESING = 5     /* singular Anderson system: the history matrix has a zero pivot (the
                          reference's GSL LU_solve reports it and GSL's handler aborts) */
};

typedef struct lfts_desc {
    int N;        /* chain length (src/lfts_params.h:25) */
    int NA;       /* A-block length; NB = N - NA */
    int m[3];     /* mesh points mx, my, mz */
    double L[3];  /* box edges in units of R0 */
    int nhMax;    /* Anderson history depth; the reference uses 10 (src/lfts_simulation.cc:34) */
} lfts_desc;

typedef struct lfts_ctx lfts_ctx;

/* Replaces the constructors diblock(NA,NB,m,L,M,Mk) (src/diblock.cc:9-25),
 * step(...) with its g(k) table (src/step.cc:10-28,69-89), anderson(M,10)
 * (src/anderson.cc:8-21) and the device state of langevin/strFunc. Returns LFTS_EINVAL
 * for N < 2, NA outside 0..N, nhMax outside 1..15, a box edge <= 0 or a mesh size of more
 * than 2047 points with a prime factor above 7 (checked before any device call), LFTS_EHIP
 * without a device. */
int lfts_create(const lfts_desc* desc, lfts_ctx** out);
int lfts_destroy(lfts_ctx* ctx);
const char* lfts_last_error(const lfts_ctx* ctx);

/* The context's HIP stream (hipStream_t), for callers that time or order work on it. */
int lfts_get_stream(lfts_ctx* ctx, void** stream);

/* Host <-> device copies of the field array w_ (src/lfts_simulation.h:21):
 * nfields = 2 copies [W-|w+], nfields = 4 copies [W-|w+|phi-|phi+]. */
int lfts_set_fields(lfts_ctx* ctx, const double* w, int nfields);
int lfts_get_fields(lfts_ctx* ctx, double* w, int nfields);
/* Device pointer to the context's resident field array (4*M doubles). */
int lfts_fields_dev(lfts_ctx* ctx, double** w_dev);

/* diblock::calc_concs(double* w) (src/diblock.cc:30-60): one propagator evaluation
 * from [W-|w+]; writes [phi-|phi+] and returns ln Q. */
int lfts_calc_concs(lfts_ctx* ctx, double* lnQ);

/* anderson::mix(diblock*, maxIter, errTol, w) (src/anderson.cc:24-93): saddle-point
 * solve for w+. *k_ret receives mix's return value (executed iterations + 1); *err_ret
 * (may be NULL) the last residual. The small LU solve runs on the host. */
int lfts_anderson_mix(lfts_ctx* ctx, int maxIter, double errTol, int* k_ret, double* err_ret);

/* anderson::LUdecomp(A, Y, X, n) (src/anderson.cc:106-127): the host-side solve of
 * A x = y for the n x n Anderson system (n <= 15), LU with partial pivoting as GSL's
 * LU_decomp/LU_solve. A is row-major and is not modified. Returns LFTS_ESING, leaving x
 * untouched, when U has a zero pivot. No context needed. */
int lfts_lu_solve(int n, const double* A, const double* y, double* x);

/* langevin(RNG, sigma, M) constructor's first draw (src/langevin.cc:7-19): sets the
 * "previous" noise eta_prev (M host doubles). */
int lfts_langevin_init(lfts_ctx* ctx, const double* eta_prev);
/* langevin::step_wm(w, RNG, XbN, sigma, dt) (src/langevin.cc:23-37) with the new noise
 * eta_new (M doubles) drawn by the caller: host pointer / device pointer variants. */
int lfts_langevin_step(lfts_ctx* ctx, const double* eta_new, double XbN, double dt);
int lfts_langevin_step_dev(lfts_ctx* ctx, const double* eta_new_dev, double XbN, double dt);
/* The noise langevin keeps between steps (src/langevin.cc:34-36), M host doubles, for a
 * checkpoint; lfts_langevin_init restores it. */
int lfts_langevin_get_prev(lfts_ctx* ctx, double* eta_prev);

/* One whole Langevin step of lfts_simulation::equilibrate()/statistics()
 * (src/lfts_simulation.cc:67-87): step_wm with device noise, mix(maxIter, errTol), then
 * the extra calc_concs that prints lnQ. Returns lnQ and the mix return value. */
int lfts_langevin_full_step_dev(lfts_ctx* ctx, const double* eta_new_dev, double XbN, double dt,
                                int maxIter, double errTol, double* lnQ, int* k_ret);

/* strFunc::sample(w) (src/strFunc.cc:36-46): S_k += |FFT(W-)_k|^2 on the half spectrum
 * k = k2 + (m2/2+1)(K1 + m1*K0). lfts_sk_get copies S (Mk doubles) and the count. */
int lfts_sk_sample(lfts_ctx* ctx);
int lfts_sk_get(lfts_ctx* ctx, double* S, int* nsamples);
int lfts_sk_reset(lfts_ctx* ctx);
/* Restores S (Mk doubles) and the sample count, for a restart from a checkpoint (no
 * reference counterpart: the reference restarts from its 6-digit field files only). */
int lfts_sk_set(lfts_ctx* ctx, const double* S, int nsamples);

/* strFunc::calcK + the stable sort of the constructor (src/strFunc.cc:27-31,80-103),
 * host-side: K (Mk doubles), wt (Mk ints), P (Mk ints). No context needed. */
int lfts_sk_bins(const int* m, const double* L, double* K, int* wt, int* P);
/* strFunc::save(fileName, dp) (src/strFunc.cc:50-76) from a host S array. */
int lfts_sk_save(const char* fileName, const int* m, const double* L, const double* S, int nsamples,
                 double CV, double chi_b, double dK, int dp);

/* lfts_simulation::getH() (src/lfts_simulation.cc:122-135): calc_concs, then
 * H = -lnQ + (sum W-^2 / XbN - sum w+) / M. */
int lfts_hamiltonian(lfts_ctx* ctx, double XbN, double* H);

/* Storage of the propagator pairs (no reference counterpart: the reference stores all
 * N pairs, src/diblock.cc:9-25). stride 0 stores every pair; stride S >= 2 keeps about
 * N/S of them, at the starts of segments chosen symmetric under bead j -> N+1-j, and
 * recomputes each segment (with its mirror) when the bead sums need it: calc_concs then
 * costs two chains instead of one, tangent and adjoint products three more. lfts_create
 * picks stride ceil(sqrt(N)) by itself when the pairs would not fit comfortably in free
 * HBM. With a stride, lfts_get_pair (i > 1) and lfts_box_derivative are unavailable and
 * the context cannot join a replica set. lfts_pair_storage reports the stride and the
 * number of pairs kept. */
int lfts_set_pair_storage(lfts_ctx* ctx, int stride);
int lfts_pair_storage(lfts_ctx* ctx, int* stride, int* nkept);

/* Testing access: stored propagator pair i (1..N) of the last evaluation as the
 * reference lays it out, [q_i | q+_{N+1-i}] (src/diblock.cc:19-24), 2*M host doubles.
 * lfts_calc_concs keeps every pair. On one context with a cubic mesh up to 64^3, an
 * Anderson iteration or lfts_hamiltonian forms the density sums during the chain and
 * keeps only pairs 1..ceil(N/2) and N; another i then returns LFTS_EINVAL. */
int lfts_get_pair(lfts_ctx* ctx, int i, double* out);
/* step::fwd(q_in, q_out, h, i) (src/step.cc:35-65) on host arrays (2*M each). */
int lfts_step_fwd(lfts_ctx* ctx, const double* q_in, double* q_out, const double* h, int i);

/* lfts_step_fwd on device arrays (q_in_dev, h_dev: 2*M doubles each; q_out_dev: 2*M,
 * distinct from both) with no allocation: no reference counterpart; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops). The
 * intermediate pair of the three axis passes lives in a fixed device array of the library
 * (32 MiB on every GPU it runs on, held until the process ends), so meshes of up to
 * 2097152 points (128^3) are served; larger ones return LFTS_EINVAL. Stream-ordered on the
 * context's stream; no synchronisation. All calls on one GPU share that array: calls that
 * overlap on one GPU race. Same results as lfts_step_fwd, bit for bit. */
int lfts_step_fwd_noalloc(lfts_ctx* ctx, const double* q_in_dev, double* q_out_dev, const double* h_dev, int i);

/* Counters: propagator evaluations and contour steps run on this context so far. */
int lfts_counters(lfts_ctx* ctx, long long* n_prop_evals, long long* n_contour_steps);

/* Live timing with HIP events recorded on the context stream (no reference counterpart;
 * the reference only prints whole-second wall times, src/fts_cpu.cc:44-57). With enable =
 * k >= 1, every k-th propagator evaluation (the first, the (k+1)-th, ...) records the
 * duration of its contour-step chain (the N-1 contour-step launches, back to back) and of
 * the whole evaluation; 0 turns the timing off. A recorded evaluation costs about 10 us of
 * host and queue time outside its chain, so k > 1 keeps that cost off the others.
 * lfts_profile_read returns the sums (ms) and counts (contour-step launches, evaluations)
 * of the recorded evaluations since the last enable. For a replica set, one contour-step
 * launch advances every member of one mesh size, and evaluations are counted per member. */
int lfts_profile(lfts_ctx* ctx, int enable);
int lfts_profile_read(lfts_ctx* ctx, double* ms_chain, long long* n_contour_steps, double* ms_pe,
                      long long* n_prop_evals);

/* Replica sets (no reference counterpart: the reference runs one system per process,
 * src/fts_cpu.cc:40). A set advances many independent contexts together: one launch per
 * contour step for every mesh size present, one host synchronisation per Anderson
 * iteration for the whole set. Every entry point has the meaning of the single-context
 * call above applied to each member, and gives each member the same result as that call
 * (the kernels and reduction orders are the same). Arrays are indexed by member; members
 * keep their own Anderson iteration counts, and a member whose loop condition ends leaves
 * the active set while the others continue. A context belongs to at most one set in use
 * at a time; the set must be destroyed before its members. */
typedef struct lfts_batch lfts_batch;
int lfts_batch_create(lfts_ctx* const* ctxs, int n, lfts_batch** out);
int lfts_batch_destroy(lfts_batch* b);
const char* lfts_batch_last_error(const lfts_batch* b);
int lfts_batch_get_stream(lfts_batch* b, void** stream);
int lfts_batch_calc_concs(lfts_batch* b, double* lnQ);
int lfts_batch_anderson_mix(lfts_batch* b, int maxIter, double errTol, int* k_ret, double* err_ret);
int lfts_batch_langevin_step_dev(lfts_batch* b, const double* const* eta_new_dev, const double* XbN,
                                 const double* dt);
int lfts_batch_full_step_dev(lfts_batch* b, const double* const* eta_new_dev, const double* XbN, const double* dt,
                             int maxIter, double errTol, double* lnQ, int* k_ret);
/* Autotuner report (no reference counterpart, in place of FFTW_PATIENT planning,
 * src/step.cc:15-16): the contour-step kernel shape chosen for each mesh size, launch
 * size and pass kind (with or without fused density sums), with the measured time per
 * pass of every candidate. A shape is a line factorisation R1 x R2 (R2 threads holding R1
 * elements of a line) and a variant: twiddle/kernel tables in registers or LDS, one or two
 * fields per workgroup. Variants of one factorisation give bit-identical results. Set
 * LFTS_SLAB_VARIANT to 0..3 to fix the variant, LFTS_SLAB_SHAPES the factorisations. The
 * text stays valid until the next call on the set. */
const char* lfts_batch_tuning(const lfts_batch* b);
const char* lfts_tuning(const lfts_ctx* ctx);
/* The line factorisations R1 x R2 of the contour-step kernel chosen so far in this process,
 * one per mesh edge, as text ("64:8x8,48:12x4"); with buf = NULL *len receives the size
 * needed (no reference counterpart: the reference's FFTW plans by timing, src/step.cc:15-16,
 * and keeps no wisdom). A factorisation is chosen by timing at an edge's first use and then
 * shared by every context and set of the process; different factorisations round
 * differently (about 1 ulp per transform), so a run that must repeat another's bits sets
 * the environment variable LFTS_SLAB_SHAPES to this text (or "default") before its first
 * evaluation. */
int lfts_slab_shapes(char* buf, long long cap, long long* len);
/* lfts_batch_full_step_dev with host noise: member id's M doubles eta_new[id] are copied
 * to its device staging buffer first (as lfts_langevin_step does for one context). */
int lfts_batch_full_step(lfts_batch* b, const double* const* eta_new, const double* XbN, const double* dt,
                         int maxIter, double errTol, double* lnQ, int* k_ret);
int lfts_batch_profile(lfts_batch* b, int enable);
int lfts_batch_profile_read(lfts_batch* b, double* ms_chain, long long* n_passes, double* ms_pe,
                            long long* n_prop_evals);

/* Derivatives of one propagator evaluation (no reference counterpart: the reference
 * computes no derivatives). Each call first runs calc_concs at the context's current
 * [W-|w+] (so [phi-|phi+] in the context are those of the current fields), then:
 *   lfts_jvp: dphi = d(phi-, phi+) along the direction (dWm, dwp) (M doubles each, either
 *             may be NULL for zero) -> 2*M doubles [dphi- | dphi+];
 *   lfts_vjp: wbar = gradient of <u-, phi-> + <u+, phi+> with respect to (W-, w+), for
 *             ubar = [u- | u+] (2*M) -> 2*M doubles [d/dW- | d/dw+].
 * Both follow the chain of src/diblock.cc:30-60 and src/step.cc:35-65 exactly (including
 * the Boltzmann factor of the chain ends for NA = 0 or NA = N). _dev variants take device
 * pointers. */
int lfts_jvp(lfts_ctx* ctx, const double* dWm, const double* dwp, double* dphi);
int lfts_vjp(lfts_ctx* ctx, const double* ubar, double* wbar);
int lfts_jvp_dev(lfts_ctx* ctx, const double* dWm_dev, const double* dwp_dev, double* dphi_dev);
int lfts_vjp_dev(lfts_ctx* ctx, const double* ubar_dev, double* wbar_dev);

/* Box derivative of H = -lnQ + (sum W-^2/XbN - sum w+)/M (src/lfts_simulation.cc:122-135)
 * with respect to the box edges L[0..2] at fixed fields (H depends on the box through
 * g(k) of src/step.cc:84-85 only): dHdL (3 doubles) = -d lnQ/dL. Also returns lnQ.
 * Requires 1 <= NA <= N-1. */
int lfts_box_derivative(lfts_ctx* ctx, double* dHdL, double* lnQ);

/* Newton-Krylov saddle-point solve for w+ at fixed W-, beside anderson::mix
 * (src/anderson.cc:24-93; no reference counterpart): inexact Newton on
 * F(w+) = phi+(w+) - 1 with restarted GMRES over lfts_jvp products, preconditioned by the
 * exact inverse response of the homogeneous melt, and backtracking. It stops when
 * err = sqrt(sum F^2 / M) (the residual of src/anderson.cc:35-40) is <= tol. On return
 * [phi-|phi+] in the context belong to the final w+. The mean of w+ is kept. */
typedef struct lfts_nk_opts {
    int maxNewton;     /* Newton iterations at most */
    double tol;        /* residual tolerance */
    int restart;       /* GMRES restart length, 1..30 (20 is typical) */
    int maxInner;      /* tangent evaluations per Newton step at most */
    double eta;        /* largest relative tolerance of a Newton step's linear solve */
    int precondition;  /* 1: homogeneous-melt preconditioner, 0: none */
} lfts_nk_opts;
typedef struct lfts_nk_info {
    int newton;        /* Newton iterations run */
    int krylov;        /* tangent evaluations (lfts_jvp products) */
    int prop_evals;    /* propagator evaluations (calc_concs) */
    double err;        /* final residual */
} lfts_nk_info;
int lfts_saddle_nk(lfts_ctx* ctx, const lfts_nk_opts* opts, lfts_nk_info* info);
/* One Langevin step with lfts_saddle_nk in place of anderson::mix: step_wm with device
 * noise, the solve, and lnQ of the final fields (src/lfts_simulation.cc:72-76 with the
 * other solver; the trajectory differs from the reference's within the tolerance). */
int lfts_langevin_full_step_nk_dev(lfts_ctx* ctx, const double* eta_new_dev, double XbN, double dt,
                                   const lfts_nk_opts* opts, double* lnQ, lfts_nk_info* info);

/* Host random stream of the reference: std::mt19937_64 seeded by the caller (the
 * reference seeds 123456789, src/lfts_simulation.cc:23) feeding one persistent
 * std::normal_distribution<double>(0, sigma) (src/langevin.cc:13,16,28); uniforms use a
 * fresh std::uniform_real_distribution<double>(0,1) per call (src/lfts_simulation.cc:151-153). */
typedef struct lfts_rng lfts_rng;
int lfts_rng_create(unsigned long long seed, double sigma, lfts_rng** out);
int lfts_rng_normals(lfts_rng* rng, double* out, long long n);
int lfts_rng_uniforms(lfts_rng* rng, double* out, long long n);
int lfts_rng_destroy(lfts_rng* rng);
/* The full state of the stream (engine, and the normal distribution with the value the
 * polar method keeps for its next call) as text: lfts_rng_save with buf = NULL sets *len
 * to the size needed; lfts_rng_load restores it, so a restarted run continues the very
 * same stream (the reference re-seeds on every start, src/lfts_simulation.cc:23). */
int lfts_rng_save(lfts_rng* rng, char* buf, long long cap, long long* len);
int lfts_rng_load(lfts_rng* rng, const char* buf, long long len);

/* lfts_params::z_inf_discrete (src/lfts_params.cc:122-142). */
double lfts_z_inf(const double* L, const int* m, int N, double Nbar, int tmax);

#ifdef __cplusplus
}
#endif
#endif /* LFTS_GPU_H */
