// Internal declarations shared by the kernels and the C-ABI layer.
//
// Every propagator evaluation, Anderson iteration and Langevin update runs over a set
// of replicas (independent systems). A single context is a set of one. Kernels receive
// the device array of replica descriptors and the list of active replicas; workgroups
// find their replica from blockIdx.
#pragma once
#include <hip/hip_runtime.h>

namespace lfts {

constexpr int MAXST = 12;     // max FFT stages per axis (generic line path)
constexpr int MAX_HIST = 15;  // max Anderson history depth supported on the device
constexpr int PSTRIDE = 16;   // doubles per workgroup in the reduction partials buffer
constexpr int SSTRIDE = 32;   // doubles of scalar results per replica
constexpr int MAX_NB = 1024;  // max workgroups per replica in the element-wise passes and reductions

// Workgroups (of 256 threads) per replica: a function of M only, so a replica's
// reduction order does not depend on which other replicas share the launch.
inline int blocks_for(long M) {
    const long nb = (M + 255) / 256;
    return (int)(nb < 1 ? 1 : (nb > MAX_NB ? MAX_NB : nb));
}

enum { LINE_CONV = 0, LINE_FFT = 1, LINE_IFFT = 2, LINE_CMUL = 3 };
// kinds of propagator chain: the plain chain, its tangent, its adjoint
enum { CHAIN_PLAIN = 0, CHAIN_TANGENT = 1, CHAIN_ADJOINT = 2 };


// Device-side descriptor of one replica.
struct Rep {
    double* w;          // 4M: W-, w+, phi-, phi+
    double* h;          // 2M: hA, hB (= propagator pair 1)
    double* q;          // pairs 2..N, pair i at (i-2)*2M: [q_i | q+_{N+1-i}]
    double* T0;         // partially convolved pair, slab passes (2M each)
    double* T1;
    const double* g[3]; // per-axis kernel in natural frequency order, g_d(k)/m_d
    double* Dh;         // (nhMax+1) M Anderson residual histories
    double* wh;         // (nhMax+1) M Anderson field histories
    double* eta_prev;   // M, previous Langevin noise
    long M;
    int N, NA, NB;
    int n;              // cubic edge when the fused slab path applies, else 0
    int id;             // index of the replica in its set (scalar results, coefficients)
    int nb;             // workgroups of the element-wise passes, blocks_for(M)
    const int* skip;    // device flag: when set, this replica's kernels return at once (or null)
};

// Per-replica coefficients of one Anderson update or Langevin step.
struct Coef {
    double lambda;
    int nh, slot;
    int nh_dots;          // history depth of the next Gram row (k_am_dots)
    double C[MAX_HIST];
    double XbN, dt;
    const double* eta_new;
    int keep_new;
};

// ---- generic line kernels (any mesh with factors 2, 3, 5, 7), one replica ----------
struct LineArgs {
    int n, nst, radix[MAXST];
    const double2* tw;    // n entries, exp(-2 pi i k/n)
    const double* gperm;  // n entries, line kernel value (g(k)/n) at DIF output position
    const int* perm;      // n entries, DIF output position of frequency k
    int axis, m0, m1, m2;
    long nlines, stride;
    int T, pitch, mode;
    const double* inA;
    const double* inB;
    double* outA;
    double* outB;
    const double* hA;
    const double* hB;
    const double* inR;
    const double2* inC;
    double2* outC;
    double* Sacc;
    // LINE_CONV with hA/hB (the finishing pass of a chain step), chain kind `chain`:
    //   CHAIN_TANGENT  out = h u + aux src;  CHAIN_ADJOINT  out = u + aux src, out2 = h out
    int chain;
    const double* auxA;
    const double* auxB;
    const double* srcA;
    const double* srcB;
    double* out2A;
    double* out2B;
    // LINE_IFFT: real part to outR instead of outC; LINE_CMUL: per-point multiplier in
    // the line's transform order, mul[line * n + e]
    double* outR;
    const double* mul;
    const int* skip;      // device flag: when set, the launch returns at once (or null)
};
hipError_t launch_line(const LineArgs& a, hipStream_t st);

// ---- fused slab pass over replicas with a cubic mesh of edge n -----------------------

// Addresses of one slab pass of one replica. Pass i finishes pair i+1 from the partial
// convolution `tin` (i >= 1) and leaves the partial convolution of the next step in
// `tout`; pass 0 (the prologue) starts from the chain's first pair in `tin`. Every
// pointer addresses pair layout [field 0 | field 1] except h/aux/src, which point at the
// M values used for field 0 (q) and field 1 (q+) separately.
struct SlabIO {
    const double* tin;
    double* tout;
    double* qout;          // pair i+1 (finishing passes)
    const double* h[2];    // Boltzmann factor per field at this step
    const double* aux[2];  // tangent: e per field; adjoint: alpha per field
    const double* src[2];  // the stored values aux multiplies, per field
    const double* g[3];    // per-axis kernels, natural frequency order
    const int* skip;       // device flag: when set, the pass returns at once (or null)
    long M;
    int finish, cont, xslab;
};

// Pass i of a replica's plain chain: pair 1 = h, pair i+1 = h_sel C pair_i
// (src/diblock.cc:40-43, src/step.cc:53-64); T0/T1 alternate as partial pairs.
__host__ __device__ inline SlabIO plain_io(const Rep& R, int i) {
    SlabIO io{};
    const long M = R.M;
    io.tin = (i == 0) ? R.h : ((i & 1) ? R.T0 : R.T1);
    io.tout = (i & 1) ? R.T1 : R.T0;
    io.qout = (i >= 1) ? R.q + 2L * (i - 1) * M : nullptr;
    io.h[0] = (i < R.NA) ? R.h : R.h + M;
    io.h[1] = (i < R.NB) ? R.h + M : R.h;
    for (int d = 0; d < 3; d++) io.g[d] = R.g[d];
    io.skip = R.skip;
    io.M = M;
    io.finish = i >= 1;
    io.cont = i < R.N - 1;
    io.xslab = (i % 2 == 0);
    return io;
}
// Step i = 0 is the prologue (x-slab pass on q_1); step i >= 1 finishes contour step i
// (y-slab pass for odd i, x-slab pass for even i). `reps` holds nact descriptors of
// replicas with edge n; when nact == 1 and `one` (a host copy of that descriptor) is
// given, the descriptor travels in the kernel arguments instead.
bool slab_supported(int n0, int n1, int n2);
// `variant` selects one of SLAB_VARIANTS kernel shapes (bit 0: twiddle/kernel tables in
// LDS instead of registers; bit 1: both fields of a plane in one workgroup). All variants
// give bit-identical results; the fastest one per mesh size and set size is picked by
// timing (the autotuner of the replica sets).
constexpr int SLAB_VARIANTS = 4;
hipError_t launch_slab(const Rep* reps, int nact, int n, int i, const double2* tw, hipStream_t st,
                       const Rep* one = nullptr, int variant = 0);
// One pass of one replica's chain of kind `mode` (CHAIN_*), over the n planes.
hipError_t launch_slab_io(int n, int mode, const SlabIO& io, const double2* tw, hipStream_t st, int variant = 0);

// ---- batched element-wise and reduction kernels ----------------------------------------
// reps: descriptors of the nact active replicas; grid (nb_max, nact).
// part: [slot][MAX_NB][PSTRIDE]; scal: [replica id][SSTRIDE]
hipError_t launch_prepare_h(const Rep* reps, int nact, int nb_max, hipStream_t st);
hipError_t launch_q_sum(const Rep* reps, int nact, int nb_max, double* part, hipStream_t st);
hipError_t launch_reduce(const Rep* reps, int nact, const double* part, int ncomp, int div_by_M, double* scal,
                         int scal_off, hipStream_t st);
hipError_t launch_phi(const Rep* reps, int nact, int nb_max, const double* scal, hipStream_t st);
hipError_t launch_am_dots(const Rep* reps, int nact, int nb_max, const Coef* coef, double* part, hipStream_t st);
hipError_t launch_am_update(const Rep* reps, int nact, int nb_max, const Coef* coef, hipStream_t st);
hipError_t launch_langevin(const Rep* reps, int nact, int nb_max, const Coef* coef, hipStream_t st);
hipError_t launch_h_sums(const Rep* reps, int nact, int nb_max, double* part, hipStream_t st);

// ---- Anderson mixing on the device (one replica) ----------------------------------------
// The loop state of anderson::mix (src/anderson.cc:24-93) kept in device memory, so that
// iterations can be enqueued ahead of the loop test: every kernel of an iteration returns
// at once when `stop` is set.
struct AmState {
    int k;             // iteration number (k of src/anderson.cc:29)
    int stop;          // loop condition failed before this iteration
    int stop_next;     // loop condition for the next iteration
    int maxIter;
    double err, tol;   // residual of the last iteration, tolerance
    double DD[(MAX_HIST + 1) * (MAX_HIST + 1)];
};
// after an iteration's Gram row landed in scal[0..nh]: residual, coefficients (LU on one
// thread, the host LU's operation order), history bookkeeping, next loop test
hipError_t launch_am_solve(AmState* s, const double* scal, Coef* coef, const double* pow09, long M, int nhMax,
                           hipStream_t st);
hipError_t launch_am_advance(AmState* s, hipStream_t st);

}  // namespace lfts
