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
constexpr int MAX_PRIME_LINE = 2047;  // longest axis with a prime factor above 7 (two LDS tiles <= 64 KB)
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
    double* phacc;      // 4M bead sums of the fused evaluation (slab path), or null
    // slab path, N >= 2: the chain's last pass stores the sum of q_N over each of its n
    // planes here (the partials Q is formed from); null: k_q_sum forms nb partials instead
    double* qpl;
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

// ---- generic line kernels (any mesh; see factorize), one replica ----------------------
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
    // LINE_CONV on a line length the axis kernel serves: fast = 1 runs the register-FFT
    // axis kernel (natural-order kernel values gnat, g(k)/n) instead of the generic one
    const double* gnat;
    int fast;
};
hipError_t launch_line(const LineArgs& a, hipStream_t st);
// the register-FFT convolution along one axis (lfts_fused.hip): every even line length
// from 16 to 256 whose prime factors are among 2, 3, 5 and 7; hipErrorNotSupported for
// any other length
hipError_t launch_axis_conv(const LineArgs& a, hipStream_t st);

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
    // plain chain, fused density sums: the field's new values times the partner values
    // (the other field of an earlier pair) are added to acc (first: acc = product)
    const double* part[2];
    double* acc[2];
    int first[2];
    double* qsum;          // last pass of a plain chain: field 0's plane sums (q_N, for Q), or null
    long M;
    int finish, cont, xslab;
};

__host__ __device__ inline const double* rep_pair(const Rep& R, int j) {
    return (j == 1) ? R.h : R.q + 2L * (j - 2) * R.M;
}

// Fused density sums (slab path, N >= 3). Pass i >= i0 = ceil(N/2) produces q_{i+1} and
// q+_{N-i}; their partners q+_{i+1} and q_{N-i} are the two fields of pair N - i, made by
// an earlier pass. So these passes add the bead products to four sums (field 0 beads
// i0+1..N ascending, field 1 beads N-i0..1 descending, each split by block A/B) instead
// of storing their pairs; only pairs 1..i0 and N are stored. For odd N the middle bead
// i0 pairs two fields of one pair and is added at the end.
// (Meshes up to 64^3: at 96^3 and 128^3 the passes fill the chip and are bandwidth-bound,
// so the sums gain nothing there, and their registers would spill.)
__host__ __device__ inline bool fused_phi(const Rep& R, int fuse) {
    return fuse && R.phacc && R.n && R.n <= 64 && R.N >= 3;
}
// which of the four sums (field 0 A, field 0 B, field 1 A, field 1 B) receive a bead
__host__ __device__ inline void phi_sums_used(int N, int NA, bool u[4]) {
    const int i0 = (N + 1) / 2;
    u[0] = i0 + 1 <= NA;
    u[1] = NA < N;
    u[2] = NA >= 1 && N - i0 >= 1;
    u[3] = N - i0 > NA;
}

// Pass i of a replica's plain chain: pair 1 = h, pair i+1 = h_sel C pair_i
// (src/diblock.cc:40-43, src/step.cc:53-64); T0/T1 alternate as partial pairs.
__host__ __device__ inline SlabIO plain_io(const Rep& R, int i, int fuse = 0) {
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
    io.qsum = (i >= 1 && i == R.N - 1) ? R.qpl : nullptr;
    io.xslab = (i % 2 == 0);
    const int i0 = (R.N + 1) / 2;
    if (fused_phi(R, fuse) && i >= i0) {
        const double* P = rep_pair(R, R.N - i);
        const int j0 = i + 1, j1 = R.N - i;  // beads of the two new fields
        const int b0 = (j0 <= R.NA) ? 0 : 1, b1 = (j1 <= R.NA) ? 0 : 1;
        io.part[0] = P + M;  // q+_{j0}
        io.part[1] = P;      // q_{j1}
        io.acc[0] = R.phacc + (long)b0 * M;
        io.acc[1] = R.phacc + (long)(2 + b1) * M;
        const int lo0 = (i0 + 1 > R.NA + 1) ? i0 + 1 : R.NA + 1;
        const int hi1 = (R.N - i0 < R.NA) ? R.N - i0 : R.NA;
        io.first[0] = (b0 == 0) ? (j0 == i0 + 1) : (j0 == lo0);
        io.first[1] = (b1 == 1) ? (j1 == R.N - i0) : (j1 == hi1);
        if (i + 1 != R.N) io.qout = nullptr;  // pair N is kept for Q
    }
    return io;
}
// Step i = 0 is the prologue (x-slab pass on q_1); step i >= 1 finishes contour step i
// (y-slab pass for odd i, x-slab pass for even i). `reps` holds nact descriptors of
// replicas with edge n; when nact == 1 and `one` (a host copy of that descriptor) is
// given, the descriptor travels in the kernel arguments instead.
bool slab_supported(int n0, int n1, int n2);
// the cubic edges the slab pass serves, largest first (one launch group each in a set)
constexpr int SLAB_EDGES[] = {128, 120, 112, 96, 80, 72, 64, 56, 48, 40, 32, 24, 16};
constexpr int N_SLAB_EDGES = 13;
// `variant` selects one of SLAB_VARIANTS kernel shapes (bit 0: twiddle/kernel tables in
// LDS instead of registers; bit 1: both fields of a plane in one workgroup). All variants
// give bit-identical results; the fastest one per mesh size and set size is picked by
// timing (the autotuner of the replica sets). Edges from 72 up have variant 1 only (the
// other shapes spill registers or exceed the LDS).
constexpr int SLAB_VARIANTS = 4;
// Line factorisations R1 x R2 of the slab pass: shape 0 is each edge's default (every chain
// kind, lfts_fused.hip); shapes 1 and up are the alternatives the autotuner also times for
// plain chains (lfts_slab_alt.hip). A variant index carries the shape: variant = 4 shape + v.
// Different factorisations round differently (about 1 ulp per transform), so a process
// uses one factorisation per edge (see slab_variant).
constexpr int MAX_SLAB_SHAPES = 3;
int slab_shape_count(int edge);
bool slab_shape_factors(int edge, int shape, int* r1, int* r2);
int slab_shape_variants(int edge, int shape);  // bit v set: variant v exists for this shape
hipError_t dispatch_edge_alt(int n, int shape, int v, const Rep* reps, int nact, int i, const SlabIO* io,
                             const double2* tw, hipStream_t st);
// fuse = 1: density sums fused into the second half of the chain (fused_phi); only for
// one replica (nact == 1 with its host descriptor `one`).
hipError_t launch_slab(const Rep* reps, int nact, int n, int i, const double2* tw, hipStream_t st,
                       const Rep* one = nullptr, int variant = 0, int fuse = 0);
// One pass of one replica's chain of kind `mode` (CHAIN_*), over the n planes.
hipError_t launch_slab_io(int n, int mode, const SlabIO& io, const double2* tw, hipStream_t st, int variant = 0);
// Pass i of the plain chains of two groups of a set in one launch (edges ea and eb, each in
// its default factorisation), bit-identical to a launch_slab of each: 64^3 with 32^3.
bool slab_pair_supported(int ea, int eb);
hipError_t launch_slab_pair(const Rep* repsA, int nA, int ea, const Rep* repsB, int nB, int eb, int i,
                            const double2* twA, const double2* twB, hipStream_t st);

// ---- batched element-wise and reduction kernels ----------------------------------------
// reps: descriptors of the nact active replicas; grid (nb_max, nact).
// part: [slot][MAX_NB][PSTRIDE]; scal: [replica id][SSTRIDE]
hipError_t launch_prepare_h(const Rep* reps, int nact, int nb_max, hipStream_t st);
// partial sums of q_N into qpart[id][workgroup] (MAX_NB doubles per member) for the members
// whose chain does not form them (qpl null)
hipError_t launch_q_sum(const Rep* reps, int nact, int nb_max, double* qpart, hipStream_t st);
// hscal (pinned host memory mapped for the device, may be null) receives the same values;
// hseq (pinned, may be null): hseq[slot] = seq once a slot's results are stored
hipError_t launch_reduce(const Rep* reps, int nact, const double* part, int ncomp, int div_by_M, double* scal,
                         int scal_off, hipStream_t st, double* hscal = nullptr, unsigned* hseq = nullptr,
                         unsigned seq = 0);
// phi from the fused sums (replicas with fused_phi(R, fuse)) or from the stored pairs,
// both in the same order of operations, with Q formed from the member's partials (qpl, or
// launch_q_sum's in qpart) and stored in scal[id][0] and hscal[id][0] (hscal as for
// launch_reduce). coef (may be null): also write each member's next Anderson Gram row
// partials into part, as launch_am_dots would; nh_one >= 0 (one member only) gives the
// depth instead of coef.
hipError_t launch_phi(const Rep* reps, int nact, int nb_max, double* scal, const double* qpart, int fuse,
                      const Coef* coef, int nh_one, double* part, hipStream_t st, double* hscal = nullptr);
hipError_t launch_am_dots(const Rep* reps, int nact, int nb_max, const Coef* coef, double* part, hipStream_t st);
// with_h: also write h = exp(-(+-W- + w+)/N) of the updated fields, as launch_prepare_h would
hipError_t launch_am_update(const Rep* reps, int nact, int nb_max, const Coef* coef, int with_h, hipStream_t st);
// the same for one member with its coefficients as a kernel argument
hipError_t launch_am_update_one(const Rep* reps, int nb_max, const Coef& c, int with_h, hipStream_t st);
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
    int singular;      // the LU met a zero pivot (the loop stops; the call reports it)
    double err, tol;   // residual of the last iteration, tolerance
    double DD[(MAX_HIST + 1) * (MAX_HIST + 1)];
};
// after an iteration's Gram row landed in scal[0..nh]: residual, coefficients (LU on one
// thread, the host LU's operation order), history bookkeeping, next loop test
hipError_t launch_am_solve(AmState* s, const double* scal, Coef* coef, const double* pow09, long M, int nhMax,
                           hipStream_t st);
hipError_t launch_am_advance(AmState* s, hipStream_t st);

}  // namespace lfts
