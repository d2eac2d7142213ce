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

enum { LINE_CONV = 0, LINE_FFT = 1 };


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
};

// Per-replica coefficients of one Anderson update or Langevin step.
struct Coef {
    double lambda;
    int nh, slot;
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
};
hipError_t launch_line(const LineArgs& a, hipStream_t st);

// ---- fused slab pass over replicas with a cubic mesh of edge n -----------------------
// Step i = 0 is the prologue (x-slab pass on q_1); step i >= 1 finishes contour step i
// (y-slab pass for odd i, x-slab pass for even i). `reps` holds nact descriptors of
// replicas with edge n; when nact == 1 and `one` (a host copy of that descriptor) is
// given, the descriptor travels in the kernel arguments instead.
bool slab_supported(int n0, int n1, int n2);
hipError_t launch_slab(const Rep* reps, int nact, int n, int i, const double2* tw, hipStream_t st,
                       const Rep* one = nullptr);

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

}  // namespace lfts
