// Internal declarations shared by the kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>

namespace lfts {

constexpr int MAXST = 12;     // max FFT stages per axis
constexpr int MAX_HIST = 15;  // max Anderson history depth supported on the device
constexpr int PSTRIDE = 16;   // doubles per workgroup in the reduction partials buffer

enum { LINE_CONV = 0, LINE_FFT = 1 };

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

// Fused slab pass (one contour step per launch), cubic 32^3 / 64^3 meshes.
struct SlabArgs {
    const double* Tin;   // partially convolved pair (2 fields of M), or pair 1 for the prologue
    double* Tout;        // partially convolved pair for the next pass
    double* qout;        // stored propagator pair q_{i+1} (finishing passes)
    const double* hq;    // Boltzmann factor for field 0 (q)
    const double* hqd;   // Boltzmann factor for field 1 (q+)
    const double* ga;    // kernel along the slab's outer axis (n entries, includes 1/n)
    const double* gb;    // kernel along z
    const double2* tw;   // exp(-2 pi i k/n), n entries
    long M;              // field stride
    long plane_stride;   // offset between consecutive slab planes
    long Sa;             // stride of the outer axis inside a plane
    int finish;          // complete the step in flight: Ca, *h, store q_{i+1}
    int cont;            // start the next step: Ca, Cz, store Tout
};

bool slab_supported(int n0, int n1, int n2);
hipError_t launch_slab(const SlabArgs& a, int n, int nplanes, hipStream_t st);

struct AmCoef {
    double lambda;
    int nh, slot;
    double C[MAX_HIST];
};

hipError_t launch_line(const LineArgs& a, hipStream_t st);
hipError_t launch_reduce_final(const double* part, int nblk, int ncomp, double div, double* out, hipStream_t st);
hipError_t launch_prepare_h(const double* w, double* h, long M, int N, int nblk, hipStream_t st);
hipError_t launch_sum_partial(const double* x, long M, double* part, int nblk, hipStream_t st);
hipError_t launch_phi(const double* h, const double* q, double* w, const double* Qp, long M, int N, int NA,
                      int nblk, hipStream_t st);
hipError_t launch_am_dots(const double* w, const double* Dh, long M, int nh, double* part, int nblk, hipStream_t st);
hipError_t launch_am_update(double* w, double* Dh, double* wh, long M, const AmCoef& c, int nblk, hipStream_t st);
hipError_t launch_langevin(double* w, double* ep, const double* en, long M, double XbN, double dt, int keep_new,
                           int nblk, hipStream_t st);
hipError_t launch_h_partial(const double* w, long M, double* part, int nblk, hipStream_t st);

}  // namespace lfts
