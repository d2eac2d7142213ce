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
