// lfts_step_fwd_noalloc: the contour step of src/step.cc:35-65 (step::fwd) on device
// arrays, for callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops). lfts_step_fwd runs the same three axis passes (z, y, then x with
// the Boltzmann factor) through an intermediate pair that it allocates on first use; here
// that pair lives in `scratch`, a file-scope device array of fixed size, so the call makes
// no hipMalloc, hipFree or other allocating call, on the device or on the host.
//
// `scratch` holds one pair (2 M doubles) for meshes of up to NOALLOC_MAX_POINTS points
// (128^3): 32 MiB, which every GPU the library runs on holds until the process ends. Each
// device has its own copy at its own address; the host takes that address once per device
// with hipGetSymbolAddress (when the first context on the device is created, or at the
// first call there) and passes it to the kernels, never the array's name. Every call on a
// GPU shares that GPU's copy: calls that overlap on one GPU (on different streams or host
// threads) race on it. Nothing in it needs to start at zero: the z pass writes the whole
// pair before the y pass reads it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "lfts_gpu.h"
#include "lfts_internal.h"
#include "lfts_state.h"

namespace {

constexpr long NOALLOC_MAX_POINTS = 2097152;  // 128^3
static __device__ __align__(256) double scratch[2 * NOALLOC_MAX_POINTS];

constexpr int MAX_DEVICES = 64;
double* g_scratch[MAX_DEVICES];        // per device: the address of its copy of `scratch`
hipError_t g_scratch_rc[MAX_DEVICES];
std::once_flag g_scratch_once[MAX_DEVICES];

// The address of device dev's copy of `scratch`, taken once per device with that device
// current (the caller's current device is restored).
double* scratch_on(int dev, hipError_t* rc) {
    std::call_once(g_scratch_once[dev], [dev]() {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e == hipSuccess && cur != dev) e = hipSetDevice(dev);
        void* p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur >= 0 && cur != dev) (void)hipSetDevice(cur);
        g_scratch_rc[dev] = e;
        g_scratch[dev] = (e == hipSuccess) ? static_cast<double*>(p) : nullptr;
    });
    *rc = g_scratch_rc[dev];
    return g_scratch[dev];
}

int fail_lit(lfts_ctx* c, int code, const char* msg) {
    c->err_lit = msg;  // static text: reporting the failure allocates nothing either
    return code;
}

}  // namespace

namespace lfts_abi {

// Set-up at context creation: the scratch address of the context's device.
int noalloc_setup(int dev) {
    if (dev < 0 || dev >= MAX_DEVICES) return LFTS_EHIP;
    hipError_t rc;
    return scratch_on(dev, &rc) ? LFTS_OK : LFTS_EHIP;
}

}  // namespace lfts_abi

using namespace lfts_abi;

extern "C" {

int lfts_step_fwd_noalloc(lfts_ctx* c, const double* q_in_dev, double* q_out_dev, const double* h_dev, int i) {
    if (!c) return LFTS_EINVAL;
    if (!q_in_dev || !q_out_dev || !h_dev)
        return fail_lit(c, LFTS_EINVAL, "step_fwd_noalloc: null device pointer");
    if (i < 1 || i >= c->N) return fail_lit(c, LFTS_EINVAL, "step_fwd_noalloc: bead index i must be in 1..N-1");
    if (c->M > NOALLOC_MAX_POINTS)
        return fail_lit(c, LFTS_EINVAL,
                        "step_fwd_noalloc: mesh larger than the 2097152 points (128^3) its fixed scratch holds; "
                        "use lfts_step_fwd");
    if (q_out_dev == q_in_dev || q_out_dev == h_dev)
        return fail_lit(c, LFTS_EINVAL, "step_fwd_noalloc: q_out must not be q_in or h");
    if (c->dev < 0 || c->dev >= MAX_DEVICES) return fail_lit(c, LFTS_EHIP, "step_fwd_noalloc: device index out of range");
    hipError_t e;
    double* tmp = scratch_on(c->dev, &e);
    if (!tmp) return fail_lit(c, LFTS_EHIP, "step_fwd_noalloc: the scratch array's address is not available");
    const long M = c->M;
    hipStream_t st = c->self->st;
    // the passes of line_contour_step, with the intermediate pair in scratch
    lfts::LineArgs z = line_args(c, 2, lfts::LINE_CONV);
    z.inA = q_in_dev; z.inB = q_in_dev + M; z.outA = tmp; z.outB = tmp + M;
    if (lfts::launch_line(z, st) != hipSuccess) return fail_lit(c, LFTS_EHIP, "step_fwd_noalloc: launch failed (z pass)");
    lfts::LineArgs y = line_args(c, 1, lfts::LINE_CONV);
    y.inA = tmp; y.inB = tmp + M; y.outA = tmp; y.outB = tmp + M;
    if (lfts::launch_line(y, st) != hipSuccess) return fail_lit(c, LFTS_EHIP, "step_fwd_noalloc: launch failed (y pass)");
    lfts::LineArgs x = line_args(c, 0, lfts::LINE_CONV);
    x.inA = tmp; x.inB = tmp + M; x.outA = q_out_dev; x.outB = q_out_dev + M;
    x.hA = (i < c->NA) ? h_dev : h_dev + M;  // q_{i+1}   (src/step.cc:53-57)
    x.hB = (i < c->NB) ? h_dev + M : h_dev;  // q+_{N-i} (src/step.cc:60-64)
    if (lfts::launch_line(x, st) != hipSuccess) return fail_lit(c, LFTS_EHIP, "step_fwd_noalloc: launch failed (x pass)");
    c->err_lit = nullptr;
    return LFTS_OK;
}

}  // extern "C"
