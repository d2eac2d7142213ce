// Fused contour step for cubic meshes (the slab pass, lfts_slab.h): the launchers, the
// dispatch over mesh edges and kernel shapes (the default line factorisation of each edge
// here, the others in lfts_slab_alt.hip), and the register-FFT axis kernel of the line path.
#include "lfts_slab.h"

namespace lfts {

// R1 x R2 per edge and shape: the default first, then the alternatives the autotuner times
namespace {
// rr: R1 x R2; mask: the variants the shape has (the others spill: 64 as 16 x 4 with the
// tables in registers, 64 as 4 x 16 with both fields in one workgroup of 1024 threads)
struct ShapeRow { int edge; int rr[MAX_SLAB_SHAPES][2]; int mask[MAX_SLAB_SHAPES]; };
constexpr ShapeRow SHAPES[] = {
    {128, {{16, 8}}, {0x2}}, {120, {{15, 8}}, {0x2}}, {112, {{14, 8}}, {0x2}}, {96, {{12, 8}}, {0x2}},
    {80, {{8, 10}}, {0x2}}, {72, {{8, 9}}, {0x2}},
    {64, {{8, 8}, {16, 4}, {4, 16}}, {0xF, 0xA, 0x3}}, {56, {{7, 8}, {14, 4}}, {0xF, 0xF}},
    {48, {{6, 8}, {12, 4}}, {0xF, 0xF}}, {40, {{5, 8}, {10, 4}}, {0xF, 0xF}}, {32, {{4, 8}, {8, 4}}, {0xF, 0xF}},
    {24, {{3, 8}, {6, 4}}, {0xF, 0xF}}, {16, {{4, 4}, {2, 8}}, {0xF, 0xF}},
};
}  // namespace

int slab_shape_variants(int edge, int shape) {
    if (shape < 0 || shape >= slab_shape_count(edge)) return 0;
    for (const ShapeRow& r : SHAPES)
        if (r.edge == edge) return r.mask[shape];
    return 0;
}

int slab_shape_count(int edge) {
    for (const ShapeRow& r : SHAPES)
        if (r.edge == edge) {
            int n = 0;
            while (n < MAX_SLAB_SHAPES && r.rr[n][0]) n++;
            return n;
        }
    return 0;
}

bool slab_shape_factors(int edge, int shape, int* r1, int* r2) {
    if (shape < 0 || shape >= slab_shape_count(edge)) return false;
    for (const ShapeRow& r : SHAPES)
        if (r.edge == edge) {
            *r1 = r.rr[shape][0];
            *r2 = r.rr[shape][1];
            return true;
        }
    return false;
}

bool slab_supported(int n0, int n1, int n2) {
    if (n0 != n1 || n1 != n2) return false;
    for (int e : SLAB_EDGES)
        if (n0 == e) return true;
    return false;
}

namespace {


hipError_t dispatch_edge(int n, int v, const Rep* reps, int nact, int i, int mode, const SlabIO* io,
                         const double2* tw, hipStream_t st) {
    if (v >= SLAB_VARIANTS) {  // an alternative factorisation (plain chains only)
        if (mode != CHAIN_PLAIN) return hipErrorInvalidValue;
        return dispatch_edge_alt(n, v / SLAB_VARIANTS, v % SLAB_VARIANTS, reps, nact, i, io, tw, st);
    }
    // Lines of 72 points and more: 9 to 16 elements per thread leave no registers for the
    // tables (the register-table shapes spill), and two planes either do not fit in LDS or
    // cap the registers so that the passes spill, so one shape (LDS tables, one field per
    // workgroup)
    if (n == 128) dispatch<128, 16, 8>(1, reps, nact, i, mode, io, tw, st);
    else if (n == 64) dispatch<64, 8, 8>(v, reps, nact, i, mode, io, tw, st);
    else if (n == 96) dispatch<96, 12, 8>(1, reps, nact, i, mode, io, tw, st);
    else if (n == 40) dispatch<40, 5, 8>(v, reps, nact, i, mode, io, tw, st);
    else if (n == 16) dispatch<16, 4, 4>(v, reps, nact, i, mode, io, tw, st);
    else if (n == 48) dispatch<48, 6, 8>(v, reps, nact, i, mode, io, tw, st);
    else if (n == 32) dispatch<32, 4, 8>(v, reps, nact, i, mode, io, tw, st);
    else if (n == 80) dispatch<80, 8, 10>(1, reps, nact, i, mode, io, tw, st);
    else if (n == 72) dispatch<72, 8, 9>(1, reps, nact, i, mode, io, tw, st);
    else if (n == 56) dispatch<56, 7, 8>(v, reps, nact, i, mode, io, tw, st);
    else if (n == 24) dispatch<24, 3, 8>(v, reps, nact, i, mode, io, tw, st);
    else if (n == 112) dispatch<112, 14, 8>(1, reps, nact, i, mode, io, tw, st);
    else if (n == 120) dispatch<120, 15, 8>(1, reps, nact, i, mode, io, tw, st);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

// The a-line accesses move two doubles at once: every field must start on 16 bytes.
bool al16(const void* p) { return (reinterpret_cast<unsigned long long>(p) & 15) == 0; }
bool io_aligned(const SlabIO& io) {
    bool ok = al16(io.tin) && al16(io.tout) && al16(io.qout);
    for (int f = 0; f < 2; f++)
        ok = ok && al16(io.h[f]) && al16(io.aux[f]) && al16(io.src[f]) && al16(io.part[f]) && al16(io.acc[f]);
    return ok;
}

}  // namespace

hipError_t launch_slab(const Rep* reps, int nact, int n, int i, const double2* tw, hipStream_t st, const Rep* one,
                       int variant, int fuse) {
    if (one) {  // host copies of the descriptors: check the alignment the kernel relies on
        for (int r = 0; r < nact; r++)
            if (!al16(one[r].h) || !al16(one[r].q) || !al16(one[r].T0) || !al16(one[r].T1) ||
                !al16(one[r].phacc) || (one[r].M & 1))
                return hipErrorInvalidValue;
    }
    if (one && nact == 1) {
        if (i > one->N - 1) return hipSuccess;
        const SlabIO io = plain_io(*one, i, fuse);
        if (io.acc[0] && n > 64) return hipErrorInvalidValue;  // no fused kernel above 64^3
        return dispatch_edge(n, variant, nullptr, 1, i, CHAIN_PLAIN, &io, tw, st);
    }
    if (fuse) return hipErrorInvalidValue;  // fused sums run only in the one-replica kernel
    return dispatch_edge(n, variant, reps, nact, i, CHAIN_PLAIN, nullptr, tw, st);
}

bool slab_pair_supported(int ea, int eb) { return ea == 64 && eb == 32; }

// 64^3 as 8 x 8 with the tables in registers, one field per workgroup, and 32^3 as 4 x 8
// with the tables in LDS, both fields per workgroup: 256 threads each
hipError_t launch_slab_pair(const Rep* repsA, int nA, int ea, const Rep* repsB, int nB, int eb, int i,
                            const double2* twA, const double2* twB, hipStream_t st) {
    if (!slab_pair_supported(ea, eb) || nA < 1 || nB < 1) return hipErrorInvalidValue;
    const dim3 grid(nA * 64 * 2 + nB * 32), block(256);
    if (i >= 1)
        hipLaunchKernelGGL((k_slab_pair<64, 8, 8, 0, 32, 4, 8, 1, 1>), grid, block, 0, st, repsA, nA, repsB, twA, twB, i);
    else
        hipLaunchKernelGGL((k_slab_pair<64, 8, 8, 0, 32, 4, 8, 1, 0>), grid, block, 0, st, repsA, nA, repsB, twA, twB, i);
    return hipGetLastError();
}

hipError_t launch_slab_io(int n, int mode, const SlabIO& io, const double2* tw, hipStream_t st, int variant) {
    if (!io_aligned(io) || (io.M & 1)) return hipErrorInvalidValue;
    if (io.acc[0] && (n > 64 || mode != CHAIN_PLAIN)) return hipErrorInvalidValue;  // no fused kernel
    return dispatch_edge(n, variant, nullptr, 1, 0, mode, &io, tw, st);
}

// ---- convolution along one axis with register FFTs (line path) ---------------------------
// For meshes the slab pass does not serve, a contour step is three passes, one per axis
// (src/step.cc:35-65 as Cz, Cy, Cx; the last multiplies by h). A pass whose line length
// launch_axis_conv serves runs here: 16 complex lines per workgroup, each line the two real
// fields of a pair at one position (field A + i field B), held by R2 threads with R1
// elements each and convolved by conv_line with the tables in LDS. For x- and y-lines the
// lanes of a wave take consecutive lines (adjacent in memory), so each load or store
// covers a 128-byte run; for z-lines (contiguous) the R2 threads of a line are adjacent.
// The finish follows the generic kernel's rules (plain, times h, tangent, adjoint).
namespace {

constexpr int AX_LINES = 16;  // 16 lines per workgroup: 128-byte runs, more workgroups per CU

__device__ __forceinline__ long axis_line_base(const LineArgs& a, long line) {
    if (a.axis == 2) return line * a.m2;
    if (a.axis == 1) {
        const long x = line / a.m2, z = line - x * a.m2;
        return x * (long)a.m1 * a.m2 + z;
    }
    return line;  // axis 0: line = y m2 + z
}

}  // namespace

template <int E, int R1, int R2, bool ZAX>
__global__ __launch_bounds__(AX_LINES * R2) void k_axis(const LineArgs a) {
    static_assert(R1 * R2 == E, "unsupported factorisation");
    // exchange row pitch (complex): odd, or for z-lines (a line's threads adjacent) the
    // conflict-free zline_pitch
    constexpr int XP = ZAX ? zline_pitch(R1, R2, (R1 * (R2 + 1)) | 1) : (R1 * (R2 + 1)) | 1;
    constexpr int XBYTES = AX_LINES * XP * 16;
    __shared__ __align__(16) unsigned char smem[XBYTES + E * 16 + E * 8];
    const int t = threadIdx.x;
    const int c = ZAX ? t / R2 : t % AX_LINES;  // line within the workgroup
    const int j = ZAX ? t % R2 : t / AX_LINES;  // lane within the line
    const long line = (long)blockIdx.x * AX_LINES + c;
    const bool live = line < a.nlines;
    const long base = live ? axis_line_base(a, line) : 0;
    const long S = a.stride;
    double2 z[R1];
#pragma unroll
    for (int k = 0; k < R1; k++) {
        const long p = base + (long)(j + R2 * k) * S;
        z[k] = live ? make_double2(a.inA[p], a.inB[p]) : make_double2(0.0, 0.0);
    }
    // the skip flag once the line loads are out (they read only the pass's own operands):
    // tested first, its argument load and wait would come before the others
    if (a.skip && *a.skip) return;
    double2* stw = reinterpret_cast<double2*>(smem + XBYTES);
    double* sg = reinterpret_cast<double*>(smem + XBYTES + E * 16);
    for (int k = t; k < E; k += AX_LINES * R2) {
        stw[k] = a.tw[k];
        sg[k] = a.gnat[k];
    }
    lds_barrier();
    LdsTab<R1, R2> tab;
    tab.load(stw, sg, j);
    conv_line<R1, R2, false>(z, reinterpret_cast<double2*>(smem) + c * XP, j, tab);  // first use
    if (!live) return;
#pragma unroll
    for (int k = 0; k < R1; k++) {
        const long p = base + (long)(j + R2 * k) * S;
        double x = z[k].x, y = z[k].y;
        if (a.hA) {
            if (a.chain == CHAIN_TANGENT) {
                x = fma(x, a.hA[p], a.auxA[p] * a.srcA[p]);
                y = fma(y, a.hB[p], a.auxB[p] * a.srcB[p]);
            } else if (a.chain == CHAIN_ADJOINT) {
                x = fma(a.auxA[p], a.srcA[p], x);
                y = fma(a.auxB[p], a.srcB[p], y);
                a.out2A[p] = x * a.hA[p];
                a.out2B[p] = y * a.hB[p];
            } else {
                x *= a.hA[p];
                y *= a.hB[p];
            }
        }
        a.outA[p] = x;
        a.outB[p] = y;
    }
}

namespace {

// (R1, R2) for z-lines; (R1X, R2X) for x- and y-lines, where a line's threads sit in
// different waves, so stage-2 DFTs over fewer threads than R2X idle whole waves instead of
// lanes and more threads per line can pay (128 = 8 x 16 there)
template <int E, int R1, int R2, int R1X = R1, int R2X = R2>
hipError_t axis_launch(const LineArgs& a, hipStream_t st) {
    const unsigned nblk = (unsigned)((a.nlines + AX_LINES - 1) / AX_LINES);
    if (a.axis == 2) hipLaunchKernelGGL((k_axis<E, R1, R2, true>), dim3(nblk), dim3(AX_LINES * R2), 0, st, a);
    else hipLaunchKernelGGL((k_axis<E, R1X, R2X, false>), dim3(nblk), dim3(AX_LINES * R2X), 0, st, a);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_axis_conv(const LineArgs& a, hipStream_t st) {
    if (a.mode != LINE_CONV || !a.gnat || !a.inA || !a.inB || !a.outA || !a.outB) return hipErrorNotSupported;
    switch (a.n) {
        case 16: return axis_launch<16, 4, 4>(a, st);
        case 24: return axis_launch<24, 3, 8>(a, st);
        case 32: return axis_launch<32, 4, 8>(a, st);
        case 40: return axis_launch<40, 5, 8>(a, st);
        case 48: return axis_launch<48, 6, 8, 3, 16>(a, st);
        case 56: return axis_launch<56, 7, 8>(a, st);
        case 64: return axis_launch<64, 8, 8>(a, st);
        case 72: return axis_launch<72, 8, 9, 8, 9>(a, st);
        case 80: return axis_launch<80, 8, 10, 5, 16>(a, st);
        case 96: return axis_launch<96, 8, 12, 6, 16>(a, st);
        case 112: return axis_launch<112, 8, 14, 7, 16>(a, st);
        case 120: return axis_launch<120, 15, 8, 8, 15>(a, st);
        case 128: return axis_launch<128, 16, 8, 8, 16>(a, st);
        // lengths the slab pass does not serve (a plane would not fit in LDS, or no plane
        // factorisation): the axis kernel keeps only a 16-line tile in LDS
        case 20: return axis_launch<20, 4, 5, 4, 5>(a, st);
        case 36: return axis_launch<36, 6, 6, 6, 6>(a, st);
        case 60: return axis_launch<60, 6, 10, 6, 10>(a, st);
        case 144: return axis_launch<144, 12, 12, 9, 16>(a, st);
        case 160: return axis_launch<160, 10, 16, 10, 16>(a, st);
        case 192: return axis_launch<192, 24, 8, 12, 16>(a, st);
        case 200: return axis_launch<200, 10, 20, 10, 20>(a, st);
        case 256: return axis_launch<256, 16, 16>(a, st);
        case 30: return axis_launch<30, 5, 6, 5, 6>(a, st);
        case 50: return axis_launch<50, 10, 5>(a, st);
        case 100: return axis_launch<100, 10, 10, 10, 10>(a, st);
        // further lengths with factors 2, 3, 5 and 7 (R1 x R2 from the register DFTs)
        case 18: return axis_launch<18, 3, 6>(a, st);
        case 28: return axis_launch<28, 4, 7, 4, 7>(a, st);
        case 42: return axis_launch<42, 6, 7, 6, 7>(a, st);
        case 54: return axis_launch<54, 6, 9, 6, 9>(a, st);
        case 70: return axis_launch<70, 7, 10>(a, st);
        case 84: return axis_launch<84, 7, 12, 7, 12>(a, st);
        case 90: return axis_launch<90, 9, 10>(a, st);
        case 98: return axis_launch<98, 14, 7>(a, st);
        case 108: return axis_launch<108, 9, 12, 9, 12>(a, st);
        case 126: return axis_launch<126, 9, 14, 9, 14>(a, st);
        case 140: return axis_launch<140, 10, 14, 10, 14>(a, st);
        case 150: return axis_launch<150, 10, 15, 10, 15>(a, st);
        case 162: return axis_launch<162, 18, 9>(a, st);
        case 168: return axis_launch<168, 12, 14, 12, 14>(a, st);
        case 180: return axis_launch<180, 15, 12, 12, 15>(a, st);
        case 196: return axis_launch<196, 14, 14>(a, st);
        case 210: return axis_launch<210, 14, 15, 14, 15>(a, st);
        case 216: return axis_launch<216, 18, 12, 12, 18>(a, st);
        case 224: return axis_launch<224, 14, 16>(a, st);
        case 240: return axis_launch<240, 15, 16>(a, st);
        case 250: return axis_launch<250, 10, 25, 10, 25>(a, st);
        case 252: return axis_launch<252, 18, 14, 14, 18>(a, st);
        default: return hipErrorNotSupported;
    }
}


}  // namespace lfts
