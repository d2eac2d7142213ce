// Fused contour step for cubic meshes of 32^3, 48^3 and 64^3 points: one launch per
// step over every active replica of that mesh size.
//
// The contour step q_{i+1} = h ⊙ C q_i (src/step.cc:35-65) uses the 3D convolution
// C = Cx Cy Cz, separable into 1D circular convolutions along x, y and z. A slab of the
// mesh holds two complete axes: an x-slab (fixed x) holds y- and z-lines, a y-slab
// (fixed y) holds x- and z-lines. Passes alternate between the two slab kinds, and each
// pass finishes the step in flight and starts the next one:
//
//   x-slab pass:  T = Cx Cz q_i (from the y-slab pass)   ->  Cy  -> *h -> q_{i+1}
//                 then Cy, Cz of q_{i+1}                  ->  T' = Cy Cz q_{i+1}
//   y-slab pass:  T = Cy Cz q_i (from the x-slab pass)    ->  Cx  -> *h -> q_{i+1}
//                 then Cx, Cz of q_{i+1}                  ->  T' = Cx Cz q_{i+1}
//
// so one launch moves one step (one global exchange per step). Both slab kinds are
// read and written in the row-major layout with runs of >= 64 contiguous bytes.
//
// One workgroup owns one slab plane of one field (q or q+). The real plane is processed
// as complex lines by pairing lines (adjacent columns 2c, 2c+1 for the a-axis, so a lane
// moves both with one 16-byte access; rows a and a + n/2 for the z-axis); a 1D
// convolution with the real, even kernel maps real lines to real lines, so real and
// imaginary parts separate exactly.
//
// A length-n line (n = R1*R2) is held by R2 threads with R1 elements each, thread j
// holding x[j + R2 k]. Forward: radix-R1 DFT in registers, twiddle w_n^{jk'}, exchange
// through LDS, radix-R2 DFTs -> X[k' + R1 f]; multiply by g; the inverse retraces the
// steps with conjugate roots. The kernel factor g includes the 1/n of each inverse.
// In LDS a line keeps element k*R2 + j at k*(R2+1) + j: the R2 threads of a line then
// read their contiguous runs from distinct banks.
//
// Outputs (pairs and partial pairs) are written with non-temporal stores: they are read
// next by workgroups on other XCDs, so they go to memory rather than stay in this L2.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "lfts_internal.h"

// Rounding is fixed by the source, not left to the compiler: fused multiply-adds are
// written out (fma) where they are wanted and nothing else is contracted, so every
// kernel shape below computes bit-identical results.
#pragma clang fp contract(off)

namespace lfts {

namespace {

__device__ __forceinline__ double2 c_add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 c_sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 c_mul(double2 a, double2 b) {
    return make_double2(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 c_mulc(double2 a, double2 b) {  // a * conj(b)
    return make_double2(fma(a.x, b.x, a.y * b.y), fma(a.y, b.x, -(a.x * b.y)));
}
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ double2 c_rot(double2 a) {
    return INV ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

typedef double d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ld2(const double* p) { return *reinterpret_cast<const double2*>(p); }
__device__ __forceinline__ void st_nt(double* p, double v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st2_nt(double* p, double2 v) {
    const d2v w = {v.x, v.y};
    __builtin_nontemporal_store(w, reinterpret_cast<d2v*>(p));
}

template <bool INV>
__device__ __forceinline__ void dft2(double2& a, double2& b) {
    double2 t = a;
    a = c_add(t, b);
    b = c_sub(t, b);
}

// natural-order radix-4 DFT of v[s], v[s+st], v[s+2st], v[s+3st]
template <bool INV>
__device__ __forceinline__ void dft4(double2& v0, double2& v1, double2& v2, double2& v3) {
    double2 s02 = c_add(v0, v2), d02 = c_sub(v0, v2);
    double2 s13 = c_add(v1, v3), d13 = c_rot<INV>(c_sub(v1, v3));
    v0 = c_add(s02, s13);
    v2 = c_sub(s02, s13);
    v1 = c_add(d02, d13);
    v3 = c_sub(d02, d13);
}

// natural-order radix-8 DFT: radix-2 split, twiddles w8^n, two radix-4 DFTs
template <bool INV>
__device__ __forceinline__ void dft8(double2* v) {
    const double r = 0.70710678118654752440;
    double2 a0 = c_add(v[0], v[4]), b0 = c_sub(v[0], v[4]);
    double2 a1 = c_add(v[1], v[5]), b1 = c_sub(v[1], v[5]);
    double2 a2 = c_add(v[2], v[6]), b2 = c_sub(v[2], v[6]);
    double2 a3 = c_add(v[3], v[7]), b3 = c_sub(v[3], v[7]);
    // b_n *= w8^n; w8 = (1 - i)/sqrt2 forward, (1 + i)/sqrt2 inverse
    b1 = INV ? make_double2(r * (b1.x - b1.y), r * (b1.x + b1.y)) : make_double2(r * (b1.x + b1.y), r * (b1.y - b1.x));
    b2 = c_rot<INV>(b2);
    b3 = INV ? make_double2(-r * (b3.x + b3.y), r * (b3.x - b3.y)) : make_double2(r * (b3.y - b3.x), -r * (b3.x + b3.y));
    dft4<INV>(a0, a1, a2, a3);
    dft4<INV>(b0, b1, b2, b3);
    v[0] = a0; v[2] = a1; v[4] = a2; v[6] = a3;
    v[1] = b0; v[3] = b1; v[5] = b2; v[7] = b3;
}

// natural-order radix-3 DFT
template <bool INV>
__device__ __forceinline__ void dft3(double2& a, double2& b, double2& c) {
    const double s3 = 0.86602540378443864676;  // sqrt(3)/2
    const double2 t = c_add(b, c), d = c_sub(b, c);
    const double2 m = make_double2(a.x - 0.5 * t.x, a.y - 0.5 * t.y);
    const double2 isd = make_double2(-s3 * d.y, s3 * d.x);  // i (sqrt3/2) d
    a = c_add(a, t);
    b = INV ? c_add(m, isd) : c_sub(m, isd);
    c = INV ? c_sub(m, isd) : c_add(m, isd);
}

// natural-order radix-12 DFT as 4 x 3: x[n2 + 3 n1] -> radix-4 over n1, twiddle
// w12^{n2 k1}, radix-3 over n2 -> X[k1 + 4 k2]
template <bool INV>
__device__ __forceinline__ void dft12(double2* v) {
    const double s3 = 0.86602540378443864676;
    double2 y[3][4];
#pragma unroll
    for (int n2 = 0; n2 < 3; n2++) {
        y[n2][0] = v[n2];
        y[n2][1] = v[n2 + 3];
        y[n2][2] = v[n2 + 6];
        y[n2][3] = v[n2 + 9];
        dft4<INV>(y[n2][0], y[n2][1], y[n2][2], y[n2][3]);
    }
    // w12^e (forward) for e = 1, 2, 3, 4, 6; the inverse uses the conjugates
    const double2 w1 = make_double2(s3, INV ? 0.5 : -0.5);
    const double2 w2 = make_double2(0.5, INV ? s3 : -s3);
    const double2 w4 = make_double2(-0.5, INV ? s3 : -s3);
    y[1][1] = c_mul(y[1][1], w1);
    y[1][2] = c_mul(y[1][2], w2);
    y[1][3] = c_rot<INV>(y[1][3]);             // w12^3 = -i
    y[2][1] = c_mul(y[2][1], w2);
    y[2][2] = c_mul(y[2][2], w4);
    y[2][3] = make_double2(-y[2][3].x, -y[2][3].y);  // w12^6 = -1
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) {
        dft3<INV>(y[0][k1], y[1][k1], y[2][k1]);
        v[k1] = y[0][k1];
        v[k1 + 4] = y[1][k1];
        v[k1 + 8] = y[2][k1];
    }
}

template <int R, bool INV>
__device__ __forceinline__ void dft(double2* v) {
    if constexpr (R == 12) dft12<INV>(v);
    else if constexpr (R == 8) dft8<INV>(v);
    else if constexpr (R == 4) dft4<INV>(v[0], v[1], v[2], v[3]);
    else if constexpr (R == 2) dft2<INV>(v[0], v[1]);
}

// Twiddles w_n^{jk} (k < R1) and kernel values g(jU + u + R1 f) of the line lane j a
// thread serves: held in registers (loaded once per pass) or read from the workgroup's
// LDS copy of the axis tables when used (fewer registers; 48-point lines need that to
// run two waves per SIMD).
template <int R1, int R2>
struct RegTab {
    double2 tw[R1];
    double g[R1];
    __device__ __forceinline__ void load(const double2* __restrict__ twt, const double* __restrict__ gt, int j) {
        constexpr int U = R1 / R2;
#pragma unroll
        for (int k = 0; k < R1; k++) tw[k] = twt[j * k];
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int f = 0; f < R2; f++) g[u * R2 + f] = gt[j * U + u + R1 * f];
    }
    __device__ __forceinline__ double2 w(int k) const { return tw[k]; }
    __device__ __forceinline__ double gv(int u, int f) const { return g[u * R2 + f]; }
};

template <int R1, int R2>
struct LdsTab {
    const double2* tw;
    const double* g;
    int j;
    __device__ __forceinline__ void load(const double2* twt, const double* gt, int jj) {
        tw = twt;
        g = gt;
        j = jj;
    }
    __device__ __forceinline__ double2 w(int k) const { return tw[j * k]; }
    __device__ __forceinline__ double gv(int u, int f) const { return g[j * (R1 / R2) + u + R1 * f]; }
};

// 1D circular convolution of one complex line of length N = R1*R2 held by R2 threads
// (this thread: index j, elements z[k] = x[j + R2 k]). lx is the line's LDS exchange row
// (R1*R2 complex); tab supplies the thread's twiddles exp(-2 pi i jk/N) and kernel values.
template <int R1, int R2, class Tab>
__device__ __forceinline__ void conv_line(double2* z, double2* lx, int j, const Tab& tab) {
    constexpr int U = R1 / R2;
    constexpr int Q = R2 + 1;  // padded run length in LDS
    dft<R1, false>(z);
#pragma unroll
    for (int k = 1; k < R1; k++) z[k] = c_mul(z[k], tab.w(k));
    __syncthreads();  // previous readers of the exchange rows are done
#pragma unroll
    for (int k = 0; k < R1; k++) lx[k * Q + j] = z[k];
    __syncthreads();
    double2 y[R1];
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
        for (int jj = 0; jj < R2; jj++) y[u * R2 + jj] = lx[(j * U + u) * Q + jj];
#pragma unroll
    for (int u = 0; u < U; u++) {
        dft<R2, false>(y + u * R2);
#pragma unroll
        for (int f = 0; f < R2; f++) {
            const double gg = tab.gv(u, f);
            y[u * R2 + f].x *= gg;
            y[u * R2 + f].y *= gg;
        }
        dft<R2, true>(y + u * R2);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
        for (int jj = 0; jj < R2; jj++) lx[(j * U + u) * Q + jj] = y[u * R2 + jj];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < R1; k++) z[k] = lx[k * Q + j];
#pragma unroll
    for (int k = 1; k < R1; k++) z[k] = c_mulc(z[k], tab.w(k));
    dft<R1, true>(z);
}

}  // namespace

// One slab pass of one replica (plane `plane`), addresses in `io`. The workgroup serves
// field blockIdx.y, or with FPW = 2 both fields (threads [0, TPF) field 0, [TPF, 2 TPF)
// field 1). TWL = 1 keeps the twiddle and kernel tables in LDS instead of registers.
// MODE selects the finishing rule of pair i+1 from u = (C x)(r):
//   CHAIN_PLAIN    out = h u                   (src/step.cc:53-64)
//   CHAIN_TANGENT  out = h u + aux src         (directional derivative of the chain)
//   CHAIN_ADJOINT  out = u + aux src, stored; the chain continues with h out
// Every variant performs the same arithmetic on every element: results are bit-identical.
template <int E, int R1, int R2, int MODE, int TWL, int FPW>
__device__ __forceinline__ void slab_pass(const SlabIO& io, int plane, const double2* __restrict__ tw) {
    static_assert(R1 * R2 == E && R1 % R2 == 0, "unsupported factorisation");
    constexpr int H = E / 2;           // complex lines per plane
    constexpr int TPF = H * R2;        // threads per field
    constexpr int XP = R1 * (R2 + 1) + 1;  // exchange row pitch (complex): padded line + 1
    constexpr int RP = E + 8;          // real plane row pitch (doubles): rows r, r+1 on other banks
    constexpr int XBYTES = H * XP * 16;
    constexpr int RBYTES = E * RP * 8;
    constexpr int BYTES = XBYTES > RBYTES ? XBYTES : RBYTES;
    constexpr int TBYTES = TWL ? E * 16 + 2 * E * 8 : 0;
    constexpr bool EXTRA = MODE != CHAIN_PLAIN;
    __shared__ __align__(16) unsigned char smem[FPW * BYTES + TBYTES];
    if (io.skip && *io.skip) return;

    const bool xslab = io.xslab != 0;
    const bool finish = io.finish != 0, cont = io.cont != 0;
    const long M = io.M;
    const long plane_stride = xslab ? (long)E * E : E;
    const long Sa = xslab ? E : (long)E * E;
    const double* ga = xslab ? io.g[1] : io.g[0];
    const double* gb = io.g[2];

    const int fl = (FPW == 2) ? (int)(threadIdx.x / TPF) : 0;
    const int t = threadIdx.x - fl * TPF;
    const int f = (FPW == 2) ? fl : (int)blockIdx.y;
    double2* xbuf = reinterpret_cast<double2*>(smem + fl * BYTES);
    double* rp = reinterpret_cast<double*>(smem + fl * BYTES);
    const long pbase = (long)plane * plane_stride;  // plane origin within a field
    const long base = (long)f * M + pbase;
    const int c = t % H, j = t / H;    // layout A: a-lines, column pair (2c, 2c + 1), a = j + R2 k
    const int r = t / R2, jb = t % R2; // layout B: z-lines, row pair (r, r + H), b = jb + R2 k

    // everything this thread reads is issued up front: the partial pair, the Boltzmann
    // factor of its field at this bead (src/step.cc:53-64), twiddles and kernel values
    const double* Tin = io.tin + (long)f * M;
    const double* hsel = f ? io.h[1] : io.h[0];
    const double* asel = f ? io.aux[1] : io.aux[0];
    const double* ssel = f ? io.src[1] : io.src[0];
    double2 z[R1], hz[R1], az[EXTRA ? R1 : 1], sz[EXTRA ? R1 : 1];
#pragma unroll
    for (int k = 0; k < R1; k++) {
        const long p = pbase + (long)(j + R2 * k) * Sa + 2 * c;
        z[k] = ld2(Tin + p);
        if (finish) {
            hz[k] = ld2(hsel + p);
            if constexpr (EXTRA) {
                az[k] = ld2(asel + p);
                sz[k] = ld2(ssel + p);
            }
        }
    }
    using Tab = typename std::conditional<TWL != 0, LdsTab<R1, R2>, RegTab<R1, R2>>::type;
    Tab tabA, tabB;
    if constexpr (TWL != 0) {
        double2* stw = reinterpret_cast<double2*>(smem + FPW * BYTES);
        double* sga = reinterpret_cast<double*>(smem + FPW * BYTES + E * 16);
        double* sgb = sga + E;
        for (int k = threadIdx.x; k < E; k += blockDim.x) {
            stw[k] = tw[k];
            sga[k] = ga[k];
            sgb[k] = gb[k];
        }
        __syncthreads();
        tabA.load(stw, sga, j);
        tabB.load(stw, sgb, jb);
    } else {
        tabA.load(tw, ga, j);
        tabB.load(tw, gb, jb);
    }

    if (finish) {
        conv_line<R1, R2>(z, xbuf + c * XP, j, tabA);
        double* qout = io.qout;
#pragma unroll
        for (int k = 0; k < R1; k++) {
            const long p = base + (long)(j + R2 * k) * Sa + 2 * c;
            if constexpr (MODE == CHAIN_PLAIN) {
                z[k].x *= hz[k].x;
                z[k].y *= hz[k].y;
            } else if constexpr (MODE == CHAIN_TANGENT) {
                z[k].x = fma(z[k].x, hz[k].x, az[k].x * sz[k].x);
                z[k].y = fma(z[k].y, hz[k].y, az[k].y * sz[k].y);
            } else {
                z[k].x = fma(az[k].x, sz[k].x, z[k].x);
                z[k].y = fma(az[k].y, sz[k].y, z[k].y);
            }
            st2_nt(qout + p, z[k]);
            if constexpr (MODE == CHAIN_ADJOINT) {
                z[k].x *= hz[k].x;
                z[k].y *= hz[k].y;
            }
        }
    }
    if (!cont) return;

    conv_line<R1, R2>(z, xbuf + c * XP, j, tabA);
    // re-pair the real plane: rows (r, r + H) become the complex z-lines
    __syncthreads();
#pragma unroll
    for (int k = 0; k < R1; k++) {
        const int ai = j + R2 * k;
        *reinterpret_cast<double2*>(rp + ai * RP + 2 * c) = z[k];
    }
    __syncthreads();
    double2 w[R1];
#pragma unroll
    for (int k = 0; k < R1; k++) {
        const int b = jb + R2 * k;
        w[k] = make_double2(rp[r * RP + b], rp[(r + H) * RP + b]);
    }
    conv_line<R1, R2>(w, xbuf + r * XP, jb, tabB);
    double* Tout = io.tout;
#pragma unroll
    for (int k = 0; k < R1; k++) {
        const long p = base + (long)r * Sa + jb + R2 * k;
        st_nt(Tout + p, w[k].x);
        st_nt(Tout + p + (long)H * Sa, w[k].y);
    }
}

// Plain chains of a replica set. Workgroup (slot * E + plane, field): slot picks the
// replica descriptor.
template <int E, int R1, int R2, int TWL, int FPW>
__global__ __launch_bounds__((E / 2) * R2 * FPW) void k_slab(const Rep* __restrict__ reps,
                                                           const double2* __restrict__ tw, int i) {
    const int slot = blockIdx.x / E, plane = blockIdx.x - slot * E;
    const Rep& R = reps[slot];  // fields read in place (a private copy would live in scratch)
    if (i > R.N - 1) return;    // this replica's chain is complete
    const SlabIO io = plain_io(R, i);
    slab_pass<E, R1, R2, CHAIN_PLAIN, TWL, FPW>(io, plane, tw);
}

// One replica, its pass addresses in the kernel arguments (no dependent descriptor load
// before the first field loads).
template <int E, int R1, int R2, int MODE, int TWL, int FPW>
__global__ __launch_bounds__((E / 2) * R2 * FPW) void k_slab_io(const SlabIO io, const double2* __restrict__ tw) {
    slab_pass<E, R1, R2, MODE, TWL, FPW>(io, blockIdx.x, tw);
}

bool slab_supported(int n0, int n1, int n2) {
    return n0 == n1 && n1 == n2 && (n0 == 32 || n0 == 48 || n0 == 64);
}

namespace {

template <int E, int R1, int R2, int TWL, int FPW>
void launch_set(const Rep* reps, int nact, int i, const double2* tw, hipStream_t st) {
    const dim3 grid(nact * E, 2 / FPW), block((E / 2) * R2 * FPW);
    hipLaunchKernelGGL((k_slab<E, R1, R2, TWL, FPW>), grid, block, 0, st, reps, tw, i);
}

template <int E, int R1, int R2, int TWL, int FPW>
void launch_one(int mode, const SlabIO& io, const double2* tw, hipStream_t st) {
    const dim3 grid(E, 2 / FPW), block((E / 2) * R2 * FPW);
    if (mode == CHAIN_PLAIN)
        hipLaunchKernelGGL((k_slab_io<E, R1, R2, CHAIN_PLAIN, TWL, FPW>), grid, block, 0, st, io, tw);
    else if (mode == CHAIN_TANGENT)
        hipLaunchKernelGGL((k_slab_io<E, R1, R2, CHAIN_TANGENT, TWL, FPW>), grid, block, 0, st, io, tw);
    else
        hipLaunchKernelGGL((k_slab_io<E, R1, R2, CHAIN_ADJOINT, TWL, FPW>), grid, block, 0, st, io, tw);
}

// variant v: bit 0 = tables in LDS, bit 1 = both fields in one workgroup
template <int E, int R1, int R2>
void dispatch(int v, const Rep* reps, int nact, int i, int mode, const SlabIO* io, const double2* tw,
              hipStream_t st) {
    switch (v) {
        case 1: io ? launch_one<E, R1, R2, 1, 1>(mode, *io, tw, st) : launch_set<E, R1, R2, 1, 1>(reps, nact, i, tw, st); break;
        case 2: io ? launch_one<E, R1, R2, 0, 2>(mode, *io, tw, st) : launch_set<E, R1, R2, 0, 2>(reps, nact, i, tw, st); break;
        case 3: io ? launch_one<E, R1, R2, 1, 2>(mode, *io, tw, st) : launch_set<E, R1, R2, 1, 2>(reps, nact, i, tw, st); break;
        default: io ? launch_one<E, R1, R2, 0, 1>(mode, *io, tw, st) : launch_set<E, R1, R2, 0, 1>(reps, nact, i, tw, st); break;
    }
}

hipError_t dispatch_edge(int n, int v, const Rep* reps, int nact, int i, int mode, const SlabIO* io,
                         const double2* tw, hipStream_t st) {
    if (n == 64) dispatch<64, 8, 8>(v, reps, nact, i, mode, io, tw, st);
    else if (n == 48) dispatch<48, 12, 4>(v, reps, nact, i, mode, io, tw, st);
    else if (n == 32) dispatch<32, 8, 4>(v, reps, nact, i, mode, io, tw, st);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

// The a-line accesses move two doubles at once: every field must start on 16 bytes.
bool al16(const void* p) { return (reinterpret_cast<unsigned long long>(p) & 15) == 0; }
bool io_aligned(const SlabIO& io) {
    bool ok = al16(io.tin) && al16(io.tout) && al16(io.qout);
    for (int f = 0; f < 2; f++) ok = ok && al16(io.h[f]) && al16(io.aux[f]) && al16(io.src[f]);
    return ok;
}

}  // namespace

hipError_t launch_slab(const Rep* reps, int nact, int n, int i, const double2* tw, hipStream_t st, const Rep* one,
                       int variant) {
    if (one) {  // host copies of the descriptors: check the alignment the kernel relies on
        for (int r = 0; r < nact; r++)
            if (!al16(one[r].h) || !al16(one[r].q) || !al16(one[r].T0) || !al16(one[r].T1) || (one[r].M & 1))
                return hipErrorInvalidValue;
    }
    if (one && nact == 1) {
        if (i > one->N - 1) return hipSuccess;
        const SlabIO io = plain_io(*one, i);
        return dispatch_edge(n, variant, nullptr, 1, i, CHAIN_PLAIN, &io, tw, st);
    }
    return dispatch_edge(n, variant, reps, nact, i, CHAIN_PLAIN, nullptr, tw, st);
}

hipError_t launch_slab_io(int n, int mode, const SlabIO& io, const double2* tw, hipStream_t st, int variant) {
    if (!io_aligned(io) || (io.M & 1)) return hipErrorInvalidValue;
    return dispatch_edge(n, variant, nullptr, 1, 0, mode, &io, tw, st);
}

}  // namespace lfts
