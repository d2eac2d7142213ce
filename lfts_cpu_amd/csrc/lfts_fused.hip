// Fused contour step for cubic meshes of 32^3 and 64^3 points: one launch per step.
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
// as complex lines by pairing lines (column b with column b + n/2 for the a-axis, row a
// with row a + n/2 for the z-axis); a 1D convolution with the real, even kernel maps
// real lines to real lines, so real and imaginary parts separate exactly.
//
// A length-n line (n = R1*R2) is held by R2 threads with R1 elements each, thread j
// holding x[j + R2 k]. Forward: radix-R1 DFT in registers, twiddle w_n^{jk'}, exchange
// through LDS, radix-R2 DFTs -> X[k' + R1 f]; multiply by g; the inverse retraces the
// steps with conjugate roots. The kernel factor g includes the 1/n of each inverse.
#include <hip/hip_runtime.h>

#include "lfts_internal.h"

namespace lfts {

namespace {

__device__ __forceinline__ double2 c_add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 c_sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 c_mul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 c_mulc(double2 a, double2 b) {  // a * conj(b)
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ double2 c_rot(double2 a) {
    return INV ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
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

template <int R, bool INV>
__device__ __forceinline__ void dft(double2* v) {
    if constexpr (R == 8) dft8<INV>(v);
    else if constexpr (R == 4) dft4<INV>(v[0], v[1], v[2], v[3]);
    else if constexpr (R == 2) dft2<INV>(v[0], v[1]);
}

// 1D circular convolution of one complex line of length N = R1*R2 held by R2 threads
// (this thread: index j, elements z[k] = x[j + R2 k]). lx is the line's LDS exchange row
// (R1*R2 complex). The thread's twiddles tw[k] = exp(-2 pi i jk/N) and kernel values
// g[u*R2 + f] = g(jU + u + R1 f) are held in registers.
template <int R1, int R2>
__device__ __forceinline__ void conv_line(double2* z, double2* lx, int j, const double2* tw, const double* g) {
    constexpr int U = R1 / R2;
    dft<R1, false>(z);
#pragma unroll
    for (int k = 1; k < R1; k++) z[k] = c_mul(z[k], tw[k]);
    __syncthreads();  // previous readers of the exchange rows are done
#pragma unroll
    for (int k = 0; k < R1; k++) lx[k * R2 + j] = z[k];
    __syncthreads();
    double2 y[R1];
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
        for (int jj = 0; jj < R2; jj++) y[u * R2 + jj] = lx[(j * U + u) * R2 + jj];
#pragma unroll
    for (int u = 0; u < U; u++) {
        dft<R2, false>(y + u * R2);
#pragma unroll
        for (int f = 0; f < R2; f++) {
            y[u * R2 + f].x *= g[u * R2 + f];
            y[u * R2 + f].y *= g[u * R2 + f];
        }
        dft<R2, true>(y + u * R2);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
        for (int jj = 0; jj < R2; jj++) lx[(j * U + u) * R2 + jj] = y[u * R2 + jj];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < R1; k++) z[k] = lx[k * R2 + j];
#pragma unroll
    for (int k = 1; k < R1; k++) z[k] = c_mulc(z[k], tw[k]);
    dft<R1, true>(z);
}

}  // namespace

template <int N, int R1, int R2>
__global__ __launch_bounds__((N / 2) * R2) void k_slab(SlabArgs a) {
    static_assert(R1 * R2 == N && R1 % R2 == 0, "unsupported factorisation");
    constexpr int H = N / 2;           // complex lines per plane
    constexpr int U = R1 / R2;
    constexpr int XP = N + 1;          // exchange row pitch (complex)
    constexpr int RP = N + 1;          // real plane row pitch (doubles)
    constexpr int XBYTES = H * XP * 16;
    constexpr int RBYTES = N * RP * 8;
    constexpr int BYTES = XBYTES > RBYTES ? XBYTES : RBYTES;
    __shared__ __align__(16) unsigned char smem[BYTES];
    double2* xbuf = reinterpret_cast<double2*>(smem);
    double* rp = reinterpret_cast<double*>(smem);

    const int t = threadIdx.x;
    const int f = blockIdx.y;
    const long pbase = (long)blockIdx.x * a.plane_stride;  // plane origin within a field
    const long base = (long)f * a.M + pbase;
    const int c = t % H, j = t / H;    // layout A: a-lines, column pair (c, c + H), a = j + R2 k
    const int r = t / R2, jb = t % R2; // layout B: z-lines, row pair (r, r + H), b = jb + R2 k

    // everything this thread reads is issued up front: the partial pair, the Boltzmann
    // factor, and its twiddles and kernel values (layout A and layout B)
    double2 z[R1], hz[R1];
    const double* hsel = f ? a.hqd : a.hq;
#pragma unroll
    for (int k = 0; k < R1; k++) {
        const long p = pbase + (long)(j + R2 * k) * a.Sa + c;
        z[k] = make_double2(a.Tin[(long)f * a.M + p], a.Tin[(long)f * a.M + p + H]);
        if (a.finish) hz[k] = make_double2(hsel[p], hsel[p + H]);
    }
    double2 twA[R1], twB[R1];
    double gA[R1], gB[R1];
#pragma unroll
    for (int k = 0; k < R1; k++) {
        twA[k] = a.tw[j * k];
        twB[k] = a.tw[jb * k];
    }
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
        for (int ff = 0; ff < R2; ff++) {
            gA[u * R2 + ff] = a.ga[j * U + u + R1 * ff];
            gB[u * R2 + ff] = a.gb[jb * U + u + R1 * ff];
        }

    if (a.finish) {
        conv_line<R1, R2>(z, xbuf + c * XP, j, twA, gA);
#pragma unroll
        for (int k = 0; k < R1; k++) {
            const long p = (long)f * a.M + pbase + (long)(j + R2 * k) * a.Sa + c;
            z[k].x *= hz[k].x;
            z[k].y *= hz[k].y;
            a.qout[p] = z[k].x;
            a.qout[p + H] = z[k].y;
        }
    }
    if (!a.cont) return;

    conv_line<R1, R2>(z, xbuf + c * XP, j, twA, gA);
    // re-pair the real plane: rows (r, r + H) become the complex z-lines
    __syncthreads();
#pragma unroll
    for (int k = 0; k < R1; k++) {
        const int ai = j + R2 * k;
        rp[ai * RP + c] = z[k].x;
        rp[ai * RP + c + H] = z[k].y;
    }
    __syncthreads();
    double2 w[R1];
#pragma unroll
    for (int k = 0; k < R1; k++) {
        const int b = jb + R2 * k;
        w[k] = make_double2(rp[r * RP + b], rp[(r + H) * RP + b]);
    }
    conv_line<R1, R2>(w, xbuf + r * XP, jb, twB, gB);
#pragma unroll
    for (int k = 0; k < R1; k++) {
        const long p = base + (long)r * a.Sa + jb + R2 * k;
        a.Tout[p] = w[k].x;
        a.Tout[p + (long)H * a.Sa] = w[k].y;
    }
}

bool slab_supported(int n0, int n1, int n2) {
    return n0 == n1 && n1 == n2 && (n0 == 32 || n0 == 64);
}

hipError_t launch_slab(const SlabArgs& a, int n, int nplanes, hipStream_t st) {
    const dim3 grid(nplanes, 2);
    if (n == 64) {
        hipLaunchKernelGGL((k_slab<64, 8, 8>), grid, dim3(32 * 8), 0, st, a);
    } else if (n == 32) {
        hipLaunchKernelGGL((k_slab<32, 8, 4>), grid, dim3(16 * 4), 0, st, a);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace lfts
