// Generic line kernels of the L-FTS contour step for gfx950 (FP64 throughout).
//
// The contour step of the reference (src/step.cc:35-65) is
//     q_{i+1} = h ⊙ c2r( g ⊙ r2c(q_i) ),   g(k) = exp(-|K|^2/(6N)) / M.
// g factorises over the axes, g = g_x(k0) g_y(k1) g_z(k2), so the 3D convolution with
// g is the composition of three 1D circular convolutions, one per axis. Each is a 1D
// FFT -> multiply -> inverse FFT on a line of the mesh, and maps real lines to real
// lines. The two real propagators of a pair (q_i and q+_{N+1-i}) are packed as the real
// and imaginary part of one complex line, so one complex transform serves both.
//
// Line transforms run in LDS: a decimation-in-frequency forward pass leaves the
// spectrum in digit-reversed order, the multiplier is applied in that order (gperm),
// and a decimation-in-time inverse pass returns natural order, so no reordering pass
// is needed.
#include <hip/hip_runtime.h>

#include "lfts_internal.h"

namespace lfts {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// Radix-R DFT of v[0..R) in registers. Forward uses exp(-2 pi i/R), inverse its
// conjugate; both unnormalised. tw[k] = exp(-2 pi i k/n) with R | n.
template <int R, bool INV>
__device__ __forceinline__ void bfly(double2* v, const double2* __restrict__ tw, int n) {
    if constexpr (R == 2) {
        double2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        double2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        double2 s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
        v[0] = cadd(s02, s13);
        v[2] = csub(s02, s13);
        if constexpr (!INV) {  // X1 = d02 - i d13, X3 = d02 + i d13
            v[1] = make_double2(d02.x + d13.y, d02.y - d13.x);
            v[3] = make_double2(d02.x - d13.y, d02.y + d13.x);
        } else {
            v[1] = make_double2(d02.x - d13.y, d02.y + d13.x);
            v[3] = make_double2(d02.x + d13.y, d02.y - d13.x);
        }
    } else if constexpr (R == 3) {
        const double s3 = 0.86602540378443864676;  // sqrt(3)/2
        double2 t = cadd(v[1], v[2]), d = csub(v[1], v[2]);
        double2 a = make_double2(v[0].x - 0.5 * t.x, v[0].y - 0.5 * t.y);
        v[0] = cadd(v[0], t);
        // forward: X1 = a - i s d, X2 = a + i s d
        double2 isd = make_double2(-s3 * d.y, s3 * d.x);
        if constexpr (!INV) {
            v[1] = csub(a, isd);
            v[2] = cadd(a, isd);
        } else {
            v[1] = cadd(a, isd);
            v[2] = csub(a, isd);
        }
    } else {
        double2 out[R];
        const int st = n / R;
#pragma unroll
        for (int s = 0; s < R; s++) {
            double2 acc = v[0];
#pragma unroll
            for (int q = 1; q < R; q++) {
                double2 w = tw[((q * s) % R) * st];
                acc = INV ? cadd(acc, cmulc(v[q], w)) : cadd(acc, cmul(v[q], w));
            }
            out[s] = acc;
        }
#pragma unroll
        for (int s = 0; s < R; s++) v[s] = out[s];
    }
}

// One in-place stage over T lines held in LDS (row pitch `pitch`). Block size B,
// radix R, sub-block S = B/R. Forward (DIF): butterfly then twiddle w_B^{j q};
// inverse (DIT): conjugate twiddle then inverse butterfly.
template <int R, bool INV>
__device__ void line_stage(double2* lds, int pitch, int T, int n, int B, const double2* __restrict__ tw) {
    const int S = B / R, nb = n / R, tstep = n / B;
    for (int idx = threadIdx.x; idx < T * nb; idx += blockDim.x) {
        const int t = idx / nb, b = idx - t * nb;
        const int blk = b / S, j = b - blk * S;
        double2* p = lds + t * pitch + blk * B + j;
        double2 v[R];
#pragma unroll
        for (int q = 0; q < R; q++) v[q] = p[q * S];
        if constexpr (!INV) {
            bfly<R, false>(v, tw, n);
#pragma unroll
            for (int q = 1; q < R; q++) v[q] = cmul(v[q], tw[j * q * tstep]);
        } else {
#pragma unroll
            for (int q = 1; q < R; q++) v[q] = cmulc(v[q], tw[j * q * tstep]);
            bfly<R, true>(v, tw, n);
        }
#pragma unroll
        for (int q = 0; q < R; q++) p[q * S] = v[q];
    }
}

// A stage of any prime radix R (direct DFT, R products per output), out of place: each
// thread forms single outputs of the butterflies into tmp (the line tile's layout), then
// copies them back. The same operations as line_stage: forward, the butterfly then the
// twiddle w_B^{j s}; inverse, the conjugate twiddle then the inverse butterfly.
template <bool INV>
__device__ void prime_stage(int R, double2* lds, double2* tmp, int pitch, int T, int n, int B,
                            const double2* __restrict__ tw) {
    const int S = B / R, tstep = n / B, st = n / R;
    for (int idx = threadIdx.x; idx < T * n; idx += blockDim.x) {
        const int t = idx / n, e = idx - t * n;
        const int blk = e / B, rem = e - blk * B;
        const int s = rem / S, j = rem - s * S;  // output s of butterfly j of block blk
        const double2* p = lds + t * pitch + blk * B + j;
        double2 acc = p[0];
        for (int q = 1; q < R; q++) {
            double2 v = p[q * S];
            const double2 w = tw[((q * s) % R) * st];
            if constexpr (INV) {
                v = cmulc(v, tw[j * q * tstep]);
                acc = cadd(acc, cmulc(v, w));
            } else {
                acc = cadd(acc, cmul(v, w));
            }
        }
        if (!INV && s > 0) acc = cmul(acc, tw[j * s * tstep]);
        tmp[t * pitch + e] = acc;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < T * n; idx += blockDim.x) {
        const int t = idx / n, e = idx - t * n;
        lds[t * pitch + e] = tmp[t * pitch + e];
    }
}

// Radices 2, 3, 4, 5 and 7 unrolled; with PRIME, the primes from 11 up as prime_stage, in a
// kernel of its own whose second tile (tmp) doubles its LDS (unrolled radix-11 and radix-13
// stages took 251 VGPRs, prime_stage keeps the kernel at four waves per SIMD).
template <bool INV, bool PRIME>
__device__ void run_stage(int r, double2* lds, int pitch, int T, int n, int B, const double2* tw) {
    if constexpr (PRIME) {
        if (r >= 11) {
            prime_stage<INV>(r, lds, lds + T * pitch, pitch, T, n, B, tw);
            return;
        }
    }
    switch (r) {
        case 2: line_stage<2, INV>(lds, pitch, T, n, B, tw); break;
        case 3: line_stage<3, INV>(lds, pitch, T, n, B, tw); break;
        case 4: line_stage<4, INV>(lds, pitch, T, n, B, tw); break;
        case 5: line_stage<5, INV>(lds, pitch, T, n, B, tw); break;
        case 7: line_stage<7, INV>(lds, pitch, T, n, B, tw); break;
        default: break;  // plans only contain the radices above (checked on the host)
    }
}

__device__ __forceinline__ long line_base(const LineArgs& a, long line) {
    if (a.axis == 2) return line * a.m2;
    if (a.axis == 1) {
        long x = line / a.m2, z = line - x * a.m2;
        return x * (long)a.m1 * a.m2 + z;
    }
    return line;  // axis 0: line = y*m2 + z
}

// Generic line kernel: T lines of length n along `axis` per workgroup.
//  mode LINE_CONV: real pair in (inA,inB) -> FFT -> *gperm -> IFFT -> [*h] -> real pair out
//  mode LINE_FFT : complex, real, or real pair (inA + i inB) in -> forward FFT -> natural
//                  order complex out, or, on the last axis, |X|^2 accumulated into the
//                  half spectrum Sacc.
//  mode LINE_IFFT: natural-order complex in -> inverse FFT -> complex out, or its real
//                  part (outR).
//  mode LINE_CMUL: complex in -> FFT -> *mul (per point) -> IFFT -> complex out.
template <bool PRIME>
__global__ __launch_bounds__(256) void k_line(LineArgs a) {
    extern __shared__ double2 lds[];
    const int n = a.n, T = a.T, pitch = a.pitch;
    const long line0 = (long)blockIdx.x * T;
    const bool zaxis = (a.axis == 2);

    for (int f = threadIdx.x; f < T * n; f += blockDim.x) {
        int t, e;
        if (zaxis) { t = f / n; e = f - t * n; } else { e = f / T; t = f - e * T; }
        const long line = line0 + t;
        double2 v = make_double2(0.0, 0.0);
        if (line < a.nlines) {
            const long p = line_base(a, line) + (long)e * a.stride;
            if (a.inA) v = make_double2(a.inA[p], a.inB[p]);
            else if (a.inR) v = make_double2(a.inR[p], 0.0);
            else v = a.inC[p];
        }
        // the inverse-only mode places frequency e where the forward pass would leave it
        lds[t * pitch + (a.mode == LINE_IFFT ? a.perm[e] : e)] = v;
    }
    // the skip flag after the line loads (tested first, its argument load and wait would
    // come before the others); the whole workgroup returns before any barrier
    if (a.skip && *a.skip) return;
    __syncthreads();

    int B = n;
    if (a.mode != LINE_IFFT) {
        for (int s = 0; s < a.nst; s++) {
            run_stage<false, PRIME>(a.radix[s], lds, pitch, T, n, B, a.tw);
            __syncthreads();
            B /= a.radix[s];
        }
    } else {
        B = 1;
    }

    if (a.mode != LINE_FFT) {
        if (a.mode == LINE_CONV || a.mode == LINE_CMUL) {
            for (int f = threadIdx.x; f < T * n; f += blockDim.x) {
                const int t = f / n, e = f - t * n;
                const long line = line0 + t;
                double g = 0.0;
                if (a.mode == LINE_CONV) g = a.gperm[e];
                else if (line < a.nlines) g = a.mul[line * n + e];
                double2 v = lds[t * pitch + e];
                lds[t * pitch + e] = make_double2(v.x * g, v.y * g);
            }
            __syncthreads();
        }
        for (int s = a.nst - 1; s >= 0; s--) {
            B *= a.radix[s];
            run_stage<true, PRIME>(a.radix[s], lds, pitch, T, n, B, a.tw);
            __syncthreads();
        }
        for (int f = threadIdx.x; f < T * n; f += blockDim.x) {
            int t, e;
            if (zaxis) { t = f / n; e = f - t * n; } else { e = f / T; t = f - e * T; }
            const long line = line0 + t;
            if (line >= a.nlines) continue;
            const long p = line_base(a, line) + (long)e * a.stride;
            double2 v = lds[t * pitch + e];
            if (a.mode != LINE_CONV) {
                if (a.outR) a.outR[p] = v.x;
                else a.outC[p] = v;
                continue;
            }
            double x = v.x, y = v.y;
            if (a.hA) {
                if (a.chain == CHAIN_TANGENT) {
                    x = x * a.hA[p] + a.auxA[p] * a.srcA[p];
                    y = y * a.hB[p] + a.auxB[p] * a.srcB[p];
                } else if (a.chain == CHAIN_ADJOINT) {
                    x += a.auxA[p] * a.srcA[p];
                    y += a.auxB[p] * a.srcB[p];
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
    } else {
        const int mk2 = a.m2 / 2 + 1;
        for (int f = threadIdx.x; f < T * n; f += blockDim.x) {
            int t, e;
            if (zaxis) { t = f / n; e = f - t * n; } else { e = f / T; t = f - e * T; }
            const long line = line0 + t;
            if (line >= a.nlines) continue;
            const double2 v = lds[t * pitch + a.perm[e]];
            if (a.Sacc) {
                // axis 0: line = K1*m2 + K2, element e = K0
                const int K1 = (int)(line / a.m2), K2 = (int)(line - (long)K1 * a.m2);
                if (K2 <= a.m2 / 2) a.Sacc[K2 + (long)mk2 * (K1 + (long)a.m1 * e)] += v.x * v.x + v.y * v.y;
            } else {
                const long p = line_base(a, line) + (long)e * a.stride;
                a.outC[p] = v;
            }
        }
    }
}

hipError_t launch_line(const LineArgs& a, hipStream_t st) {
    if (a.mode == LINE_CONV && a.fast) {
        const hipError_t e = launch_axis_conv(a, st);
        if (e != hipErrorNotSupported) return e;
    }
    const long nblk = (a.nlines + a.T - 1) / a.T;
    const size_t lds = sizeof(double2) * (size_t)a.T * a.pitch;
    bool prime = false;  // a stage of radix 11 or more
    for (int s = 0; s < a.nst; s++) prime = prime || a.radix[s] > 7;
    if (prime) hipLaunchKernelGGL(k_line<true>, dim3((unsigned)nblk), dim3(256), 2 * lds, st, a);
    else hipLaunchKernelGGL(k_line<false>, dim3((unsigned)nblk), dim3(256), lds, st, a);
    return hipGetLastError();
}

}  // namespace lfts
