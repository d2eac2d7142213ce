// Device kernels of the L-FTS Langevin step for gfx950 (FP64 throughout).
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

template <bool INV>
__device__ void run_stage(int r, double2* lds, int pitch, int T, int n, int B, const double2* tw) {
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
//  mode LINE_FFT : complex (or real) in -> forward FFT -> natural order complex out, or,
//                  on the last axis, |X|^2 accumulated into the half spectrum Sacc.
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
            if (a.mode == LINE_CONV) v = make_double2(a.inA[p], a.inB[p]);
            else if (a.inR) v = make_double2(a.inR[p], 0.0);
            else v = a.inC[p];
        }
        lds[t * pitch + e] = v;
    }
    __syncthreads();

    int B = n;
    for (int s = 0; s < a.nst; s++) {
        run_stage<false>(a.radix[s], lds, pitch, T, n, B, a.tw);
        __syncthreads();
        B /= a.radix[s];
    }

    if (a.mode == LINE_CONV) {
        for (int f = threadIdx.x; f < T * n; f += blockDim.x) {
            const int t = f / n, e = f - t * n;
            const double g = a.gperm[e];
            double2 v = lds[t * pitch + e];
            lds[t * pitch + e] = make_double2(v.x * g, v.y * g);
        }
        __syncthreads();
        for (int s = a.nst - 1; s >= 0; s--) {
            B *= a.radix[s];
            run_stage<true>(a.radix[s], lds, pitch, T, n, B, a.tw);
            __syncthreads();
        }
        for (int f = threadIdx.x; f < T * n; f += blockDim.x) {
            int t, e;
            if (zaxis) { t = f / n; e = f - t * n; } else { e = f / T; t = f - e * T; }
            const long line = line0 + t;
            if (line >= a.nlines) continue;
            const long p = line_base(a, line) + (long)e * a.stride;
            double2 v = lds[t * pitch + e];
            double x = v.x, y = v.y;
            if (a.hA) { x *= a.hA[p]; y *= a.hB[p]; }
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

// ------------------------------------------------------------------------------------
// Reductions: per-workgroup partials in a fixed order, then one workgroup sums them,
// so results are identical from run to run.
// ------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Sum of v over the workgroup (256 threads); valid in thread 0.
__device__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wv] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) s += red[i];
    return s;
}

// out[c] = (sum_b part[b*PSTRIDE + c]) / div, c < ncomp
__global__ __launch_bounds__(256) void k_reduce_final(const double* __restrict__ part, int nblk, int ncomp,
                                                      double div, double* __restrict__ out) {
    __shared__ double red[8];
    for (int c = 0; c < ncomp; c++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += part[(long)b * PSTRIDE + c];
        s = block_sum(s, red);
        if (threadIdx.x == 0) out[c] = s / div;
    }
}

// src/diblock.cc:64-72 (prepare_h): hA = exp(-(W-+w+)/N), hB = exp(-(-W-+w+)/N)
__global__ __launch_bounds__(256) void k_prepare_h(const double* __restrict__ w, double* __restrict__ h, long M, int N) {
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < M; r += (long)gridDim.x * blockDim.x) {
        const double wm = w[r], wp = w[r + M];
        h[r] = exp(-(wm + wp) / N);
        h[r + M] = exp(-(-wm + wp) / N);
    }
}

// partial sums of x[0..M)
__global__ __launch_bounds__(256) void k_sum_partial(const double* __restrict__ x, long M, double* __restrict__ part) {
    __shared__ double red[8];
    double s = 0.0;
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < M; r += (long)gridDim.x * blockDim.x) s += x[r];
    s = block_sum(s, red);
    if (threadIdx.x == 0) part[(long)blockIdx.x * PSTRIDE] = s;
}

__device__ __forceinline__ const double* pair_ptr(const double* h, const double* q, int i, long M) {
    return (i == 1) ? h : q + 2L * (i - 2) * M;
}

// src/diblock.cc:51-53,76-95: phiA = sum_{i<=NA} q_i q+_i, phiB = sum_{i>NA} q_i q+_i,
// normalised by N Q h, written as phi- = phiA - phiB, phi+ = phiA + phiB.
__global__ __launch_bounds__(256) void k_phi(const double* __restrict__ h, const double* __restrict__ q,
                                             double* __restrict__ w, const double* __restrict__ Qp, long M,
                                             int N, int NA) {
    const double Q = Qp[0];
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < M; r += (long)gridDim.x * blockDim.x) {
        double pA = 0.0, pB = 0.0;
        for (int i = 1; i <= NA; i++)
            pA += pair_ptr(h, q, i, M)[r] * pair_ptr(h, q, N + 1 - i, M)[M + r];
        for (int i = NA + 1; i <= N; i++)
            pB += pair_ptr(h, q, i, M)[r] * pair_ptr(h, q, N + 1 - i, M)[M + r];
        const double phA = pA / (N * Q * h[r]);
        const double phB = pB / (N * Q * h[r + M]);
        w[2 * M + r] = phA - phB;
        w[3 * M + r] = phA + phB;
    }
}

// src/anderson.cc:35-55: DD[0][m] = <Dh0, Dh_m>, Dh0 = phi+ - 1 (m = 0 uses Dh0 itself)
__global__ __launch_bounds__(256) void k_am_dots(const double* __restrict__ w, const double* __restrict__ Dh,
                                                 long M, int nh, double* __restrict__ part) {
    __shared__ double red[8];
    double acc[MAX_HIST + 1];
#pragma unroll
    for (int m = 0; m <= MAX_HIST; m++) acc[m] = 0.0;
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < M; r += (long)gridDim.x * blockDim.x) {
        const double d0 = w[3 * M + r] - 1.0;
        acc[0] += d0 * d0;
#pragma unroll
        for (int m = 1; m <= MAX_HIST; m++)
            if (m <= nh) acc[m] += d0 * Dh[(long)m * M + r];
    }
#pragma unroll
    for (int m = 0; m <= MAX_HIST; m++) {
        if (m > nh) break;
        const double s = block_sum(acc[m], red);
        if (threadIdx.x == 0) part[(long)blockIdx.x * PSTRIDE + m] = s;
    }
}

// src/anderson.cc:57-89: w+ += lambda Dh0 + sum_n C_n((wh_n + lambda Dh_n) - (wh0 + lambda Dh0));
// then (Dh0, wh0) become history slot `slot`.
__global__ __launch_bounds__(256) void k_am_update(double* __restrict__ w, double* __restrict__ Dh,
                                                   double* __restrict__ wh, long M, AmCoef c) {
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < M; r += (long)gridDim.x * blockDim.x) {
        const double d0 = w[3 * M + r] - 1.0;
        const double w0 = w[M + r];
        double wn = w0 + c.lambda * d0;
        const double base = w0 + c.lambda * d0;
        for (int n = 1; n <= c.nh; n++)
            wn += c.C[n - 1] * ((wh[(long)n * M + r] + c.lambda * Dh[(long)n * M + r]) - base);
        w[M + r] = wn;
        Dh[(long)c.slot * M + r] = d0;
        wh[(long)c.slot * M + r] = w0;
    }
}

// src/langevin.cc:31: W- += -(phi- + 2 W-/XbN) dt + 0.5 (eta_prev + eta_new)
// When keep_new is set, eta_new is copied over eta_prev in the same pass (the swap of
// src/langevin.cc:34-36 for noise that lives in a caller-owned buffer).
__global__ __launch_bounds__(256) void k_langevin(double* __restrict__ w, double* __restrict__ ep,
                                                  const double* __restrict__ en, long M, double XbN, double dt,
                                                  int keep_new) {
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < M; r += (long)gridDim.x * blockDim.x) {
        const double wm = w[r];
        const double a = ep[r], b = en[r];
        w[r] = wm + (-(w[2 * M + r] + 2 * wm / XbN) * dt + 0.5 * (a + b));
        if (keep_new) ep[r] = b;
    }
}

// src/lfts_simulation.cc:130-131: partials of sum w+ and sum W-^2
__global__ __launch_bounds__(256) void k_h_partial(const double* __restrict__ w, long M, double* __restrict__ part) {
    __shared__ double red[8];
    double sp = 0.0, s2 = 0.0;
    for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < M; r += (long)gridDim.x * blockDim.x) {
        sp += w[M + r];
        s2 += w[r] * w[r];
    }
    const double a = block_sum(sp, red);
    const double b = block_sum(s2, red);
    if (threadIdx.x == 0) {
        part[(long)blockIdx.x * PSTRIDE] = a;
        part[(long)blockIdx.x * PSTRIDE + 1] = b;
    }
}

// ------------------------------------------------------------------------------------
// launch wrappers (host side)
// ------------------------------------------------------------------------------------
hipError_t launch_line(const LineArgs& a, hipStream_t st) {
    const long nblk = (a.nlines + a.T - 1) / a.T;
    const size_t lds = sizeof(double2) * (size_t)a.T * a.pitch;
    hipLaunchKernelGGL(k_line, dim3((unsigned)nblk), dim3(256), lds, st, a);
    return hipGetLastError();
}

hipError_t launch_reduce_final(const double* part, int nblk, int ncomp, double div, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, st, part, nblk, ncomp, div, out);
    return hipGetLastError();
}

hipError_t launch_prepare_h(const double* w, double* h, long M, int N, int nblk, hipStream_t st) {
    hipLaunchKernelGGL(k_prepare_h, dim3(nblk), dim3(256), 0, st, w, h, M, N);
    return hipGetLastError();
}

hipError_t launch_sum_partial(const double* x, long M, double* part, int nblk, hipStream_t st) {
    hipLaunchKernelGGL(k_sum_partial, dim3(nblk), dim3(256), 0, st, x, M, part);
    return hipGetLastError();
}

hipError_t launch_phi(const double* h, const double* q, double* w, const double* Qp, long M, int N, int NA,
                      int nblk, hipStream_t st) {
    hipLaunchKernelGGL(k_phi, dim3(nblk), dim3(256), 0, st, h, q, w, Qp, M, N, NA);
    return hipGetLastError();
}

hipError_t launch_am_dots(const double* w, const double* Dh, long M, int nh, double* part, int nblk, hipStream_t st) {
    hipLaunchKernelGGL(k_am_dots, dim3(nblk), dim3(256), 0, st, w, Dh, M, nh, part);
    return hipGetLastError();
}

hipError_t launch_am_update(double* w, double* Dh, double* wh, long M, const AmCoef& c, int nblk, hipStream_t st) {
    hipLaunchKernelGGL(k_am_update, dim3(nblk), dim3(256), 0, st, w, Dh, wh, M, c);
    return hipGetLastError();
}

hipError_t launch_langevin(double* w, double* ep, const double* en, long M, double XbN, double dt, int keep_new,
                           int nblk, hipStream_t st) {
    hipLaunchKernelGGL(k_langevin, dim3(nblk), dim3(256), 0, st, w, ep, en, M, XbN, dt, keep_new);
    return hipGetLastError();
}

hipError_t launch_h_partial(const double* w, long M, double* part, int nblk, hipStream_t st) {
    hipLaunchKernelGGL(k_h_partial, dim3(nblk), dim3(256), 0, st, w, M, part);
    return hipGetLastError();
}

}  // namespace lfts
