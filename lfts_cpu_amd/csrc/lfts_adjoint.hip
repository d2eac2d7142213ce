// Derivatives of one propagator evaluation (src/diblock.cc:30-60) on the GPU.
//
// The chain of the reference is q_1 = hA, q_j = f_j C q_{j-1} and q+_N = hB,
// q+_j = p_j C q+_{j+1} (src/diblock.cc:40-43, src/step.cc:53-64), where f_j / p_j are
// the Boltzmann factors the two chains apply at bead j, and C is the convolution with
// g(k) (real, even, so C is symmetric). With e_X = dh_X / h_X:
//
//   tangent:  dq_j = f_j C dq_{j-1} + e_{f_j} q_j   (a plain chain plus a source term);
//   adjoint:  qbar_j = alpha q+_j + C(f_{j+1} qbar_{j+1})   (a chain run from the other
//             end: convolve, add the source, store, multiply by the next bead's factor).
//
// Both run through the same slab (or line) passes as the plain chain, stored in the
// reference's pair layout: tangent pair j = [dq_j | dq+_{N+1-j}], adjoint pair m =
// [qbar_{N+1-m} | q+bar_m]. The adjoint pass m uses the Boltzmann factors of forward
// pass N-1-m in the same field positions, so no transposed copies are made.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "lfts_gpu.h"
#include "lfts_internal.h"
#include "lfts_state.h"

namespace lfts {

namespace {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Sum of v over a 256-thread workgroup; valid in thread 0.
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

#define FOR_R(r, n) for (long r = blockIdx.x * (long)blockDim.x + threadIdx.x; r < (n); r += (long)gridDim.x * blockDim.x)

__device__ __forceinline__ const double* pair_at(const double* p1, const double* rest, long M, int j) {
    return (j == 1) ? p1 : rest + 2L * (j - 2) * M;
}

}  // namespace

// part[m][blockIdx.x] = sum over this block's points of x[r] * Y[m * ystride + r]
// (x or Y null: 1), for the eight components m = 8 blockIdx.y + 0..7 below ny.
__global__ __launch_bounds__(256) void k_multi_dot(const double* __restrict__ x, const double* __restrict__ Y,
                                                   long ystride, int ny, long n, double* __restrict__ part) {
    __shared__ double red[8];
    const int m0 = 8 * blockIdx.y;
    const int cnt = min(8, ny - m0);
    double acc[8];
#pragma unroll
    for (int m = 0; m < 8; m++) acc[m] = 0.0;
    FOR_R(r, n) {
        const double xv = x ? x[r] : 1.0;
#pragma unroll
        for (int m = 0; m < 8; m++)
            if (m < cnt) acc[m] += xv * (Y ? Y[(long)(m0 + m) * ystride + r] : 1.0);
    }
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const double s = block_sum(acc[m], red);
        if (threadIdx.x == 0 && m < cnt) part[(long)(m0 + m) * MAX_NB + blockIdx.x] = s;
    }
}

// out[c] = scale * sum_b part[c][b], b < nb, one workgroup per component, fixed order.
__global__ __launch_bounds__(256) void k_final(const double* __restrict__ part, int nb, double scale,
                                               double* __restrict__ out) {
    __shared__ double red[8];
    const int c = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) s += part[(long)c * MAX_NB + b];
    s = block_sum(s, red);
    if (threadIdx.x == 0) out[c] = s * scale;
}

// Tangent sources: e = [eA | eB] = [-(u + v)/N | -(v - u)/N] for (dW-, dw+) = (u, v),
// and the first tangent pair dh = [eA hA | eB hB] (src/diblock.cc:64-72, 40).
__global__ __launch_bounds__(256) void k_tan_prep(long M, int N, const double* __restrict__ h,
                                                  const double* __restrict__ du, const double* __restrict__ dv,
                                                  double* __restrict__ e, double* __restrict__ d1) {
    FOR_R(r, M) {
        const double u = du ? du[r] : 0.0, v = dv ? dv[r] : 0.0;
        const double eA = -(u + v) / N, eB = -(v - u) / N;
        e[r] = eA;
        e[r + M] = eB;
        d1[r] = eA * h[r];
        d1[r + M] = eB * h[r + M];
    }
}

struct TanPhiArgs {
    long M;
    int N, NA;
    const double* h;    // [hA | hB]
    const double* q;    // forward pairs 2..N (pair 1 = h)
    const double* d1;   // tangent pair 1
    const double* dq;   // tangent pairs 2..N
    const double* e;    // [eA | eB]
    const double* Q;    // device scalar Q
    const double* dQ;   // device scalar dQ
    double* dphi;       // out [dphi- | dphi+]
};

// d phi (src/diblock.cc:51-53, 85-95 differentiated): with S_X = sum q_i q+_i over the
// beads of block X, phi_X = S_X / (N Q h_X), so
// d phi_X = dS_X / (N Q h_X) - phi_X (dQ/Q + e_X).
__global__ __launch_bounds__(256) void k_tan_phi(const TanPhiArgs a) {
    const long M = a.M;
    const int N = a.N;
    const double Q = *a.Q, dQ = *a.dQ;
    FOR_R(r, M) {
        double SA = 0.0, SB = 0.0, dSA = 0.0, dSB = 0.0;
        for (int i = 1; i <= N; i++) {
            const double* P = pair_at(a.h, a.q, M, i);
            const double* Pd = pair_at(a.h, a.q, M, N + 1 - i);
            const double* D = pair_at(a.d1, a.dq, M, i);
            const double* Dd = pair_at(a.d1, a.dq, M, N + 1 - i);
            const double q = P[r], qd = Pd[M + r], dq = D[r], dqd = Dd[M + r];
            if (i <= a.NA) {
                SA += q * qd;
                dSA += dq * qd + q * dqd;
            } else {
                SB += q * qd;
                dSB += dq * qd + q * dqd;
            }
        }
        const double hA = a.h[r], hB = a.h[M + r];
        const double pA = SA / (N * Q * hA), pB = SB / (N * Q * hB);
        const double dpA = dSA / (N * Q * hA) - pA * (dQ / Q + a.e[r]);
        const double dpB = dSB / (N * Q * hB) - pB * (dQ / Q + a.e[M + r]);
        a.dphi[r] = dpA - dpB;
        a.dphi[M + r] = dpA + dpB;
    }
}

struct AdjArgs {
    long M;
    int N, NA, NB;
    const double* h;     // [hA | hB]
    const double* q;     // forward pairs 2..N
    const double* w;     // fields [W- | w+ | phi- | phi+] at the forward state
    const double* ubar;  // [u- | u+]
    const double* Q;     // device scalar Q
    const double* S;     // device scalar <u-, phi-> + <u+, phi+>
    double* alpha;       // [alphaA | alphaB]
    double* a1;          // adjoint pair 1 (stored values)
    double* b1;          // adjoint pair 1 times the next factors (chain input)
    const double* adq;   // adjoint pairs 2..N (final pass)
    double* wbar;        // out [dL/dW- | dL/dw+]
};

// Adjoint sources of L = <uA, phiA> + <uB, phiB> (uA = u+ + u-, uB = u+ - u-):
// dL/dS_X = alpha_X = u_X / (N Q h_X), dL/dQ = -S/Q; adjoint pair 1 =
// [alpha_{X(N)} q+_N + Qbar/M | alpha_{X(1)} q_1] (q_1 = hA, q+_N = hB), and its product
// with the factors of forward pass N-1.
__global__ __launch_bounds__(256) void k_adj_prep(const AdjArgs a) {
    const long M = a.M;
    const int N = a.N;
    const double Q = *a.Q;
    const double Qbar = -(*a.S) / Q;
    FOR_R(r, M) {
        const double hA = a.h[r], hB = a.h[M + r];
        const double uA = a.ubar[M + r] + a.ubar[r], uB = a.ubar[M + r] - a.ubar[r];
        const double aA = uA / (N * Q * hA), aB = uB / (N * Q * hB);
        a.alpha[r] = aA;
        a.alpha[M + r] = aB;
        const double s0 = ((N <= a.NA) ? aA : aB) * hB + Qbar / M;
        const double s1 = ((1 <= a.NA) ? aA : aB) * hA;
        a.a1[r] = s0;
        a.a1[M + r] = s1;
        a.b1[r] = ((N - 1 < a.NA) ? hA : hB) * s0;
        a.b1[M + r] = ((N - 1 < a.NB) ? hB : hA) * s1;
    }
}

// dL/dh_X h_X = sum over beads j with factor X of qbar_j q_j (q chain) and q+bar_j q+_j
// (q+ chain), minus u_X phi_X (the normalisation); dL/dw = -(1/N) sum_X (+-) h_X dL/dh_X.
__global__ __launch_bounds__(256) void k_adj_final(const AdjArgs a) {
    const long M = a.M;
    const int N = a.N, NA = a.NA;
    FOR_R(r, M) {
        const double uA = a.ubar[M + r] + a.ubar[r], uB = a.ubar[M + r] - a.ubar[r];
        const double phA = 0.5 * (a.w[3 * M + r] + a.w[2 * M + r]);
        const double phB = 0.5 * (a.w[3 * M + r] - a.w[2 * M + r]);
        double GA = -uA * phA, GB = -uB * phB;
        for (int m = 1; m <= N; m++) {
            const int j = N + 1 - m;
            const double* Am = pair_at(a.a1, a.adq, M, m);
            const double* Pj = pair_at(a.h, a.q, M, j);
            const double x = Am[r] * Pj[r];          // qbar_j q_j
            const double y = Am[M + r] * Pj[M + r];  // q+bar_m q+_m
            if (j == 1 || j <= NA) GA += x; else GB += x;
            if (m != N && m <= NA) GA += y; else GB += y;
        }
        a.wbar[r] = -(GA - GB) / N;
        a.wbar[M + r] = -(GA + GB) / N;
    }
}

// acc[k] += Re(X(k) conj Y(k)) for Z = F(x + i y): = Im(Z(k) Z(-k)) / 2.
__global__ __launch_bounds__(256) void k_stress_acc(const double2* __restrict__ Z, double* __restrict__ acc, int m0,
                                                    int m1, int m2) {
    const long M = (long)m0 * m1 * m2;
    FOR_R(r, M) {
        const int K2 = (int)(r % m2);
        const long t = r / m2;
        const int K1 = (int)(t % m1), K0 = (int)(t / m1);
        const long rn = ((long)((m0 - K0) % m0) * m1 + (m1 - K1) % m1) * m2 + (m2 - K2) % m2;
        const double2 a = Z[r], b = Z[rn];
        acc[r] += 0.5 * (a.x * b.y + a.y * b.x);
    }
}

struct StressArgs {
    const double* acc;
    const double* g[3];
    const double* kap[3];
    int m0, m1, m2;
    double* part;
};

// part[d][block] = sum_k g(k) kap_d(k_d) acc[k], g = g_x g_y g_z
__global__ __launch_bounds__(256) void k_stress_sum(const StressArgs a) {
    __shared__ double red[8];
    const long M = (long)a.m0 * a.m1 * a.m2;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    FOR_R(r, M) {
        const int K2 = (int)(r % a.m2);
        const long t = r / a.m2;
        const int K1 = (int)(t % a.m1), K0 = (int)(t / a.m1);
        const double v = a.g[0][K0] * a.g[1][K1] * a.g[2][K2] * a.acc[r];
        s0 += v * a.kap[0][K0];
        s1 += v * a.kap[1][K1];
        s2 += v * a.kap[2][K2];
    }
    const double t0 = block_sum(s0, red);
    const double t1 = block_sum(s1, red);
    const double t2 = block_sum(s2, red);
    if (threadIdx.x == 0) {
        a.part[blockIdx.x] = t0;
        a.part[MAX_NB + blockIdx.x] = t1;
        a.part[2 * MAX_NB + blockIdx.x] = t2;
    }
}

__global__ __launch_bounds__(256) void k_fill(double* __restrict__ x, long n, double v) {
    FOR_R(r, n) x[r] = v;
}

// y = a x + b y (y is not read when b = 0)
__global__ __launch_bounds__(256) void k_axpby(double* __restrict__ y, double a, const double* __restrict__ x, double b,
                                               long n) {
    if (b == 0.0) {
        FOR_R(r, n) y[r] = a * x[r];
    } else {
        FOR_R(r, n) y[r] = a * x[r] + b * y[r];
    }
}

// b = 1 - phi+ (minus the saddle-point residual of src/anderson.cc:35-36)
__global__ __launch_bounds__(256) void k_residual(const double* __restrict__ phip, double* __restrict__ b, long n) {
    FOR_R(r, n) b[r] = 1.0 - phip[r];
}

// y = s y + sum_i coef[i] V_i (coef in device memory), i < nv
__global__ __launch_bounds__(256) void k_basis_comb(double* __restrict__ y, double s, const double* __restrict__ V,
                                                    long stride, const double* __restrict__ coef, int nv,
                                                    double sign, long n) {
    FOR_R(r, n) {
        double acc = 0.0;
        for (int i = 0; i < nv; i++) acc += coef[i] * V[(long)i * stride + r];
        y[r] = s * y[r] + sign * acc;
    }
}

}  // namespace lfts

namespace lfts_abi {

using lfts::LineArgs;
using lfts::SlabIO;

namespace {

constexpr int RED_RES = 32 * lfts::MAX_NB;  // offset of the results in dred
constexpr int NRES = 128;                   // result slots: Krylov coefficients at 0/40/80,
constexpr int R_DQ = 120, R_S = 121, R_BOX = 122, R_NRM = 126;  // scalars of single calls

double* D1(lfts_ctx* c) { return c->dbuf; }
double* D2(lfts_ctx* c) { return c->dbuf + 2 * c->M; }
double* AUX(lfts_ctx* c) { return c->dbuf + 4 * c->M; }
double* IO(lfts_ctx* c) { return c->dbuf + 6 * c->M; }
double* RES(lfts_ctx* c) { return c->dred + RED_RES; }

int ensure_deriv(lfts_ctx* c) {
    const long M = c->M;
    if (!c->dq && !c->ck_stride) HIPCHK(c, hipMalloc(&c->dq, sizeof(double) * 2 * M * (size_t)(c->N - 1)));
    if (!c->dbuf) HIPCHK(c, hipMalloc(&c->dbuf, sizeof(double) * 10 * M));
    if (!c->dred) HIPCHK(c, hipMalloc(&c->dred, sizeof(double) * (RED_RES + NRES)));
    if (!c->hred) HIPCHK(c, hipHostMalloc(&c->hred, sizeof(double) * NRES, hipHostMallocDefault));
    return LFTS_OK;
}

int grid_for(long n) { return lfts::blocks_for(n); }

// res[off + m] = scale * sum_r x[r] Y_m[r] (m < ny), deterministic two-stage sum
int multi_dot(lfts_ctx* c, const double* x, const double* Y, long ystride, int ny, long n, double scale, int off) {
    hipStream_t st = c->self->st;
    const int nb = grid_for(n);
    hipLaunchKernelGGL(lfts::k_multi_dot, dim3(nb, (ny + 7) / 8), dim3(256), 0, st, x, Y, ystride, ny, n, c->dred);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(lfts::k_final, dim3(ny), dim3(256), 0, st, (const double*)c->dred, nb, scale, RES(c) + off);
    HIPCHK(c, hipGetLastError());
    return LFTS_OK;
}

}  // namespace

int dev_sum(lfts_ctx* c, const double* x, const double* y, long n, double scale, double* out) {
    if (int rc = ensure_deriv(c)) return rc;
    hipStream_t st = c->self->st;
    const int nb = grid_for(n);
    hipLaunchKernelGGL(lfts::k_multi_dot, dim3(nb, 1), dim3(256), 0, st, x, y, 0L, 1, n, c->dred);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(lfts::k_final, dim3(1), dim3(256), 0, st, (const double*)c->dred, nb, scale, out);
    HIPCHK(c, hipGetLastError());
    return LFTS_OK;
}

namespace {

// The forward evaluation at the current fields: pairs, Q (scal slot 0) and phi.
int forward(lfts_ctx* c) {
    lfts_batch* b = c->self;
    if (int rc = set_active(b, all_ids(b))) return rc;
    return calc_concs_async(b, true);
}

const double* Qdev(lfts_ctx* c) { return c->self->d_scal; }

// Factor pointer of field f at pass i of a chain of kind `mode` over base [A | B].
const double* factor(const lfts_ctx* c, int mode, int i, int f, const double* base) {
    const long M = c->M;
    bool A;
    if (mode == lfts::CHAIN_ADJOINT) {
        const int s = c->N - 1 - i;
        A = f ? !(s < c->NB) : (s < c->NA);
    } else {
        A = f ? !(i < c->NB) : (i < c->NA);
    }
    return A ? base : base + M;
}

// alpha of field f at adjoint pass i: field 0 sums bead N-i, field 1 bead i+1
const double* adj_alpha(const lfts_ctx* c, int i, int f, const double* base) {
    const int bead = f ? i + 1 : c->N - i;
    return (bead <= c->NA) ? base : base + c->M;
}

}  // namespace

// Passes i0 .. i1-1 of every chain in `chains`, interleaved pass by pass (a derived chain
// reads the forward pair its pass needs after the forward chain's pass wrote it).
int run_lockstep(lfts_ctx* c, int i0, int i1, ChainSpec* const* chains, int n) {
    if (i1 <= i0) return LFTS_OK;
    hipStream_t st = c->self->st;
    const long M = c->M;
    auto src = [&](const ChainSpec& ch, int i, int f) -> const double* {
        const double* P = ch.fwd(i + 1);
        if (ch.mode == lfts::CHAIN_ADJOINT) return f ? P : P + M;
        return f ? P + M : P;
    };
    auto auxp = [&](const ChainSpec& ch, int i, int f) -> const double* {
        return (ch.mode == lfts::CHAIN_ADJOINT) ? adj_alpha(c, i, f, ch.aux) : factor(c, ch.mode, i, f, ch.aux);
    };
    if (c->fused) {
        // the kernel shape the context's own plain chain was tuned to (every shape gives
        // the same bits)
        // the tuned variant of the plain chain, on the default factorisation (the derivative
        // chains have no other)
        const auto tv = c->self->tuned.find(std::make_tuple(c->d.m[0], 1, 0));
        const int variant = (tv != c->self->tuned.end()) ? tv->second % lfts::SLAB_VARIANTS : 0;
        for (int p = i0 - 1; p < i1; p++)
            for (int k = 0; k < n; k++) {
                const ChainSpec& ch = *chains[k];
                SlabIO io{};
                io.tin = (p == i0 - 1) ? ch.start : ((p & 1) ? ch.T0 : ch.T1);
                io.tout = (p & 1) ? ch.T1 : ch.T0;
                const bool fin = p >= i0;
                io.qout = fin ? ch.out(p + 1) : nullptr;
                if (fin)
                    for (int f = 0; f < 2; f++) {
                        io.h[f] = factor(c, ch.mode, p, f, c->h);
                        if (ch.mode != lfts::CHAIN_PLAIN) {
                            io.aux[f] = auxp(ch, p, f);
                            io.src[f] = src(ch, p, f);
                        }
                    }
                for (int d = 0; d < 3; d++) io.g[d] = c->ax[d].gnat;
                io.M = M;
                io.finish = fin;
                io.cont = p < i1 - 1;
                io.xslab = (p % 2 == 0);
                HIPCHK(c, lfts::launch_slab_io(c->d.m[0], ch.mode, io, c->ax[0].tw, st, variant));
            }
    } else {
        for (int p = i0; p < i1; p++)
            for (int k = 0; k < n; k++) {
                const ChainSpec& ch = *chains[k];
                const double* in;
                if (p == i0) in = ch.start;
                else in = (ch.mode == lfts::CHAIN_ADJOINT) ? ch.post : ch.out(p);
                double* out = ch.out(p + 1);
                double* tmp = ch.T0;
                LineArgs z = line_args(c, 2, lfts::LINE_CONV);
                z.inA = in; z.inB = in + M; z.outA = tmp; z.outB = tmp + M;
                HIPCHK(c, lfts::launch_line(z, st));
                LineArgs y = line_args(c, 1, lfts::LINE_CONV);
                y.inA = tmp; y.inB = tmp + M; y.outA = tmp; y.outB = tmp + M;
                HIPCHK(c, lfts::launch_line(y, st));
                LineArgs x = line_args(c, 0, lfts::LINE_CONV);
                x.inA = tmp; x.inB = tmp + M; x.outA = out; x.outB = out + M;
                x.hA = factor(c, ch.mode, p, 0, c->h);
                x.hB = factor(c, ch.mode, p, 1, c->h);
                x.chain = ch.mode;
                if (ch.mode != lfts::CHAIN_PLAIN) {
                    x.auxA = auxp(ch, p, 0);
                    x.auxB = auxp(ch, p, 1);
                    x.srcA = src(ch, p, 0);
                    x.srcB = src(ch, p, 1);
                }
                if (ch.mode == lfts::CHAIN_ADJOINT) {
                    x.out2A = ch.post;
                    x.out2B = ch.post + M;
                }
                HIPCHK(c, lfts::launch_line(x, st));
            }
    }
    return LFTS_OK;
}

namespace {

// A tangent or adjoint chain over all beads from `pin` (pair 1 for the tangent, the
// multiplied pair 1 for the adjoint), storing pairs 2..N in c->dq.
int run_chain(lfts_ctx* c, int mode, const double* pin) {
    const long M = c->M;
    ChainSpec ch;
    ch.mode = mode;
    ch.start = pin;
    ch.T0 = c->T0;
    ch.T1 = c->T1;
    ch.post = D2(c);
    ch.aux = AUX(c);
    ch.out = [c, M](int j) { return c->dq + 2L * (j - 2) * M; };
    ch.fwd = [c](int j) -> const double* { return pair_ptr(c, j); };
    ChainSpec* cs[1] = {&ch};
    return run_lockstep(c, 1, c->N, cs, 1);
}

}  // namespace

// Tangent of (phi-, phi+) at the forward state along (du, dv) (device pointers, may be
// null), into dphi (2M device doubles). `fresh`: run the forward evaluation first.
int jvp_dev(lfts_ctx* c, const double* du, const double* dv, double* dphi, bool fresh) {
    if (int rc = ensure_deriv(c)) return rc;
    if (fresh)
        if (int rc = forward(c)) return rc;
    hipStream_t st = c->self->st;
    const long M = c->M;
    const int nb = grid_for(M);
    hipLaunchKernelGGL(lfts::k_tan_prep, dim3(nb), dim3(256), 0, st, M, c->N, (const double*)c->h, du, dv, AUX(c),
                       D1(c));
    HIPCHK(c, hipGetLastError());
    if (c->ck_stride) {
        if (int rc = ckpt_derived(c, lfts::CHAIN_TANGENT, D1(c), D1(c))) return rc;
        if (int rc = multi_dot(c, ckpt_pair_N(c, true), nullptr, 0, 1, M, 1.0 / (double)M, R_DQ)) return rc;
        if (int rc = ckpt_tan_finish(c, AUX(c), RES(c) + R_DQ, dphi)) return rc;
        c->n_tan++;
        return LFTS_OK;
    }
    if (int rc = run_chain(c, lfts::CHAIN_TANGENT, D1(c))) return rc;
    const double* dqN = c->dq + 2L * (c->N - 2) * M;
    if (int rc = multi_dot(c, dqN, nullptr, 0, 1, M, 1.0 / (double)M, R_DQ)) return rc;
    lfts::TanPhiArgs a{};
    a.M = M;
    a.N = c->N;
    a.NA = c->NA;
    a.h = c->h;
    a.q = c->q;
    a.d1 = D1(c);
    a.dq = c->dq;
    a.e = AUX(c);
    a.Q = Qdev(c);
    a.dQ = RES(c) + R_DQ;
    a.dphi = dphi;
    hipLaunchKernelGGL(lfts::k_tan_phi, dim3(nb), dim3(256), 0, st, a);
    HIPCHK(c, hipGetLastError());
    c->n_tan++;
    c->n_cs += c->N - 1;
    return LFTS_OK;
}

// Gradient of <u-, phi-> + <u+, phi+> with respect to (W-, w+) at the forward state.
int vjp_dev(lfts_ctx* c, const double* ubar, double* wbar, bool fresh) {
    if (int rc = ensure_deriv(c)) return rc;
    if (fresh)
        if (int rc = forward(c)) return rc;
    hipStream_t st = c->self->st;
    const long M = c->M;
    const int nb = grid_for(M);
    if (int rc = multi_dot(c, ubar, c->w + 2 * M, 0, 1, 2 * M, 1.0, R_S)) return rc;
    lfts::AdjArgs a{};
    a.M = M;
    a.N = c->N;
    a.NA = c->NA;
    a.NB = c->NB;
    a.h = c->h;
    a.q = c->q;
    a.w = c->w;
    a.ubar = ubar;
    a.Q = Qdev(c);
    a.S = RES(c) + R_S;
    a.alpha = AUX(c);
    a.a1 = D1(c);
    a.b1 = D2(c);
    a.adq = c->dq;
    a.wbar = wbar;
    hipLaunchKernelGGL(lfts::k_adj_prep, dim3(nb), dim3(256), 0, st, a);
    HIPCHK(c, hipGetLastError());
    if (c->ck_stride) {
        if (int rc = ckpt_derived(c, lfts::CHAIN_ADJOINT, D1(c), D2(c))) return rc;
        if (int rc = ckpt_adj_finish(c, ubar, wbar)) return rc;
        c->n_adj++;
        return LFTS_OK;
    }
    if (int rc = run_chain(c, lfts::CHAIN_ADJOINT, D2(c))) return rc;
    hipLaunchKernelGGL(lfts::k_adj_final, dim3(nb), dim3(256), 0, st, a);
    HIPCHK(c, hipGetLastError());
    c->n_adj++;
    c->n_cs += c->N - 1;
    return LFTS_OK;
}

// ---- Newton-Krylov saddle-point solve --------------------------------------------------

namespace {

// Preconditioner multiplier: the inverse of d phi+/d w+ at w = 0 (every h = 1, Q = 1),
// which is diagonal in k. There both chains give dq_j = e sum_{t<j} G^t with
// G = exp(-K^2/(6N)) the mode's factor per step and e = -v/N, so d phi+ = lambda v with
// lambda(k) = -(1/N) ((2/N) sum_{j=1..N} sum_{t<j} G^t - 1), from -1 at k = 0 to about
// -1/N at large k. P = 1/lambda for k != 0 and 0 at k = 0 (w+ keeps its mean, as under
// Anderson mixing); 1/M undoes the unnormalised transforms.
int ensure_precond(lfts_ctx* c) {
    if (c->pmul) return LFTS_OK;
    const int* m = c->d.m;
    const long M = c->M;
    const int N = c->N;
    std::vector<double> G1[3];
    for (int d = 0; d < 3; d++) {
        G1[d].resize(m[d]);
        for (int k = 0; k < m[d]; k++) {
            const int f = (k <= m[d] / 2) ? k : k - m[d];
            const double K = 2.0 * M_PI * (double)f / c->d.L[d];
            G1[d][k] = std::exp(-K * K / (6.0 * N));
        }
    }
    std::vector<int> rad, freq(m[0]);
    factorize(m[0], rad);
    for (int k = 0; k < m[0]; k++) freq[dif_pos(k, rad, 0, m[0])] = k;
    std::vector<double> mul(M);
    for (int K1 = 0; K1 < m[1]; K1++)
        for (int K2 = 0; K2 < m[2]; K2++)
            for (int e = 0; e < m[0]; e++) {
                const int K0 = freq[e];
                double v = 0.0;
                if (K0 || K1 || K2) {
                    const double G = G1[0][K0] * G1[1][K1] * G1[2][K2];
                    double s = 0.0;
                    if (1.0 - G > 1e-2) {
                        // sum_{j=1..N} (1 - G^j)/(1 - G) = (N - G (1 - G^N)/(1 - G)) / (1 - G)
                        const double u = 1.0 - G;
                        s = (N - G * (1.0 - std::pow(G, N)) / u) / u;
                    } else {
                        double part = 0.0, gt = 1.0;
                        for (int j = 1; j <= N; j++) {
                            part += gt;
                            gt *= G;
                            s += part;
                        }
                    }
                    const double lam = -(1.0 / N) * ((2.0 / N) * s - 1.0);
                    v = 1.0 / (lam * (double)M);
                }
                mul[((long)K1 * m[2] + K2) * m[0] + e] = v;
            }
    HIPCHK(c, hipMalloc(&c->pmul, sizeof(double) * M));
    HIPCHK(c, hipMemcpy(c->pmul, mul.data(), sizeof(double) * M, hipMemcpyHostToDevice));
    return LFTS_OK;
}

// out = P in: z and y transforms, the x transform with the multiplier, inverse y and z
int apply_precond(lfts_ctx* c, const double* in, double* out) {
    hipStream_t st = c->self->st;
    if (!c->zk) HIPCHK(c, hipMalloc(&c->zk, sizeof(double2) * c->M));
    LineArgs z = line_args(c, 2, lfts::LINE_FFT);
    z.inR = in;
    z.outC = c->zk;
    HIPCHK(c, lfts::launch_line(z, st));
    LineArgs y = line_args(c, 1, lfts::LINE_FFT);
    y.inC = c->zk;
    y.outC = c->zk;
    HIPCHK(c, lfts::launch_line(y, st));
    LineArgs x = line_args(c, 0, lfts::LINE_CMUL);
    x.inC = c->zk;
    x.outC = c->zk;
    x.mul = c->pmul;
    HIPCHK(c, lfts::launch_line(x, st));
    LineArgs yi = line_args(c, 1, lfts::LINE_IFFT);
    yi.inC = c->zk;
    yi.outC = c->zk;
    HIPCHK(c, lfts::launch_line(yi, st));
    LineArgs zi = line_args(c, 2, lfts::LINE_IFFT);
    zi.inC = c->zk;
    zi.outR = out;
    HIPCHK(c, lfts::launch_line(zi, st));
    return LFTS_OK;
}

int launch_axpby(lfts_ctx* c, double* y, double a, const double* x, double b) {
    hipLaunchKernelGGL(lfts::k_axpby, dim3(grid_for(c->M)), dim3(256), 0, c->self->st, y, a, x, b, c->M);
    HIPCHK(c, hipGetLastError());
    return LFTS_OK;
}

int read_res(lfts_ctx* c, int off, int n) {
    HIPCHK(c, hipMemcpyAsync(c->hred + off, RES(c) + off, sizeof(double) * n, hipMemcpyDeviceToHost, c->self->st));
    return sync_batch(c->self);
}

struct NkWork {
    double* V;     // (m + 1) M basis
    double* DPHI;  // 2M tangent output
    double* Z;     // M
    double* X;     // M Newton step
    double* B;     // M right-hand side 1 - phi+
};

// Forward evaluation at the current fields, B = 1 - phi+ and err = sqrt(sum B^2 / M)
// (the residual of src/anderson.cc:35-40).
int eval_residual(lfts_ctx* c, const NkWork& W, double* err, int* npe) {
    const long M = c->M;
    if (int rc = forward(c)) return rc;
    (*npe)++;
    hipLaunchKernelGGL(lfts::k_residual, dim3(grid_for(M)), dim3(256), 0, c->self->st, (const double*)(c->w + 3 * M),
                       W.B, M);
    HIPCHK(c, hipGetLastError());
    if (int rc = multi_dot(c, W.B, W.B, 0, 1, M, 1.0, R_NRM)) return rc;
    if (int rc = read_res(c, R_NRM, 1)) return rc;
    *err = std::sqrt(c->hred[R_NRM] / (double)M);
    return std::isfinite(*err) ? LFTS_OK : fail(c, LFTS_ENUM, "saddle residual is not finite");
}

// Restarted GMRES(m) with right preconditioning for J P y = B, X = P y, relative
// tolerance rtol, at most maxit tangent evaluations. Classical Gram-Schmidt, twice.
int gmres(lfts_ctx* c, const NkWork& W, int m, int maxit, double rtol, double beta0, bool prec, int* njvp) {
    hipStream_t st = c->self->st;
    const long M = c->M;
    const int nb = grid_for(M);
    constexpr int O1 = 0, O2 = 40, ON = 80;
    hipLaunchKernelGGL(lfts::k_fill, dim3(nb), dim3(256), 0, st, W.X, M, 0.0);
    HIPCHK(c, hipGetLastError());
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
    auto Hij = [&](int i, int j) -> double& { return H[(size_t)i * m + j]; };
    int total = 0;
    bool first = true;
    while (total < maxit) {
        double beta;
        if (first) {
            if (int rc = launch_axpby(c, W.V, 1.0, W.B, 0.0)) return rc;
            beta = beta0;
        } else {
            if (int rc = jvp_dev(c, nullptr, W.X, W.DPHI, false)) return rc;
            (*njvp)++;
            if (int rc = launch_axpby(c, W.V, 1.0, W.B, 0.0)) return rc;
            if (int rc = launch_axpby(c, W.V, -1.0, W.DPHI + M, 1.0)) return rc;
            if (int rc = multi_dot(c, W.V, W.V, 0, 1, M, 1.0, R_NRM)) return rc;
            if (int rc = read_res(c, R_NRM, 1)) return rc;
            beta = std::sqrt(c->hred[R_NRM]);
        }
        first = false;
        if (!(beta > rtol * beta0)) break;
        if (int rc = launch_axpby(c, W.V, 1.0 / beta, W.V, 0.0)) return rc;
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int jj = 0;
        bool done = false;
        for (int j = 0; j < m && total < maxit; j++) {
            double* Vj = W.V + (long)j * M;
            const double* in = Vj;
            if (prec) {
                if (int rc = apply_precond(c, Vj, W.Z)) return rc;
                in = W.Z;
            }
            if (int rc = jvp_dev(c, nullptr, in, W.DPHI, false)) return rc;
            (*njvp)++;
            total++;
            double* w = W.DPHI + M;
            if (int rc = multi_dot(c, w, W.V, M, j + 1, M, 1.0, O1)) return rc;
            hipLaunchKernelGGL(lfts::k_basis_comb, dim3(nb), dim3(256), 0, st, w, 1.0, (const double*)W.V, M,
                               (const double*)(RES(c) + O1), j + 1, -1.0, M);
            HIPCHK(c, hipGetLastError());
            if (int rc = multi_dot(c, w, W.V, M, j + 1, M, 1.0, O2)) return rc;
            hipLaunchKernelGGL(lfts::k_basis_comb, dim3(nb), dim3(256), 0, st, w, 1.0, (const double*)W.V, M,
                               (const double*)(RES(c) + O2), j + 1, -1.0, M);
            HIPCHK(c, hipGetLastError());
            if (int rc = multi_dot(c, w, w, 0, 1, M, 1.0, ON)) return rc;
            if (int rc = read_res(c, 0, ON + 1)) return rc;
            for (int i = 0; i <= j; i++) Hij(i, j) = c->hred[O1 + i] + c->hred[O2 + i];
            const double hn = std::sqrt(c->hred[ON]);
            if (!std::isfinite(hn)) return fail(c, LFTS_ENUM, "GMRES: non-finite Krylov vector");
            if (hn > 0.0)
                if (int rc = launch_axpby(c, W.V + (long)(j + 1) * M, 1.0 / hn, w, 0.0)) return rc;
            for (int i = 0; i < j; i++) {
                const double t = cs[i] * Hij(i, j) + sn[i] * Hij(i + 1, j);
                Hij(i + 1, j) = -sn[i] * Hij(i, j) + cs[i] * Hij(i + 1, j);
                Hij(i, j) = t;
            }
            const double den = std::hypot(Hij(j, j), hn);
            cs[j] = Hij(j, j) / den;
            sn[j] = hn / den;
            Hij(j, j) = den;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            jj = j + 1;
            if (std::fabs(g[j + 1]) <= rtol * beta0 || hn == 0.0) {
                done = true;
                break;
            }
        }
        if (jj == 0) break;
        for (int i = jj - 1; i >= 0; i--) {
            double s = g[i];
            for (int k = i + 1; k < jj; k++) s -= Hij(i, k) * y[k];
            y[i] = s / Hij(i, i);
        }
        // X += P (V y)
        for (int i = 0; i < jj; i++) c->hred[O1 + i] = y[i];
        HIPCHK(c, hipMemcpyAsync(RES(c) + O1, c->hred + O1, sizeof(double) * jj, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(lfts::k_basis_comb, dim3(nb), dim3(256), 0, st, W.DPHI, 0.0, (const double*)W.V, M,
                           (const double*)(RES(c) + O1), jj, 1.0, M);
        HIPCHK(c, hipGetLastError());
        if (prec) {
            if (int rc = apply_precond(c, W.DPHI, W.DPHI + M)) return rc;
            if (int rc = launch_axpby(c, W.X, 1.0, W.DPHI + M, 1.0)) return rc;
        } else {
            if (int rc = launch_axpby(c, W.X, 1.0, W.DPHI, 1.0)) return rc;
        }
        if (int rc = sync_batch(c->self)) return rc;  // hred is rewritten next cycle
        if (done) break;
    }
    return LFTS_OK;
}

}  // namespace

// Inexact Newton on F(w+) = phi+(w+) - 1 at fixed W-: J dx = -F by GMRES on tangent
// evaluations, forcing term eta_k = min(eta, 0.9 (err_k/err_{k-1})^2) (Eisenstat and
// Walker's second choice), backtracking on the residual norm.
int saddle_nk(lfts_ctx* c, const lfts_nk_opts& o, lfts_nk_info* info) {
    const long M = c->M;
    const int m = o.restart;
    if (int rc = ensure_deriv(c)) return rc;
    if (o.precondition)
        if (int rc = ensure_precond(c)) return rc;
    if (c->kry_m < m) {
        (void)hipFree(c->kry);
        c->kry = nullptr;
        HIPCHK(c, hipMalloc(&c->kry, sizeof(double) * (size_t)(m + 6) * M));
        c->kry_m = m;
    }
    NkWork W;
    W.V = c->kry;
    W.DPHI = W.V + (long)(m + 1) * M;
    W.Z = W.DPHI + 2 * M;
    W.X = W.Z + M;
    W.B = W.X + M;
    double* wp = c->w + M;
    int npe = 0, njvp = 0, newton = 0;
    double err = 0.0;
    if (int rc = eval_residual(c, W, &err, &npe)) return rc;
    double err_prev = err, eta = o.eta;
    while (newton < o.maxNewton && err > o.tol) {
        if (newton > 0) eta = std::min(o.eta, std::max(0.9 * (err / err_prev) * (err / err_prev), 0.5 * o.tol / err));
        const double beta0 = err * std::sqrt((double)M);
        if (int rc = gmres(c, W, m, o.maxInner, eta, beta0, o.precondition != 0, &njvp)) return rc;
        if (int rc = launch_axpby(c, wp, 1.0, W.X, 1.0)) return rc;
        double lam = 1.0, err_new = 0.0;
        if (int rc = eval_residual(c, W, &err_new, &npe)) return rc;
        for (int bt = 0; bt < 6 && !(err_new <= (1.0 - 1e-4 * lam) * err); bt++) {
            lam *= 0.5;
            if (int rc = launch_axpby(c, wp, -lam, W.X, 1.0)) return rc;
            if (int rc = eval_residual(c, W, &err_new, &npe)) return rc;
        }
        err_prev = err;
        err = err_new;
        newton++;
    }
    if (info) {
        info->newton = newton;
        info->krylov = njvp;
        info->prop_evals = npe;
        info->err = err;
    }
    return LFTS_OK;
}

}  // namespace lfts_abi

using namespace lfts_abi;

extern "C" {

int lfts_saddle_nk(lfts_ctx* c, const lfts_nk_opts* opts, lfts_nk_info* info) {
    if (!c || !opts) return fail(c, LFTS_EINVAL, "saddle_nk: null argument");
    if (opts->restart < 1 || opts->restart > 30 || opts->maxInner < 1 || opts->maxNewton < 0 || !(opts->eta > 0.0))
        return fail(c, LFTS_EINVAL, "saddle_nk: restart must be 1..30, maxInner >= 1, eta > 0");
    return saddle_nk(c, *opts, info);
}

// A Langevin step with the Newton-Krylov solve in place of anderson::mix: step_wm, the
// solve, and lnQ of the final fields, which the solve's last propagator evaluation
// already computed (the separate calc_concs of src/lfts_simulation.cc:76 is not needed).
int lfts_langevin_full_step_nk_dev(lfts_ctx* c, const double* eta_new_dev, double XbN, double dt,
                                   const lfts_nk_opts* opts, double* lnQ, lfts_nk_info* info) {
    if (!c || !eta_new_dev || !opts) return fail(c, LFTS_EINVAL, "full_step_nk_dev: null argument");
    if (int rc = lfts_langevin_step_dev(c, eta_new_dev, XbN, dt)) return rc;
    if (int rc = lfts_saddle_nk(c, opts, info)) return rc;
    HIPCHK(c, hipMemcpyAsync(c->hred + R_NRM + 1, Qdev(c), sizeof(double), hipMemcpyDeviceToHost, c->self->st));
    HIPCHK(c, hipStreamSynchronize(c->self->st));
    const double Q = c->hred[R_NRM + 1];
    if (lnQ) *lnQ = std::log(Q);
    return (std::isfinite(Q) && Q > 0) ? LFTS_OK : fail(c, LFTS_ENUM, "partition function Q is not finite and positive");
}

int lfts_jvp_dev(lfts_ctx* c, const double* dWm, const double* dwp, double* dphi) {
    if (!c || !dphi) return fail(c, LFTS_EINVAL, "jvp: null argument");
    if (int rc = jvp_dev(c, dWm, dwp, dphi, true)) return rc;
    HIPCHK(c, hipStreamSynchronize(c->self->st));
    return LFTS_OK;
}

int lfts_vjp_dev(lfts_ctx* c, const double* ubar, double* wbar) {
    if (!c || !ubar || !wbar) return fail(c, LFTS_EINVAL, "vjp: null argument");
    if (int rc = vjp_dev(c, ubar, wbar, true)) return rc;
    HIPCHK(c, hipStreamSynchronize(c->self->st));
    return LFTS_OK;
}

int lfts_jvp(lfts_ctx* c, const double* dWm, const double* dwp, double* dphi) {
    if (!c || !dphi) return fail(c, LFTS_EINVAL, "jvp: null argument");
    if (int rc = ensure_deriv(c)) return rc;
    hipStream_t st = c->self->st;
    const long M = c->M;
    double* io = IO(c);  // [u | v | dphi (2M)]
    if (dWm) HIPCHK(c, hipMemcpyAsync(io, dWm, sizeof(double) * M, hipMemcpyHostToDevice, st));
    if (dwp) HIPCHK(c, hipMemcpyAsync(io + M, dwp, sizeof(double) * M, hipMemcpyHostToDevice, st));
    if (int rc = jvp_dev(c, dWm ? io : nullptr, dwp ? io + M : nullptr, io + 2 * M, true)) return rc;
    HIPCHK(c, hipMemcpyAsync(dphi, io + 2 * M, sizeof(double) * 2 * M, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    return LFTS_OK;
}

int lfts_vjp(lfts_ctx* c, const double* ubar, double* wbar) {
    if (!c || !ubar || !wbar) return fail(c, LFTS_EINVAL, "vjp: null argument");
    if (int rc = ensure_deriv(c)) return rc;
    hipStream_t st = c->self->st;
    const long M = c->M;
    double* io = IO(c);  // [ubar (2M) | wbar (2M)]
    HIPCHK(c, hipMemcpyAsync(io, ubar, sizeof(double) * 2 * M, hipMemcpyHostToDevice, st));
    if (int rc = vjp_dev(c, io, io + 2 * M, true)) return rc;
    HIPCHK(c, hipMemcpyAsync(wbar, io + 2 * M, sizeof(double) * 2 * M, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    return LFTS_OK;
}

// dH/dL_d = -d lnQ/dL_d at fixed fields (H of src/lfts_simulation.cc:122-135 depends on
// the box only through g(k), src/step.cc:84-85): dQ/dL_d = (1/M) sum_{i<N} sum_k g(k)
// kap_d(k) Re(X_i conj Y_{i+1}), X_i = F q_i, Y_i = F q+_i, from one complex transform of
// q_i + i q+_{i+1} per bead.
int lfts_box_derivative(lfts_ctx* c, double* dHdL, double* lnQ) {
    if (!c || !dHdL) return fail(c, LFTS_EINVAL, "box_derivative: null argument");
    if (c->NA < 1 || c->NA > c->N - 1)
        return fail(c, LFTS_EINVAL, "box_derivative: needs 1 <= NA <= N-1 (the q+ chain is then the q chain's adjoint)");
    if (int rc = ensure_deriv(c)) return rc;
    hipStream_t st = c->self->st;
    const long M = c->M;
    if (!c->zk) HIPCHK(c, hipMalloc(&c->zk, sizeof(double2) * M));
    if (!c->acc) HIPCHK(c, hipMalloc(&c->acc, sizeof(double) * M));
    if (int rc = forward(c)) return rc;
    const int nb = grid_for(M);
    hipLaunchKernelGGL(lfts::k_fill, dim3(nb), dim3(256), 0, st, c->acc, M, 0.0);
    HIPCHK(c, hipGetLastError());
    // bead i: one complex transform of q_i + i q+_{i+1}, then the |k|-weighted products
    auto bead = [&](int, const double* qi, const double* qd) -> int {
        LineArgs z = line_args(c, 2, lfts::LINE_FFT);
        z.inA = qi;
        z.inB = qd;
        z.outC = c->zk;
        HIPCHK(c, lfts::launch_line(z, st));
        LineArgs y = line_args(c, 1, lfts::LINE_FFT);
        y.inC = c->zk;
        y.outC = c->zk;
        HIPCHK(c, lfts::launch_line(y, st));
        LineArgs x = line_args(c, 0, lfts::LINE_FFT);
        x.inC = c->zk;
        x.outC = c->zk;
        HIPCHK(c, lfts::launch_line(x, st));
        hipLaunchKernelGGL(lfts::k_stress_acc, dim3(nb), dim3(256), 0, st, (const double2*)c->zk, c->acc,
                           c->d.m[0], c->d.m[1], c->d.m[2]);
        HIPCHK(c, hipGetLastError());
        return LFTS_OK;
    };
    if (c->ck_stride) {  // pairs recomputed segment by segment from the checkpoints
        if (int rc = ckpt_for_bead_pairs(c, bead)) return rc;
    } else {
        for (int i = 1; i < c->N; i++)
            if (int rc = bead(i, pair_ptr(c, i), pair_ptr(c, c->N - i) + M)) return rc;  // q+_{i+1}
    }
    lfts::StressArgs a{};
    a.acc = c->acc;
    for (int d = 0; d < 3; d++) {
        a.g[d] = c->ax[d].gnat;
        a.kap[d] = c->ax[d].kap;
    }
    a.m0 = c->d.m[0];
    a.m1 = c->d.m[1];
    a.m2 = c->d.m[2];
    a.part = c->dred;
    hipLaunchKernelGGL(lfts::k_stress_sum, dim3(nb), dim3(256), 0, st, a);
    HIPCHK(c, hipGetLastError());
    hipLaunchKernelGGL(lfts::k_final, dim3(3), dim3(256), 0, st, (const double*)c->dred, nb, 1.0 / (double)M,
                       RES(c) + R_BOX);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(c->hred, RES(c) + R_BOX, sizeof(double) * 3, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipMemcpyAsync(c->hred + 3, Qdev(c), sizeof(double), hipMemcpyDeviceToHost, st));
    if (int rc = sync_batch(c->self)) return rc;
    const double Q = c->hred[3];
    for (int d = 0; d < 3; d++) dHdL[d] = -c->hred[d] / Q;
    if (lnQ) *lnQ = std::log(Q);
    return std::isfinite(Q) && Q > 0 ? LFTS_OK : fail(c, LFTS_ENUM, "partition function Q is not finite and positive");
}

}  // extern "C"
