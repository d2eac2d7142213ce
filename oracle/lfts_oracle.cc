// =====================================================================================
// ORACLE — test infrastructure only.
//
// A CPU restatement of the arithmetic of tmbeardsley/lfts_cpu's Langevin step, written
// from the reference's behaviour (cited file:line below), used as the parity checker for
// the HIP path. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load this library, and only as the checker or as the timed CPU baseline. The product
// (lfts_cpu_amd/) never links, loads or calls it.
//
// The reference itself cannot be built here: it needs FFTW3 and GSL, which are absent, so
// there is no oracle/_ref build. In their place this file carries:
//   * its own mixed-radix complex FFT (Stockham autosort, radices 4/2/3/5/any prime),
//     standing in for FFTW's r2c/c2r plans (src/step.cc:15-16, src/strFunc.cc:15);
//   * its own LU with partial pivoting, standing in for GSL's gsl_linalg_LU_decomp /
//     gsl_linalg_LU_solve (src/anderson.cc:118-126).
// The batched-2 real transforms of a contour step are evaluated with the two-for-one
// identity: z = q + i*q_dagger is transformed once, multiplied by the real, even kernel
// g(k) and transformed back; Re/Im are then exactly the two real convolutions that
// r2c -> *g -> c2r gives (src/step.cc:35-65). Reductions keep the reference's serial
// summation order.
//
// Pinning (what checks each function; tests/test_oracle_*.py):
//   fft            brute-force O(n^2) DFT, numpy.fft, Parseval, plane waves
//   g table        closed form on tiny meshes, Nyquist convention
//   calc_concs     closed forms (w=0; uniform w+; uniform W-), the invariant that
//                  (1/M) sum_r q_i q+_i / h is the same for every bead i, mean(phi+)=1,
//                  an independent numpy restatement that uses numpy.fft.rfftn/irfftn
//   anderson       every iteration's update against its least-squares statement
//                  (numpy.linalg.lstsq over the previous nh iterations' histories,
//                  which the traced slot writes must hold), the closed forms of the
//                  first two iterations, the lambda/nh/slot schedules across a ring
//                  wrap, the loop's stop rule; each of six planted misreadings fails
//                  one of these; converged residual, fixed point; LU vs scipy
//   langevin       closed-form update; the noise stream against the C++ standard's
//                  mt19937_64 known answer and a Python restatement of the polar method
//   strFunc        single cosine mode (|w_k|^2 = (A M/2)^2), bins vs numpy, weights;
//                  save: shell averages of numpy's full-spectrum |FFT|^2 over exact
//                  |k| shells (even mz), every empty shell -0.5/chi_b, the odd-mz
//                  weight quirk; each planted misreading of save fails
//   getH           closed form for uniform fields
//   jvp, vjp       central finite differences of calc_concs; <u, J v> = <J^T u, v>
//   dlnQ/dL        central finite differences of calc_concs over the box edges
//   z_inf          mpmath at 50 digits
// =====================================================================================
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <memory>
#include <numeric>
#include <random>
#include <string>
#include <vector>

typedef std::complex<double> cplx;

// -------------------------------------------------------------------------------------
// FFT: unnormalised complex DFT, X[k] = sum_n x[n] exp(-2 pi i n k / n), any length.
// Stockham autosort: after the stage that completes length-L sub-DFTs, y[j*(n/L)+k]
// holds the length-L DFT (index j) of the stride-(n/L) subsequence starting at k.
// -------------------------------------------------------------------------------------
struct fft_plan1d {
    int n = 1;
    std::vector<int> radix;
    std::vector<cplx> w;  // w[k] = exp(-2 pi i k / n)
};

static void make_plan(fft_plan1d& P, int n) {
    P.n = n;
    P.radix.clear();
    int r = n;
    while (r % 4 == 0) { P.radix.push_back(4); r /= 4; }
    while (r % 2 == 0) { P.radix.push_back(2); r /= 2; }
    for (int p = 3; p * p <= r; p += 2)
        while (r % p == 0) { P.radix.push_back(p); r /= p; }
    if (r > 1) P.radix.push_back(r);
    P.w.resize(n);
    for (int k = 0; k < n; k++) {
        double a = -2.0 * M_PI * (double)k / (double)n;
        P.w[k] = cplx(std::cos(a), std::sin(a));
    }
}

// forward transform of x (length n) in place, using scratch y (length n)
static void fft_forward(const fft_plan1d& P, cplx* x, cplx* y) {
    const int n = P.n;
    int L = 1;
    cplx* src = x;
    cplx* dst = y;
    cplx a[64], b[64];
    std::vector<cplx> abig, bbig;
    for (int p : P.radix) {
        cplx* A = a;
        cplx* B = b;
        if (p > 64) { abig.resize(p); bbig.resize(p); A = abig.data(); B = bbig.data(); }
        const int mp = n / (L * p);
        for (int j = 0; j < L; j++) {
            for (int k = 0; k < mp; k++) {
                for (int q = 0; q < p; q++) {
                    cplx v = src[j * p * mp + k + q * mp];
                    long tw = ((long)q * j * mp) % n;
                    A[q] = (tw == 0) ? v : v * P.w[tw];
                }
                for (int s = 0; s < p; s++) {
                    cplx acc = A[0];
                    for (int q = 1; q < p; q++) {
                        long e = ((long)q * s) % p;
                        acc += (e == 0) ? A[q] : A[q] * P.w[e * (n / p)];
                    }
                    B[s] = acc;
                }
                for (int s = 0; s < p; s++) dst[(j + s * L) * mp + k] = B[s];
            }
        }
        std::swap(src, dst);
        L *= p;
    }
    if (src != x) std::memcpy(x, src, sizeof(cplx) * n);
}

static void fft_1d(const fft_plan1d& P, cplx* x, cplx* y, int sign) {
    if (sign < 0) { fft_forward(P, x, y); return; }
    for (int i = 0; i < P.n; i++) x[i] = std::conj(x[i]);
    fft_forward(P, x, y);
    for (int i = 0; i < P.n; i++) x[i] = std::conj(x[i]);
}

// 3D transform of a row-major m[0] x m[1] x m[2] complex array (last index fastest)
// Test hook: the same arithmetic in another rounding order, to measure how far two
// correct evaluations of one run drift apart (never used for parity or timing).
// bit 0: the 3D transforms run along x, y, z instead of z, y, x; bit 1: the reductions
// (Q, the Anderson residual and Gram row, the sums of getH) run from the last point to
// the first.
static int g_order = 0;

// sum of a[0..n) in the reference's serial order, or from the last element (g_order)
static double osum(const double* a, long n) {
    double s = 0.0;
    if (g_order & 2)
        for (long r = n - 1; r >= 0; r--) s += a[r];
    else
        for (long r = 0; r < n; r++) s += a[r];
    return s;
}

static void fft_3d(const fft_plan1d* P, const int* m, cplx* a, int sign) {
    const long M = (long)m[0] * m[1] * m[2];
    for (int step = 0; step < 3; step++) {
        const int ax = (g_order & 1) ? step : 2 - step;
        const int n = m[ax];
        if (n == 1) continue;
        long stride = 1;
        for (int d = ax + 1; d < 3; d++) stride *= m[d];
        const long nlines = M / n;
        #pragma omp parallel
        {
            std::vector<cplx> buf(n), scr(n);
            #pragma omp for schedule(static)
            for (long l = 0; l < nlines; l++) {
                long outer = l / stride, inner = l % stride;
                long base = outer * stride * n + inner;
                for (int e = 0; e < n; e++) buf[e] = a[base + e * stride];
                fft_1d(P[ax], buf.data(), scr.data(), sign);
                for (int e = 0; e < n; e++) a[base + e * stride] = buf[e];
            }
        }
    }
}

// -------------------------------------------------------------------------------------
// step: g table and one contour step (src/step.cc:10-89)
// -------------------------------------------------------------------------------------
struct orc_step {
    int NA, NB, N, M, Mk;
    int m[3];
    double L[3];
    fft_plan1d P[3];
    std::vector<double> g;     // half spectrum, reference index k2 + (m2/2+1)(K1 + m1 K0)
    std::vector<double> gfull; // full spectrum, row-major (K0,K1,K2)
    std::vector<cplx> z;
};

// src/step.cc:69-89 (update_g_lookup)
static void step_g_lookup(orc_step& s) {
    const int* m = s.m;
    const double* L = s.L;
    int N = s.NA + s.NB;
    for (int k0 = -(m[0] - 1) / 2; k0 <= m[0] / 2; k0++) {
        int K0 = (k0 < 0) ? (k0 + m[0]) : k0;
        double kx_sq = k0 * k0 / (L[0] * L[0]);
        for (int k1 = -(m[1] - 1) / 2; k1 <= m[1] / 2; k1++) {
            int K1 = (k1 < 0) ? (k1 + m[1]) : k1;
            double ky_sq = k1 * k1 / (L[1] * L[1]);
            for (int k2 = 0; k2 <= m[2] / 2; k2++) {
                double kz_sq = k2 * k2 / (L[2] * L[2]);
                int k = k2 + (m[2] / 2 + 1) * (K1 + m[1] * K0);
                double K = 2 * M_PI * pow(kx_sq + ky_sq + kz_sq, 0.5);
                s.g[k] = exp(-K * K / (6.0 * N)) / s.M;
            }
        }
    }
    // Full-spectrum view used by the complex two-for-one transform: the value at
    // K2 > m2/2 is the value at the conjugate point, which depends only on squares.
    const int mk2 = m[2] / 2 + 1;
    for (int K0 = 0; K0 < m[0]; K0++)
        for (int K1 = 0; K1 < m[1]; K1++)
            for (int K2 = 0; K2 < m[2]; K2++) {
                int k2 = (K2 <= m[2] / 2) ? K2 : m[2] - K2;
                s.gfull[((long)K0 * m[1] + K1) * m[2] + K2] = s.g[k2 + mk2 * (K1 + m[1] * K0)];
            }
}

static void step_init(orc_step& s, int NA, int NB, const int* m, const double* L) {
    s.NA = NA; s.NB = NB; s.N = NA + NB;
    for (int d = 0; d < 3; d++) { s.m[d] = m[d]; s.L[d] = L[d]; make_plan(s.P[d], m[d]); }
    s.M = m[0] * m[1] * m[2];
    s.Mk = m[0] * m[1] * (m[2] / 2 + 1);
    s.g.assign(s.Mk, 0.0);
    s.gfull.assign(s.M, 0.0);
    s.z.assign(s.M, cplx(0, 0));
    step_g_lookup(s);
}

// src/step.cc:35-65 (step::fwd): q_out = h_sel * c2r(g * r2c(q_in)) for both halves
static void step_fwd(orc_step& s, const double* q_in, double* q_out, const double* h, int i) {
    const int M = s.M;
    for (int r = 0; r < M; r++) s.z[r] = cplx(q_in[r], q_in[r + M]);
    fft_3d(s.P, s.m, s.z.data(), -1);
    for (int r = 0; r < M; r++) s.z[r] *= s.gfull[r];
    fft_3d(s.P, s.m, s.z.data(), +1);
    const double* hq = (i < s.NA) ? h : h + M;
    const double* hqd = (i < s.NB) ? h + M : h;
    for (int r = 0; r < M; r++) {
        q_out[r] = s.z[r].real() * hq[r];
        q_out[r + M] = s.z[r].imag() * hqd[r];
    }
}

// -------------------------------------------------------------------------------------
// diblock (src/diblock.cc:9-95)
// -------------------------------------------------------------------------------------
struct orc_diblock {
    int M, NA, N;
    std::vector<double> qr;  // 2(N+1)M; pair i at offset 2iM: [q_i | q+_{N+1-i}]
    std::vector<double> h;   // [hA | hB]
    orc_step st;
    double* q1(int i) { return qr.data() + 2L * i * M; }
    double* q2(int j) { return qr.data() + (2L * (N + 1 - j) + 1) * M; }
};

static double diblock_calc_concs(orc_diblock& d, double* w) {
    const int M = d.M, N = d.N, NA = d.NA;
    double* phiA = w + 2 * M;
    double* phiB = w + 3 * M;
    double* hA = d.h.data();
    double* hB = d.h.data() + M;
    // prepare_h (src/diblock.cc:64-72)
    for (int r = 0; r < M; r++) {
        hA[r] = exp(-(w[r] + w[r + M]) / N);
        hB[r] = exp(-(-w[r] + w[r + M]) / N);
    }
    // q_1 = hA, q+_N = hB (src/diblock.cc:40)
    for (int r = 0; r < 2 * M; r++) d.q1(1)[r] = d.h[r];
    // chain (src/diblock.cc:43)
    for (int i = 1; i < N; i++) step_fwd(d.st, d.q1(i), d.q1(i + 1), d.h.data(), i);
    // Q (src/diblock.cc:46-48)
    double Q = osum(d.q1(N), M);
    Q /= M;
    // phi sums (src/diblock.cc:51-53, 76-80)
    for (int r = 0; r < 2 * M; r++) phiA[r] = 0.0;
    for (int i = 1; i <= NA; i++) {
        const double* a = d.q1(i); const double* b = d.q2(i);
        for (int r = 0; r < M; r++) phiA[r] += a[r] * b[r];
    }
    for (int i = NA + 1; i <= N; i++) {
        const double* a = d.q1(i); const double* b = d.q2(i);
        for (int r = 0; r < M; r++) phiB[r] += a[r] * b[r];
    }
    // normalize_phi (src/diblock.cc:85-95)
    for (int r = 0; r < M; r++) {
        double pA = phiA[r] / (N * Q * hA[r]);
        double pB = phiB[r] / (N * Q * hB[r]);
        phiA[r] = pA - pB;
        phiB[r] = pA + pB;
    }
    return log(Q);
}

// -------------------------------------------------------------------------------------
// Derivatives of calc_concs (no reference counterpart; pinned by central finite
// differences of diblock_calc_concs and by the identity <u, J v> = <J^T u, v>).
//
// Chain of the reference (src/diblock.cc:40-43, src/step.cc:53-64): q_1 = hA,
// q_j = f_j C q_{j-1}; q+_N = hB, q+_j = p_j C q+_{j+1}, with the Boltzmann factor of
// bead j being f_j (q chain: A for j = 1, else A iff j <= NA) and p_j (q+ chain: B for
// j = N, else B iff j > NA). phiA/phiB sum beads j <= NA / j > NA (src/diblock.cc:51-53).
// -------------------------------------------------------------------------------------
static inline bool fq_is_A(int j, int NA) { return j == 1 || j <= NA; }
static inline bool pq_is_A(int j, int N, int NA) { return j != N && j <= NA; }

// (oa, ob) = (C a, C b) with the step's kernel g (src/step.cc:44-50)
static void conv2(orc_step& s, const double* a, const double* b, double* oa, double* ob) {
    const int M = s.M;
    for (int r = 0; r < M; r++) s.z[r] = cplx(a[r], b[r]);
    fft_3d(s.P, s.m, s.z.data(), -1);
    for (int r = 0; r < M; r++) s.z[r] *= s.gfull[r];
    fft_3d(s.P, s.m, s.z.data(), +1);
    for (int r = 0; r < M; r++) { oa[r] = s.z[r].real(); ob[r] = s.z[r].imag(); }
}

// Directional derivative of (phi-, phi+) at w along (dW-, dw+) (either may be null).
static void diblock_jvp(orc_diblock& d, const double* w_in, const double* dwm, const double* dwp, double* dphi) {
    const int M = d.M, N = d.N, NA = d.NA, NB = N - NA;
    std::vector<double> w(w_in, w_in + 4L * M);
    diblock_calc_concs(d, w.data());
    const double* hA = d.h.data();
    const double* hB = d.h.data() + M;
    double Q = 0.0;
    for (int r = 0; r < M; r++) Q += d.q1(N)[r];
    Q /= M;
    std::vector<double> eA(M), eB(M);
    for (int r = 0; r < M; r++) {
        const double u = dwm ? dwm[r] : 0.0, v = dwp ? dwp[r] : 0.0;
        eA[r] = -(u + v) / N;
        eB[r] = -(v - u) / N;
    }
    // dp[i] = [dq_i | dq+_{N+1-i}], same layout as the pairs
    std::vector<double> dp(2L * (N + 1) * M, 0.0);
    auto D = [&](int i) { return dp.data() + 2L * i * M; };
    for (int r = 0; r < M; r++) { D(1)[r] = eA[r] * hA[r]; D(1)[r + M] = eB[r] * hB[r]; }
    std::vector<double> ca(M), cb(M);
    for (int i = 1; i < N; i++) {
        conv2(d.st, D(i), D(i) + M, ca.data(), cb.data());
        const double* hq = (i < NA) ? hA : hB;
        const double* eq = (i < NA) ? eA.data() : eB.data();
        const double* hqd = (i < NB) ? hB : hA;
        const double* eqd = (i < NB) ? eB.data() : eA.data();
        const double* P = d.q1(i + 1);
        for (int r = 0; r < M; r++) {
            D(i + 1)[r] = ca[r] * hq[r] + eq[r] * P[r];
            D(i + 1)[r + M] = cb[r] * hqd[r] + eqd[r] * P[r + M];
        }
    }
    double dQ = 0.0;
    for (int r = 0; r < M; r++) dQ += D(N)[r];
    dQ /= M;
    for (int r = 0; r < M; r++) {
        double SA = 0, SB = 0, dSA = 0, dSB = 0;
        for (int i = 1; i <= N; i++) {
            const double q = d.q1(i)[r], qd = d.q2(i)[r];
            const double dq = D(i)[r], dqd = D(N + 1 - i)[r + M];
            if (i <= NA) { SA += q * qd; dSA += dq * qd + q * dqd; }
            else { SB += q * qd; dSB += dq * qd + q * dqd; }
        }
        const double pA = SA / (N * Q * hA[r]), pB = SB / (N * Q * hB[r]);
        const double dpA = dSA / (N * Q * hA[r]) - pA * (dQ / Q + eA[r]);
        const double dpB = dSB / (N * Q * hB[r]) - pB * (dQ / Q + eB[r]);
        dphi[r] = dpA - dpB;
        dphi[r + M] = dpA + dpB;
    }
}

// Gradient (dW-, dw+) of L = <u-, phi-> + <u+, phi+> at w, for ubar = [u- | u+].
static void diblock_vjp(orc_diblock& d, const double* w_in, const double* ubar, double* wbar) {
    const int M = d.M, N = d.N, NA = d.NA;
    std::vector<double> w(w_in, w_in + 4L * M);
    diblock_calc_concs(d, w.data());
    const double* hA = d.h.data();
    const double* hB = d.h.data() + M;
    double Q = 0.0;
    for (int r = 0; r < M; r++) Q += d.q1(N)[r];
    Q /= M;
    std::vector<double> phA(M, 0.0), phB(M, 0.0), uA(M), uB(M), aA(M), aB(M);
    double Sdot = 0.0;
    for (int r = 0; r < M; r++) {
        double SA = 0, SB = 0;
        for (int i = 1; i <= N; i++) {
            const double v = d.q1(i)[r] * d.q2(i)[r];
            if (i <= NA) SA += v; else SB += v;
        }
        phA[r] = SA / (N * Q * hA[r]);
        phB[r] = SB / (N * Q * hB[r]);
        uA[r] = ubar[r + M] + ubar[r];
        uB[r] = ubar[r + M] - ubar[r];
        Sdot += uA[r] * phA[r] + uB[r] * phB[r];
        aA[r] = uA[r] / (N * Q * hA[r]);
        aB[r] = uB[r] / (N * Q * hB[r]);
    }
    const double Qbar = -Sdot / Q;
    auto alpha = [&](int j) { return (j <= NA) ? aA.data() : aB.data(); };
    auto fq = [&](int j) { return fq_is_A(j, NA) ? hA : hB; };
    auto pq = [&](int j) { return pq_is_A(j, N, NA) ? hA : hB; };
    // qb[j] = dL/dq_j (j = N..1), qdb[j] = dL/dq+_j (j = 1..N)
    std::vector<double> qb((size_t)(N + 1) * M), qdb((size_t)(N + 1) * M), ta(M), tb(M), ca(M), cb(M);
    auto QB = [&](int j) { return qb.data() + (size_t)j * M; };
    auto QDB = [&](int j) { return qdb.data() + (size_t)j * M; };
    for (int r = 0; r < M; r++) {
        QB(N)[r] = alpha(N)[r] * d.q2(N)[r] + Qbar / M;
        QDB(1)[r] = alpha(1)[r] * d.q1(1)[r];
    }
    for (int m = 1; m < N; m++) {
        const int j = N - m, jd = m + 1;  // qb[j] from qb[j+1]; qdb[jd] from qdb[jd-1]
        for (int r = 0; r < M; r++) {
            ta[r] = fq(j + 1)[r] * QB(j + 1)[r];
            tb[r] = pq(jd - 1)[r] * QDB(jd - 1)[r];
        }
        conv2(d.st, ta.data(), tb.data(), ca.data(), cb.data());
        for (int r = 0; r < M; r++) {
            QB(j)[r] = alpha(j)[r] * d.q2(j)[r] + ca[r];
            QDB(jd)[r] = alpha(jd)[r] * d.q1(jd)[r] + cb[r];
        }
    }
    for (int r = 0; r < M; r++) {
        double GA = -uA[r] * phA[r], GB = -uB[r] * phB[r];
        for (int j = 1; j <= N; j++) {
            const double a = QB(j)[r] * d.q1(j)[r];
            if (fq_is_A(j, NA)) GA += a; else GB += a;
            const double b = QDB(j)[r] * d.q2(j)[r];
            if (pq_is_A(j, N, NA)) GA += b; else GB += b;
        }
        wbar[r] = -(GA - GB) / N;
        wbar[r + M] = -(GA + GB) / N;
    }
}

// d lnQ / d L_d at fixed fields through g(k) = exp(-K^2/(6N))/M (src/step.cc:84-85):
// dQ/dL_d = (1/M) sum_{i<N} sum_k g(k) kap_d(k) Re(X_i(k) conj(Y_{i+1}(k))), X_i = F q_i,
// Y_i = F q+_i, kap_d = (2 pi k_d)^2 / (3 N L_d^3). Valid when the q+ chain is the
// adjoint of the q chain (1 <= NA <= N-1). Returns ln Q.
static double diblock_dlnQ_dL(orc_diblock& d, const double* w_in, double* out) {
    const int M = d.M, N = d.N;
    std::vector<double> w(w_in, w_in + 4L * M);
    const double lnQ = diblock_calc_concs(d, w.data());
    const double Q = exp(lnQ);
    const int* m = d.st.m;
    std::vector<double> acc(M, 0.0);
    std::vector<cplx> X(M), Y(M);
    for (int i = 1; i < N; i++) {
        for (int r = 0; r < M; r++) { X[r] = cplx(d.q1(i)[r], 0.0); Y[r] = cplx(d.q2(i + 1)[r], 0.0); }
        fft_3d(d.st.P, m, X.data(), -1);
        fft_3d(d.st.P, m, Y.data(), -1);
        for (int r = 0; r < M; r++) acc[r] += (X[r] * std::conj(Y[r])).real();
    }
    for (int dd = 0; dd < 3; dd++) {
        double s = 0.0;
        for (int K0 = 0; K0 < m[0]; K0++)
            for (int K1 = 0; K1 < m[1]; K1++)
                for (int K2 = 0; K2 < m[2]; K2++) {
                    const int Kd = (dd == 0) ? K0 : (dd == 1 ? K1 : K2);
                    const int f = (Kd <= m[dd] / 2) ? Kd : Kd - m[dd];
                    const double L = d.st.L[dd];
                    const double kap = (2 * M_PI * f) * (2 * M_PI * f) / (3.0 * N * L * L * L);
                    const long k = ((long)K0 * m[1] + K1) * m[2] + K2;
                    s += d.st.gfull[k] * kap * acc[k];
                }
        out[dd] = s / M / Q;
    }
    return lnQ;
}

// -------------------------------------------------------------------------------------
// LU with partial pivoting (stands in for GSL, src/anderson.cc:106-127): row of the
// largest |a_ij| in the column is swapped up; L unit lower, U upper; then solve.
// -------------------------------------------------------------------------------------
static void lu_solve(int n, const double* A_in, const double* b, double* x) {
    std::vector<double> A(A_in, A_in + (size_t)n * n);
    std::vector<int> p(n);
    for (int i = 0; i < n; i++) p[i] = i;
    for (int j = 0; j < n - 1; j++) {
        int ip = j;
        double amax = fabs(A[(size_t)j * n + j]);
        for (int i = j + 1; i < n; i++) {
            double v = fabs(A[(size_t)i * n + j]);
            if (v > amax) { amax = v; ip = i; }
        }
        if (ip != j) {
            for (int k = 0; k < n; k++) std::swap(A[(size_t)j * n + k], A[(size_t)ip * n + k]);
            std::swap(p[j], p[ip]);
        }
        double ajj = A[(size_t)j * n + j];
        if (ajj != 0.0) {
            for (int i = j + 1; i < n; i++) {
                double aij = A[(size_t)i * n + j] / ajj;
                A[(size_t)i * n + j] = aij;
                for (int k = j + 1; k < n; k++) A[(size_t)i * n + k] -= aij * A[(size_t)j * n + k];
            }
        }
    }
    std::vector<double> y(n);
    for (int i = 0; i < n; i++) {
        double s = b[p[i]];
        for (int k = 0; k < i; k++) s -= A[(size_t)i * n + k] * y[k];
        y[i] = s;
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = y[i];
        for (int k = i + 1; k < n; k++) s -= A[(size_t)i * n + k] * x[k];
        x[i] = s / A[(size_t)i * n + i];
    }
}

// -------------------------------------------------------------------------------------
// anderson (src/anderson.cc:8-93)
// -------------------------------------------------------------------------------------
struct orc_anderson {
    int M, nhMax;
    std::vector<double> DD;   // (nhMax+1)^2, DD[a*(nhMax+1)+b]
    std::vector<double> U, V, C;
    std::vector<double> Dh, wh; // (nhMax+1) x M
    double last_err = 0.0;
    // Test hooks. trace: one record per iteration (layout at am_trace_record) for the
    // least-squares pin of the iterate sequence. fault: a planted misreading of
    // src/anderson.cc (1..6, see anderson_mix), used only to show that the pins catch it.
    double* trace = nullptr;
    int trace_cap = 0, trace_n = 0;
    int fault = 0;
};

// Trace record of one iteration: [k, nh, lambda, err, slot, C[0..nhMax-1], Dh0 (M),
// wh0 (M), w+ after the update (M)]; slot is the history row the iteration's Dh0/wh0
// are copied into at its end.
static size_t am_trace_len(const orc_anderson& a) { return 5 + (size_t)a.nhMax + 3 * (size_t)a.M; }

static int anderson_mix(orc_anderson& a, orc_diblock* dbc, int maxIter, double errTol, double* w) {
    const int M = a.M, H = a.nhMax + 1;
    double lambda, err = 1.0, S1;
    int k, m, n, nh;
    auto DD = [&](int i, int j) -> double& { return a.DD[(size_t)i * H + j]; };
    auto Dh = [&](int i) { return a.Dh.data() + (size_t)i * M; };
    auto wh = [&](int i) { return a.wh.data() + (size_t)i * M; };
    // planted faults (tests only): 1 lambda = 1-0.9^(k+1); 2 ring slot 1+k%nhMax;
    // 3 the DD entries past the slot are not refreshed; 4 U laid out with stride nhMax;
    // 5 nh capped at nhMax-1; 6 the lambda*Dh_n term dropped from the history sum
    const int F = a.fault;
    for (k = 1; k < maxIter && err > errTol; k++) {
        diblock_calc_concs(*dbc, w);
        for (int r = 0; r < M; r++) { Dh(0)[r] = w[r + 3 * M] - 1.0; wh(0)[r] = w[r + M]; }
        const bool rev = (g_order & 2) != 0;
        S1 = 0.0;
        for (int t = 0; t < M; t++) { const int r = rev ? M - 1 - t : t; S1 += pow(Dh(0)[r], 2.0); }
        err = pow(S1 / M, 0.5);
        lambda = 1.0 - pow(0.9, double(k + (F == 1)));
        nh = (k < a.nhMax + 1) ? k - 1 : a.nhMax - (F == 5);
        const int us = (F == 4) ? a.nhMax : nh;  // U's row stride
        for (m = 0; m <= nh; m++) {
            double s = 0.0;
            for (int t = 0; t < M; t++) { const int r = rev ? M - 1 - t : t; s += Dh(0)[r] * Dh(m)[r]; }
            DD(0, m) = s;
        }
        if (k < 2) {
            for (int r = 0; r < M; r++) w[r + M] += lambda * Dh(0)[r];
        } else {
            for (m = 1; m <= nh; m++) {
                a.V[m - 1] = DD(0, 0) - DD(0, m);
                for (n = 1; n <= m; n++)
                    a.U[(m - 1) * us + n - 1] = a.U[(n - 1) * us + m - 1] =
                        DD(0, 0) - DD(0, m) - DD(0, n) + DD(n, m);
            }
            lu_solve(nh, a.U.data(), a.V.data(), a.C.data());
            for (int r = 0; r < M; r++) w[r + M] += lambda * Dh(0)[r];
            const double lh = (F == 6) ? 0.0 : lambda;
            for (n = 1; n <= nh; n++) {
                const double* dn = Dh(n); const double* wn = wh(n);
                const double* d0 = Dh(0); const double* w0 = wh(0);
                for (int r = 0; r < M; r++)
                    w[r + M] += a.C[n - 1] * ((wn[r] + lh * dn[r]) - (w0[r] + lambda * d0[r]));
            }
        }
        n = 1 + (k - 1 + (F == 2)) % (a.nhMax);
        if (a.trace && a.trace_n < a.trace_cap) {
            double* t = a.trace + (size_t)a.trace_n++ * am_trace_len(a);
            t[0] = k; t[1] = nh; t[2] = lambda; t[3] = err; t[4] = n;
            for (m = 0; m < a.nhMax; m++) t[5 + m] = (k >= 2 && m < nh) ? a.C[m] : 0.0;
            t += 5 + a.nhMax;
            std::memcpy(t, Dh(0), sizeof(double) * M);
            std::memcpy(t + M, wh(0), sizeof(double) * M);
            std::memcpy(t + 2 * (size_t)M, w + M, sizeof(double) * M);
        }
        std::memcpy(Dh(n), Dh(0), sizeof(double) * M);
        std::memcpy(wh(n), wh(0), sizeof(double) * M);
        DD(n, n) = DD(0, 0);
        for (m = 1; m < n; m++) DD(m, n) = DD(0, m);
        if (F != 3)
            for (m = n + 1; m <= nh; m++) DD(n, m) = DD(0, m);
    }
    a.last_err = err;
    return k;
}

// -------------------------------------------------------------------------------------
// langevin update (src/langevin.cc:23-37); the noise arrays come from the caller
// -------------------------------------------------------------------------------------
static void langevin_update(double* w, const double* eta_prev, const double* eta_new, int M,
                            double XbN, double dt) {
    for (int r = 0; r < M; r++)
        w[r] += -(w[r + 2 * M] + 2 * w[r] / XbN) * dt + 0.5 * (eta_prev[r] + eta_new[r]);
}

// -------------------------------------------------------------------------------------
// strFunc (src/strFunc.cc:9-103)
// -------------------------------------------------------------------------------------
struct orc_strfunc {
    int m[3], M, Mk;
    double dK, coeff, chi_b;
    int nsamples = 0;
    std::vector<double> S, K;
    std::vector<int> wt, P;
    fft_plan1d PL[3];
    std::vector<cplx> z;
    // planted misreading of src/strFunc.cc (tests only; 0 = none): 1 weight 2 on the
    // k2 = mz/2 plane; 2 n_same not reset between bins; 3 the -0.5/chi_b term dropped;
    // 4 bins split on every change of K (dK ignored)
    int fault = 0;
};

static void sk_init(orc_strfunc& s, const int* m, const double* L, double CV, double chi_b, double dK) {
    for (int d = 0; d < 3; d++) { s.m[d] = m[d]; make_plan(s.PL[d], m[d]); }
    s.M = m[0] * m[1] * m[2];
    s.Mk = m[0] * m[1] * (m[2] / 2 + 1);
    s.dK = dK;
    s.chi_b = chi_b;
    s.nsamples = 0;
    s.coeff = CV / (chi_b * chi_b * s.M * s.M);
    s.S.assign(s.Mk, 0.0);
    s.K.assign(s.Mk, 0.0);
    s.wt.assign(s.Mk, 2);
    s.P.resize(s.Mk);
    s.z.resize(s.M);
    // calcK (src/strFunc.cc:80-103)
    for (int k0 = -(m[0] - 1) / 2; k0 <= m[0] / 2; k0++) {
        int K0 = (k0 < 0) ? (k0 + m[0]) : k0;
        double kx_sq = k0 * k0 / (L[0] * L[0]);
        for (int k1 = -(m[1] - 1) / 2; k1 <= m[1] / 2; k1++) {
            int K1 = (k1 < 0) ? (k1 + m[1]) : k1;
            double ky_sq = k1 * k1 / (L[1] * L[1]);
            for (int k2 = 0; k2 <= m[2] / 2; k2++) {
                double kz_sq = k2 * k2 / (L[2] * L[2]);
                int k = k2 + (m[2] / 2 + 1) * (K1 + m[1] * K0);
                s.K[k] = 2 * M_PI * pow(kx_sq + ky_sq + kz_sq, 0.5);
                if ((k2 == 0) || (k2 == m[2] / 2)) s.wt[k] = 1;
            }
        }
    }
    std::iota(s.P.begin(), s.P.end(), 0);
    const double* Kp = s.K.data();
    std::stable_sort(s.P.begin(), s.P.end(), [Kp](size_t i, size_t j) { return Kp[i] < Kp[j]; });
}

// src/strFunc.cc:36-46: S_k += |r2c(W-)_k|^2 over the half spectrum
static void sk_sample(orc_strfunc& s, const double* wm) {
    for (int r = 0; r < s.M; r++) s.z[r] = cplx(wm[r], 0.0);
    fft_3d(s.PL, s.m, s.z.data(), -1);
    const int mk2 = s.m[2] / 2 + 1;
    for (int K0 = 0; K0 < s.m[0]; K0++)
        for (int K1 = 0; K1 < s.m[1]; K1++)
            for (int k2 = 0; k2 < mk2; k2++)
                s.S[k2 + mk2 * (K1 + s.m[1] * K0)] += std::norm(s.z[((long)K0 * s.m[1] + K1) * s.m[2] + k2]);
    s.nsamples++;
}

// src/strFunc.cc:50-76
static void sk_save(orc_strfunc& s, const std::string& fileName, int dp) {
    double S_sum = 0.0;
    int n_same = 0;
    std::ofstream out(fileName);
    out.precision(dp);
    out.setf(std::ios::fixed, std::ios::floatfield);
    const int F = s.fault;
    const int mk2 = s.m[2] / 2 + 1;
    for (int k = 0; k < s.Mk; k++) {
        int wt = s.wt[s.P[k]];
        if (F == 1 && s.m[2] % 2 == 0 && s.P[k] % mk2 == mk2 - 1) wt = 2;
        S_sum += wt * ((s.coeff / s.nsamples) * s.S[s.P[k]] - (F == 3 ? 0.0 : 0.5 / s.chi_b));
        n_same += wt;
        const double dK = (F == 4) ? 0.0 : s.dK;
        if ((k == s.Mk - 1) || (fabs(s.K[s.P[k + 1]] - s.K[s.P[k]]) > dK)) {
            out << s.K[s.P[k]] << "\t" << S_sum / n_same << std::endl;
            S_sum = 0.0;
            if (F != 2) n_same = 0;
        }
    }
}

// -------------------------------------------------------------------------------------
// parameters (src/lfts_params.cc:73-142)
// -------------------------------------------------------------------------------------
struct orc_params {
    int N, NA, NB, isXeN, m[3], equil_its, sim_its, sample_freq, save_freq, loadType, M, Mk;
    double chi_b, C, dt, L[3], V, sigma, XeN;
};

static double z_inf_discrete(const double* L, const int* m, int N, double Nbar, int tmax) {
    double sum, prod = 1.0, X, ell, R0, R0dL[3];
    sum = 0.5;
    R0 = sqrt((double)N);
    for (int i = 0; i < 3; i++) { R0dL[i] = R0 * L[i] / m[i]; prod *= R0dL[i]; }
    ell = pow(prod, 1.0 / 3.0);
    for (int t = 1; t <= tmax; t++) {
        X = (M_PI / ell) * sqrt(t / 6.0);
        prod = 1.0;
        for (int i = 0; i < 3; i++) prod *= erf(X * ell / R0dL[i]);
        sum += pow(sqrt(M_PI) / (2 * X), 3.0) * prod;
    }
    sum *= 2 * R0 / (pow(ell, 3.0) * sqrt(Nbar));
    X = (M_PI / ell) * sqrt((0.5 + tmax) / 6.0);
    return 1.0 - sum - 3 * R0 / (ell * sqrt(M_PI * Nbar) * X);
}

static bool read_params(orc_params& p, const std::string& fileName) {
    std::ifstream in(fileName);
    if (in.fail()) return false;
    in >> p.N >> p.NA >> p.chi_b >> p.C >> p.dt >> p.isXeN;
    in >> p.m[0] >> p.m[1] >> p.m[2] >> p.L[0] >> p.L[1] >> p.L[2];
    in >> p.equil_its >> p.sim_its >> p.sample_freq >> p.save_freq >> p.loadType;
    p.equil_its = (p.equil_its / p.sample_freq) * p.sample_freq;
    p.sim_its = (p.sim_its / p.sample_freq) * p.sample_freq;
    double z_inf = z_inf_discrete(p.L, p.m, p.N, p.C * p.C, 100);
    if (p.isXeN == 1) { p.XeN = p.chi_b; p.chi_b = p.chi_b / z_inf; }
    else p.XeN = p.chi_b * z_inf;
    p.M = p.m[0] * p.m[1] * p.m[2];
    p.Mk = p.m[0] * p.m[1] * (p.m[2] / 2 + 1);
    p.V = p.L[0] * p.L[1] * p.L[2];
    p.sigma = sqrt(2.0 * p.M * p.dt / (p.C * p.V));
    p.NB = p.N - p.NA;
    return true;
}

static void save_params(const orc_params& p, const std::string& fileName) {
    double XN_out = (p.isXeN == 1) ? p.XeN : p.chi_b;
    std::ofstream o(fileName);
    o << p.N << " " << p.NA << " " << XN_out << " " << p.C << " " << p.dt << " " << p.isXeN << std::endl;
    o << p.m[0] << " " << p.m[1] << " " << p.m[2] << " " << p.L[0] << " " << p.L[1] << " " << p.L[2] << std::endl;
    o << p.equil_its << " " << p.sim_its << " " << p.sample_freq << " " << p.save_freq << " " << 1 << std::endl;
}

static void save_array(const double* a, const std::string& fileName, int n, bool append) {
    std::ofstream o;
    if (append) o.open(fileName, std::ios_base::app); else o.open(fileName);
    for (int r = 0; r < n; r++) o << a[r] << std::endl;
}

static void read_array(double* a, const std::string& fileName, int n, int nIgnore) {
    std::ifstream in(fileName);
    for (int i = 0; i < nIgnore; i++) in.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
    for (int r = 0; r < n; r++) in >> a[r];
}

// -------------------------------------------------------------------------------------
// simulation driver (src/lfts_simulation.cc:8-157)
// -------------------------------------------------------------------------------------
struct orc_sim {
    orc_params p;
    std::mt19937_64 rng;
    std::normal_distribution<double> nrm;
    std::vector<double> w, noise;
    double* eta_prev;
    double* eta_new;
    std::unique_ptr<orc_diblock> dbc;
    orc_anderson am;
    orc_strfunc sk;
    std::string outdir;
    bool write_files = true;
    bool verbose = false;
    int last_nAM = 0;
    std::vector<int> nAM_log;
    std::string path(const std::string& f) const { return outdir.empty() ? f : outdir + "/" + f; }
};

static void sim_save_std(orc_sim& s, const std::string& f) {
    save_params(s.p, s.path(f));
    save_array(s.w.data(), s.path(f), 2 * s.p.M, true);
}

extern "C" {

// ---- FFT and step -------------------------------------------------------------------
void orc_fft1d(int n, double* data, int sign) {
    fft_plan1d P; make_plan(P, n);
    std::vector<cplx> scr(n);
    fft_1d(P, reinterpret_cast<cplx*>(data), scr.data(), sign);
}

void orc_fft3d(const int* m, double* data, int sign) {
    fft_plan1d P[3];
    for (int d = 0; d < 3; d++) make_plan(P[d], m[d]);
    fft_3d(P, m, reinterpret_cast<cplx*>(data), sign);
}

void* orc_create(int N, int NA, const int* m, const double* L) {
    orc_diblock* d = new orc_diblock();
    d->NA = NA; d->N = N;
    d->M = m[0] * m[1] * m[2];
    d->qr.assign(2L * (N + 1) * d->M, 0.0);
    d->h.assign(2L * d->M, 0.0);
    step_init(d->st, NA, N - NA, m, L);
    return d;
}

void orc_destroy(void* h) { delete static_cast<orc_diblock*>(h); }

void orc_ghalf(void* h, double* out) {
    orc_diblock* d = static_cast<orc_diblock*>(h);
    std::memcpy(out, d->st.g.data(), sizeof(double) * d->st.Mk);
}

double orc_calc_concs(void* h, double* w) { return diblock_calc_concs(*static_cast<orc_diblock*>(h), w); }

// pair i (1..N) as stored by the reference: [q_i | q+_{N+1-i}]
void orc_get_pair(void* h, int i, double* out) {
    orc_diblock* d = static_cast<orc_diblock*>(h);
    std::memcpy(out, d->q1(i), sizeof(double) * 2 * d->M);
}

void orc_jvp(void* h, const double* w, const double* dwm, const double* dwp, double* dphi) {
    diblock_jvp(*static_cast<orc_diblock*>(h), w, dwm, dwp, dphi);
}
void orc_vjp(void* h, const double* w, const double* ubar, double* wbar) {
    diblock_vjp(*static_cast<orc_diblock*>(h), w, ubar, wbar);
}
double orc_dlnQ_dL(void* h, const double* w, double* out) {
    return diblock_dlnQ_dL(*static_cast<orc_diblock*>(h), w, out);
}

void orc_step_fwd(void* h, const double* q_in, double* q_out, const double* hh, int i) {
    orc_diblock* d = static_cast<orc_diblock*>(h);
    step_fwd(d->st, q_in, q_out, hh, i);
}

// ---- Anderson ----------------------------------------------------------------------
void* orc_am_create(int M, int nhMax) {
    orc_anderson* a = new orc_anderson();
    a->M = M; a->nhMax = nhMax;
    a->DD.assign((size_t)(nhMax + 1) * (nhMax + 1), 0.0);
    a->U.assign((size_t)nhMax * nhMax, 0.0);
    a->V.assign(nhMax, 0.0);
    a->C.assign(nhMax, 0.0);
    a->Dh.assign((size_t)(nhMax + 1) * M, 0.0);
    a->wh.assign((size_t)(nhMax + 1) * M, 0.0);
    return a;
}
void orc_am_destroy(void* a) { delete static_cast<orc_anderson*>(a); }
int orc_am_mix(void* a, void* dbc, int maxIter, double tol, double* w) {
    return anderson_mix(*static_cast<orc_anderson*>(a), static_cast<orc_diblock*>(dbc), maxIter, tol, w);
}
double orc_am_last_err(void* a) { return static_cast<orc_anderson*>(a)->last_err; }
// Test hooks: record every iteration of the following mix calls into trace (cap records
// of am_trace_len doubles; NULL stops recording), plant a fault (0 = none).
void orc_am_set_trace(void* ap, double* trace, int cap) {
    orc_anderson* a = static_cast<orc_anderson*>(ap);
    a->trace = trace; a->trace_cap = trace ? cap : 0; a->trace_n = 0;
}
int orc_am_trace_count(void* a) { return static_cast<orc_anderson*>(a)->trace_n; }
void orc_am_set_fault(void* a, int fault) { static_cast<orc_anderson*>(a)->fault = fault; }

void orc_lu_solve(int n, const double* A, const double* b, double* x) { lu_solve(n, A, b, x); }

// ---- Langevin and noise --------------------------------------------------------------
void orc_langevin_update(double* w, const double* eta_prev, const double* eta_new, int M, double XbN, double dt) {
    langevin_update(w, eta_prev, eta_new, M, XbN, dt);
}

// n draws of std::normal_distribution<double>(0, sigma) from std::mt19937_64(seed)
void orc_normals(unsigned long long seed, double sigma, int n, double* out) {
    std::mt19937_64 r; r.seed(seed);
    std::normal_distribution<double> d(0.0, sigma);
    for (int i = 0; i < n; i++) out[i] = d(r);
}
void orc_mt64(unsigned long long seed, int n, unsigned long long* out) {
    std::mt19937_64 r; r.seed(seed);
    for (int i = 0; i < n; i++) out[i] = r();
}

// ---- structure function ------------------------------------------------------------
void* orc_sk_create(const int* m, const double* L, double CV, double chi_b, double dK) {
    orc_strfunc* s = new orc_strfunc();
    sk_init(*s, m, L, CV, chi_b, dK);
    return s;
}
void orc_sk_destroy(void* s) { delete static_cast<orc_strfunc*>(s); }
void orc_sk_sample(void* s, const double* wm) { sk_sample(*static_cast<orc_strfunc*>(s), wm); }
void orc_sk_get(void* sp, double* S, double* K, int* wt, int* P, int* ns) {
    orc_strfunc* s = static_cast<orc_strfunc*>(sp);
    if (S) std::memcpy(S, s->S.data(), sizeof(double) * s->Mk);
    if (K) std::memcpy(K, s->K.data(), sizeof(double) * s->Mk);
    if (wt) std::memcpy(wt, s->wt.data(), sizeof(int) * s->Mk);
    if (P) std::memcpy(P, s->P.data(), sizeof(int) * s->Mk);
    if (ns) *ns = s->nsamples;
}
void orc_sk_save(void* s, const char* fileName, int dp) { sk_save(*static_cast<orc_strfunc*>(s), fileName, dp); }
void orc_sk_set_fault(void* s, int fault) { static_cast<orc_strfunc*>(s)->fault = fault; }

// ---- parameters --------------------------------------------------------------------
double orc_z_inf(const double* L, const int* m, int N, double Nbar, int tmax) {
    return z_inf_discrete(L, m, N, Nbar, tmax);
}

// ints: N NA NB isXeN m0 m1 m2 equil sim sample save loadType M Mk (14)
// dbls: chi_b C dt L0 L1 L2 V sigma XeN (9)
int orc_read_params(const char* fileName, int* ints, double* dbls) {
    orc_params p;
    if (!read_params(p, fileName)) return 1;
    int iv[14] = {p.N, p.NA, p.NB, p.isXeN, p.m[0], p.m[1], p.m[2], p.equil_its, p.sim_its,
                  p.sample_freq, p.save_freq, p.loadType, p.M, p.Mk};
    double dv[9] = {p.chi_b, p.C, p.dt, p.L[0], p.L[1], p.L[2], p.V, p.sigma, p.XeN};
    std::memcpy(ints, iv, sizeof iv);
    std::memcpy(dbls, dv, sizeof dv);
    return 0;
}

// ---- whole simulation --------------------------------------------------------------
// src/lfts_simulation.cc:8-57 (constructor): init order and RNG consumption order.
// flags: bit 0 writes the output files; bit 1 skips the initial mix (for timing a
// Langevin step from a state that is set afterwards with orc_sim_set_w).
void* orc_sim_create(const char* inputFile, const char* outdir, int flags) {
    const int write_files = flags & 1;
    orc_sim* s = new orc_sim();
    if (!read_params(s->p, inputFile)) { delete s; return nullptr; }
    s->outdir = outdir ? outdir : "";
    s->write_files = write_files != 0;
    const int M = s->p.M;
    s->rng.seed(123456789);
    s->w.assign(4L * M, 0.0);
    s->dbc.reset(static_cast<orc_diblock*>(orc_create(s->p.N, s->p.NA, s->p.m, s->p.L)));
    orc_anderson* a = static_cast<orc_anderson*>(orc_am_create(M, 10));
    s->am = *a; delete a;
    s->nrm = std::normal_distribution<double>(0.0, s->p.sigma);
    s->noise.assign(2L * M, 0.0);
    for (int r = 0; r < M; r++) s->noise[r] = s->nrm(s->rng);
    s->eta_prev = s->noise.data();
    s->eta_new = s->noise.data() + M;
    sk_init(s->sk, s->p.m, s->p.L, s->p.C * s->p.V, s->p.chi_b, 1e-5);
    if (s->p.loadType == 1) {
        read_array(s->w.data(), inputFile, 2 * M, 3);
    } else if (s->p.loadType == 2) {
        // field_generator::create_lamellar (src/field_generator.h:11-24), kx=3
        const int* m = s->p.m;
        for (int x = 0; x < m[0]; x++)
            for (int y = 0; y < m[1]; y++)
                for (int z = 0; z < m[2]; z++) {
                    int r = m[2] * (x * m[1] + y) + z;
                    s->w[r] = s->p.chi_b * cos(2.0 * M_PI * (3 * ((double)x) / m[0] + 0 * ((double)y) / m[1] + 0 * ((double)z) / m[2]));
                    s->w[r + M] = 0.0;
                }
    } else {
        std::uniform_real_distribution<double> u(0.0, 1.0);
        for (int r = 0; r < M; r++) { s->w[r] = s->p.chi_b * (u(s->rng) - 0.5); s->w[r + M] = 0.0; }
    }
    if (flags & 2) return s;
    s->last_nAM = anderson_mix(s->am, s->dbc.get(), 200, 1e-4, s->w.data());
    if (s->write_files) {
        sim_save_std(*s, "w_0");
        save_array(s->w.data() + 2 * M, s->path("phi_0"), 2 * M, false);
    }
    return s;
}

void orc_sim_destroy(void* s) { delete static_cast<orc_sim*>(s); }

void orc_sim_get_w(void* sp, double* w) {
    orc_sim* s = static_cast<orc_sim*>(sp);
    std::memcpy(w, s->w.data(), sizeof(double) * 4 * s->p.M);
}
void orc_sim_set_w(void* sp, const double* w) {
    orc_sim* s = static_cast<orc_sim*>(sp);
    std::memcpy(s->w.data(), w, sizeof(double) * 4 * s->p.M);
}
int orc_sim_last_nAM(void* sp) { return static_cast<orc_sim*>(sp)->last_nAM; }

// One Langevin step of equilibrate()/statistics() (src/lfts_simulation.cc:67-118),
// without file output: step_wm, mix, calc_concs (lnQ), and S(k) sample if asked.
double orc_sim_step(void* sp, int sample) {
    orc_sim* s = static_cast<orc_sim*>(sp);
    const int M = s->p.M;
    for (int r = 0; r < M; r++) s->eta_new[r] = s->nrm(s->rng);
    langevin_update(s->w.data(), s->eta_prev, s->eta_new, M, s->p.chi_b, s->p.dt);
    std::swap(s->eta_prev, s->eta_new);
    s->last_nAM = anderson_mix(s->am, s->dbc.get(), 200, 1e-4, s->w.data());
    double lnQ = diblock_calc_concs(*s->dbc, s->w.data());
    if (sample) sk_sample(s->sk, s->w.data());
    return lnQ;
}

static void sim_loop(orc_sim* s, bool stats) {
    const int M = s->p.M;
    const int nits = stats ? s->p.sim_its : s->p.equil_its;
    const std::string tag = stats ? "st_" : "eq_";
    int it;
    for (it = 1; it <= nits; it++) {
        double lnQ = orc_sim_step(s, stats && (it % s->p.sample_freq == 0));
        std::cout << "lnQ = " << lnQ << std::endl;
        if (it % s->p.save_freq == 0 && s->write_files) {
            sim_save_std(*s, "w_" + tag + std::to_string(it));
            save_array(s->w.data() + 2 * M, s->path("phi_" + tag + std::to_string(it)), 2 * M, false);
            if (stats) sk_save(s->sk, s->path("struct_st_" + std::to_string(it)), 8);
        }
    }
    if (s->write_files) {
        sim_save_std(*s, "w_" + tag + std::to_string(it - 1));
        save_array(s->w.data() + 2 * M, s->path("phi_" + tag + std::to_string(it - 1)), 2 * M, false);
        if (stats) sk_save(s->sk, s->path("struct_st_" + std::to_string(it - 1)), 8);
    }
}

void orc_sim_equilibrate(void* sp) { sim_loop(static_cast<orc_sim*>(sp), false); }
void orc_sim_statistics(void* sp) { sim_loop(static_cast<orc_sim*>(sp), true); }

// src/lfts_simulation.cc:122-135
double orc_sim_getH(void* sp) {
    orc_sim* s = static_cast<orc_sim*>(sp);
    const int M = s->p.M;
    double lnQ = diblock_calc_concs(*s->dbc, s->w.data());
    std::vector<double> sq(M);
    for (int r = 0; r < M; r++) sq[r] = pow(s->w[r], 2.0);
    const double wp_sum = osum(s->w.data() + M, M), wm2_sum = osum(sq.data(), M);
    return -lnQ + (wm2_sum / s->p.chi_b - wp_sum) / M;
}

void orc_set_order(int order) { g_order = order; }

void orc_sim_sk_get(void* sp, double* S, int* ns) {
    orc_sim* s = static_cast<orc_sim*>(sp);
    orc_sk_get(&s->sk, S, nullptr, nullptr, nullptr, ns);
}

}  // extern "C"
