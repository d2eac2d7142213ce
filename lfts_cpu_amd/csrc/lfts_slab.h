// Device code of the fused slab pass, shared by the translation units that instantiate its
// kernel shapes: lfts_fused.hip (the default line factorisation of every slab edge, all
// chain kinds) and lfts_slab_alt.hip (the other factorisations the kernel-shape autotuner
// times, plain chains).
//
// Fused contour step for cubic meshes of 16^3, 24^3, 32^3, 40^3, 48^3, 56^3, 64^3, 72^3,
// 80^3, 96^3, 112^3, 120^3 and 128^3 points: one launch per step over every active replica
// of that mesh size.
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
// Default factorisations: 16 = 4 x 4, 72 = 8 x 9, 80 = 8 x 10, and R1 x 8 for 24, 32, 40,
// 48, 56, 64, 96, 112, 120, 128 (R1 = 3, 4, 5, 6, 7, 8, 12, 14, 15, 16); the autotuner
// also times R1 x 4 (and 2 x 8, 4 x 16) shapes for the edges up to 64. With R1 < R2 some
// threads idle in the stage-2 DFTs; with R1 > R2 some take two. One plane of 128 points
// fills 149 KB of the 160 KB LDS. Edges from 72 up have the LDS-table, one-field shape only.
// In LDS a line keeps element k*R2 + j at k*(R2+1) + j: the R2 threads of a line then
// read their contiguous runs from distinct banks.
//
// Outputs (pairs and, for one replica, partial pairs) are written with non-temporal
// stores: they are read next by workgroups on other XCDs, so they go to memory rather than
// stay in this L2 (a replica set's partial pairs use ordinary stores, see slab_pass).
#ifndef LFTS_SLAB_H
#define LFTS_SLAB_H

#include <hip/hip_runtime.h>

#include <type_traits>

#include "lfts_internal.h"

// Rounding is fixed by the source, not left to the compiler: fused multiply-adds are
// written out (fma) where they are wanted and nothing else is contracted (this file is
// built with -ffp-contract=on, under which the pragma holds), so every kernel below,
// whatever its shape or the code around the butterflies, computes bit-identical results.
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

// Workgroup barrier for LDS traffic only. The waves of a slab pass share data through
// LDS alone (exchange rows, the re-paired plane, the tables), so the barrier waits for
// LDS operations but not for global loads still in flight (the operands of later steps)
// or for stores (pairs written while the next convolution runs), as __syncthreads would.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// a*b rounded once, as an fma with a +0 addend: the compiler may not merge it into a
// following add (it may for a plain product, and whether it does depends on the code
// around it), so every kernel shape continues the chain from exactly the stored value.
__device__ __forceinline__ double rmul(double a, double b) { return fma(a, b, 0.0); }

// Every global access of the slab pass names the global address space. Its pointers may
// come from descriptors read from memory, where the compiler cannot tell and would emit
// flat accesses: slower, and counted with the LDS operations, so that an LDS barrier
// would wait for them.
typedef double d2v __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) const d2v gcd2v;
typedef __attribute__((address_space(1))) d2v gd2v;
typedef __attribute__((address_space(1))) const double gcd;
typedef __attribute__((address_space(1))) double gd;
__device__ __forceinline__ double2 ld2(const double* p) {
    const d2v w = *(gcd2v*)(p);
    return make_double2(w.x, w.y);
}
__device__ __forceinline__ double2 ld2(const double2* p) { return ld2(reinterpret_cast<const double*>(p)); }
__device__ __forceinline__ double ldg(const double* p) { return *(gcd*)(p); }
// non-temporal: data read once and not kept in L2 (a cached line the pass then
// overwrites would be left dirty, to be written back when the kernel ends)
__device__ __forceinline__ double2 ld2_nt(const double* p) {
    const d2v w = __builtin_nontemporal_load((gcd2v*)(p));
    return make_double2(w.x, w.y);
}
__device__ __forceinline__ void st_nt(double* p, double v) { __builtin_nontemporal_store(v, (gd*)(p)); }
__device__ __forceinline__ void st_g(double* p, double v) { *(gd*)(p) = v; }
__device__ __forceinline__ void st2_nt(double* p, double2 v) {
    const d2v w = {v.x, v.y};
    __builtin_nontemporal_store(w, (gd2v*)(p));
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

// radix-8 DFT after its first radix-2 stage (a_n = v_n + v_{n+4}, b_n = v_n - v_{n+4}):
// twiddles w8^n on b_n and two radix-4 DFTs, outputs in natural order. The w8 and w8^3
// twiddles share the factor r = 1/sqrt2: with b1 w8 = r P and b3 w8^3 = r Q, the b-DFT
// needs only r (P + Q) and r rot(P - Q), each folded into its final add as an fma.
template <bool INV>
__device__ __forceinline__ void dft8_tail(double2* a, double2* b, double2* v) {
    const double r = 0.70710678118654752440;
    // w8 = (1 - i)/sqrt2 forward, (1 + i)/sqrt2 inverse; w8^3 = (-1 - i)/sqrt2, (-1 + i)/sqrt2
    const double2 P = INV ? make_double2(b[1].x - b[1].y, b[1].x + b[1].y) : make_double2(b[1].x + b[1].y, b[1].y - b[1].x);
    const double2 Q = INV ? make_double2(-(b[3].x + b[3].y), b[3].x - b[3].y) : make_double2(b[3].y - b[3].x, -(b[3].x + b[3].y));
    const double2 S = c_add(P, Q), D = c_rot<INV>(c_sub(P, Q));
    const double2 b2 = c_rot<INV>(b[2]);
    const double2 s02 = c_add(b[0], b2), d02 = c_sub(b[0], b2);
    dft4<INV>(a[0], a[1], a[2], a[3]);
    v[0] = a[0]; v[2] = a[1]; v[4] = a[2]; v[6] = a[3];
    v[1] = make_double2(fma(r, S.x, s02.x), fma(r, S.y, s02.y));
    v[5] = make_double2(fma(-r, S.x, s02.x), fma(-r, S.y, s02.y));
    v[3] = make_double2(fma(r, D.x, d02.x), fma(r, D.y, d02.y));
    v[7] = make_double2(fma(-r, D.x, d02.x), fma(-r, D.y, d02.y));
}

// natural-order radix-8 DFT: radix-2 split, twiddles w8^n, two radix-4 DFTs
template <bool INV>
__device__ __forceinline__ void dft8(double2* v) {
    double2 a[4], b[4];
#pragma unroll
    for (int n = 0; n < 4; n++) {
        a[n] = c_add(v[n], v[n + 4]);
        b[n] = c_sub(v[n], v[n + 4]);
    }
    dft8_tail<INV>(a, b, v);
}

// natural-order radix-5 DFT: with a = x1 + x4, b = x2 + x3, c = x1 - x4, d = x2 - x3,
// X1,4 = x0 + c1 a + c2 b -/+ i (s1 c + s2 d) and X2,3 = x0 + c2 a + c1 b -/+ i (s2 c - s1 d)
// (forward; the inverse swaps the signs of the i terms), c_k, s_k = cos, sin(2 pi k / 5)
template <bool INV>
__device__ __forceinline__ void dft5(double2* v) {
    const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
    const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;
    const double2 a = c_add(v[1], v[4]), b = c_add(v[2], v[3]);
    const double2 c = c_sub(v[1], v[4]), d = c_sub(v[2], v[3]);
    const double2 x0 = v[0];
    const double2 t1 = make_double2(fma(c2, b.x, fma(c1, a.x, x0.x)), fma(c2, b.y, fma(c1, a.y, x0.y)));
    const double2 t2 = make_double2(fma(c1, b.x, fma(c2, a.x, x0.x)), fma(c1, b.y, fma(c2, a.y, x0.y)));
    const double2 u1 = make_double2(fma(s2, d.x, s1 * c.x), fma(s2, d.y, s1 * c.y));
    const double2 u2 = make_double2(fma(-s1, d.x, s2 * c.x), fma(-s1, d.y, s2 * c.y));
    v[0] = c_add(x0, c_add(a, b));
    // t - i u = (t.x + u.y, t.y - u.x); t + i u = (t.x - u.y, t.y + u.x)
    const double2 t1m = make_double2(t1.x + u1.y, t1.y - u1.x), t1p = make_double2(t1.x - u1.y, t1.y + u1.x);
    const double2 t2m = make_double2(t2.x + u2.y, t2.y - u2.x), t2p = make_double2(t2.x - u2.y, t2.y + u2.x);
    v[1] = INV ? t1p : t1m;
    v[4] = INV ? t1m : t1p;
    v[2] = INV ? t2p : t2m;
    v[3] = INV ? t2m : t2p;
}

// natural-order radix-3 DFT; the 1/2 and sqrt(3)/2 products are folded into their adds
template <bool INV>
__device__ __forceinline__ void dft3(double2& a, double2& b, double2& c) {
    const double s3 = 0.86602540378443864676;  // sqrt(3)/2
    const double2 t = c_add(b, c), d = c_sub(b, c);
    const double2 m = make_double2(fma(-0.5, t.x, a.x), fma(-0.5, t.y, a.y));
    a = c_add(a, t);
    // i (sqrt3/2) d = (-s3 d.y, s3 d.x); forward: b = m - i s3 d, c = m + i s3 d
    const double2 mi = make_double2(fma(s3, d.y, m.x), fma(-s3, d.x, m.y));  // m - i s3 d
    const double2 pi = make_double2(fma(-s3, d.y, m.x), fma(s3, d.x, m.y));  // m + i s3 d
    b = INV ? pi : mi;
    c = INV ? mi : pi;
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

// x * w16^2 and x * w16^6 (w16 = exp(-2 pi i/16) forward, its conjugate inverse)
template <bool INV>
__device__ __forceinline__ double2 mul_w16_2(double2 a) {
    const double r = 0.70710678118654752440;
    return INV ? make_double2(r * (a.x - a.y), r * (a.x + a.y)) : make_double2(r * (a.x + a.y), r * (a.y - a.x));
}
template <bool INV>
__device__ __forceinline__ double2 mul_w16_6(double2 a) {
    const double r = 0.70710678118654752440;
    return INV ? make_double2(-r * (a.x + a.y), r * (a.x - a.y)) : make_double2(r * (a.y - a.x), -r * (a.x + a.y));
}

// natural-order radix-16 DFT as 4 x 4: x[n2 + 4 n1] -> radix-4 over n1, twiddle
// w16^{n2 k1}, radix-4 over n2 -> X[k1 + 4 k2]
template <bool INV>
__device__ __forceinline__ void dft16(double2* v) {
    const double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173;  // cos, sin(pi/8)
    const double sg = INV ? 1.0 : -1.0;
    double2 y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) {
        y[n2][0] = v[n2];
        y[n2][1] = v[n2 + 4];
        y[n2][2] = v[n2 + 8];
        y[n2][3] = v[n2 + 12];
        dft4<INV>(y[n2][0], y[n2][1], y[n2][2], y[n2][3]);
    }
    const double2 w1 = make_double2(c1, sg * s1), w3 = make_double2(s1, sg * c1), w9 = make_double2(-c1, -sg * s1);
    y[1][1] = c_mul(y[1][1], w1);
    y[1][2] = mul_w16_2<INV>(y[1][2]);
    y[1][3] = c_mul(y[1][3], w3);
    y[2][1] = mul_w16_2<INV>(y[2][1]);
    y[2][2] = c_rot<INV>(y[2][2]);  // w16^4 = -i forward
    y[2][3] = mul_w16_6<INV>(y[2][3]);
    y[3][1] = c_mul(y[3][1], w3);
    y[3][2] = mul_w16_6<INV>(y[3][2]);
    y[3][3] = c_mul(y[3][3], w9);
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) {
        dft4<INV>(y[0][k1], y[1][k1], y[2][k1], y[3][k1]);
        v[k1] = y[0][k1];
        v[k1 + 4] = y[1][k1];
        v[k1 + 8] = y[2][k1];
        v[k1 + 12] = y[3][k1];
    }
}

// natural-order radix-7 DFT: with a_n = x_n + x_(7-n), b_n = x_n - x_(7-n) (n = 1, 2, 3),
// X_k = x0 + sum_n cos(2 pi n k/7) a_n -/+ i sum_n sin(2 pi n k/7) b_n for X_k / X_(7-k)
// (forward; the inverse swaps the signs of the i terms)
template <bool INV>
__device__ __forceinline__ void dft7(double2* v) {
    const double c1 = 0.623489801858733530525, c2 = -0.222520933956314404289, c3 = -0.900968867902419126236;
    const double s1 = 0.781831482468029808708, s2 = 0.974927912181823607018, s3 = 0.433883739117558120476;
    const double2 x0 = v[0];
    const double2 a1 = c_add(v[1], v[6]), a2 = c_add(v[2], v[5]), a3 = c_add(v[3], v[4]);
    const double2 b1 = c_sub(v[1], v[6]), b2 = c_sub(v[2], v[5]), b3 = c_sub(v[3], v[4]);
    // cos(2 pi n k/7) and sin(2 pi n k/7) for k = 1, 2, 3 and n = 1, 2, 3
    const double ck[3][3] = {{c1, c2, c3}, {c2, c3, c1}, {c3, c1, c2}};
    const double sk[3][3] = {{s1, s2, s3}, {s2, -s3, -s1}, {s3, -s1, s2}};
    v[0] = c_add(x0, c_add(c_add(a1, a2), a3));
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double2 t = make_double2(fma(ck[k][2], a3.x, fma(ck[k][1], a2.x, fma(ck[k][0], a1.x, x0.x))),
                                       fma(ck[k][2], a3.y, fma(ck[k][1], a2.y, fma(ck[k][0], a1.y, x0.y))));
        const double2 u = make_double2(fma(sk[k][2], b3.x, fma(sk[k][1], b2.x, sk[k][0] * b1.x)),
                                       fma(sk[k][2], b3.y, fma(sk[k][1], b2.y, sk[k][0] * b1.y)));
        const double2 tm = make_double2(t.x + u.y, t.y - u.x), tp = make_double2(t.x - u.y, t.y + u.x);
        v[k + 1] = INV ? tp : tm;  // t - i u forward
        v[6 - k] = INV ? tm : tp;
    }
}

// cos and sin of 2 pi e / R (0 <= e < R) for the twiddles of the composite lengths; e is a
// compile-time constant wherever these are used (unrolled loops), so they fold to literals
template <int R>
__device__ __forceinline__ double2 root_cs(int e) {
    if constexpr (R == 9) {
        constexpr double c[9] = {
            1.0, 0.766044443118978035202, 0.173648177666930348852, -0.5,
            -0.939692620785908384054, -0.939692620785908384054, -0.5, 0.173648177666930348852,
            0.766044443118978035202};
        constexpr double s[9] = {
            0.0, 0.642787609686539326323, 0.984807753012208059367, 0.866025403784438646764,
            0.342020143325668733044, -0.342020143325668733044, -0.866025403784438646764, -0.984807753012208059367,
            -0.642787609686539326323};
        return make_double2(c[e], s[e]);
    } else if constexpr (R == 10) {
        constexpr double c[10] = {
            1.0, 0.809016994374947424102, 0.309016994374947424102, -0.309016994374947424102,
            -0.809016994374947424102, -1.0, -0.809016994374947424102, -0.309016994374947424102,
            0.309016994374947424102, 0.809016994374947424102};
        constexpr double s[10] = {
            0.0, 0.587785252292473129169, 0.951056516295153572116, 0.951056516295153572116,
            0.587785252292473129169, 0.0, -0.587785252292473129169, -0.951056516295153572116,
            -0.951056516295153572116, -0.587785252292473129169};
        return make_double2(c[e], s[e]);
    } else if constexpr (R == 14) {
        constexpr double c[14] = {
            1.0, 0.900968867902419126236, 0.623489801858733530525, 0.222520933956314404289,
            -0.222520933956314404289, -0.623489801858733530525, -0.900968867902419126236, -1.0,
            -0.900968867902419126236, -0.623489801858733530525, -0.222520933956314404289, 0.222520933956314404289,
            0.623489801858733530525, 0.900968867902419126236};
        constexpr double s[14] = {
            0.0, 0.433883739117558120476, 0.781831482468029808708, 0.974927912181823607018,
            0.974927912181823607018, 0.781831482468029808708, 0.433883739117558120476, 0.0,
            -0.433883739117558120476, -0.781831482468029808708, -0.974927912181823607018, -0.974927912181823607018,
            -0.781831482468029808708, -0.433883739117558120476};
        return make_double2(c[e], s[e]);
    } else if constexpr (R == 15) {
        constexpr double c[15] = {
            1.0, 0.913545457642600895502, 0.669130606358858213826, 0.309016994374947424102,
            -0.1045284632676534714, -0.5, -0.809016994374947424102, -0.978147600733805637929,
            -0.978147600733805637929, -0.809016994374947424102, -0.5, -0.1045284632676534714,
            0.309016994374947424102, 0.669130606358858213826, 0.913545457642600895502};
        constexpr double s[15] = {
            0.0, 0.406736643075800207754, 0.743144825477394235015, 0.951056516295153572116,
            0.994521895368273336923, 0.866025403784438646764, 0.587785252292473129169, 0.207911690817759337102,
            -0.207911690817759337102, -0.587785252292473129169, -0.866025403784438646764, -0.994521895368273336923,
            -0.951056516295153572116, -0.743144825477394235015, -0.406736643075800207754};
        return make_double2(c[e], s[e]);
    } else if constexpr (R == 18) {
        constexpr double c[18] = {
            1.0, 0.939692620785908384054, 0.766044443118978035202, 0.5,
            0.173648177666930348852, -0.173648177666930348852, -0.5, -0.766044443118978035202,
            -0.939692620785908384054, -1.0, -0.939692620785908384054, -0.766044443118978035202,
            -0.5, -0.173648177666930348852, 0.173648177666930348852, 0.5,
            0.766044443118978035202, 0.939692620785908384054};
        constexpr double s[18] = {
            0.0, 0.342020143325668733044, 0.642787609686539326323, 0.866025403784438646764,
            0.984807753012208059367, 0.984807753012208059367, 0.866025403784438646764, 0.642787609686539326323,
            0.342020143325668733044, 0.0, -0.342020143325668733044, -0.642787609686539326323,
            -0.866025403784438646764, -0.984807753012208059367, -0.984807753012208059367, -0.866025403784438646764,
            -0.642787609686539326323, -0.342020143325668733044};
        return make_double2(c[e], s[e]);
    } else if constexpr (R == 20) {
        constexpr double c[20] = {
            1.0, 0.951056516295153572116, 0.809016994374947424102, 0.587785252292473129169,
            0.309016994374947424102, 0.0, -0.309016994374947424102, -0.587785252292473129169,
            -0.809016994374947424102, -0.951056516295153572116, -1.0, -0.951056516295153572116,
            -0.809016994374947424102, -0.587785252292473129169, -0.309016994374947424102, 0.0,
            0.309016994374947424102, 0.587785252292473129169, 0.809016994374947424102, 0.951056516295153572116};
        constexpr double s[20] = {
            0.0, 0.309016994374947424102, 0.587785252292473129169, 0.809016994374947424102,
            0.951056516295153572116, 1.0, 0.951056516295153572116, 0.809016994374947424102,
            0.587785252292473129169, 0.309016994374947424102, 0.0, -0.309016994374947424102,
            -0.587785252292473129169, -0.809016994374947424102, -0.951056516295153572116, -1.0,
            -0.951056516295153572116, -0.809016994374947424102, -0.587785252292473129169, -0.309016994374947424102};
        return make_double2(c[e], s[e]);
    } else if constexpr (R == 24) {
        constexpr double c[24] = {
            1.0, 0.96592582628906828675, 0.866025403784438646764, 0.707106781186547524401,
            0.5, 0.258819045102520762349, 0.0, -0.258819045102520762349,
            -0.5, -0.707106781186547524401, -0.866025403784438646764, -0.96592582628906828675,
            -1.0, -0.96592582628906828675, -0.866025403784438646764, -0.707106781186547524401,
            -0.5, -0.258819045102520762349, 0.0, 0.258819045102520762349,
            0.5, 0.707106781186547524401, 0.866025403784438646764, 0.96592582628906828675};
        constexpr double s[24] = {
            0.0, 0.258819045102520762349, 0.5, 0.707106781186547524401,
            0.866025403784438646764, 0.96592582628906828675, 1.0, 0.96592582628906828675,
            0.866025403784438646764, 0.707106781186547524401, 0.5, 0.258819045102520762349,
            0.0, -0.258819045102520762349, -0.5, -0.707106781186547524401,
            -0.866025403784438646764, -0.96592582628906828675, -1.0, -0.96592582628906828675,
            -0.866025403784438646764, -0.707106781186547524401, -0.5, -0.258819045102520762349};
        return make_double2(c[e], s[e]);
    } else if constexpr (R == 25) {
        constexpr double c[25] = {
            1.0, 0.96858316112863111949, 0.876306680043863587308, 0.728968627421411523147,
            0.535826794978996618271, 0.309016994374947424102, 0.0627905195293133760762, -0.187381314585724630543,
            -0.425779291565072648863, -0.637423989748689710177, -0.809016994374947424102, -0.929776485888251403661,
            -0.99211470131447783105, -0.99211470131447783105, -0.929776485888251403661, -0.809016994374947424102,
            -0.637423989748689710177, -0.425779291565072648863, -0.187381314585724630543, 0.0627905195293133760762,
            0.309016994374947424102, 0.535826794978996618271, 0.728968627421411523147, 0.876306680043863587308,
            0.96858316112863111949};
        constexpr double s[25] = {
            0.0, 0.248689887164854788242, 0.481753674101715274987, 0.684547105928688673732,
            0.844327925502015078549, 0.951056516295153572116, 0.998026728428271561952, 0.982287250728688681086,
            0.904827052466019527714, 0.770513242775789230803, 0.587785252292473129169, 0.368124552684677959157,
            0.125333233564304245373, -0.125333233564304245373, -0.368124552684677959157, -0.587785252292473129169,
            -0.770513242775789230803, -0.904827052466019527714, -0.982287250728688681086, -0.998026728428271561952,
            -0.951056516295153572116, -0.844327925502015078549, -0.684547105928688673732, -0.481753674101715274987,
            -0.248689887164854788242};
        return make_double2(c[e], s[e]);
    } else if constexpr (R == 6) {
        constexpr double c[6] = {
            1.0, 0.5, -0.5, -1.0,
            -0.5, 0.5};
        constexpr double s[6] = {
            0.0, 0.866025403784438646764, 0.866025403784438646764, 0.0,
            -0.866025403784438646764, -0.866025403784438646764};
        return make_double2(c[e], s[e]);
    } else {
        static_assert(R == 9, "no twiddle table for this length");
        return make_double2(1.0, 0.0);
    }
}

template <int R, bool INV>
__device__ __forceinline__ void dft(double2* v);

// natural-order DFT of length A B (Cooley-Tukey): x[n2 + B n1] -> length-A DFTs over n1,
// twiddle w_(AB)^(n2 k1), length-B DFTs over n2 -> X[k1 + A k2]
template <int A, int B, bool INV>
__device__ __forceinline__ void dft_ct(double2* v) {
    double2 y[B][A];
#pragma unroll
    for (int n2 = 0; n2 < B; n2++) {
#pragma unroll
        for (int n1 = 0; n1 < A; n1++) y[n2][n1] = v[n2 + B * n1];
        dft<A, INV>(y[n2]);
    }
#pragma unroll
    for (int n2 = 1; n2 < B; n2++)
#pragma unroll
        for (int k1 = 1; k1 < A; k1++) {
            const double2 w = root_cs<A * B>(n2 * k1);
            y[n2][k1] = INV ? c_mul(y[n2][k1], w) : c_mulc(y[n2][k1], w);
        }
#pragma unroll
    for (int k1 = 0; k1 < A; k1++) {
        double2 z[B];
#pragma unroll
        for (int n2 = 0; n2 < B; n2++) z[n2] = y[n2][k1];
        dft<B, INV>(z);
#pragma unroll
        for (int k2 = 0; k2 < B; k2++) v[k1 + A * k2] = z[k2];
    }
}

// Inverse DFT of length R of v_n = y_n g_n (g real). For R = 4 and 8 the products are
// folded into the first butterfly stage (pairs n, n + R/2) as fma, then the DFT continues
// as dft<R, true>; other lengths multiply first.
template <int R>
__device__ __forceinline__ void inv_dft_g(double2* y, const double* g) {
    if constexpr (R != 4 && R != 8) {
#pragma unroll
        for (int n = 0; n < R; n++) y[n] = make_double2(y[n].x * g[n], y[n].y * g[n]);
        dft<R, true>(y);
    } else {
        constexpr int Hh = R / 2;
        double2 a[Hh], b[Hh];
#pragma unroll
        for (int n = 0; n < Hh; n++) {
            const double2 t = make_double2(y[n + Hh].x * g[n + Hh], y[n + Hh].y * g[n + Hh]);
            a[n] = make_double2(fma(y[n].x, g[n], t.x), fma(y[n].y, g[n], t.y));
            b[n] = make_double2(fma(y[n].x, g[n], -t.x), fma(y[n].y, g[n], -t.y));
        }
        if constexpr (R == 8) {
            dft8_tail<true>(a, b, y);
        } else {
            // radix 4: a = (s02, s13), b = (d02, v1 - v3)
            const double2 d13 = c_rot<true>(b[1]);
            y[0] = c_add(a[0], a[1]);
            y[2] = c_sub(a[0], a[1]);
            y[1] = c_add(b[0], d13);
            y[3] = c_sub(b[0], d13);
        }
    }
}

template <int R, bool INV>
__device__ __forceinline__ void dft(double2* v) {
    if constexpr (R == 25) dft_ct<5, 5, INV>(v);
    else if constexpr (R == 24) dft_ct<8, 3, INV>(v);
    else if constexpr (R == 20) dft_ct<4, 5, INV>(v);
    else if constexpr (R == 18) dft_ct<2, 9, INV>(v);
    else if constexpr (R == 16) dft16<INV>(v);
    else if constexpr (R == 15) dft_ct<3, 5, INV>(v);
    else if constexpr (R == 14) dft_ct<2, 7, INV>(v);
    else if constexpr (R == 12) dft12<INV>(v);
    else if constexpr (R == 10) dft_ct<2, 5, INV>(v);
    else if constexpr (R == 9) dft_ct<3, 3, INV>(v);
    else if constexpr (R == 8) dft8<INV>(v);
    else if constexpr (R == 7) dft7<INV>(v);
    else if constexpr (R == 6) dft_ct<2, 3, INV>(v);
    else if constexpr (R == 5) dft5<INV>(v);
    else if constexpr (R == 4) dft4<INV>(v[0], v[1], v[2], v[3]);
    else if constexpr (R == 3) dft3<INV>(v[0], v[1], v[2]);
    else if constexpr (R == 2) dft2<INV>(v[0], v[1]);
    else static_assert(R == 2, "no register DFT of this length");
}

// Stage-2 DFTs (of length R2, R1 of them) per thread: up to ceil(R1 / R2). When R2
// divides R1, thread j takes DFTs j U .. j U + U - 1; otherwise DFT u R2 + j for each u
// with u R2 + j < R1 (some threads take one fewer, or none).
template <int R1, int R2>
constexpr int stage2_per_thread() { return (R1 + R2 - 1) / R2; }
template <int R1, int R2>
__device__ __forceinline__ int stage2_dft(int j, int u) {
    return (R1 % R2 == 0) ? j * (R1 / R2) + u : u * R2 + j;
}

// Twiddles w_n^{jk} (k < R1) and kernel values g(jU + u + R1 f) of the line lane j a
// thread serves: held in registers (loaded once per pass) or read from the workgroup's
// LDS copy of the axis tables when used (fewer registers, so more waves per SIMD where a
// pass is register-bound).
template <int R1, int R2>
struct RegTab {
    static constexpr int U = stage2_per_thread<R1, R2>();
    double2 tw[R1];
    double g[U * R2];
    __device__ __forceinline__ void load(const double2* __restrict__ twt, const double* __restrict__ gt, int j) {
#pragma unroll
        for (int k = 0; k < R1; k++) tw[k] = ld2(twt + j * k);
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int d0 = stage2_dft<R1, R2>(j, u);
            const int d = d0 < R1 ? d0 : 0;  // a slot without a DFT reads DFT 0's values
#pragma unroll
            for (int f = 0; f < R2; f++) g[u * R2 + f] = ldg(gt + d + R1 * f);
        }
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
    __device__ __forceinline__ double gv(int u, int f) const { return g[stage2_dft<R1, R2>(j, u) + R1 * f]; }
};

// 1D circular convolution of one complex line of length N = R1*R2 held by R2 threads
// (this thread: index j, elements z[k] = x[j + R2 k]). lx is the line's LDS exchange row
// (R1*R2 complex); tab supplies the thread's twiddles exp(-2 pi i jk/N) and kernel values.
// The thread writes its column (entries k Q + j), reads back rows, rewrites them and reads
// its column again. SYNC_IN: wait for the workgroup before the first write, because the
// exchange rows were last used otherwise; without it the last LDS accesses to them must
// have been the same threads' reads of their own columns (a convolution of the same lines)
// or none.
template <int R1, int R2, bool SYNC_IN, class Tab>
__device__ __forceinline__ void conv_line(double2* z, double2* lx, int j, const Tab& tab) {
    constexpr int U = stage2_per_thread<R1, R2>();
    constexpr int Q = R2 + 1;  // padded run length in LDS
    dft<R1, false>(z);
#pragma unroll
    for (int k = 1; k < R1; k++) z[k] = c_mul(z[k], tab.w(k));
    if constexpr (SYNC_IN) lds_barrier();  // previous readers of the exchange rows are done
#pragma unroll
    for (int k = 0; k < R1; k++) lx[k * Q + j] = z[k];
    lds_barrier();
    double2 y[U * R2];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int d = stage2_dft<R1, R2>(j, u);
        if (R1 % R2 == 0 || d < R1) {
#pragma unroll
            for (int jj = 0; jj < R2; jj++) y[u * R2 + jj] = lx[d * Q + jj];
            dft<R2, false>(y + u * R2);
            double gl[R2];
#pragma unroll
            for (int f = 0; f < R2; f++) gl[f] = tab.gv(u, f);
            inv_dft_g<R2>(y + u * R2, gl);
        }
    }
    // no barrier: each row d is read and rewritten by one thread only (d = stage2_dft(j, u)
    // takes every row once over the line's threads), and a wavefront's LDS accesses are
    // performed in order
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int d = stage2_dft<R1, R2>(j, u);
        if (R1 % R2 == 0 || d < R1)
#pragma unroll
            for (int jj = 0; jj < R2; jj++) lx[d * Q + jj] = y[u * R2 + jj];
    }
    lds_barrier();
#pragma unroll
    for (int k = 0; k < R1; k++) z[k] = lx[k * Q + j];
#pragma unroll
    for (int k = 1; k < R1; k++) z[k] = c_mulc(z[k], tab.w(k));
    dft<R1, true>(z);
}

// Sum of one plane of field 0 (the chain's last pass: q_N, whose sum gives Q) into *out,
// in an order fixed by the field-local thread index t alone, so every kernel and variant
// (one or two fields per workgroup) gives the same bits: each thread has added its R1
// values in k order (s), threads t < 32 add every 32nd thread's sum through LDS (red: TPF
// doubles of the field's own plane buffer), and a width-32 shuffle tree adds theirs.
// Every thread of the workgroup takes part in the two barriers.
template <int TPF>
__device__ __forceinline__ void plane_qsum(double s, double* red, int t, int f, double* out) {
    static_assert(TPF >= 32, "a field has at least 32 threads");
    lds_barrier();  // the last readers of the exchange rows are done
    red[t] = s;
    lds_barrier();
    if (f == 0 && t < 32) {  // threads 0..31 of field 0: one wavefront's lanes 0..31
        double u = red[t];
        for (int p = t + 32; p < TPF; p += 32) u += red[p];
#pragma unroll
        for (int o = 16; o > 0; o >>= 1) u += __shfl_down(u, o, 32);
        if (t == 0) *out = u;
    }
}

}  // namespace

// Row pitch (complex entries) of the exchange rows of lines whose R2 = 8 threads are
// consecutive lanes: the first pitch from the line's extent, (R1 - 1)(R2 + 1) + R2, that is
// 8 mod 16, so neighbouring lines start 32 banks apart (an odd pitch put them on the same
// banks); other shapes keep the odd pitch xp.
constexpr int zline_pitch(int r1, int r2, int xp) {
    return r2 == 8 ? ((r1 - 1) * (r2 + 1) + r2) + ((24 - ((r1 - 1) * (r2 + 1) + r2) % 16) % 16) : xp;
}

// LDS of one slab pass shape: FPW exchange/plane buffers (the padded exchange rows of the
// line DFTs, reused for the re-paired real plane) and, with TWL = 1, the tables. The
// kernel owns the array (a kernel that runs several shapes sizes it for the largest) and
// hands it to slab_pass.
template <int E, int R1, int R2, int TWL, int FPW>
struct SlabLds {
    static constexpr int H = E / 2;                      // complex lines per plane
    static constexpr int XP = (R1 * (R2 + 1)) | 1;       // exchange row pitch (complex): padded line, odd
    // z-line exchange rows (zline_pitch): 40 for 24 and 32 points, 56 for 40 and 48, 72 for
    // 56 and 64, within the plane buffer's size; 120, 136, 136 and 152 for 96, 112, 120 and
    // 128 points, which grows the largest workgroup's LDS to 159744 bytes (128 points,
    // tables included).
    static constexpr int XPZ = zline_pitch(R1, R2, XP);
    static constexpr int RP = E + ((E + 8) % 16 == 8 ? 8 : 16);  // real plane row pitch (doubles), = 8 mod
                                                                 // 16: rows r, r+1 on other banks
    static constexpr int XBYTES = H * (XP > XPZ ? XP : XPZ) * 16;
    static constexpr int RBYTES = E * RP * 8;
    static constexpr int BYTES = XBYTES > RBYTES ? XBYTES : RBYTES;
    static constexpr int TBYTES = TWL ? E * 16 + 2 * E * 8 : 0;
    static constexpr int TOTAL = FPW * BYTES + TBYTES;
    static constexpr int THREADS = H * R2 * FPW;        // workgroup size of the shape
};

// One slab pass of one replica (plane `plane`), addresses in `io`, LDS `smem` (SlabLds
// TOTAL bytes). The workgroup serves field `fidx`, or with FPW = 2 both fields (threads
// [0, TPF) field 0, [TPF, 2 TPF) field 1). TWL = 1 keeps the twiddle and kernel tables in
// LDS instead of registers.
// MODE selects the finishing rule of pair i+1 from u = (C x)(r):
//   CHAIN_PLAIN    out = h u                   (src/step.cc:53-64)
//   CHAIN_TANGENT  out = h u + aux src         (directional derivative of the chain)
//   CHAIN_ADJOINT  out = u + aux src, stored; the chain continues with h out
// Every variant performs the same arithmetic on every element: results are bit-identical.
// FUSE = 1 (plain chain, io.acc set): the pass also adds its new values times the
// partner values to the density sums (acc = fma(new, partner, acc), or the product on
// first touch). Only the one-replica kernel uses it: it runs one wave per SIMD, so the
// extra live registers cost nothing, while the replica-set kernel needs them for
// occupancy (and gains nothing there: the bytes moved are the same either way).
// The loads are unconditional (a pass without a finish reads its own input in place of
// h and the derivative factors), so the compiler can count them exactly and the first
// convolution waits only for what it reads.
// FIN = 0 only for a chain's first pass (no step to finish): fixed per kernel, so the
// code is straight and the compiler's wait counts stay exact (a run-time branch around
// the finish would make it wait for every store before the next convolution).
// TNT: the partially convolved pair T goes out with non-temporal stores (one replica:
// a cached line left dirty in L2 is written back when the kernel ends, which lengthens
// the gap to the dependent next pass) or with ordinary stores (replica set: its T of all
// members, 122 MB for the 64-member workload, is read back by the next launch, and
// cached stores made the 64^3 and 48^3 group passes 9-13 % faster).
// LATE_SKIP: test the skip flag once the pass's loads are issued (they read only its own
// in-bounds operands), so that the kernel's first argument loads are not waited for before
// the rest; otherwise before anything (a replica set's launch, whose skipped members should
// not read their fields).
template <int E, int R1, int R2, int MODE, int TWL, int FPW, int FUSE, int FIN, bool TNT = true,
          bool LATE_SKIP = false>
__device__ __forceinline__ void slab_pass(const SlabIO& io, int plane, const double2* __restrict__ tw,
                                          unsigned char* smem, int fidx) {
    static_assert(R1 * R2 == E, "unsupported factorisation");
    using Lds = SlabLds<E, R1, R2, TWL, FPW>;
    constexpr int H = Lds::H;
    constexpr int TPF = H * R2;        // threads per field
    constexpr int XP = Lds::XP;
    constexpr int XPZ = Lds::XPZ;
    constexpr int RP = Lds::RP;
    constexpr int BYTES = Lds::BYTES;
    constexpr bool EXTRA = MODE != CHAIN_PLAIN;
    constexpr bool PLAIN = MODE == CHAIN_PLAIN && FUSE;
    if constexpr (!LATE_SKIP)
        if (io.skip && *io.skip) return;

    const bool xslab = io.xslab != 0;
    constexpr bool finish = FIN != 0;
    const bool cont = io.cont != 0;
    const long M = io.M;
    const long plane_stride = xslab ? (long)E * E : E;
    const long Sa = xslab ? E : (long)E * E;
    const double* ga = xslab ? io.g[1] : io.g[0];
    const double* gb = io.g[2];

    const int fl = (FPW == 2) ? (int)(threadIdx.x / TPF) : 0;
    const int t = threadIdx.x - fl * TPF;
    const int f = (FPW == 2) ? fl : fidx;
    double2* xbuf = reinterpret_cast<double2*>(smem + fl * BYTES);
    double* rp = reinterpret_cast<double*>(smem + fl * BYTES);
    const long pbase = (long)plane * plane_stride;  // plane origin within a field
    const long base = (long)f * M + pbase;
    const int c = t % H, j = t / H;    // layout A: a-lines, column pair (2c, 2c + 1), a = j + R2 k
    const int r = t / R2, jb = t % R2; // layout B: z-lines, row pair (r, r + H), b = jb + R2 k

    // everything this thread reads is issued up front, in the order it is needed (loads
    // complete in order): the partial pair, the twiddle and kernel tables, then for the
    // finish the Boltzmann factor of its field at this bead (src/step.cc:53-64), the
    // derivative factors and the density-sum operands
    const double* Tin = io.tin + (long)f * M;
    const double* hsel = finish ? (f ? io.h[1] : io.h[0]) : Tin;
    const double* asel = finish ? (f ? io.aux[1] : io.aux[0]) : Tin;
    const double* ssel = finish ? (f ? io.src[1] : io.src[0]) : Tin;
    double2 z[R1], hz[R1], az[EXTRA ? R1 : 1], sz[EXTRA ? R1 : 1];
    // fused density sums: partner values and running sums of this field (a pass that
    // fuses always finishes a step)
    const double* psel = nullptr;
    double* accp = nullptr;
    bool first = false;
    if constexpr (PLAIN) {
        psel = f ? io.part[1] : io.part[0];
        accp = f ? io.acc[1] : io.acc[0];
        first = f ? io.first[1] != 0 : io.first[0] != 0;
    }
    double2 pz[PLAIN ? R1 : 1], oz[PLAIN ? R1 : 1];
    // the partial pair is read once, non-temporally (5 % shorter 64^3 passes for one
    // replica, 2.6 % more set throughput, than through the caches)
#pragma unroll
    for (int k = 0; k < R1; k++) z[k] = ld2_nt(Tin + pbase + (long)(j + R2 * k) * Sa + 2 * c);
    using Tab = typename std::conditional<TWL != 0, LdsTab<R1, R2>, RegTab<R1, R2>>::type;
    Tab tabA, tabB;
    // LDS tables: one entry per thread (the workgroup has at least E threads)
    const bool tl = TWL != 0 && threadIdx.x < E;
    double2 tv{};
    double gav = 0.0, gbv = 0.0;
    if constexpr (TWL != 0) {
        // (inside the branch the compiler loads the table pointers after the partial pair's
        // loads are out: loading every argument first, with unconditional table loads,
        // measured 8.54 against 8.37 us per contour step)
        if (tl) {
            tv = ld2(tw + threadIdx.x);
            gav = ldg(ga + threadIdx.x);
            gbv = ldg(gb + threadIdx.x);
        }
    } else {
        tabA.load(tw, ga, j);
        tabB.load(tw, gb, jb);
    }
    // operands of the finish: issued now, or (derivative passes on 16-element lines,
    // where registers are short) after the first convolution
    constexpr bool EARLY = R1 <= 12 || (MODE == CHAIN_PLAIN && !FUSE);
    auto load_finish = [&]() {
#pragma unroll
        for (int k = 0; k < R1; k++) {
            const long p = pbase + (long)(j + R2 * k) * Sa + 2 * c;
            hz[k] = ld2(hsel + p);
            if constexpr (EXTRA) {
                az[k] = ld2(asel + p);
                sz[k] = ld2(ssel + p);
            }
        }
        if constexpr (PLAIN) {
#pragma unroll
            for (int k = 0; k < R1; k++) {
                const long p = pbase + (long)(j + R2 * k) * Sa + 2 * c;
                pz[k] = ld2_nt(psel + p);
                oz[k] = ld2_nt(accp + p);  // unused on first touch
            }
        }
    };
    if constexpr (EARLY) load_finish();
    if constexpr (LATE_SKIP)
        if (io.skip && *io.skip) return;  // the whole workgroup: no LDS written, no barrier taken
    if constexpr (TWL != 0) {
        double2* stw = reinterpret_cast<double2*>(smem + FPW * BYTES);
        double* sga = reinterpret_cast<double*>(smem + FPW * BYTES + E * 16);
        double* sgb = sga + E;
        if (tl) {
            stw[threadIdx.x] = tv;
            sga[threadIdx.x] = gav;
            sgb[threadIdx.x] = gbv;
        }
        lds_barrier();
        tabA.load(stw, sga, j);
        tabB.load(stw, sgb, jb);
    }

    // the chain's last pass adds up its new field-0 values for Q (plane_qsum)
    const bool qsum = MODE == CHAIN_PLAIN && finish && !cont && io.qsum;
    double qs = 0.0;
    if constexpr (finish) {
        conv_line<R1, R2, false>(z, xbuf + c * XP, j, tabA);  // the exchange rows' first use
        if constexpr (!EARLY) load_finish();
        double* qout = io.qout;
#pragma unroll
        for (int k = 0; k < R1; k++) {
            const long p = base + (long)(j + R2 * k) * Sa + 2 * c;
            if constexpr (MODE == CHAIN_PLAIN) {
                z[k].x = rmul(z[k].x, hz[k].x);
                z[k].y = rmul(z[k].y, hz[k].y);
                if constexpr (PLAIN) {
                    const double ax = z[k].x * pz[k].x, ay = z[k].y * pz[k].y;
                    const double bx = fma(z[k].x, pz[k].x, oz[k].x), by = fma(z[k].y, pz[k].y, oz[k].y);
                    st2_nt(accp + (p - (long)f * M), first ? make_double2(ax, ay) : make_double2(bx, by));
                }
            } else if constexpr (MODE == CHAIN_TANGENT) {
                z[k].x = fma(z[k].x, hz[k].x, az[k].x * sz[k].x);
                z[k].y = fma(z[k].y, hz[k].y, az[k].y * sz[k].y);
            } else {
                z[k].x = fma(az[k].x, sz[k].x, z[k].x);
                z[k].y = fma(az[k].y, sz[k].y, z[k].y);
            }
            if (!PLAIN || qout) st2_nt(qout + p, z[k]);
            if (qsum) {
                qs = (k == 0) ? z[k].x : qs + z[k].x;
                qs += z[k].y;
            }
            if constexpr (MODE == CHAIN_ADJOINT) {
                z[k].x *= hz[k].x;
                z[k].y *= hz[k].y;
            }
        }
    }
    if (!cont) {
        if (qsum) plane_qsum<TPF>(qs, rp, t, f, io.qsum + plane);
        return;
    }

    conv_line<R1, R2, false>(z, xbuf + c * XP, j, tabA);  // first use, or the same lines again
    // re-pair the real plane: rows (r, r + H) become the complex z-lines
    lds_barrier();
#pragma unroll
    for (int k = 0; k < R1; k++) {
        const int ai = j + R2 * k;
        *reinterpret_cast<double2*>(rp + ai * RP + 2 * c) = z[k];
    }
    lds_barrier();
    double2 w[R1];
#pragma unroll
    for (int k = 0; k < R1; k++) {
        const int b = jb + R2 * k;
        w[k] = make_double2(rp[r * RP + b], rp[(r + H) * RP + b]);
    }
    conv_line<R1, R2, true>(w, xbuf + r * XPZ, jb, tabB);  // after the re-pairing reads
    double* Tout = io.tout;
#pragma unroll
    for (int k = 0; k < R1; k++) {
        const long p = base + (long)r * Sa + jb + R2 * k;
        if constexpr (TNT) {
            st_nt(Tout + p, w[k].x);
            st_nt(Tout + p + (long)H * Sa, w[k].y);
        } else {
            st_g(Tout + p, w[k].x);
            st_g(Tout + p + (long)H * Sa, w[k].y);
        }
    }
}

// Plain chains of a replica set. Workgroup (slot * E + plane, field): slot picks the
// replica descriptor.
template <int E, int R1, int R2, int TWL, int FPW, int FIN>
__global__ __launch_bounds__((E / 2) * R2 * FPW) void k_slab(const Rep* __restrict__ reps,
                                                           const double2* __restrict__ tw, int i) {
    __shared__ __align__(16) unsigned char smem[SlabLds<E, R1, R2, TWL, FPW>::TOTAL];
    const int slot = blockIdx.x / E, plane = blockIdx.x - slot * E;
    const Rep& R = reps[slot];  // fields read in place (a private copy would live in scratch)
    if (i > R.N - 1) return;    // this replica's chain is complete
    const SlabIO io = plain_io(R, i);
    slab_pass<E, R1, R2, CHAIN_PLAIN, TWL, FPW, 0, FIN, false>(io, plane, tw, smem, (int)blockIdx.y);
}

// Plain chains of two groups of a replica set in one launch: group A (edge EA, one field
// per workgroup: workgroup (slot * EA + plane) * 2 + field) first, then group B (edge EB,
// both fields per workgroup: workgroup slot * EB + plane). The two shapes have the same
// workgroup size, so every thread of a workgroup runs its group's pass and its barriers.
// Each plane runs the same slab_pass as in k_slab: bit-identical results. A small group
// whose passes are latency-bound (32^3) then rides in the launches of a large,
// bandwidth-bound one (64^3) instead of costing a launch of its own per contour step.
template <int EA, int RA1, int RA2, int TWLA, int EB, int RB1, int RB2, int TWLB, int FIN>
__global__ __launch_bounds__((EA / 2) * RA2) void k_slab_pair(const Rep* __restrict__ repsA, int nA,
                                                             const Rep* __restrict__ repsB,
                                                             const double2* __restrict__ twA,
                                                             const double2* __restrict__ twB, int i) {
    using LA = SlabLds<EA, RA1, RA2, TWLA, 1>;
    using LB = SlabLds<EB, RB1, RB2, TWLB, 2>;
    static_assert(LA::THREADS == LB::THREADS, "the two shapes must have the same workgroup size");
    __shared__ __align__(16) unsigned char smem[LA::TOTAL > LB::TOTAL ? LA::TOTAL : LB::TOTAL];
    const int nwa = nA * EA * 2;
    if ((int)blockIdx.x < nwa) {
        const int slot = blockIdx.x / (EA * 2), rest = blockIdx.x - slot * (EA * 2);
        const Rep& R = repsA[slot];
        if (i > R.N - 1) return;  // whole workgroup: no barrier is left waiting
        const SlabIO io = plain_io(R, i);
        slab_pass<EA, RA1, RA2, CHAIN_PLAIN, TWLA, 1, 0, FIN, false>(io, rest >> 1, twA, smem, rest & 1);
    } else {
        const int b = blockIdx.x - nwa;
        const int slot = b / EB, plane = b - slot * EB;
        const Rep& R = repsB[slot];
        if (i > R.N - 1) return;
        const SlabIO io = plain_io(R, i);
        slab_pass<EB, RB1, RB2, CHAIN_PLAIN, TWLB, 2, 0, FIN, false>(io, plane, twB, smem, 0);
    }
}

// One replica, its pass addresses in the kernel arguments (no dependent descriptor load
// before the first field loads).
// One replica's pass has E * 2 / FPW workgroups, at most one per CU: the compiler may
// schedule for one wave per SIMD (registers are not the limit there).
template <int E, int R1, int R2, int MODE, int TWL, int FPW, int FUSE, int FIN>
__global__ __launch_bounds__((E / 2) * R2 * FPW) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_slab_io(const SlabIO io, const double2* __restrict__ tw) {
    __shared__ __align__(16) unsigned char smem[SlabLds<E, R1, R2, TWL, FPW>::TOTAL];
    slab_pass<E, R1, R2, MODE, TWL, FPW, FUSE, FIN, true, true>(io, blockIdx.x, tw, smem, (int)blockIdx.y);
}

namespace {

template <int E, int R1, int R2, int TWL, int FPW>
void launch_set(const Rep* reps, int nact, int i, const double2* tw, hipStream_t st) {
    const dim3 grid(nact * E, 2 / FPW), block((E / 2) * R2 * FPW);
    if (i >= 1) hipLaunchKernelGGL((k_slab<E, R1, R2, TWL, FPW, 1>), grid, block, 0, st, reps, tw, i);
    else hipLaunchKernelGGL((k_slab<E, R1, R2, TWL, FPW, 0>), grid, block, 0, st, reps, tw, i);
}

template <int E, int R1, int R2, int TWL, int FPW>
void launch_one(int mode, const SlabIO& io, const double2* tw, hipStream_t st) {
    const dim3 grid(E, 2 / FPW), block((E / 2) * R2 * FPW);
    const bool fin = io.finish != 0;
    bool fused = false;
    if constexpr (E <= 64) {  // density sums are fused on meshes up to 64^3 only
        if (mode == CHAIN_PLAIN && io.acc[0] && fin) {
            hipLaunchKernelGGL((k_slab_io<E, R1, R2, CHAIN_PLAIN, TWL, FPW, 1, 1>), grid, block, 0, st, io, tw);
            fused = true;
        }
    }
    if (fused)
        return;
    if (mode == CHAIN_PLAIN && fin)
        hipLaunchKernelGGL((k_slab_io<E, R1, R2, CHAIN_PLAIN, TWL, FPW, 0, 1>), grid, block, 0, st, io, tw);
    else if (mode == CHAIN_PLAIN)
        hipLaunchKernelGGL((k_slab_io<E, R1, R2, CHAIN_PLAIN, TWL, FPW, 0, 0>), grid, block, 0, st, io, tw);
    else if (mode == CHAIN_TANGENT && fin)
        hipLaunchKernelGGL((k_slab_io<E, R1, R2, CHAIN_TANGENT, TWL, FPW, 0, 1>), grid, block, 0, st, io, tw);
    else if (mode == CHAIN_TANGENT)
        hipLaunchKernelGGL((k_slab_io<E, R1, R2, CHAIN_TANGENT, TWL, FPW, 0, 0>), grid, block, 0, st, io, tw);
    else if (fin)
        hipLaunchKernelGGL((k_slab_io<E, R1, R2, CHAIN_ADJOINT, TWL, FPW, 0, 1>), grid, block, 0, st, io, tw);
    else
        hipLaunchKernelGGL((k_slab_io<E, R1, R2, CHAIN_ADJOINT, TWL, FPW, 0, 0>), grid, block, 0, st, io, tw);
}

// plain chains only (the alternative factorisations: contour steps of calc_concs and the
// Anderson iterations, not the derivative chains)
template <int E, int R1, int R2, int TWL, int FPW>
void launch_one_plain(const SlabIO& io, const double2* tw, hipStream_t st) {
    const dim3 grid(E, 2 / FPW), block((E / 2) * R2 * FPW);
    if (io.acc[0] && io.finish)
        hipLaunchKernelGGL((k_slab_io<E, R1, R2, CHAIN_PLAIN, TWL, FPW, 1, 1>), grid, block, 0, st, io, tw);
    else if (io.finish)
        hipLaunchKernelGGL((k_slab_io<E, R1, R2, CHAIN_PLAIN, TWL, FPW, 0, 1>), grid, block, 0, st, io, tw);
    else
        hipLaunchKernelGGL((k_slab_io<E, R1, R2, CHAIN_PLAIN, TWL, FPW, 0, 0>), grid, block, 0, st, io, tw);
}

// MASK: the variants this shape has (bit v); the others spill registers and are not built
template <int E, int R1, int R2, int MASK>
hipError_t dispatch_plain(int v, const Rep* reps, int nact, int i, const SlabIO* io, const double2* tw,
                          hipStream_t st) {
    if (v < 0 || v > 3 || !((MASK >> v) & 1)) return hipErrorInvalidValue;
    if constexpr ((MASK & 1) != 0)
        if (v == 0) io ? launch_one_plain<E, R1, R2, 0, 1>(*io, tw, st) : launch_set<E, R1, R2, 0, 1>(reps, nact, i, tw, st);
    if constexpr ((MASK & 2) != 0)
        if (v == 1) io ? launch_one_plain<E, R1, R2, 1, 1>(*io, tw, st) : launch_set<E, R1, R2, 1, 1>(reps, nact, i, tw, st);
    if constexpr ((MASK & 4) != 0)
        if (v == 2) io ? launch_one_plain<E, R1, R2, 0, 2>(*io, tw, st) : launch_set<E, R1, R2, 0, 2>(reps, nact, i, tw, st);
    if constexpr ((MASK & 8) != 0)
        if (v == 3) io ? launch_one_plain<E, R1, R2, 1, 2>(*io, tw, st) : launch_set<E, R1, R2, 1, 2>(reps, nact, i, tw, st);
    return hipGetLastError();
}

// variant v: bit 0 = tables in LDS, bit 1 = both fields in one workgroup
template <int E, int R1, int R2>
void dispatch(int v, const Rep* reps, int nact, int i, int mode, const SlabIO* io, const double2* tw,
              hipStream_t st) {
    if constexpr (E >= 72) {  // one shape only (see dispatch_edge)
        (void)v;
        io ? launch_one<E, R1, R2, 1, 1>(mode, *io, tw, st) : launch_set<E, R1, R2, 1, 1>(reps, nact, i, tw, st);
    } else {
        switch (v) {
            case 1: io ? launch_one<E, R1, R2, 1, 1>(mode, *io, tw, st) : launch_set<E, R1, R2, 1, 1>(reps, nact, i, tw, st); break;
            case 2: io ? launch_one<E, R1, R2, 0, 2>(mode, *io, tw, st) : launch_set<E, R1, R2, 0, 2>(reps, nact, i, tw, st); break;
            case 3: io ? launch_one<E, R1, R2, 1, 2>(mode, *io, tw, st) : launch_set<E, R1, R2, 1, 2>(reps, nact, i, tw, st); break;
            default: io ? launch_one<E, R1, R2, 0, 1>(mode, *io, tw, st) : launch_set<E, R1, R2, 0, 1>(reps, nact, i, tw, st); break;
        }
    }
}

}  // namespace

}  // namespace lfts

#endif  // LFTS_SLAB_H
