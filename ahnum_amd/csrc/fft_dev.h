// Device code of the gfx950 squaring engine: FP64 butterflies, Stockham FFTs in LDS, the
// real-split square/multiply, the balanced carry with the top fold, and the three pass kernels
// as templates over the shape. A template argument of 0 means "read the size from the shape at
// run time" (the generic kernels); a non-zero size makes every loop bound, stride, division and
// radix a compile-time constant (the fast kernels, instantiated in kernels_fast.hip).
//
//   rows pass      per row pair (k_r, R-k_r): row FFT (length C) -> real-split square or
//                  multiply -> inverse row FFT   (the reference's pass 2: src/neon.c:987-1054)
//   cols_inv pass  per column group: conj four-step twiddle -> inverse column FFT (length R) ->
//                  scale k/(4H w_j), round, max round-off, int64 carry over 16-digit sub-chunks,
//                  the 128-bit top fold, absorption of each sub-chunk's carry by the next one of
//                  the row; writes int32 digits and one carry per row chunk
//   cols_fwd pass  per column group: absorb the previous row chunk's carry into the first
//                  sub-chunk (its last digit keeps any remainder), weight, forward column FFT,
//                  four-step twiddle   (together the reference's pass 1: src/neon.c:811-877)
#ifndef KITE_FFT_DEV_H
#define KITE_FFT_DEV_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "engine_dev.h"

namespace kite {

typedef unsigned __int128 u128;
typedef __int128 i128;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // a * conj(b)
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
template <bool INV> __device__ __forceinline__ double2 mul_mi(double2 a) {  // * -i (fwd) or * +i (inv)
    return INV ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

// ---- small DFTs, y_k = sum_j x_j e^{-+2 pi i jk/p} ----------------------------------------------

template <int P, bool INV> struct Dft;

template <bool INV> struct Dft<2, INV> {
    static __device__ __forceinline__ void run(double2 *v) {
        double2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <bool INV> struct Dft<4, INV> {
    static __device__ __forceinline__ void run(double2 *v) {
        double2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        double2 s13 = cadd(v[1], v[3]), d13 = mul_mi<INV>(csub(v[1], v[3]));
        v[0] = cadd(s02, s13);
        v[2] = csub(s02, s13);
        v[1] = cadd(d02, d13);
        v[3] = csub(d02, d13);
    }
};

template <bool INV> struct Dft<8, INV> {
    static __device__ __forceinline__ void run(double2 *v) {
        const double h = 0.70710678118654752440;
        double2 e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4, INV>::run(e);
        Dft<4, INV>::run(o);
        // o_k *= w8^k, w8 = e^{-+ i pi/4}
        double2 o1 = INV ? make_double2(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y))
                         : make_double2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
        double2 o2 = mul_mi<INV>(o[2]);
        double2 o3 = INV ? make_double2(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y))
                         : make_double2(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
};

// v * w16^E (w16 = e^{-+2 pi i/16}) for a compile-time E: exact for multiples of 4, two real
// multiplications for the odd multiples of 2, a complex multiplication otherwise.
template <int E, bool INV> __device__ __forceinline__ double2 mul_w16(double2 v) {
    constexpr int e = E & 15;
    constexpr double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173, h = 0.70710678118654752440;
    if constexpr (e == 0) {
        return v;
    } else if constexpr (e == 4) {
        return mul_mi<INV>(v);
    } else if constexpr (e == 8) {
        return make_double2(-v.x, -v.y);
    } else if constexpr (e == 12) {
        return mul_mi<!INV>(v);
    } else if constexpr (e % 2 == 0) {
        // (cos, sin) of 2 pi e/16 for e = 2, 6, 10, 14 are (+-h, +-h)
        constexpr double c = (e == 2 || e == 14) ? h : -h, s = (e == 2 || e == 6) ? h : -h;
        constexpr double sg = INV ? s : -s;  // multiplier c + i sg
        if constexpr (c == sg) return make_double2(c * (v.x - v.y), c * (v.x + v.y));
        else return make_double2(c * (v.x + v.y), c * (v.y - v.x));
    } else {
        constexpr int q = e % 4;  // 1: angle pi/8 off an axis, 3: 3 pi/8
        constexpr double ca = q == 1 ? c1 : s1, sa = q == 1 ? s1 : c1;
        // cos/sin of 2 pi e/16 from the quadrant of e
        constexpr int quad = e / 4;
        constexpr double c = quad == 0 ? ca : quad == 1 ? -sa : quad == 2 ? -ca : sa;
        constexpr double s = quad == 0 ? sa : quad == 1 ? ca : quad == 2 ? -sa : -ca;
        constexpr double sg = INV ? s : -s;
        return make_double2(v.x * c - v.y * sg, v.x * sg + v.y * c);
    }
}

// 16-point DFT as 4 x 4: n = 4a + b, k = k1 + 4 k2; y[b][k1] = DFT4_a(v[4a+b]) w16^{b k1}, then
// v[k1 + 4 k2] = DFT4_b(y[b][k1]).
template <bool INV> struct Dft<16, INV> {
    template <int B> static __device__ __forceinline__ void inner(const double2 *v, double2 *y) {
        double2 t[4] = {v[B], v[4 + B], v[8 + B], v[12 + B]};
        Dft<4, INV>::run(t);
        y[4 * B + 0] = t[0];
        y[4 * B + 1] = mul_w16<B * 1, INV>(t[1]);
        y[4 * B + 2] = mul_w16<B * 2, INV>(t[2]);
        y[4 * B + 3] = mul_w16<B * 3, INV>(t[3]);
    }
    static __device__ __forceinline__ void run(double2 *v) {
        double2 y[16];
        inner<0>(v, y);
        inner<1>(v, y);
        inner<2>(v, y);
        inner<3>(v, y);
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) {
            double2 t[4] = {y[k1], y[4 + k1], y[8 + k1], y[12 + k1]};
            Dft<4, INV>::run(t);
#pragma unroll
            for (int k2 = 0; k2 < 4; k2++) v[k1 + 4 * k2] = t[k2];
        }
    }
};

template <int P> struct OddRoots;
template <> struct OddRoots<3> {
    static __device__ __forceinline__ double c(int) { return -0.5; }
    static __device__ __forceinline__ double s(int q) { return q == 1 ? 0.86602540378443864676 : -0.86602540378443864676; }
};
template <> struct OddRoots<5> {
    static __device__ __forceinline__ double c(int q) {
        return (q == 1 || q == 4) ? 0.30901699437494742410 : -0.80901699437494742410;
    }
    static __device__ __forceinline__ double s(int q) {
        return q == 1 ? 0.95105651629515357212 : q == 2 ? 0.58778525229247312917
             : q == 3 ? -0.58778525229247312917 : -0.95105651629515357212;
    }
};
template <> struct OddRoots<7> {
    static __device__ __forceinline__ double c(int q) {
        return (q == 1 || q == 6) ? 0.62348980185873353053
             : (q == 2 || q == 5) ? -0.22252093395631440429 : -0.90096886790241912624;
    }
    static __device__ __forceinline__ double s(int q) {
        return q == 1 ? 0.78183148246802980871 : q == 2 ? 0.97492791218182360702
             : q == 3 ? 0.43388373911755812048 : q == 4 ? -0.43388373911755812048
             : q == 5 ? -0.97492791218182360702 : -0.78183148246802980871;
    }
};

// odd p: pairs (x_j, x_{p-j}) as s = sum, d = difference; y_k = A -+ iB, y_{p-k} = A +- iB
template <int P, bool INV> struct DftOdd {
    static __device__ __forceinline__ void run(double2 *v) {
        constexpr int M = (P - 1) / 2;
        double2 sp[M], dm[M];
        double2 y0 = v[0];
#pragma unroll
        for (int j = 1; j <= M; j++) {
            sp[j - 1] = cadd(v[j], v[P - j]);
            dm[j - 1] = csub(v[j], v[P - j]);
            y0 = cadd(y0, sp[j - 1]);
        }
        double2 out[P];
        out[0] = y0;
#pragma unroll
        for (int k = 1; k <= M; k++) {
            double2 A = v[0], B = make_double2(0, 0);
#pragma unroll
            for (int j = 1; j <= M; j++) {
                const int q = (j * k) % P;
                const double cq = OddRoots<P>::c(q), sq = OddRoots<P>::s(q);
                A.x += sp[j - 1].x * cq;
                A.y += sp[j - 1].y * cq;
                B.x += dm[j - 1].x * sq;
                B.y += dm[j - 1].y * sq;
            }
            double2 miB = make_double2(B.y, -B.x);
            out[k] = INV ? csub(A, miB) : cadd(A, miB);
            out[P - k] = INV ? cadd(A, miB) : csub(A, miB);
        }
#pragma unroll
        for (int k = 0; k < P; k++) v[k] = out[k];
    }
};
template <bool INV> struct Dft<3, INV> : DftOdd<3, INV> {};
template <bool INV> struct Dft<5, INV> : DftOdd<5, INV> {};
template <bool INV> struct Dft<7, INV> : DftOdd<7, INV> {};

// Composite lengths N1 N2 with coprime factors by the prime-factor (Good-Thomas) mapping, no twiddle
// multiplications: input n = (N2 n1 + N1 n2) mod N, output k = (N2 q2 k1 + N1 q1 k2) mod N with
// q2 = N2^-1 mod N1, q1 = N1^-1 mod N2; then n k = N2 q2 N2 n1 k1 + N1 q1 N1 n2 k2 = N2 n1 k1 +
// N1 n2 k2 (mod N), so the DFT is a DFT-N1 over n1 of DFT-N2s over n2.
__host__ __device__ constexpr int inv_mod(int a, int m) {
    for (int q = 1; q < m; q++)
        if ((a * q) % m == 1) return q;
    return 1;
}
template <int N1, int N2, bool INV> struct DftPfa {
    static __device__ __forceinline__ void run(double2 *v) {
        constexpr int N = N1 * N2, Q2 = inv_mod(N2 % N1, N1), Q1 = inv_mod(N1 % N2, N2);
        double2 t[N1][N2];
#pragma unroll
        for (int n1 = 0; n1 < N1; n1++) {
#pragma unroll
            for (int n2 = 0; n2 < N2; n2++) t[n1][n2] = v[(N2 * n1 + N1 * n2) % N];
            Dft<N2, INV>::run(t[n1]);
        }
#pragma unroll
        for (int k2 = 0; k2 < N2; k2++) {
            double2 u[N1];
#pragma unroll
            for (int n1 = 0; n1 < N1; n1++) u[n1] = t[n1][k2];
            Dft<N1, INV>::run(u);
#pragma unroll
            for (int k1 = 0; k1 < N1; k1++) v[(N2 * Q2 * k1 + N1 * Q1 * k2) % N] = u[k1];
        }
    }
};
template <bool INV> struct Dft<6, INV> : DftPfa<2, 3, INV> {};
template <bool INV> struct Dft<10, INV> : DftPfa<2, 5, INV> {};
template <bool INV> struct Dft<12, INV> : DftPfa<4, 3, INV> {};
template <bool INV> struct Dft<14, INV> : DftPfa<2, 7, INV> {};

// ---- Stockham stages in LDS -----------------------------------------------------------------
// `how` sequences of length n; element (seq, idx) at buf[idx*is + seq*ss]. A stage of radix P
// after Ns points have been combined: butterfly j reads idx j + r*n/P, twiddles input r by
// w_{Ns P}^{(j mod Ns) r} = tw[(j mod Ns) r n/(Ns P)], writes idx (j/Ns) Ns P + (j mod Ns) + r Ns.

template <int P, bool INV>
__device__ __forceinline__ void butterfly(const double2 *__restrict__ src, double2 *__restrict__ dst, int j,
                                          int seq, int n, int Ns, int is, int ss, const double2 *__restrict__ tw) {
    const int nb = n / P, ts = n / (Ns * P), jm = j % Ns;
    double2 v[P];
#pragma unroll
    for (int r = 0; r < P; r++) v[r] = src[(j + r * nb) * is + seq * ss];
    if (jm) {
#pragma unroll
        for (int r = 1; r < P; r++) {
            const double2 t = tw[jm * r * ts];
            v[r] = INV ? cmulc(v[r], t) : cmul(v[r], t);
        }
    }
    Dft<P, INV>::run(v);
    const int base = (j / Ns) * Ns * P + jm;
#pragma unroll
    for (int r = 0; r < P; r++) dst[(base + r * Ns) * is + seq * ss] = v[r];
}

// run-time plan
template <int P, bool INV>
__device__ __forceinline__ void stage_rt(const double2 *src, double2 *dst, int n, int Ns, int how, int is, int ss,
                                         bool seqfast, const double2 *tw) {
    const int nb = n / P, total = nb * how;
    for (int w = threadIdx.x; w < total; w += kNT) {
        const int seq = seqfast ? w % how : w / nb, j = seqfast ? w / how : w % nb;
        butterfly<P, INV>(src, dst, j, seq, n, Ns, is, ss, tw);
    }
}

template <bool INV>
__device__ double2 *fft_rt(double2 *a, double2 *b, const Plan &pl, int how, int is, int ss, bool seqfast,
                           const double2 *tw) {
    int Ns = 1;
    for (int s = 0; s < pl.nst; s++) {
        const int p = pl.rad[s];
        switch (p) {
        case 2: stage_rt<2, INV>(a, b, pl.n, Ns, how, is, ss, seqfast, tw); break;
        case 3: stage_rt<3, INV>(a, b, pl.n, Ns, how, is, ss, seqfast, tw); break;
        case 4: stage_rt<4, INV>(a, b, pl.n, Ns, how, is, ss, seqfast, tw); break;
        case 5: stage_rt<5, INV>(a, b, pl.n, Ns, how, is, ss, seqfast, tw); break;
        case 7: stage_rt<7, INV>(a, b, pl.n, Ns, how, is, ss, seqfast, tw); break;
        default: stage_rt<8, INV>(a, b, pl.n, Ns, how, is, ss, seqfast, tw); break;
        }
        __syncthreads();
        double2 *t = a;
        a = b;
        b = t;
        Ns *= p;
    }
    return a;
}

// compile-time plan: the same radix order as the host's make_plan (8s, then 4, 2, 3, 5, 7)
template <int N, int NS, int I, int HOW, int IS, int SS, bool SEQFAST, bool INV>
__device__ __forceinline__ double2 *fft_ct(double2 *a, double2 *b, const double2 *tw) {
    if constexpr (I == ct_nstages(N)) {
        return a;
    } else {
        constexpr int P = ct_radix(N, I), NB = N / P, TOTAL = NB * HOW;
#pragma unroll
        for (int w0 = 0; w0 < TOTAL; w0 += kNT) {
            const int w = w0 + (int)threadIdx.x;
            if (TOTAL % kNT == 0 || w < TOTAL) {
                const int seq = SEQFAST ? w % HOW : w / NB, j = SEQFAST ? w / HOW : w % NB;
                butterfly<P, INV>(a, b, j, seq, N, NS, IS, SS, tw);
            }
        }
        __syncthreads();
        return fft_ct<N, NS * P, I + 1, HOW, IS, SS, SEQFAST, INV>(b, a, tw);
    }
}

// Loads and stores of device memory through global-address-space pointers: global (not flat)
// instructions, which wait on the memory counter only.
typedef double dv2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 gld(const double2 *p) {
    const dv2_t v = *(const __attribute__((address_space(1))) dv2_t *)p;
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ void gst(double2 *p, double2 v) {
    dv2_t w;
    w.x = v.x;
    w.y = v.y;
    *(__attribute__((address_space(1))) dv2_t *)p = w;
}
// Scalars and int4 through global-address-space pointers (the tables a kernel copies into LDS: a
// flat load could alias LDS, so the compiler would wait for each one before the next LDS write).
template <class T> __device__ __forceinline__ T ldg(const T *p) { return *(const __attribute__((address_space(1))) T *)p; }
template <class T> __device__ __forceinline__ void stg(T *p, T v) { *(__attribute__((address_space(1))) T *)p = v; }
typedef int iv4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ int4 gld4(const int32_t *p) {
    const iv4_t v = *(const __attribute__((address_space(1))) iv4_t *)p;
    return make_int4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gst4(int32_t *p, int4 v) {
    iv4_t w;
    w.x = v.x;
    w.y = v.y;
    w.z = v.z;
    w.w = v.w;
    *(__attribute__((address_space(1))) iv4_t *)p = w;
}

// w_H^t for t < H from the two-level table
__device__ __forceinline__ double2 tw_h(const ShapeDev &sh, int t) {
    return cmul(sh.thi[t >> sh.sshift], sh.tlo[t & ((1 << sh.sshift) - 1)]);
}
// The same through the slot's copy of the table pointers (global loads). The compile-time kernels
// use it in place of their H-sized tables (four-step and split twiddles in access order) when those
// are too large to stay in L2 (R >= 128 columns, C >= 512 rows: megabytes per shape, read from HBM
// on every pass, more than the residue itself at the longest lengths); the two parts are a few KB.
// The product of the two correctly rounded factors is within two roundings of w_H^t.
__device__ __forceinline__ double2 tw_h2(const SlotDev &s, int t) {
    return cmul(gld(s.thi + (t >> s.sshift)), gld(s.tlo + (t & ((1 << s.sshift) - 1))));
}
template <int R> constexpr bool cols_tw2() { return R >= 128; }
template <int C> constexpr bool rows_tw2() { return C >= 512; }

// ---- real-split pointwise square / multiply ----------------------------------------------------
// a = U_k, b = U_{H-k}, t = w_H^k; e = a + conj(b), o = -i (a - conj(b)).
// square: X = e^2 + t o^2, Y = 2 e o; multiply by the operand's (ya, yb): X = e ey + t o oy,
// Y = e oy + o ey. New a = X + iY, new b = conj(X) + i conj(Y) (4x the packed product's DFT).
__device__ __forceinline__ void split_sq(double2 &a, double2 &b, double2 t) {
    double2 e = make_double2(a.x + b.x, a.y - b.y);
    double2 o = make_double2(a.y + b.y, b.x - a.x);
    double2 e2 = make_double2(e.x * e.x - e.y * e.y, 2 * e.x * e.y);
    double2 o2 = make_double2(o.x * o.x - o.y * o.y, 2 * o.x * o.y);
    double2 X = cadd(e2, cmul(t, o2));
    double2 Y = make_double2(2 * (e.x * o.x - e.y * o.y), 2 * (e.x * o.y + e.y * o.x));
    a = make_double2(X.x - Y.y, X.y + Y.x);
    b = make_double2(X.x + Y.y, Y.x - X.y);
}

__device__ __forceinline__ void split_mul(double2 &a, double2 &b, double2 t, double2 ya, double2 yb) {
    double2 e = make_double2(a.x + b.x, a.y - b.y);
    double2 o = make_double2(a.y + b.y, b.x - a.x);
    double2 ey = make_double2(ya.x + yb.x, ya.y - yb.y);
    double2 oy = make_double2(ya.y + yb.y, yb.x - ya.x);
    double2 X = cadd(cmul(e, ey), cmul(t, cmul(o, oy)));
    double2 Y = cadd(cmul(e, oy), cmul(o, ey));
    a = make_double2(X.x - Y.y, X.y + Y.x);
    b = make_double2(X.x + Y.y, Y.x - X.y);
}

// floor(t / k) for t < 2^64, 0 < k < 2^32, inv = floor((2^64 - 1) / k): inv >= (2^64 - k) / k, so
// the estimate hi64(t inv) is floor(t/k) or one less; one correction step (Barrett reduction)
__device__ __forceinline__ uint64_t divk(uint64_t t, uint32_t k, uint64_t inv, uint64_t &r) {
    uint64_t q = __umul64hi(t, inv);
    r = t - q * k;
    if (r >= k) {
        q++;
        r -= k;
    }
    return q;
}

// floor(h / k) for 0 < k < 2^32, by 64/32-bit long division in base 2^32 (each step's dividend
// (r << 32 | next word) is below k 2^32 <= 2^64)
__device__ __forceinline__ i128 floordiv_small(i128 h, uint32_t k, uint64_t inv) {
    const bool neg = h < 0;
    u128 a = neg ? (u128)(-h) : (u128)h;
    uint64_t hi = (uint64_t)(a >> 64), lo = (uint64_t)a, r;
    const uint64_t qh = divk(hi, k, inv, r);
    const uint64_t q1 = divk((r << 32) | (lo >> 32), k, inv, r);
    const uint64_t q0 = divk((r << 32) | (lo & 0xffffffffull), k, inv, r);
    u128 q = ((u128)qh << 64) | ((u128)q1 << 32) | q0;
    if (!neg) return (i128)q;
    return -(i128)q - (r ? 1 : 0);
}

// Fold the top region (digits top_t..L-1, within the last sub-chunk, plus the carry at 2^{s_L}):
// H = q*k + r  ->  q into digit 0, r*2^m rebalanced into the top digits (src/ctx.c:345-375).
__device__ __forceinline__ int64_t fold_top_dev(const SlotDev &s, int32_t *d, int64_t carry) {
    const int t0 = s.top_i0, nt = s.top_n, m = s.top_m;
    // the stored value is V = x 2^n: x + xadd is V + xadd 2^n, i.e. xadd 2^m above the top region's base
    i128 Hh = ((i128)carry << s.top_shift) + ((i128)s.xadd << m);
    int pos = 0;
    for (int i = 0; i < nt; i++) {
        Hh += (i128)d[t0 + i] << pos;
        pos += s.b0 + (int)((s.top_big >> i) & 1);
    }
    const i128 K = (i128)s.k;
    const i128 lo = Hh & (((i128)1 << m) - 1), hi = Hh >> m;
    i128 q = floordiv_small(hi, (uint32_t)s.k, s.kinv);
    i128 top = lo + ((hi - q * K) << m);
    if (2 * top >= (K << m)) {
        top -= K << m;
        q += 1;
    }
    for (int i = 0; i + 1 < nt; i++) {
        const int b = s.b0 + (int)((s.top_big >> i) & 1);
        const i128 qq = (top + ((i128)1 << (b - 1))) >> b;
        d[t0 + i] = (int32_t)(top - (qq << b));
        top = qq;
    }
    if (nt > 0) d[t0 + nt - 1] = (int32_t)top;
    return (int64_t)q;
}

__device__ __forceinline__ double pow2i(int e) { return e == 0 ? 1.0 : e == 1 ? 2.0 : 4.0; }
__device__ __forceinline__ double pow2ni(int e) { return e == 0 ? 1.0 : e == 1 ? 0.5 : 0.25; }

// Balanced split in exact FP64 arithmetic: x is an integer with |x| < 2^52, so x + 2^(b-1), the
// scaling by 2^-b, floor and x - q 2^b are all exact. Returns the carry q, sets the digit d.
__device__ __forceinline__ double bal_f(double x, double base, double ibase, double half, double &d) {
    const double q = floor((x + half) * ibase);
    d = fma(-q, base, x);
    return q;
}

// balanced split of v into a digit of b bits and the carry
__device__ __forceinline__ int64_t bal(int64_t v, int b, int32_t &d) {
    const int64_t q = (v + ((int64_t)1 << (b - 1))) >> b;
    d = (int32_t)(v - (q << b));
    return q;
}

// ---- rows pass ---------------------------------------------------------------------------------
// Row job u covers pairs [u*PPW, (u+1)*PPW): pair 0 is (0, R/2) (both self-paired; R/2 absent
// for odd R), pair p > 0 is (p, R-p). LDS holds 2*PPW rows of C (row slot 2i+h) twice.

template <int CC>
__device__ __forceinline__ void rows_body(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs, int mode,
                                          int step) {
    extern __shared__ double2 lds[];
    const int2 jb = jobs ? jobs[blockIdx.x] : make_int2(0, (int)blockIdx.x);
    const SlotDev &s = slots[jb.x];
    if (step >= (int)s.count) return;  // this slot sits out this launch (uniform per work-group)
    const ShapeDev &sh = *s.shp;
    constexpr int PPW_CT = CC ? ppw_for_C(CC) : 1;
    const int R = sh.R, C = CC ? CC : sh.C, PPW = PPW_CT, NROW = 2 * PPW;
    const int npairs = R % 2 == 0 ? R / 2 : (R + 1) / 2;
    double2 *A = lds, *B = lds + NROW * C;
    const double2 *twC = sh.twC;
    if (CC && CC <= kTwLdsMaxC) {
        double2 *T = lds + 2 * NROW * C;
        for (int i = threadIdx.x; i < C; i += kNT) T[i] = sh.twC[i];
        twC = T;
    }
    const double2 *x = mode == kModeRowFft ? s.y_src : s.x;
    // row index of slot (i, h), or -1
    auto row_of = [&](int i, int h) -> int {
        const int p = jb.y * PPW + i;
        if (p >= npairs) return -1;
        if (p == 0) return h == 0 ? 0 : (R % 2 == 0 && R > 1 ? R / 2 : -1);
        return h == 0 ? p : R - p;
    };
    for (int e = threadIdx.x; e < NROW * C; e += kNT) {
        const int slot = e / C, c = e - slot * C;
        const int row = row_of(slot >> 1, slot & 1);
        A[e] = row >= 0 ? x[(size_t)row * C + c] : make_double2(0, 0);
    }
    __syncthreads();
    double2 *F;
    if constexpr (CC != 0) F = fft_ct<CC, 1, 0, 2 * PPW_CT, 1, CC, false, false>(A, B, twC);
    else F = fft_rt<false>(A, B, sh.pc, NROW, 1, C, false, twC);
    if (mode == kModeRowFft) {  // operand preparation: store the row spectra
        for (int e = threadIdx.x; e < NROW * C; e += kNT) {
            const int slot = e / C, c = e - slot * C;
            const int row = row_of(slot >> 1, slot & 1);
            if (row >= 0) s.yfreq[(size_t)row * C + c] = F[e];
        }
        return;
    }
    const bool mul = mode == kModeMul;
    const double2 *Y = s.yfreq;
    for (int w = threadIdx.x; w < NROW * C; w += kNT) {
        const int slot = w / C, kc = w - slot * C, i = slot >> 1, h = slot & 1;
        const int p = jb.y * PPW + i;
        if (p >= npairs) continue;
        int ka, kb, pa, pb;  // rows and columns of the pair's two elements
        if (p == 0) {
            ka = row_of(0, h);
            if (ka < 0) continue;
            kb = ka;
            pa = kc;
            pb = h == 0 ? (C - kc) % C : C - 1 - kc;
            if (pb < kc) continue;
        } else {
            if (h) continue;
            ka = p;
            kb = R - p;
            pa = kc;
            pb = C - 1 - kc;
        }
        double2 *ea = &F[slot * C + pa];
        double2 *eb = &F[(p == 0 ? slot : slot + 1) * C + pb];
        double2 a = *ea, b = *eb;
        const double2 t = tw_h(sh, ka + R * kc);
        if (mul) split_mul(a, b, t, Y[(size_t)ka * C + pa], Y[(size_t)kb * C + pb]);
        else split_sq(a, b, t);
        *ea = a;
        if (eb != ea) *eb = b;
    }
    __syncthreads();
    double2 *Gb;
    if constexpr (CC != 0) Gb = fft_ct<CC, 1, 0, 2 * PPW_CT, 1, CC, false, true>(F, F == A ? B : A, twC);
    else Gb = fft_rt<true>(F, F == A ? B : A, sh.pc, NROW, 1, C, false, twC);
    for (int e = threadIdx.x; e < NROW * C; e += kNT) {
        const int slot = e / C, c = e - slot * C;
        const int row = row_of(slot >> 1, slot & 1);
        if (row >= 0) s.x[(size_t)row * C + c] = Gb[e];
    }
}

// ---- column passes -----------------------------------------------------------------------------
// LDS: A, B (R*G complex each), then the column twiddles (R complex) and one int64 carry per
// sub-chunk. Digits are staged as int32 [r][2G] in whichever of A/B is free.

template <int RC>
__device__ __forceinline__ void cols_inv_body(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs, int mode,
                                              int step) {
    extern __shared__ double2 lds[];
    const int2 jb = jobs ? jobs[blockIdx.x] : make_int2(0, (int)blockIdx.x);
    const SlotDev &s = slots[jb.x];
    if (step >= (int)s.count) return;
    const ShapeDev &sh = *s.shp;
    constexpr int G_CT = RC ? g_for_R(RC) : 0;
    const int R = RC ? RC : sh.R, G = RC ? G_CT : sh.G, C = sh.C, NG = C / G, g = jb.y, c0 = g * G;
    const int SUB = G < kSub ? G : kSub, NSUB = G / SUB;
    double2 *A = lds, *B = lds + R * G, *T = lds + 2 * R * G;
    int64_t *SC = reinterpret_cast<int64_t *>(T + R);
    const double2 *twR = sh.twR;
    if (RC) {
        for (int i = threadIdx.x; i < R; i += kNT) T[i] = sh.twR[i];
        twR = T;
    }
    for (int i = threadIdx.x; i < R * G; i += kNT) {
        const int kr = i / G, l = i - kr * G;
        A[i] = cmulc(s.x[(size_t)kr * C + c0 + l], tw_h(sh, (c0 + l) * kr));
    }
    __syncthreads();
    double2 *F;
    if constexpr (RC != 0) F = fft_ct<RC, 1, 0, G_CT, G_CT, 1, true, true>(A, B, twR);
    else F = fft_rt<true>(A, B, sh.pr, G, G, 1, true, twR);
    int32_t *D = reinterpret_cast<int32_t *>(F == A ? B : A);
    const bool sq = mode != kModeDigits;
    const double sc0 = sq ? s.kscale : s.rscale;
    const uint8_t *code = s.dcode + (size_t)g * R * G;
    double lmax = 0;
    const double bb = (double)((int64_t)1 << s.b0), ib = 1.0 / bb;
    for (int w = threadIdx.x; w < R * NSUB; w += kNT) {
        const int r = w / NSUB, q = w - r * NSUB;
        const double tap = s.ta[R + r] * sc0;
        // all loads and the rounding first (independent), then the serial carry in registers
        double rx[2 * kSub];
        int bigs[2 * kSub];
#pragma unroll
        for (int l = 0; l < kSub; l++) {
            if (l >= SUB) break;
            const int col = q * SUB + l;
            const double2 v = F[r * G + col];
            const int cd = code[r * G + col];
            const int cb = 2 * (c0 + col);
            const double x0 = v.x * (tap * s.tb[2 * C + cb] * pow2ni(cd & 3));
            const double x1 = v.y * (tap * s.tb[2 * C + cb + 1] * pow2ni((cd >> 3) & 3));
            rx[2 * l] = rint(x0);
            rx[2 * l + 1] = rint(x1);
            lmax = fmax(lmax, fmax(fabs(x0 - rx[2 * l]), fabs(x1 - rx[2 * l + 1])));
            bigs[2 * l] = (cd >> 2) & 1;
            bigs[2 * l + 1] = (cd >> 5) & 1;
        }
        int32_t *dr = D + r * 2 * G + 2 * q * SUB;
        double cy = 0;
#pragma unroll
        for (int i = 0; i < 2 * kSub; i++) {
            if (i >= 2 * SUB) break;
            if (sq) {
                const double base = bigs[i] ? 2 * bb : bb, ibase = bigs[i] ? 0.5 * ib : ib;
                double d;
                cy = bal_f(rx[i] + cy, base, ibase, 0.5 * base, d);
                dr[i] = (int32_t)d;
            } else {
                dr[i] = (int32_t)rx[i];
            }
        }
        SC[w] = (int64_t)cy;
    }
    __syncthreads();
    if (sq) {
        if (g == NG - 1 && threadIdx.x == 0) {  // the top fold on the number's last sub-chunk
            const int w = R * NSUB - 1;
            SC[w] = fold_top_dev(s, D + (R - 1) * 2 * G + 2 * (NSUB - 1) * SUB, SC[w]);
        }
        __syncthreads();
        // each sub-chunk after the first of its row absorbs its left neighbour's carry
        for (int w = threadIdx.x; w < R * NSUB; w += kNT) {
            const int r = w / NSUB, q = w - r * NSUB;
            if (q == 0) continue;
            int64_t cy = SC[w - 1];
            int32_t *dr = D + r * 2 * G + 2 * q * SUB;
            for (int i = 0; i < 2 * SUB && cy; i++) {
                if (i == 2 * SUB - 1) {
                    dr[i] = (int32_t)(dr[i] + cy);
                    break;
                }
                const int cd = code[r * G + q * SUB + (i >> 1)];
                const int big = (i & 1) ? (cd >> 5) & 1 : (cd >> 2) & 1;
                cy = bal((int64_t)dr[i] + cy, s.b0 + big, dr[i]);
            }
        }
        for (int r = threadIdx.x; r < R; r += kNT) s.cout[g * R + r] = SC[r * NSUB + NSUB - 1];
    }
    for (int off = 32; off > 0; off >>= 1) lmax = fmax(lmax, __shfl_xor(lmax, off));
    if ((threadIdx.x & 63) == 0 && lmax > 0)
        atomicMax(s.maxerr, (unsigned long long)__double_as_longlong(lmax));
    __syncthreads();
    int32_t *out = s.dig + (size_t)g * R * 2 * G;
    for (int i = threadIdx.x; i < R * 2 * G; i += kNT) out[i] = D[i];
}

template <int RC>
__device__ __forceinline__ void cols_fwd_body(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs, int mode,
                                              int step) {
    extern __shared__ double2 lds[];
    const int2 jb = jobs ? jobs[blockIdx.x] : make_int2(0, (int)blockIdx.x);
    const SlotDev &s = slots[jb.x];
    if (step >= (int)s.count) return;
    const ShapeDev &sh = *s.shp;
    constexpr int G_CT = RC ? g_for_R(RC) : 0;
    const int R = RC ? RC : sh.R, G = RC ? G_CT : sh.G, C = sh.C, NG = C / G, g = jb.y, c0 = g * G;
    const int SUB = G < kSub ? G : kSub;
    double2 *A = lds, *B = lds + R * G, *T = lds + 2 * R * G;
    const double2 *twR = sh.twR;
    if (RC) {
        for (int i = threadIdx.x; i < R; i += kNT) T[i] = sh.twR[i];
        twR = T;
    }
    int32_t *D = reinterpret_cast<int32_t *>(B);
    const int32_t *in = s.dig + (size_t)g * R * 2 * G;
    for (int i = threadIdx.x; i < R * 2 * G; i += kNT) D[i] = in[i];
    __syncthreads();
    const uint8_t *code = s.dcode + (size_t)g * R * G;
    if (mode != kModeDigits) {
        for (int r = threadIdx.x; r < R; r += kNT) {
            int pidx;  // row chunk before (g, r) in digit order, cyclically
            if (g > 0) pidx = (g - 1) * R + r;
            else if (r > 0) pidx = (NG - 1) * R + r - 1;
            else pidx = (NG - 1) * R + R - 1;
            int64_t cy = s.cout[pidx];
            int32_t *dr = D + r * 2 * G;
            for (int i = 0; i < 2 * SUB && cy; i++) {
                if (i == 2 * SUB - 1) {
                    dr[i] = (int32_t)(dr[i] + cy);
                    break;
                }
                const int cd = code[r * G + (i >> 1)];
                const int big = (i & 1) ? (cd >> 5) & 1 : (cd >> 2) & 1;
                cy = bal((int64_t)dr[i] + cy, s.b0 + big, dr[i]);
            }
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < R * G; i += kNT) {
        const int r = i / G, l = i - r * G;
        const int cd = code[i];
        const double tam = s.ta[r];
        const int cb = 2 * (c0 + l);
        const double w0 = tam * s.tb[cb] * pow2i(cd & 3);
        const double w1 = tam * s.tb[cb + 1] * pow2i((cd >> 3) & 3);
        A[i] = make_double2((double)D[r * 2 * G + 2 * l] * w0, (double)D[r * 2 * G + 2 * l + 1] * w1);
    }
    __syncthreads();
    double2 *F;
    if constexpr (RC != 0) F = fft_ct<RC, 1, 0, G_CT, G_CT, 1, true, false>(A, B, twR);
    else F = fft_rt<false>(A, B, sh.pr, G, G, 1, true, twR);
    for (int i = threadIdx.x; i < R * G; i += kNT) {
        const int kr = i / G, l = i - kr * G;
        s.x[(size_t)kr * C + c0 + l] = cmul(F[i], tw_h(sh, (c0 + l) * kr));
    }
}

// ================================================================================================
// Compile-time kernels: single-buffer Stockham stages (a stage's inputs are held in registers
// across one barrier, halving LDS per work-group), one thread per 16-digit carry sub-chunk working
// in registers and moving its digits straight between HBM and registers, the full w_H table.
// ================================================================================================

// The work-item id through an opaque copy: in a kernel that loops over jobs, everything derived
// from the id would otherwise be hoisted out of the loop as invariant and kept live across the
// whole job body (which spilled hundreds of bytes per thread); each job reads its own copy. (The
// one-job kernels read threadIdx.x directly: the opaque copy costs them about 3%.)
__device__ __forceinline__ int tid_x() {
    int t = (int)threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

// LDS index with PAD: one padding element after every 2^PAD (PAD = 0: none), which spreads the
// strided writes of the early Stockham stages over distinct LDS bank groups.
template <int PAD> __device__ __forceinline__ int padi(int i) { return PAD ? i + (i >> PAD) : i; }

// Radix plan of fft_ip: lengths of 1024 and more start with radix-16 stages (1024 = 16 x 8 x 8,
// 2048 = 16 x 16 x 8), then the plan of the other kernels: three stages instead of four, without
// the radix-2 stage (N/2 butterflies) of 8 x 8 x 8 x 2.
__host__ __device__ constexpr int ip_lead16(int n) {
    return n >= 2048 && n % 256 == 0 ? 2 : n >= 1024 && n % 16 == 0 ? 1 : 0;
}
__host__ __device__ constexpr int ip_nstages(int n) {
    return ip_lead16(n) + ct_nstages(n >> (4 * ip_lead16(n)));
}
__host__ __device__ constexpr int ip_radix(int n, int idx) {
    return idx < ip_lead16(n) ? 16 : ct_radix(n >> (4 * ip_lead16(n)), idx - ip_lead16(n));
}

// TPAD: the twiddle table is padded like the data (entry t at padi<TPAD>(t)).
template <int N, int NS, int I, int HOW, int IS, int SS, bool SEQFAST, bool INV, int PAD = 0, int NT = kNT,
          int TPAD = 0>
__device__ __forceinline__ void fft_ip(double2 *a, const double2 *tw) {
    if constexpr (I < ip_nstages(N)) {
        constexpr int P = ip_radix(N, I), NB = N / P, TOTAL = NB * HOW;
        constexpr int ITER = (TOTAL + NT - 1) / NT, TS = N / (NS * P);
        double2 v[ITER][P];
        int ob[ITER];
#pragma unroll
        for (int it = 0; it < ITER; it++) {
            const int w = it * NT + (int)threadIdx.x;
            if (TOTAL % NT == 0 || w < TOTAL) {
                const int seq = SEQFAST ? w % HOW : w / NB, j = SEQFAST ? w / HOW : w % NB;
                const int jm = j % NS;
#pragma unroll
                for (int r = 0; r < P; r++) v[it][r] = a[padi<PAD>((j + r * NB) * IS + seq * SS)];
                if (NS > 1 && jm) {
#pragma unroll
                    for (int r = 1; r < P; r++) {
                        const double2 t = tw[padi<TPAD>(jm * r * TS)];
                        v[it][r] = INV ? cmulc(v[it][r], t) : cmul(v[it][r], t);
                    }
                }
                Dft<P, INV>::run(v[it]);
                ob[it] = ((j / NS) * NS * P + jm) * IS + seq * SS;
            }
        }
        // in place: reads before writes, except in the last stage (NS P = N), where every thread
        // writes back the elements it read
        if constexpr (NS * P != N) __syncthreads();
#pragma unroll
        for (int it = 0; it < ITER; it++) {
            const int w = it * NT + (int)threadIdx.x;
            if (TOTAL % NT == 0 || w < TOTAL) {
#pragma unroll
                for (int r = 0; r < P; r++) a[padi<PAD>(ob[it] + r * NS * IS)] = v[it][r];
            }
        }
        __syncthreads();
        fft_ip<N, NS * P, I + 1, HOW, IS, SS, SEQFAST, INV, PAD, NT, TPAD>(a, tw);
    }
}

// fft_ip for HOW column sequences (SEQFAST, SS = 1, no padding function) whose first stage may take
// its inputs from ld(it, r, row, col) instead of the buffer (LDIN) and whose last stage may hand its
// outputs to st(it, r, row, col, v) instead of the buffer (STOUT): a load or store phase fused into
// the FFT saves a pass through LDS and its barriers. (it, r) name the thread's iteration and the
// butterfly input/output, (row, col) the element.
// Radix plan of fft_io: two stages (N/8, then 8) when N/8 is a length of Dft, which halves the
// LDS passes and barriers of the column lengths up to 128; else the plan of the other kernels.
// For the column lengths up to 128 whose R/8 is no DFT length here, the most balanced split into
// two DFT lengths (100 = 10 x 10, 36 = 6 x 6, 120 = 12 x 10), else the plan of the other kernels.
// Longer columns, and the shorter ones without such a split (io_long), where the plan of the other
// kernels has a radix-2 stage or could have fewer stages: the fewest stages of radices up to
// kIoMaxRadix, then the smallest largest radix, then the fewest butterflies (sum of N/P), radices
// in descending order (108 = 6 x 6 x 3, 144 = 6 x 6 x 4, 216 = 6 x 6 x 6, 400 = 10 x 8 x 5). The
// radix stays small because the fused column pass is latency-bound: more, shorter butterflies per
// stage beat fewer stages (144 as 12 x 12 and 192 as 16 x 12 measured 10-15% slower).
constexpr int kIoMaxRadix = 10;
__host__ __device__ constexpr bool dft_len(int n) {
    return n == 2 || n == 3 || n == 4 || n == 5 || n == 6 || n == 7 || n == 8 || n == 10 || n == 12 || n == 14 ||
           n == 16;
}
__host__ __device__ constexpr int io_first(int n) {
    if (n % 8 == 0 && dft_len(n / 8)) return n / 8;
    int best = 0, bd = 0;
    if (n <= 128)
        for (int a = 2; a <= 16; a++)
            if (n % a == 0 && dft_len(a) && dft_len(n / a)) {
                const int d = a > n / a ? a - n / a : n / a - a;
                if (!best || d < bd) {
                    best = a;
                    bd = d;
                }
            }
    return best;
}
__host__ __device__ constexpr bool io_two(int n) { return io_first(n) != 0; }
struct IoSplit {
    int ns, r[4];
};
__host__ __device__ constexpr bool ct_has2(int n) {
    for (int i = 0; i < ct_nstages(n); i++)
        if (ct_radix(n, i) == 2) return true;
    return false;
}
__host__ __device__ constexpr IoSplit io_long(int n) {
    IoSplit best{0, {0, 0, 0, 0}};
    int bmax = 0, bsum = 0;
    auto take = [&](int ns, int a, int b, int c, int d) {
        const int sum = n / a + n / b + (c ? n / c : 0) + (d ? n / d : 0);
        if (best.ns == 0 || ns < best.ns || (ns == best.ns && (a < bmax || (a == bmax && sum < bsum)))) {
            best = IoSplit{ns, {a, b, c, d}};
            bmax = a;
            bsum = sum;
        }
    };
    auto ok = [](int r) { return r >= 2 && r <= kIoMaxRadix && dft_len(r); };
    for (int a = kIoMaxRadix; a >= 2; a--) {
        if (!ok(a) || n % a) continue;
        const int m = n / a;
        if (m <= a && ok(m)) take(2, a, m, 0, 0);
        for (int b = a; b >= 2; b--) {
            if (!ok(b) || m % b) continue;
            const int m2 = m / b;
            if (m2 <= b && ok(m2)) take(3, a, b, m2, 0);
            for (int c = b; c >= 2; c--)
                if (ok(c) && m2 % c == 0 && m2 / c <= c && ok(m2 / c)) take(4, a, b, c, m2 / c);
        }
    }
    // only where it beats the plan of the other kernels: fewer stages, or no radix-2 stage (which
    // has N/2 butterflies)
    if (best.ns && best.ns >= ct_nstages(n) && !ct_has2(n)) best.ns = 0;
    return best;
}
__host__ __device__ constexpr int io_nstages(int n) {
    return io_two(n) ? 2 : io_long(n).ns ? io_long(n).ns : ct_nstages(n);
}
__host__ __device__ constexpr int io_radix(int n, int idx) {
    if (io_two(n)) return idx == 0 ? io_first(n) : n / io_first(n);
    return io_long(n).ns ? io_long(n).r[idx] : ct_radix(n, idx);
}

struct NoPre {
    __device__ void operator()() const {}
};

// pre() runs once, after the first stage's LDS writes and before its barrier: a kernel stores
// register-held tables to LDS there, after the first stage's HBM loads have been issued.
// OT: the work-item id through tid_x (for kernels that loop over jobs); NT: threads per work-group.
// A compile-time bound the compiler may not use to drop a phase's thread test (the test keeps the
// phase's code predicated, which at the fused kernel's register limit schedules without spills).
__device__ __forceinline__ int opaque_bound(int n) {
    asm volatile("" : "+s"(n));
    return n;
}

template <int N, int NS, int I, int HOW, int IS, bool INV, bool LDIN, bool STOUT, bool OT = false, int NT = kNT,
          class LD, class ST, class PRE = NoPre>
__device__ __forceinline__ void fft_io(double2 *a, const double2 *tw, LD &ld, ST &st, PRE pre = PRE()) {
    constexpr int NST = io_nstages(N);
    if constexpr (I < NST) {
        const int tid = OT ? tid_x() : (int)threadIdx.x;
        constexpr int P = io_radix(N, I), NB = N / P, TOTAL = NB * HOW;
        constexpr int ITER = (TOTAL + NT - 1) / NT, TS = N / (NS * P);
        constexpr bool FROM_LD = LDIN && I == 0, TO_ST = STOUT && I == NST - 1;
        double2 v[ITER][P];
        int orow[ITER], seqs[ITER];
#pragma unroll
        for (int it = 0; it < ITER; it++) {
            const int w = it * NT + tid;
            if (w < opaque_bound(TOTAL)) {
                const int seq = w % HOW, j = w / HOW, jm = j % NS;
#pragma unroll
                for (int r = 0; r < P; r++) {
                    if constexpr (FROM_LD) v[it][r] = ld(it, r, j + r * NB, seq);
                    else v[it][r] = a[(j + r * NB) * IS + seq];
                }
                if (NS > 1 && jm) {
#pragma unroll
                    for (int r = 1; r < P; r++) {
                        const double2 t = tw[jm * r * TS];
                        v[it][r] = INV ? cmulc(v[it][r], t) : cmul(v[it][r], t);
                    }
                }
                Dft<P, INV>::run(v[it]);
                orow[it] = (j / NS) * NS * P + jm;
                seqs[it] = seq;
            }
        }
        if constexpr (TO_ST) {
#pragma unroll
            for (int it = 0; it < ITER; it++) {
                const int w = it * NT + tid;
                if (w < opaque_bound(TOTAL)) {
#pragma unroll
                    for (int r = 0; r < P; r++) st(it, r, orow[it] + r * NS, seqs[it], v[it][r]);
                }
            }
        } else {
            // the stage is in place: reads before writes, except in the last stage, whose outputs
            // (rows j + r NS, NS = N / P = NB) are the very elements each thread read
            if constexpr (!FROM_LD && I != NST - 1) __syncthreads();
#pragma unroll
            for (int it = 0; it < ITER; it++) {
                const int w = it * NT + tid;
                if (w < opaque_bound(TOTAL)) {
#pragma unroll
                    for (int r = 0; r < P; r++) a[(orow[it] + r * NS) * IS + seqs[it]] = v[it][r];
                }
            }
            if constexpr (I == 0) pre();
            __syncthreads();
            fft_io<N, NS * P, I + 1, HOW, IS, INV, LDIN, STOUT, OT, NT>(a, tw, ld, st);
        }
    }
}

// The last stage's thread mapping of fft_io (for prefetching what its outputs need).
template <int N, int HOW, int NT = kNT> struct LastStage {
    static constexpr int P = io_radix(N, io_nstages(N) - 1), NB = N / P, TOTAL = NB * HOW;
    static constexpr int ITER = (TOTAL + NT - 1) / NT;
};

// Rows: 2*PPW rows of C in LDS (row slot 2i+h = side h of local pair i), twiddles w_C after them.
// PPW_T pairs per work-group: ppw_for_C(C) fills every stage's threads; 1 spreads one test's
// rows over more work-groups (lower latency when a single test runs alone).
template <int C, int PPW_T = ppw_for_C(C)>
__device__ __forceinline__ void rows_fast(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs, int mode,
                                          int step) {
    extern __shared__ double2 lds[];
    const int2 jb = jobs ? jobs[blockIdx.x] : make_int2(0, (int)blockIdx.x);
    const SlotDev &s = slots[jb.x];
    if (step >= (int)s.count) return;
    constexpr int PPW = PPW_T, NROW = 2 * PPW, NE = NROW * C, NT = rows_nt(C);
    constexpr bool TWL = C <= kTwLdsMaxC;
    constexpr int PD = rows_pad(C), TPD = rows_twpad(C);  // row buffer index i lives at padi<PD>(i)
    const int R = s.R, npairs = R % 2 == 0 ? R / 2 : (R + 1) / 2;
    double2 *A = lds;
    const double2 *twC = s.twC;
    if (TWL) {
        double2 *T = lds + padi<PD>(NE);
        for (int i = threadIdx.x; i < C; i += NT) T[padi<TPD>(i)] = gld(s.twC + i);
        twC = T;
    }
    auto row_of = [&](int i, int h) -> int {
        const int p = jb.y * PPW + i;
        if (p >= npairs) return -1;
        if (p == 0) return h == 0 ? 0 : (R % 2 == 0 && R > 1 ? R / 2 : -1);
        return h == 0 ? p : R - p;
    };
    const double2 *x = mode == kModeRowFft ? s.y_src : s.x;
    for (int e = threadIdx.x; e < NE; e += NT) {
        const int slot = e / C, c = e - slot * C;
        const int row = row_of(slot >> 1, slot & 1);
        A[padi<PD>(e)] = row >= 0 ? gld(x + (size_t)row * C + c) : make_double2(0, 0);
    }
    // split items: one per (local pair, k_c); the twiddle of item kc of pair p > 0 (and of row 0 of
    // pair 0) is fetched before the FFT
    constexpr int NIT = PPW * C, ITEMS = (NIT + NT - 1) / NT;
    const bool split = mode == kModeSquare || mode == kModeMul;
    double2 tw_[ITEMS];
#pragma unroll
    for (int it = 0; it < ITEMS; it++) {
        const int w = it * NT + (int)threadIdx.x;
        const int i = w / C, kc = w - i * C, p = jb.y * PPW + i;
        const bool need = split && (NIT % NT == 0 || w < NIT) && p < npairs;
        if constexpr (rows_tw2<C>()) tw_[it] = need ? tw_h2(s, p + R * kc) : make_double2(1, 0);
        else tw_[it] = need ? gld(s.twRows + (size_t)p * C + kc) : make_double2(1, 0);
    }
    __syncthreads();
    fft_ip<C, 1, 0, NROW, 1, C, false, false, PD, NT, TPD>(A, twC);
    if (mode == kModeRowFft) {  // operand preparation: store the row spectra
        for (int e = threadIdx.x; e < NE; e += NT) {
            const int slot = e / C, c = e - slot * C;
            const int row = row_of(slot >> 1, slot & 1);
            if (row >= 0) gst(s.yfreq + (size_t)row * C + c, A[padi<PD>(e)]);
        }
        return;
    }
    const bool mul = mode == kModeMul;
#pragma unroll
    for (int it = 0; it < ITEMS; it++) {
        const int w = it * NT + (int)threadIdx.x;
        if (NIT % NT != 0 && w >= NIT) continue;
        const int i = w / C, kc = w - i * C, p = jb.y * PPW + i;
        if (p >= npairs) continue;
        if (p > 0) {  // (p, kc) with (R-p, C-1-kc)
            const int pb = C - 1 - kc;
            const int ea = padi<PD>(2 * i * C + kc), eb = padi<PD>((2 * i + 1) * C + pb);
            double2 a = A[ea], b = A[eb];
            if (mul) split_mul(a, b, tw_[it], gld(s.yfreq + (size_t)p * C + kc), gld(s.yfreq + (size_t)(R - p) * C + pb));
            else split_sq(a, b, tw_[it]);
            A[ea] = a;
            A[eb] = b;
        } else {
            // row 0: kc with (C-kc) mod C; row R/2 (if present): kc with C-1-kc
            int pb = (C - kc) % C;
            if (pb >= kc) {
                double2 a = A[padi<PD>(kc)], b = A[padi<PD>(pb)];
                if (mul) split_mul(a, b, tw_[it], gld(s.yfreq + kc), gld(s.yfreq + pb));
                else split_sq(a, b, tw_[it]);
                A[padi<PD>(kc)] = a;
                if (pb != kc) A[padi<PD>(pb)] = b;
            }
            const int rh = row_of(0, 1);
            pb = C - 1 - kc;
            if (rh >= 0 && pb >= kc) {
                const double2 t = rows_tw2<C>() ? tw_h2(s, rh + R * kc) : gld(s.twRows + (size_t)rh * C + kc);
                double2 a = A[padi<PD>(C + kc)], b = A[padi<PD>(C + pb)];
                if (mul) split_mul(a, b, t, gld(s.yfreq + (size_t)rh * C + kc), gld(s.yfreq + (size_t)rh * C + pb));
                else split_sq(a, b, t);
                A[padi<PD>(C + kc)] = a;
                if (pb != kc) A[padi<PD>(C + pb)] = b;
            }
        }
    }
    __syncthreads();
    fft_ip<C, 1, 0, NROW, 1, C, false, true, PD, NT, TPD>(A, twC);
    for (int e = threadIdx.x; e < NE; e += NT) {
        const int slot = e / C, c = e - slot * C;
        const int row = row_of(slot >> 1, slot & 1);
        if (row >= 0) gst(s.x + (size_t)row * C + c, A[padi<PD>(e)]);
    }
}

// Square-mode rows for C = 256 held in registers. A work-group is two waves of four 16-lane groups;
// each group takes one row: group i of wave 0 row k_r and group i of wave 1 row R - k_r of local
// pair i (four pairs per work-group, the job granularity of the LDS row kernel), so the side of a
// pair is uniform over a wave. Lane t loads elements 16a + t (a < 16),
// so every load and store moves 256 contiguous bytes per group. Forward row FFT as 16 x 16: DFT16
// over a in registers, twiddle w_256^{t k1}, a transpose through the group's LDS block (slot
// k1*16 + (b ^ k1): conflict-free both ways), DFT16 over b. Lane t then holds k_c = t + 16 k2 of
// row k_r and k_c = (15 - t) + 16 k2 of row R - k_r. For the real split each group copies its
// row to LDS in natural order and every lane reads the partners of its own elements
// (k_c <-> C - 1 - k_c in the partner row; within the row for the self-paired rows 0 and R/2)
// and computes its own half of each pair. The inverse runs the forward steps backwards.
constexpr int kRowsRegC = 256;


// The half of the real-split square (split_sq) that belongs to `u`: u is side a (partner w is
// b) or, with isb, side b; tw = w_H^k of side a's frequency k.
__device__ __forceinline__ double2 split_sq_half(double2 u, double2 w, double2 tw, bool isb) {
    const double2 a = isb ? w : u, b = isb ? u : w;
    const double2 e = make_double2(a.x + b.x, a.y - b.y);
    const double2 o = make_double2(a.y + b.y, b.x - a.x);
    const double2 e2 = make_double2(e.x * e.x - e.y * e.y, 2 * e.x * e.y);
    const double2 o2 = make_double2(o.x * o.x - o.y * o.y, 2 * o.x * o.y);
    const double2 X = cadd(e2, cmul(tw, o2));
    const double2 Y = make_double2(2 * (e.x * o.x - e.y * o.y), 2 * (e.x * o.y + e.y * o.x));
    return isb ? make_double2(X.x + Y.y, Y.x - X.y) : make_double2(X.x - Y.y, X.y + Y.x);
}

// w_H^{p + R k_c} = T0 w16^{k2} for side a's k_c = t + 16 k2 (H = 256 R, T0 = w_H^{p + R t}); the
// side-b element k2 pairs with side a's k2' = 15 - k2 of the same t.
// One round of the real split for this lane's elements k2 in [K0, K0 + 8) (pairs k_r, R - k_r with
// k_r > 0): the partner of k_c = km + 16 k2 is the partner row's element C - 1 - k_c, which the
// partner group has just written to its LDS block at natural position mod 128. Results go to
// out[k2 - K0]; side a's twiddle w_H^{k_r + R k_c} is T0 w16^{k2} (T0 = w_H^{k_r + R t}).
template <int K0, int K2, bool SIDE_B>
__device__ __forceinline__ void rows256_split_half(const double2 *v, double2 *out, const double2 *Xp, int km,
                                                   double2 T0) {
    if constexpr (K2 < K0 + 8) {
        const double2 w = Xp[(kRowsRegC - 1 - km - 16 * K2) & 127];
        out[K2 - K0] = split_sq_half(v[K2], w, mul_w16<SIDE_B ? 15 - K2 : K2, false>(T0), SIDE_B);
        rows256_split_half<K0, K2 + 1, SIDE_B>(v, out, Xp, km, T0);
    }
}

// The same round for the first four pairs, where rows 0 and R/2 pair within themselves
// (k_c <-> (C - k_c) mod C and k_c <-> C - 1 - k_c): exact split twiddles from the table.
template <int K0>
__device__ __forceinline__ void rows256_split_half0(const double2 *v, double2 *out, const double2 *blk, int km,
                                                    bool special, int side, int row, int p,
                                                    const double2 *twRows) {
    constexpr int C = kRowsRegC;
#pragma unroll
    for (int k2 = K0; k2 < K0 + 8; k2++) {
        const int kc = km + 16 * k2;
        const int pb = special ? (side == 0 ? (C - kc) & (C - 1) : C - 1 - kc) : C - 1 - kc;
        const bool isb = special ? kc > pb : side == 1;
        const int twi = (special ? row : p) * C + (isb ? pb : kc);
        const double2 w = pb == kc ? v[k2] : blk[pb & 127];
        out[k2 - K0] = split_sq_half(v[k2], w, gld(twRows + twi), isb);
    }
}

// Transpose of 16 values per lane through a group's LDS block of 256 doubles, real and imaginary
// parts in turn: value j of this lane goes to slot wr(j), and value j comes back from slot rd(j).
template <class WR, class RD>
__device__ __forceinline__ void xpose_ri(double2 *v, double *X, WR wr, RD rd) {
#pragma unroll
    for (int j = 0; j < 16; j++) X[wr(j)] = v[j].x;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 16; j++) v[j].x = X[rd(j)];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 16; j++) X[wr(j)] = v[j].y;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 16; j++) v[j].y = X[rd(j)];
    __syncthreads();
}

template <bool OT = false>  // OT: the work-item id through tid_x (a kernel looping over jobs)
__device__ __forceinline__ void rows256_sq(const SlotDev *__restrict__ slots, int2 jb, int step) {
    constexpr int C = kRowsRegC;
    __shared__ double xs[8 * 256];  // one 2 KB block per 16-lane group
    const SlotDev &s = slots[jb.x];
    {
        // the slot fields of the prologue in one scalar round trip
        const uint32_t count = s.count;
        const int R0 = s.R;
        const void *p0 = s.x, *p1 = s.tw16, *p2 = s.twRows;
        asm volatile("" ::"s"(count), "s"(R0), "s"(p0), "s"(p1), "s"(p2));
        if (step >= (int)count) return;
    }
    const int lane = OT ? tid_x() : (int)threadIdx.x, g = lane >> 4, t = lane & 15;
    const int side = g >> 2, p = jb.y * 4 + (g & 3);
    const int R = s.R, npairs = R % 2 == 0 ? R / 2 : (R + 1) / 2;
    const bool special = p == 0;  // rows 0 and R/2, each paired with itself
    int row = -1;
    if (p < npairs) row = side == 0 ? p : (special ? (R % 2 == 0 ? R / 2 : -1) : R - p);
    const bool act = row >= 0;
    const int km = (side == 1 && !special) ? 15 - t : t;  // k1 held after the transpose
    double *X = xs + g * 256;
    double2 *xr = s.x + (size_t)(act ? row : 0) * C + t;
    // the row twiddles w_256^{t j} ([j][t], 4 KB) in LDS: L2 round trips in the FFTs cost more
    // (its loads are issued first and stored to LDS after the row's loads are issued, so the row
    // does not wait for the table's round trip)
    __shared__ double2 tws[256];
    static_assert(kRowsSqThreads * 2 == 256, "two twiddles per thread");
    const double2 tw_a = gld(s.tw16 + lane), tw_b = gld(s.tw16 + lane + kRowsSqThreads);
    // the split's twiddle base (side a's w_H^{p + R t}), fetched with the row: issued after the FFT
    // its round trip stood between the FFT and the split
    const bool first = jb.y == 0;  // uniform over the work-group: holds rows 0 and R/2
    const double2 T0 = (act && !first) ? gld(s.twRows + (size_t)p * C + t) : make_double2(0, 0);
    double2 v[16];
#pragma unroll
    for (int i = 0; i < 16; i++) v[i] = act ? gld(xr + 16 * i) : make_double2(0, 0);
    tws[lane] = tw_a;
    tws[lane + kRowsSqThreads] = tw_b;
    __syncthreads();
    const double2 *tw16 = tws;
    // forward: DFT16 over a, twiddle w_256^{t k1}, transpose, DFT16 over b
    Dft<16, false>::run(v);
    // twiddles w_256^{t j} from the [j][t] table
#pragma unroll
    for (int j = 1; j < 16; j++) {
        if (j % 4 == 1) __builtin_amdgcn_sched_barrier(0);
        v[j] = cmul(v[j], tw16[16 * j + t]);
    }
    xpose_ri(v, X, [&](int j) { return j * 16 + (t ^ j); }, [&](int j) { return km * 16 + (j ^ km); });
    Dft<16, false>::run(v);  // v[k2] = U[row][km + 16 k2]
    // real-split square: the partner of element k2 (k_c = km + 16 k2) is the partner row's element
    // 15 - k2 held by lane t of the partner group (whose km is 15 - km); two rounds of 8 through LDS
    // real-split square in two rounds through LDS: every partner of an element with k_c < 128 has
    // k_c >= 128 and the other way round (or is the element itself), so each group first publishes
    // its upper half (k2 = 8..15) and combines its lower half, then publishes its lower half and
    // combines its upper half in place
    double2 *Xc = reinterpret_cast<double2 *>(X);
    const double2 *blk = (first && special) ? Xc : reinterpret_cast<const double2 *>(xs + (g ^ 4) * 256);
    double2 lo[8];
#pragma unroll
    for (int k2 = 8; k2 < 16; k2++) Xc[(km + 16 * k2) & 127] = v[k2];
    __syncthreads();
    if (!first) {
        if (side == 0) rows256_split_half<0, 0, false>(v, lo, blk, km, T0);
        else rows256_split_half<0, 0, true>(v, lo, blk, km, T0);
    } else if (act) {
        rows256_split_half0<0>(v, lo, blk, km, special, side, row, p, s.twRows);
    }
    __syncthreads();
#pragma unroll
    for (int k2 = 0; k2 < 8; k2++) Xc[km + 16 * k2] = v[k2];
    __syncthreads();
    if (!first) {
        if (side == 0) rows256_split_half<8, 8, false>(v, v + 8, blk, km, T0);
        else rows256_split_half<8, 8, true>(v, v + 8, blk, km, T0);
    } else if (act) {
        rows256_split_half0<8>(v, v + 8, blk, km, special, side, row, p, s.twRows);
    }
    __syncthreads();
#pragma unroll
    for (int k2 = 0; k2 < 8; k2++) v[k2] = lo[k2];
    // inverse: DFT16 over k2, twiddle w_256^{-k1 b}, transpose, DFT16 over k1
    Dft<16, true>::run(v);
#pragma unroll
    for (int j = 1; j < 16; j++) {
        if (j % 4 == 1) __builtin_amdgcn_sched_barrier(0);
        v[j] = cmulc(v[j], tw16[16 * j + km]);
    }
    xpose_ri(v, X, [&](int j) { return j * 16 + (km ^ j); }, [&](int j) { return t * 16 + (j ^ t); });
    Dft<16, true>::run(v);
    if (act) {
#pragma unroll
        for (int i = 0; i < 16; i++) gst(xr + 16 * i, v[i]);
    }
}

// Column passes: LDS = R*G complex, the column twiddles (R complex), one int64 carry per
// sub-chunk, and a 64-digit scratch for the top fold.
template <int R> struct ColGeom {
    static constexpr int G = g_for_R(R), SUB = G < kSub ? G : kSub, NSUB = G / SUB, NE = R * G;
    static constexpr int NSC = R * NSUB, ITSC = (NSC + kNT - 1) / kNT;
    static constexpr size_t lds_bytes() { return (size_t)16 * (NE + R) + (size_t)8 * NSC + 4 * 64; }
};

template <int R>
__device__ __forceinline__ void cols_inv_fast(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs, int mode,
                                              int step) {
    extern __shared__ double2 lds[];
    using Gm = ColGeom<R>;
    constexpr int G = Gm::G, SUB = Gm::SUB, NSUB = Gm::NSUB, NE = Gm::NE, NSC = Gm::NSC, ITSC = Gm::ITSC;
    const int2 jb = jobs ? jobs[blockIdx.x] : make_int2(0, (int)blockIdx.x);
    const SlotDev &s = slots[jb.x];
    if (step >= (int)s.count) return;
    const int C = s.C, NG = C / G, g = jb.y, c0 = g * G;
    double2 *A = lds, *T = lds + NE;
    int64_t *SC = reinterpret_cast<int64_t *>(T + R);
    int32_t *FD = reinterpret_cast<int32_t *>(SC + NSC);
    for (int i = threadIdx.x; i < R; i += kNT) T[i] = gld(s.twR + i);
    const double2 *tg = s.twCols + (size_t)g * NE;
    for (int i = threadIdx.x; i < NE; i += kNT) {
        const int kr = i / G, l = i - kr * G;
        A[i] = cmulc(gld(s.x + (size_t)kr * C + c0 + l), gld(tg + i));
    }
    __syncthreads();
    fft_ip<R, 1, 0, G, G, 1, true, true>(A, T);
    const bool sq = mode != kModeDigits;
    const double sc0 = sq ? s.kscale : s.rscale;
    const double bb = (double)((int64_t)1 << s.b0), ibb = 1.0 / bb;
    const uint8_t *code = s.dcode + (size_t)g * R * G;
    double lmax = 0;
    int32_t d[ITSC][2 * SUB];
    uint32_t bg[ITSC];  // bit i: digit i of the sub-chunk has b0+1 bits
#pragma unroll
    for (int it = 0; it < ITSC; it++) {
        const int w = it * kNT + (int)threadIdx.x;
        if (NSC % kNT != 0 && w >= NSC) continue;
        const int r = w / NSUB, q = w - r * NSUB;
        const double tap = ldg(s.ta + R + r) * sc0;
        double cy = 0;
        bg[it] = 0;
#pragma unroll
        for (int l = 0; l < SUB; l++) {
            const int col = q * SUB + l;
            const double2 v = A[r * G + col];
            const int cd = ldg(code + r * G + col);
            const int cb = 2 * (c0 + col);
            const double x0 = v.x * (tap * ldg(s.tb + 2 * C + cb) * pow2ni(cd & 3));
            const double x1 = v.y * (tap * ldg(s.tb + 2 * C + cb + 1) * pow2ni((cd >> 3) & 3));
            const double r0 = rint(x0), r1 = rint(x1);
            lmax = fmax(lmax, fmax(fabs(x0 - r0), fabs(x1 - r1)));
            const int b0f = (cd >> 2) & 1, b1f = (cd >> 5) & 1;
            bg[it] |= (uint32_t)(b0f | (b1f << 1)) << (2 * l);
            if (sq) {
                double dd;
                cy = bal_f(r0 + cy, b0f ? 2 * bb : bb, b0f ? 0.5 * ibb : ibb, b0f ? bb : 0.5 * bb, dd);
                d[it][2 * l] = (int32_t)dd;
                cy = bal_f(r1 + cy, b1f ? 2 * bb : bb, b1f ? 0.5 * ibb : ibb, b1f ? bb : 0.5 * bb, dd);
                d[it][2 * l + 1] = (int32_t)dd;
            } else {
                d[it][2 * l] = (int32_t)r0;
                d[it][2 * l + 1] = (int32_t)r1;
            }
        }
        int64_t co = (int64_t)cy;
        if (sq && g == NG - 1 && w == NSC - 1) {  // the top fold on the number's last sub-chunk
#pragma unroll
            for (int i = 0; i < 2 * SUB; i++) FD[i] = d[it][i];
            co = fold_top_dev(s, FD, co);
#pragma unroll
            for (int i = 0; i < 2 * SUB; i++) d[it][i] = FD[i];
        }
        SC[w] = co;
    }
    __syncthreads();
    int32_t *out = s.dig + (size_t)g * R * 2 * G;
#pragma unroll
    for (int it = 0; it < ITSC; it++) {
        const int w = it * kNT + (int)threadIdx.x;
        if (NSC % kNT != 0 && w >= NSC) continue;
        const int r = w / NSUB, q = w - r * NSUB;
        if (sq && q > 0) {  // absorb the left neighbour's carry; the last digit keeps the rest
            int64_t cy = SC[w - 1];
#pragma unroll
            for (int i = 0; i < 2 * SUB; i++) {
                if (!cy) break;
                if (i == 2 * SUB - 1) {
                    d[it][i] = (int32_t)(d[it][i] + cy);
                    break;
                }
                cy = bal((int64_t)d[it][i] + cy, s.b0 + (int)((bg[it] >> i) & 1), d[it][i]);
            }
        }
        int32_t *o4 = out + r * 2 * G + 2 * q * SUB;
#pragma unroll
        for (int i = 0; i < 2 * SUB / 4; i++)
            gst4(o4 + 4 * i, make_int4(d[it][4 * i], d[it][4 * i + 1], d[it][4 * i + 2], d[it][4 * i + 3]));
        if (sq && q == NSUB - 1) stg(s.cout + g * R + r, SC[w]);
    }
    for (int off = 32; off > 0; off >>= 1) lmax = fmax(lmax, __shfl_xor(lmax, off));
    if ((threadIdx.x & 63) == 0 && lmax > 0)
        atomicMax(s.maxerr, (unsigned long long)__double_as_longlong(lmax));
}

template <int R>
__device__ __forceinline__ void cols_fwd_fast(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs, int mode,
                                              int step) {
    extern __shared__ double2 lds[];
    using Gm = ColGeom<R>;
    constexpr int G = Gm::G, SUB = Gm::SUB, NSUB = Gm::NSUB, NE = Gm::NE, NSC = Gm::NSC, ITSC = Gm::ITSC;
    constexpr int ITE = (NE + kNT - 1) / kNT;
    const int2 jb = jobs ? jobs[blockIdx.x] : make_int2(0, (int)blockIdx.x);
    const SlotDev &s = slots[jb.x];
    if (step >= (int)s.count) return;
    const int C = s.C, NG = C / G, g = jb.y, c0 = g * G;
    double2 *A = lds, *T = lds + NE;
    for (int i = threadIdx.x; i < R; i += kNT) T[i] = gld(s.twR + i);
    // four-step twiddles for the store, fetched early
    double2 tws[ITE];
#pragma unroll
    for (int it = 0; it < ITE; it++) {
        const int i = it * kNT + (int)threadIdx.x;
        if (NE % kNT != 0 && i >= NE) continue;
        tws[it] = gld(s.twCols + (size_t)g * NE + i);
    }
    const int32_t *in = s.dig + (size_t)g * R * 2 * G;
    const uint8_t *code = s.dcode + (size_t)g * R * G;
    const bool sq = mode != kModeDigits;
#pragma unroll
    for (int it = 0; it < ITSC; it++) {
        const int w = it * kNT + (int)threadIdx.x;
        if (NSC % kNT != 0 && w >= NSC) continue;
        const int r = w / NSUB, q = w - r * NSUB;
        int32_t d[2 * SUB];
        const int32_t *i4 = in + r * 2 * G + 2 * q * SUB;
#pragma unroll
        for (int i = 0; i < 2 * SUB / 4; i++) {
            const int4 v = gld4(i4 + 4 * i);
            d[4 * i] = v.x;
            d[4 * i + 1] = v.y;
            d[4 * i + 2] = v.z;
            d[4 * i + 3] = v.w;
        }
        int cds[SUB];
#pragma unroll
        for (int l = 0; l < SUB; l++) cds[l] = ldg(code + r * G + q * SUB + l);
        if (sq && q == 0) {  // the previous row chunk's carry, absorbed within this sub-chunk
            int pidx;
            if (g > 0) pidx = (g - 1) * R + r;
            else if (r > 0) pidx = (NG - 1) * R + r - 1;
            else pidx = (NG - 1) * R + R - 1;
            int64_t cy = ldg(s.cout + pidx);
#pragma unroll
            for (int i = 0; i < 2 * SUB; i++) {
                if (!cy) break;
                if (i == 2 * SUB - 1) {
                    d[i] = (int32_t)(d[i] + cy);
                    break;
                }
                const int big = (i & 1) ? (cds[i >> 1] >> 5) & 1 : (cds[i >> 1] >> 2) & 1;
                cy = bal((int64_t)d[i] + cy, s.b0 + big, d[i]);
            }
        }
        const double tam = ldg(s.ta + r);
#pragma unroll
        for (int l = 0; l < SUB; l++) {
            const int cb = 2 * (c0 + q * SUB + l);
            const double w0 = tam * ldg(s.tb + cb) * pow2i(cds[l] & 3);
            const double w1 = tam * ldg(s.tb + cb + 1) * pow2i((cds[l] >> 3) & 3);
            A[r * G + q * SUB + l] = make_double2((double)d[2 * l] * w0, (double)d[2 * l + 1] * w1);
        }
    }
    __syncthreads();
    fft_ip<R, 1, 0, G, G, 1, true, false>(A, T);
#pragma unroll
    for (int it = 0; it < ITE; it++) {
        const int i = it * kNT + (int)threadIdx.x;
        if (NE % kNT != 0 && i >= NE) continue;
        const int kr = i / G, l = i - kr * G;
        gst(s.x + (size_t)kr * C + c0 + l, cmul(A[i], tws[it]));
    }
}

// Fused column pass (square mode): cols_inv_fast and cols_fwd_fast in one work-group, the digits
// kept in registers. The carry out of each row chunk (g, r) goes to the work-group of the next
// chunk through flags[g R + r] = seq << 48 | carry (low 48 bits, two's complement), written and
// polled with agent-scope atomics: a single 64-bit word holds both, so no fence is needed. The
// work-group of chunk (g, r) waits for chunk (g-1, r) (g > 0) or (NG-1, r-1) (g = 0; (NG-1, R-1)
// holds the fold); every work-group publishes before it waits, and work-groups are dispatched in
// order, so every wait ends. A wait is bounded: past kSpinLimit polls the slot's round-off is set
// to infinity (the host reports the failure) and the carry is taken as 0.
//
// seq only has to differ from what the word held before this launch (the previous launch of the
// same engine wrote seq - 1), so 16 bits suffice. A carry is below 2^52 / 2^(b0 - 1) in magnitude
// (the digits before the carry are exact FP64 integers); the host keeps b0 >= 6 on this path, so
// |carry| < 2^47 fits the 48-bit field. A carry that does not (only possible if a digit left the
// exact range, which the round-off check reports anyway) also sets the round-off to infinity.
constexpr uint32_t kSeqMask = 0xFFFF;
constexpr int kFlagShift = 48;
constexpr int64_t kFlagCarryMax = ((int64_t)1 << 47) - 1;
constexpr int kSpinLimit = 1 << 22;

__device__ __forceinline__ unsigned long long pack_flag(uint32_t seq, int64_t carry) {
    return ((unsigned long long)(seq & kSeqMask) << kFlagShift) |
           ((unsigned long long)carry & ((1ull << kFlagShift) - 1));
}

__device__ __forceinline__ bool flag_carry_fits(int64_t carry) {
    return carry <= kFlagCarryMax && carry >= -kFlagCarryMax;
}

__device__ __forceinline__ int64_t wait_flag(unsigned long long *f, uint32_t seq, unsigned long long *maxerr) {
    const unsigned long long want = (unsigned long long)(seq & kSeqMask);
    for (int spin = 0;; spin++) {
        const unsigned long long v = __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((v >> kFlagShift) == want)
            return (int64_t)(v << (64 - kFlagShift)) >> (64 - kFlagShift);  // sign-extend the carry
        if (spin >= kSpinLimit) {
            atomicMax(maxerr, 0x7FF0000000000000ull);  // +inf
            return 0;
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

// Balanced split by exponents: x an integer with |x| < 2^52, b its digit width; x 2^-b, + 1/2,
// floor and x - q 2^b are exact. Returns the carry q, sets the digit d.
__device__ __forceinline__ double bal_e(double x, int b, double &d) {
    const double q = floor(ldexp(x, -b) + 0.5);
    d = x - ldexp(q, b);
    return q;
}

// Threads per work-group of the fused column pass at column length R: the most any of its phases
// uses (the FFT stages, one thread per carry sub-chunk), in whole waves, at most kNT. At R = 96
// (G = 16) every phase needs at most 192 threads: a fourth wave would hold registers and idle, and
// three-wave work-groups fit five to a CU where four-wave ones fit four.
template <int N, int HOW, int I = 0> constexpr int io_max_total() {
    if constexpr (I >= io_nstages(N)) return 0;
    else {
        constexpr int t = N / io_radix(N, I) * HOW, r = io_max_total<N, HOW, I + 1>();
        return t > r ? t : r;
    }
}
template <int R> constexpr int fused_nt() {
    constexpr int G = g_for_R(R), SUB = G < kSub ? G : kSub, NSC = R * (G / SUB);
    constexpr int m = io_max_total<R, G>() > NSC ? io_max_total<R, G>() : NSC;
    constexpr int w = (m + 63) / 64 * 64;
    return w < kNT ? w : kNT;
}

template <int R, int NT = kNT> struct FusedGeom {
    static constexpr int G = g_for_R(R), GP = G + 1, SUB = G < kSub ? G : kSub, NSUB = G / SUB;
    static constexpr int NE = R * G, NSC = R * NSUB, ITSC = (NSC + NT - 1) / NT, ITE = (NE + NT - 1) / NT;
    // LDS: R x G elements with row stride GP (the padding spreads the sub-chunk reads over the
    // banks), the column twiddles, one carry per sub-chunk, the group's weight factors (inverse and
    // forward: 2G column and R row values each), a 64-digit scratch for the fold. (The digit codes
    // stay in registers: at R = 96, G = 16 this is 31.5 KB, five work-groups per CU.)
    static constexpr size_t lds_bytes() {
        return (size_t)16 * (R * GP + R) + (size_t)8 * NSC + (size_t)8 * (4 * G + 2 * R) + 4 * 64;
    }
};

// OT: the work-item id through tid_x (for the persistent kernels, which loop over jobs); NT: threads
// per work-group (any multiple of 64: phases needing more threads iterate)
template <int R, bool OT = false, int NT = kNT>
__device__ __forceinline__ void cols_fused_fast(const SlotDev &s, int g, int step) {
    extern __shared__ double2 lds[];
    using Gm = FusedGeom<R, NT>;
    constexpr int G = Gm::G, GP = Gm::GP, SUB = Gm::SUB, NSUB = Gm::NSUB, NE = Gm::NE, NSC = Gm::NSC;
    constexpr int ITSC = Gm::ITSC, ITE = Gm::ITE;
    {
        // the slot fields of the prologue in one scalar round trip: issued together here, not one
        // by one after the step test
        const uint32_t count = s.count;
        const int C0 = s.C;
        const void *p0 = s.x, *p1 = s.twR, *p2 = s.twCols, *p3 = s.ta, *p4 = s.tb, *p5 = s.dcode;
        asm volatile("" ::"s"(count), "s"(C0), "s"(p0), "s"(p1), "s"(p2), "s"(p3), "s"(p4), "s"(p5));
        if (step >= (int)count) return;
    }
    const int tid = OT ? tid_x() : (int)threadIdx.x;
    const int C = s.C, NG = C / G, c0 = g * G;
    const uint32_t seq = s.seq + (uint32_t)step + 1;
    double2 *A = lds, *T = lds + R * GP;
    int64_t *SC = reinterpret_cast<int64_t *>(T + R);
    double *WB = reinterpret_cast<double *>(SC + NSC);  // [2G] inverse column, [2G] forward column
    double *WA = WB + 4 * G;                             // [R] inverse row (with k/(4H)), [R] forward row
    int32_t *FD = reinterpret_cast<int32_t *>(WA + 2 * R);
    // The group's tables (column twiddles, weight factors) are loaded into registers here and
    // stored to LDS by pre() below, once the first FFT stage has issued its HBM loads: their round
    // trips overlap instead of preceding the residue's.
    constexpr int NTR = (R + NT - 1) / NT, NTB = (2 * G + NT - 1) / NT;
    double2 rT[NTR];
    double rWA[NTR], rWA2[NTR], rWB[NTB], rWB2[NTB];
    // The digit codes of the thread's own sub-chunks (SUB bytes each: row r, columns q SUB ..),
    // fetched now and kept in registers for the carry and the weighting.
    static_assert(SUB == 4 || SUB == 8, "digit codes are fetched as 4 or 8 bytes");
    uint64_t cdw[ITSC];
#pragma unroll
    for (int it = 0; it < ITSC; it++) {
        const int w = it * NT + tid;
        cdw[it] = 0;
        if (w >= opaque_bound(NSC)) continue;
        const int q = w / R, r = w - q * R;
        const uint8_t *cp = s.dcode + ((size_t)g * R + r) * G + q * SUB;
        if constexpr (SUB == 8) cdw[it] = ldg(reinterpret_cast<const unsigned long long *>(cp));
        else cdw[it] = ldg(reinterpret_cast<const uint32_t *>(cp));
    }
#pragma unroll
    for (int q = 0; q < NTR; q++) {
        const int i = q * NT + tid;
        if (R % NT == 0 || i < R) {
            rT[q] = gld(s.twR + i);
            rWA[q] = ldg(s.ta + R + i);
            rWA2[q] = ldg(s.ta + i);
        }
    }
#pragma unroll
    for (int q = 0; q < NTB; q++) {
        const int i = q * NT + tid;
        if ((2 * G) % NT == 0 || i < 2 * G) {
            rWB[q] = ldg(s.tb + 2 * C + 2 * c0 + i);
            rWB2[q] = ldg(s.tb + 2 * c0 + i);
        }
    }
    const double kscale = s.kscale;
    auto pre = [&]() {
#pragma unroll
        for (int q = 0; q < NTR; q++) {
            const int i = q * NT + tid;
            if (R % NT == 0 || i < R) {
                T[i] = rT[q];
                WA[i] = rWA[q] * kscale;
                WA[R + i] = rWA2[q];
            }
        }
#pragma unroll
        for (int q = 0; q < NTB; q++) {
            const int i = q * NT + tid;
            if ((2 * G) % NT == 0 || i < 2 * G) {
                WB[i] = rWB[q];
                WB[2 * G + i] = rWB2[q];
            }
        }
    };
    const double2 *tg = s.twCols + (size_t)g * NE;
    // inverse column FFT, its first stage reading the residue (times the conjugate four-step
    // twiddle) straight from HBM; the tables above are published by its first barrier
    {
        const double2 *xin = s.x + c0;
        auto ld = [&](int, int, int row, int col) {
            if constexpr (cols_tw2<R>()) return cmulc(gld(xin + (size_t)row * C + col), tw_h2(s, (c0 + col) * row));
            else return cmulc(gld(xin + (size_t)row * C + col), gld(tg + row * G + col));
        };
        auto st = [](int, int, int, int, double2) {};
        fft_io<R, 1, 0, G, GP, true, true, false, OT, NT>(A, T, ld, st, pre);
    }
    const int b0 = s.b0;
    double lmax = 0;
    int32_t d[ITSC][2 * SUB];
    uint32_t bg[ITSC];  // bit i: digit i of the sub-chunk has b0+1 bits
#pragma unroll
    for (int it = 0; it < ITSC; it++) {
        const int w = it * NT + tid;
        if (w >= opaque_bound(NSC)) continue;
        // sub-chunk (r, q) is w = q R + r: the lanes of a wave take consecutive rows, so their
        // reads of A (row stride G + 1) fall in distinct LDS banks
        const int q = w / R, r = w - q * R;
        const double tap = WA[r];
        double cy = 0;
        bg[it] = 0;
#pragma unroll
        for (int l = 0; l < SUB; l++) {
            const int col = q * SUB + l;
            const double2 v = A[r * GP + col];
            const int cd = (int)(cdw[it] >> (8 * l)) & 0xFF;
            const double x0 = ldexp(v.x * (tap * WB[2 * col]), -(cd & 3));
            const double x1 = ldexp(v.y * (tap * WB[2 * col + 1]), -((cd >> 3) & 3));
            const double r0 = rint(x0), r1 = rint(x1);
            lmax = fmax(lmax, fmax(fabs(x0 - r0), fabs(x1 - r1)));
            const int b0f = (cd >> 2) & 1, b1f = (cd >> 5) & 1;
            bg[it] |= (uint32_t)(b0f | (b1f << 1)) << (2 * l);
            double dd;
            cy = bal_e(r0 + cy, b0 + b0f, dd);
            d[it][2 * l] = (int32_t)dd;
            cy = bal_e(r1 + cy, b0 + b1f, dd);
            d[it][2 * l + 1] = (int32_t)dd;
            // the chain is sequential: reading further ahead only costs registers
            if (l % 2 == 1) __builtin_amdgcn_sched_barrier(0);
        }
        const int64_t co = (int64_t)cy;
        const bool folder = g == NG - 1 && w == NSC - 1;  // holds the number's last sub-chunk
        if (folder) {
#pragma unroll
            for (int i = 0; i < 2 * SUB; i++) FD[i] = d[it][i];
        }
        SC[w] = co;
        if (q == NSUB - 1 && !folder) {  // this row chunk's carry, for the next chunk's work-group
            if (!flag_carry_fits(co)) atomicMax(s.maxerr, 0x7FF0000000000000ull);
            __hip_atomic_store(s.flags + g * R + r, pack_flag(seq, co), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    __syncthreads();
    // the top fold, here where few registers are live (its 128-bit arithmetic inside the loop above
    // made every thread spill); the last sub-chunk's carry goes nowhere inside the work-group
    if (g == NG - 1 && tid == (NSC - 1) % NT) {
        constexpr int it = ITSC - 1;
        const int64_t co = fold_top_dev(s, FD, SC[NSC - 1]);
        SC[NSC - 1] = co;
        if (!flag_carry_fits(co)) atomicMax(s.maxerr, 0x7FF0000000000000ull);
        __hip_atomic_store(s.flags + g * R + R - 1, pack_flag(seq, co), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
        for (int i = 0; i < 2 * SUB; i++) d[it][i] = FD[i];
    }
    // absorb the left neighbour's carry (within the work-group, or the previous chunk's through its
    // flag); the last digit keeps the rest. Then weight into A for the forward column FFT.
#pragma unroll
    for (int it = 0; it < ITSC; it++) {
        const int w = it * NT + tid;
        if (w >= opaque_bound(NSC)) continue;
        const int q = w / R, r = w - q * R;
        int64_t cy;
        if (q > 0) {
            cy = SC[w - R];
        } else {
            const int pidx = g > 0 ? (g - 1) * R + r : (r > 0 ? (NG - 1) * R + r - 1 : (NG - 1) * R + R - 1);
            cy = wait_flag(s.flags + pidx, seq, s.maxerr);
        }
#pragma unroll
        for (int i = 0; i < 2 * SUB; i++) {
            if (!cy) break;
            if (i == 2 * SUB - 1) {
                d[it][i] = (int32_t)(d[it][i] + cy);
                break;
            }
            cy = bal((int64_t)d[it][i] + cy, b0 + (int)((bg[it] >> i) & 1), d[it][i]);
        }
        const double tam = WA[R + r];
#pragma unroll
        for (int l = 0; l < SUB; l++) {
            const int col = q * SUB + l;
            const int cd = (int)(cdw[it] >> (8 * l)) & 0xFF;
            const double w0 = ldexp(tam * WB[2 * G + 2 * col], cd & 3);
            const double w1 = ldexp(tam * WB[2 * G + 2 * col + 1], (cd >> 3) & 3);
            A[r * GP + col] = make_double2((double)d[it][2 * l] * w0, (double)d[it][2 * l + 1] * w1);
        }
    }
    // forward column FFT, its last stage writing the residue (times the four-step twiddle) straight
    // to HBM. The twiddles of the thread's last-stage outputs are fetched now, during the FFT; the
    // bases are laundered so that the inverse side's 64-bit addresses are not kept live (spilled)
    // through the whole kernel for reuse here.
    using LS = LastStage<R, G, NT>;
    constexpr bool TW2 = cols_tw2<R>();
    double2 tws[TW2 ? 1 : LS::ITER][LS::P];  // the table's twiddles (two-level: formed at the store)
    const double2 *tgs = tg;
    double2 *xo = s.x + c0;
    int Co = C;
    asm volatile("" : "+v"(tgs), "+v"(xo), "+v"(Co));
    if constexpr (!TW2) {
#pragma unroll
        for (int it = 0; it < LS::ITER; it++) {
            const int w = it * NT + tid;
            if (w < opaque_bound(LS::TOTAL)) {
#pragma unroll
                for (int r = 0; r < LS::P; r++) tws[it][r] = gld(tgs + (w / G + r * LS::NB) * G + w % G);
            }
        }
    }
    __syncthreads();
    {
        auto ld = [](int, int, int, int) { return make_double2(0, 0); };
        auto st = [&](int it, int r, int row, int col, double2 v) {
            if constexpr (TW2) gst(xo + row * Co + col, cmul(v, tw_h2(s, (c0 + col) * row)));
            else gst(xo + row * Co + col, cmul(v, tws[it][r]));
        };
        fft_io<R, 1, 0, G, GP, false, false, true, OT, NT>(A, T, ld, st);
    }
    for (int off = 32; off > 0; off >>= 1) lmax = fmax(lmax, __shfl_xor(lmax, off));
    if ((tid & 63) == 0 && lmax > 0)
        atomicMax(s.maxerr, (unsigned long long)__double_as_longlong(lmax));
}

}  // namespace kite

#endif
