// gfx950 squaring engine: one squaring mod N = k*2^n - 1 as three kernels over the four-step
// layout of the packed complex transform (H = L/2 = R x C complex doubles, row-major).
//
//   k_rows      per row pair (k_r, R-k_r): row FFT (length C) -> real-split square or multiply
//               -> inverse row FFT                                   (replaces the reference's
//               pass 2: src/neon.c:987-1054, src/fast2p.c:545-617)
//   k_cols_inv  per column group: conj four-step twiddle -> inverse column FFT (length R) ->
//               scale k/(4H w_j), round, int64 carry along each row chunk of 2G digits, the
//               128-bit top fold on the last chunk; writes int32 digits and one carry per chunk
//   k_cols_fwd  per column group: absorb the previous chunk's carry (the last digit of a chunk
//               keeps any remainder), weight, forward column FFT, four-step twiddle
//               (together the reference's pass 1: src/neon.c:811-877, src/fast2p.c:453-491)
//
// The reference carries sequentially across column groups with a deferred first group; here
// every chunk's carry-out is independent of its carry-in (one-step absorption), so all column
// groups of all tests run in parallel and the carries cross work-groups through a kernel
// boundary. FFTs are Stockham autosort in LDS (natural order in and out) with ping-pong buffers,
// radix 2/3/4/5/7/8 butterflies in registers, FP64 throughout; every twiddle and weight comes
// from host tables computed with exact argument reduction.
//
// Work is described by per-work-group job records (slot, unit) over a table of slots, so the
// same kernels run one test (kite.h API) or many tests of different lengths in one launch.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "engine_dev.h"
#include "kite_batch.h"
#include "kite_internal.h"

typedef unsigned __int128 u128;
typedef __int128 i128;

namespace kite {

static thread_local std::string g_err;
const char *engine_last_error() { return g_err.c_str(); }

#define HIP_OK(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            g_err = std::string(#expr) + ": " + hipGetErrorString(e_);                    \
            fprintf(stderr, "kite: %s\n", g_err.c_str());                                 \
            return -1;                                                                    \
        }                                                                                 \
    } while (0)

// ================================================================================================
// Device code
// ================================================================================================

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // a * conj(b)
    return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// multiply by -i (forward) or +i (inverse)
template <bool INV> __device__ __forceinline__ double2 mul_mi(double2 a) {
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

// odd p: pairs (x_j, x_{p-j}) with s = sum, d = difference; y_k = A -+ iB, y_{p-k} = A +- iB
template <int P> struct OddRoots;
template <> struct OddRoots<3> {
    static __device__ __forceinline__ double c(int q) { return -0.5; }
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
            // forward: y_k = A - iB, y_{P-k} = A + iB
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

// ---- Stockham stage in LDS ------------------------------------------------------------------
// `howmany` sequences of length n; element (seq, idx) at buf[idx*is + seq*ss]. Stage with radix P
// after Ns points have been combined: butterfly j reads idx j + r*n/P, twiddles by
// w_{Ns P}^{(j mod Ns) r} = tw[(j mod Ns) r n/(Ns P)], writes idx (j/Ns) Ns P + (j mod Ns) + r Ns.
template <int P, bool INV>
__device__ __forceinline__ void fft_stage(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                          int n, int Ns, int howmany, int is, int ss, bool seqfast,
                                          const double2 *__restrict__ tw) {
    const int nb = n / P, total = nb * howmany, ts = n / (Ns * P);
    for (int w = threadIdx.x; w < total; w += blockDim.x) {
        int seq, j;
        if (seqfast) {
            seq = w % howmany;
            j = w / howmany;
        } else {
            j = w % nb;
            seq = w / nb;
        }
        const int jm = j % Ns;
        double2 v[P];
#pragma unroll
        for (int r = 0; r < P; r++) v[r] = src[(j + r * nb) * is + seq * ss];
        if (jm) {
#pragma unroll
            for (int r = 1; r < P; r++) {
                double2 t = tw[jm * r * ts];
                v[r] = INV ? cmulc(v[r], t) : cmul(v[r], t);
            }
        }
        Dft<P, INV>::run(v);
        const int base = (j / Ns) * Ns * P + jm;
#pragma unroll
        for (int r = 0; r < P; r++) dst[(base + r * Ns) * is + seq * ss] = v[r];
    }
}

template <bool INV>
__device__ double2 *lds_fft(double2 *a, double2 *b, const Plan &pl, int howmany, int is, int ss,
                            bool seqfast, const double2 *tw) {
    int Ns = 1;
    for (int s = 0; s < pl.nst; s++) {
        const int p = pl.rad[s];
        switch (p) {
        case 2: fft_stage<2, INV>(a, b, pl.n, Ns, howmany, is, ss, seqfast, tw); break;
        case 3: fft_stage<3, INV>(a, b, pl.n, Ns, howmany, is, ss, seqfast, tw); break;
        case 4: fft_stage<4, INV>(a, b, pl.n, Ns, howmany, is, ss, seqfast, tw); break;
        case 5: fft_stage<5, INV>(a, b, pl.n, Ns, howmany, is, ss, seqfast, tw); break;
        case 7: fft_stage<7, INV>(a, b, pl.n, Ns, howmany, is, ss, seqfast, tw); break;
        default: fft_stage<8, INV>(a, b, pl.n, Ns, howmany, is, ss, seqfast, tw); break;
        }
        __syncthreads();
        double2 *t = a;
        a = b;
        b = t;
        Ns *= p;
    }
    return a;
}

// w_H^t for t < H from the two-level table
__device__ __forceinline__ double2 tw_h(const ShapeDev &sh, int t) {
    return cmul(sh.thi[t >> sh.sshift], sh.tlo[t & ((1 << sh.sshift) - 1)]);
}

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

__device__ __forceinline__ const SlotDev &pick_slot(const SlotDev *slots, const SlotDev &one, int i) {
    return slots ? slots[i] : one;
}

// rows handled by row job u: (0, R/2) for u = 0 (both self-paired), else (u, R-u)
__device__ __forceinline__ void row_pair(int R, int u, int &ra, int &rb) {
    if (u == 0) {
        ra = 0;
        rb = (R % 2 == 0 && R > 1) ? R / 2 : -1;
    } else {
        ra = u;
        rb = R - u;
    }
}

// ---- kernels -----------------------------------------------------------------------------------

extern "C" __global__ void __launch_bounds__(kNT) k_rows(const SlotDev *slots, SlotDev one,
                                                      const int2 *jobs, int mode, int step) {
    extern __shared__ double2 lds[];
    const int2 jb = jobs ? jobs[blockIdx.x] : make_int2(0, (int)blockIdx.x);
    const SlotDev &s = pick_slot(slots, one, jb.x);
    if (step >= (int)s.count) return;  // this slot sits out this launch (uniform per work-group)
    const ShapeDev &sh = *s.shp;
    const int R = sh.R, C = sh.C;
    int ra, rb;
    row_pair(R, jb.y, ra, rb);
    const int nrows = rb >= 0 ? 2 : 1;
    double2 *A = lds, *B = lds + 2 * C;
    double2 *x = mode == kModeRowFft ? const_cast<double2 *>(s.y_src) : s.x;
    for (int i = threadIdx.x; i < nrows * C; i += blockDim.x) {
        const int sr = i >= C, c = i - sr * C;
        A[i] = x[(size_t)(sr ? rb : ra) * C + c];
    }
    __syncthreads();
    double2 *F = lds_fft<false>(A, B, sh.pc, nrows, 1, C, false, sh.twC);
    if (mode == kModeRowFft) {  // operand preparation: store the row spectra
        for (int i = threadIdx.x; i < nrows * C; i += blockDim.x) {
            const int sr = i >= C, c = i - sr * C;
            s.yfreq[(size_t)(sr ? rb : ra) * C + c] = F[i];
        }
        return;
    }
    const bool mul = mode == kModeMul;
    const double2 *Y = s.yfreq;
    if (jb.y == 0) {
        // self-paired rows: row 0 pairs k_c with (C-k_c) mod C, row R/2 pairs k_c with C-1-k_c
        for (int i = threadIdx.x; i < nrows * C; i += blockDim.x) {
            const int sr = i >= C, kc = i - sr * C;
            const int pc = sr ? C - 1 - kc : (C - kc) % C;
            if (pc < kc) continue;
            const int kr = sr ? rb : 0;
            double2 a = F[sr * C + kc], b = F[sr * C + pc];
            const double2 t = tw_h(sh, kr + R * kc);
            if (mul) split_mul(a, b, t, Y[(size_t)kr * C + kc], Y[(size_t)kr * C + pc]);
            else split_sq(a, b, t);
            F[sr * C + kc] = a;
            if (pc != kc) F[sr * C + pc] = b;
        }
    } else {
        for (int kc = threadIdx.x; kc < C; kc += blockDim.x) {
            const int pc = C - 1 - kc;
            double2 a = F[kc], b = F[C + pc];
            const double2 t = tw_h(sh, ra + R * kc);
            if (mul) split_mul(a, b, t, Y[(size_t)ra * C + kc], Y[(size_t)rb * C + pc]);
            else split_sq(a, b, t);
            F[kc] = a;
            F[C + pc] = b;
        }
    }
    __syncthreads();
    double2 *Gb = lds_fft<true>(F, F == A ? B : A, sh.pc, nrows, 1, C, false, sh.twC);
    for (int i = threadIdx.x; i < nrows * C; i += blockDim.x) {
        const int sr = i >= C, c = i - sr * C;
        s.x[(size_t)(sr ? rb : ra) * C + c] = Gb[i];
    }
}

// floor(h / k) for 0 < k < 2^32
__device__ __forceinline__ i128 floordiv_small(i128 h, uint32_t k) {
    const bool neg = h < 0;
    u128 a = neg ? (u128)(-h) : (u128)h;
    uint64_t hi = (uint64_t)(a >> 64), lo = (uint64_t)a;
    uint64_t qh = hi / k, r = hi % k;
    uint64_t t = (r << 32) | (lo >> 32);
    uint64_t q1 = t / k;
    r = t % k;
    t = (r << 32) | (lo & 0xffffffffull);
    uint64_t q0 = t / k;
    r = t % k;
    u128 q = ((u128)qh << 64) | ((u128)q1 << 32) | q0;
    if (!neg) return (i128)q;
    return -(i128)q - (r ? 1 : 0);
}

// Fold the top region (digits top_t..L-1 of the last chunk plus the carry at 2^{s_L}):
// H = q*k + r  ->  q into digit 0, r*2^m rebalanced into the top digits (src/ctx.c:345-375).
__device__ int64_t fold_top_dev(const SlotDev &s, int32_t *d, int64_t carry) {
    const int t0 = s.top_i0, nt = s.top_n, m = s.top_m;
    i128 Hh = (i128)carry << s.top_shift;
    int pos = 0;
    for (int i = 0; i < nt; i++) {
        Hh += (i128)d[t0 + i] << pos;
        pos += s.top_bits[i];
    }
    const i128 K = (i128)s.k;
    const i128 lo = Hh & (((i128)1 << m) - 1), hi = Hh >> m;
    i128 q = floordiv_small(hi, (uint32_t)s.k);
    i128 top = lo + ((hi - q * K) << m);
    if (2 * top >= (K << m)) {
        top -= K << m;
        q += 1;
    }
    for (int i = 0; i + 1 < nt; i++) {
        const int b = s.top_bits[i];
        const i128 qq = (top + ((i128)1 << (b - 1))) >> b;
        d[t0 + i] = (int32_t)(top - (qq << b));
        top = qq;
    }
    if (nt > 0) d[t0 + nt - 1] = (int32_t)top;
    return (int64_t)q;
}

__device__ __forceinline__ double pow2i(int e) { return e == 0 ? 1.0 : e == 1 ? 2.0 : 4.0; }
__device__ __forceinline__ double pow2ni(int e) { return e == 0 ? 1.0 : e == 1 ? 0.5 : 0.25; }

extern "C" __global__ void __launch_bounds__(kNT) k_cols_inv(const SlotDev *slots, SlotDev one,
                                                      const int2 *jobs, int mode, int step) {
    extern __shared__ double2 lds[];
    const int2 jb = jobs ? jobs[blockIdx.x] : make_int2(0, (int)blockIdx.x);
    const SlotDev &s = pick_slot(slots, one, jb.x);
    if (step >= (int)s.count) return;  // this slot sits out this launch (uniform per work-group)
    const ShapeDev &sh = *s.shp;
    const int R = sh.R, C = sh.C, G = sh.G, NG = sh.NG, g = jb.y, c0 = g * G;
    double2 *A = lds, *B = lds + R * G;
    for (int i = threadIdx.x; i < R * G; i += blockDim.x) {
        const int kr = i / G, l = i - kr * G;
        A[i] = cmulc(s.x[(size_t)kr * C + c0 + l], tw_h(sh, (c0 + l) * kr));
    }
    __syncthreads();
    double2 *F = lds_fft<true>(A, B, sh.pr, G, G, 1, true, sh.twR);
    int32_t *D = reinterpret_cast<int32_t *>(F == A ? B : A);
    const bool sq = mode != kModeDigits;
    const double sc0 = sq ? s.kscale : s.rscale;
    const uint8_t *code = s.dcode + (size_t)g * R * G;
    double lmax = 0;
    for (int r = threadIdx.x; r < R; r += blockDim.x) {
        const double tap = s.ta[R + r];
        int64_t cy = 0;
        int32_t *dr = D + r * 2 * G;
        for (int l = 0; l < G; l++) {
            const double2 v = F[r * G + l];
            const int cd = code[r * G + l];
            const int cb = 2 * (c0 + l);
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int ce = h ? (cd >> 3) & 3 : cd & 3;
                const int big = h ? (cd >> 5) & 1 : (cd >> 2) & 1;
                const double x = (h ? v.y : v.x) * (sc0 * tap * s.tb[2 * C + cb + h] * pow2ni(ce));
                const double rx = rint(x);
                lmax = fmax(lmax, fabs(x - rx));
                if (sq) {
                    const int b = s.b0 + big;
                    const int64_t t = (int64_t)rx + cy;
                    const int64_t q = (t + ((int64_t)1 << (b - 1))) >> b;
                    dr[2 * l + h] = (int32_t)(t - (q << b));
                    cy = q;
                } else {
                    dr[2 * l + h] = (int32_t)rx;
                }
            }
        }
        if (sq) {
            if (g == NG - 1 && r == R - 1) cy = fold_top_dev(s, dr, cy);
            s.cout[g * R + r] = cy;
        }
    }
    for (int off = 32; off > 0; off >>= 1) lmax = fmax(lmax, __shfl_xor(lmax, off));
    if ((threadIdx.x & 63) == 0 && lmax > 0)
        atomicMax(s.maxerr, (unsigned long long)__double_as_longlong(lmax));
    __syncthreads();
    int32_t *out = s.dig + (size_t)g * R * 2 * G;
    for (int i = threadIdx.x; i < R * 2 * G; i += blockDim.x) out[i] = D[i];
}

extern "C" __global__ void __launch_bounds__(kNT) k_cols_fwd(const SlotDev *slots, SlotDev one,
                                                      const int2 *jobs, int mode, int step) {
    extern __shared__ double2 lds[];
    const int2 jb = jobs ? jobs[blockIdx.x] : make_int2(0, (int)blockIdx.x);
    const SlotDev &s = pick_slot(slots, one, jb.x);
    if (step >= (int)s.count) return;  // this slot sits out this launch (uniform per work-group)
    const ShapeDev &sh = *s.shp;
    const int R = sh.R, C = sh.C, G = sh.G, NG = sh.NG, g = jb.y, c0 = g * G;
    double2 *A = lds, *B = lds + R * G;
    int32_t *D = reinterpret_cast<int32_t *>(B);
    const int32_t *in = s.dig + (size_t)g * R * 2 * G;
    for (int i = threadIdx.x; i < R * 2 * G; i += blockDim.x) D[i] = in[i];
    __syncthreads();
    const uint8_t *code = s.dcode + (size_t)g * R * G;
    if (mode != kModeDigits) {
        for (int r = threadIdx.x; r < R; r += blockDim.x) {
            int pidx;  // chunk before (g, r) in digit order, cyclically
            if (g > 0) pidx = (g - 1) * R + r;
            else if (r > 0) pidx = (NG - 1) * R + r - 1;
            else pidx = (NG - 1) * R + R - 1;
            int64_t cy = s.cout[pidx];
            int32_t *dr = D + r * 2 * G;
            for (int i = 0; i < 2 * G && cy; i++) {
                if (i == 2 * G - 1) {
                    dr[i] = (int32_t)(dr[i] + cy);
                    break;
                }
                const int cd = code[r * G + (i >> 1)];
                const int big = (i & 1) ? (cd >> 5) & 1 : (cd >> 2) & 1;
                const int b = s.b0 + big;
                const int64_t t = (int64_t)dr[i] + cy;
                const int64_t q = (t + ((int64_t)1 << (b - 1))) >> b;
                dr[i] = (int32_t)(t - (q << b));
                cy = q;
            }
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < R * G; i += blockDim.x) {
        const int r = i / G, l = i - r * G;
        const int cd = code[i];
        const double tam = s.ta[r];
        const int cb = 2 * (c0 + l);
        const double w0 = tam * s.tb[cb] * pow2i(cd & 3);
        const double w1 = tam * s.tb[cb + 1] * pow2i((cd >> 3) & 3);
        A[i] = make_double2((double)D[r * 2 * G + 2 * l] * w0, (double)D[r * 2 * G + 2 * l + 1] * w1);
    }
    __syncthreads();
    double2 *F = lds_fft<false>(A, B, sh.pr, G, G, 1, true, sh.twR);
    for (int i = threadIdx.x; i < R * G; i += blockDim.x) {
        const int kr = i / G, l = i - kr * G;
        s.x[(size_t)kr * C + c0 + l] = cmul(F[i], tw_h(sh, (c0 + l) * kr));
    }
}

extern "C" __global__ void k_addsub(double2 *x, const double2 *y, int n, double sign) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        double2 a = x[i], b = y[i];
        x[i] = make_double2(a.x + sign * b.x, a.y + sign * b.y);
    }
}

extern "C" __global__ void k_set_u64(unsigned long long *p, unsigned long long v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ================================================================================================
// Host side
// ================================================================================================

// cos/sin(2 pi j / n): reduce j/n to [0, 1/8] with exact integer steps, evaluate in long double
static void root_of_unity(uint64_t j, uint64_t n, double *c, double *s) {
    j %= n;
    bool ns = false, nc = false, sw = false;
    uint64_t num = j, den = n;
    if (2 * num > den) { num = den - num; ns = true; }          // 2pi - x
    if (4 * num > den) { num = den - 2 * num; den *= 2; nc = true; }  // pi - x
    if (8 * num > den) { num = den - 4 * num; den *= 4; sw = true; }  // pi/2 - x
    const long double x = 6.283185307179586476925286766559005768L * (long double)num / (long double)den;
    long double cc = cosl(x), ss = sinl(x);
    if (sw) std::swap(cc, ss);
    if (nc) cc = -cc;
    if (ns) ss = -ss;
    *c = (double)cc;
    *s = (double)ss;
}

// forward roots w_n^t = e^{-2 pi i t/n}
static std::vector<double2> roots_table(uint64_t n, uint64_t count, uint64_t step) {
    std::vector<double2> t(count);
    for (uint64_t i = 0; i < count; i++) {
        double c, s;
        root_of_unity(i * step % n, n, &c, &s);
        t[i] = make_double2(c, -s);
    }
    return t;
}

static bool make_plan(uint32_t n, Plan *pl) {
    pl->n = (int)n;
    pl->nst = 0;
    uint32_t m = n;
    while (m % 8 == 0 && pl->nst < kMaxStages) { pl->rad[pl->nst++] = 8; m /= 8; }
    for (uint32_t p : {4u, 2u, 3u, 5u, 7u})
        while (m % p == 0 && pl->nst < kMaxStages) { pl->rad[pl->nst++] = (int)p; m /= p; }
    return m == 1;
}

static bool smooth7(uint32_t n) {
    for (uint32_t p : {2u, 3u, 5u, 7u})
        while (n % p == 0) n /= p;
    return n == 1;
}

// Choose H = R x C and the column group width G for length L (top_digits = L - top_t, 0 if
// unknown). Rows pass: 2 rows of C in LDS (ping-pong: 64 C bytes); column pass: G columns of R
// (32 G R bytes). Preference: C about 4R (balanced passes), larger G for coalesced strips.
struct ShapeChoice {
    uint32_t R = 0, C = 0, G = 0;
};

static bool choose_shape(uint32_t L, uint32_t top_digits, uint32_t variant, ShapeChoice *out) {
    if (L % 2 || L < 64) return false;
    const uint32_t H = L / 2;
    std::vector<ShapeChoice> cands;
    std::vector<double> costs;
    for (uint32_t C = 16; C <= kMaxC && C <= H; C++) {
        if (H % C || C % 4 || !smooth7(C)) continue;
        const uint32_t R = H / C;
        if (R > kMaxR || !smooth7(R)) continue;
        uint32_t G = 0;
        for (uint32_t g = 32; g >= 4; g /= 2)
            if (C % g == 0 && (uint64_t)g * R * 32 <= kMaxColLds) { G = g; break; }
        if (!G) continue;
        if (top_digits && 2 * G < top_digits) continue;
        double cost = fabs(log2((double)C) - log2((double)R) - 2.0);
        if (G < 8) cost += 2;
        cands.push_back({R, C, G});
        costs.push_back(cost);
    }
    if (cands.empty()) return false;
    // rank by cost; `variant` picks the (variant)-th best (autotuner), 0 = best
    std::vector<size_t> idx(cands.size());
    for (size_t i = 0; i < idx.size(); i++) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return costs[a] < costs[b]; });
    size_t pick = variant < idx.size() ? variant : 0;
    *out = cands[idx[pick]];
    return true;
}

bool engine_shape_ok(uint32_t L, uint32_t top_digits) {
    ShapeChoice sc;
    return choose_shape(L, top_digits, 0, &sc);
}

// Shapes (twiddle tables) are shared by every context of the same (R, C, G) on a device.
struct ShapeEntry {
    ShapeChoice ch;
    int device = 0;
    ShapeDev host{};
    ShapeDev *dev = nullptr;
    std::vector<void *> bufs;
};

static std::mutex g_shape_mu;
static std::vector<ShapeEntry *> g_shapes;

template <class T> static int upload(const std::vector<T> &v, T **out, std::vector<void *> &bufs) {
    HIP_OK(hipMalloc((void **)out, v.size() * sizeof(T) + 16));
    HIP_OK(hipMemcpy(*out, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    bufs.push_back(*out);
    return 0;
}

static ShapeEntry *get_shape(const ShapeChoice &ch) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(g_shape_mu);
    for (ShapeEntry *e : g_shapes)
        if (e->device == dev && e->ch.R == ch.R && e->ch.C == ch.C && e->ch.G == ch.G) return e;
    auto *e = new ShapeEntry();
    e->ch = ch;
    e->device = dev;
    ShapeDev &h = e->host;
    h.R = (int)ch.R;
    h.C = (int)ch.C;
    h.G = (int)ch.G;
    h.NG = (int)(ch.C / ch.G);
    const uint64_t H = (uint64_t)ch.R * ch.C;
    if (!make_plan(ch.R, &h.pr) || !make_plan(ch.C, &h.pc)) {
        delete e;
        return nullptr;
    }
    int sh = 0;
    while ((1ull << (2 * sh)) < H) sh++;
    h.sshift = sh;
    const uint64_t S = 1ull << sh, nhi = (H + S - 1) / S;
    double2 *twR, *twC, *tlo, *thi;
    if (upload(roots_table(ch.R, ch.R, 1), &twR, e->bufs) || upload(roots_table(ch.C, ch.C, 1), &twC, e->bufs) ||
        upload(roots_table(H, S, 1), &tlo, e->bufs) || upload(roots_table(H, nhi, S), &thi, e->bufs)) {
        delete e;
        return nullptr;
    }
    h.twR = twR;
    h.twC = twC;
    h.tlo = tlo;
    h.thi = thi;
    std::vector<ShapeDev> one(1, h);
    if (upload(one, &e->dev, e->bufs)) {
        delete e;
        return nullptr;
    }
    g_shapes.push_back(e);
    return e;
}

struct Engine {
    hipStream_t stream = nullptr;
    ShapeEntry *shape = nullptr;
    SlotDev slot{};
    double *ta = nullptr, *tb = nullptr;
    uint8_t *dcode = nullptr;
    int32_t *dig = nullptr;
    int64_t *cout = nullptr;
    double2 *yfreq = nullptr;
    unsigned long long *maxerr = nullptr;
    std::vector<int32_t> hdig;
    int nrowjobs = 0, ncoljobs = 0;
    size_t lds_rows = 0, lds_cols = 0;
    uint32_t H = 0;
};

static std::once_flag g_attr_once;

static void set_lds_limits() {
    std::call_once(g_attr_once, [] {
        hipFuncSetAttribute((const void *)k_rows, hipFuncAttributeMaxDynamicSharedMemorySize, kMaxLds);
        hipFuncSetAttribute((const void *)k_cols_inv, hipFuncAttributeMaxDynamicSharedMemorySize, kMaxLds);
        hipFuncSetAttribute((const void *)k_cols_fwd, hipFuncAttributeMaxDynamicSharedMemorySize, kMaxLds);
    });
}

// Per-test tables: weights factor over digit j = 2rC + b (b = 2c + h < 2C) as
// w_j = 2^{ceil(fa+fb)} 2^{-fa} 2^{-fb}, fa = frac(2rC T'/L), fb = frac(b T'/L), with the integer
// ceil(fa + fb) in {0,1,2} decided exactly from the 128-bit numerators.
int build_slot_tables(kite_ctx *ctx, const ShapeDev &sh, std::vector<double> &ta, std::vector<double> &tb,
                      std::vector<uint8_t> &dcode) {
    const uint32_t R = sh.R, C = sh.C, G = sh.G, NG = sh.NG;
    const u128 D = (u128)ctx->L << 64;
    std::vector<u128> na(R), nb(2 * C);
    ta.assign(2 * R, 0);
    tb.assign(4 * C, 0);
    for (uint32_t r = 0; r < R; r++) {
        na[r] = phase_num(ctx, 2ull * r * C);
        ta[r] = exp2_neg_frac(ctx, na[r]);
        long double f = (long double)na[r] / ((long double)ctx->L * 18446744073709551616.0L);
        ta[R + r] = (double)exp2l(f);
    }
    for (uint32_t b = 0; b < 2 * C; b++) {
        nb[b] = phase_num(ctx, b);
        tb[b] = exp2_neg_frac(ctx, nb[b]);
        long double f = (long double)nb[b] / ((long double)ctx->L * 18446744073709551616.0L);
        tb[2 * C + b] = (double)exp2l(f);
    }
    dcode.assign((size_t)R * C, 0);
    for (uint32_t g = 0; g < NG; g++)
        for (uint32_t r = 0; r < R; r++)
            for (uint32_t l = 0; l < G; l++) {
                uint8_t code = 0;
                for (int h = 0; h < 2; h++) {
                    const uint32_t b = 2 * (g * G + l) + h;
                    const uint64_t j = 2ull * r * C + b;
                    const u128 sum = na[r] + nb[b];
                    const int ce = sum == 0 ? 0 : (sum <= D ? 1 : 2);
                    const int big = (int)ctx->bits[j] - (int)ctx->b0;
                    if (big < 0 || big > 1) return -1;
                    code |= (uint8_t)((ce | (big << 2)) << (3 * h));
                }
                dcode[((size_t)g * R + r) * G + l] = code;
            }
    return 0;
}

int fill_slot_fold(kite_ctx *ctx, const ShapeDev &sh, SlotDev *s) {
    const uint32_t L = ctx->L, chunk0 = L - 2 * sh.G;
    s->k = ctx->k;
    s->b0 = (int)ctx->b0;
    s->top_n = (int)(L - ctx->top_t);
    if (s->top_n > 2 * sh.G || s->top_n > kMaxTop) return -1;
    s->top_i0 = (int)(ctx->top_t - chunk0);
    if (ctx->top_t == L) s->top_i0 = 2 * sh.G;
    s->top_m = (int)ctx->top_m;
    s->top_shift = (int)(ctx->sL - ctx->top_st);
    for (int i = 0; i < s->top_n; i++) s->top_bits[i] = ctx->bits[ctx->top_t + i];
    s->kscale = (double)ctx->k / (2.0 * L);  // k / (4H)
    s->rscale = 1.0 / sh.R;
    return 0;
}

int engine_init(kite_ctx *ctx) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        g_err = "no HIP device";
        return -1;
    }
    set_lds_limits();
    ShapeChoice ch;
    if (!choose_shape(ctx->L, ctx->L - ctx->top_t, ctx->plan_order, &ch)) {
        g_err = "no shape for this length";
        return -1;
    }
    Engine *E = new Engine();
    ctx->eng = E;
    E->shape = get_shape(ch);
    if (!E->shape) return -1;
    const ShapeDev &sh = E->shape->host;
    E->H = ctx->L / 2;
    std::vector<double> ta, tb;
    std::vector<uint8_t> dcode;
    if (build_slot_tables(ctx, sh, ta, tb, dcode)) {
        g_err = "digit widths outside {b0, b0+1}";
        return -1;
    }
    HIP_OK(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking));
    HIP_OK(hipMalloc(&E->ta, ta.size() * 8));
    HIP_OK(hipMemcpy(E->ta, ta.data(), ta.size() * 8, hipMemcpyHostToDevice));
    HIP_OK(hipMalloc(&E->tb, tb.size() * 8));
    HIP_OK(hipMemcpy(E->tb, tb.data(), tb.size() * 8, hipMemcpyHostToDevice));
    HIP_OK(hipMalloc(&E->dcode, dcode.size()));
    HIP_OK(hipMemcpy(E->dcode, dcode.data(), dcode.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMalloc(&E->dig, (size_t)ctx->L * 4));
    HIP_OK(hipMalloc(&E->cout, (size_t)sh.NG * sh.R * 8));
    HIP_OK(hipMalloc(&E->yfreq, (size_t)E->H * sizeof(double2)));
    HIP_OK(hipMalloc(&E->maxerr, 16));
    HIP_OK(hipMemset(E->maxerr, 0, 16));
    SlotDev &s = E->slot;
    s.shp = E->shape->dev;
    s.ta = E->ta;
    s.tb = E->tb;
    s.dcode = E->dcode;
    s.dig = E->dig;
    s.cout = E->cout;
    s.yfreq = E->yfreq;
    s.maxerr = E->maxerr;
    s.count = 1;
    if (fill_slot_fold(ctx, sh, &s)) {
        g_err = "top region does not fit the last chunk";
        return -1;
    }
    E->nrowjobs = sh.R % 2 == 0 ? sh.R / 2 : (sh.R + 1) / 2;
    E->ncoljobs = sh.NG;
    E->lds_rows = (size_t)64 * sh.C;
    E->lds_cols = (size_t)32 * sh.G * sh.R;
    E->hdig.resize(ctx->L);
    return 0;
}

void engine_free(kite_ctx *ctx) {
    Engine *E = ctx->eng;
    if (!E) return;
    if (E->stream) hipStreamSynchronize(E->stream);
    void *ps[] = {E->ta, E->tb, E->dcode, E->dig, E->cout, E->yfreq, E->maxerr};
    for (void *p : ps)
        if (p) hipFree(p);
    if (E->stream) hipStreamDestroy(E->stream);
    delete E;
    ctx->eng = nullptr;
}

void *engine_num_alloc(kite_ctx *ctx) {
    void *p = nullptr;
    if (hipMalloc(&p, (size_t)ctx->eng->H * sizeof(double2)) != hipSuccess) return nullptr;
    hipMemsetAsync(p, 0, (size_t)ctx->eng->H * sizeof(double2), ctx->eng->stream);
    return p;
}

void engine_num_free(kite_ctx *ctx, void *p) {
    if (!p) return;
    hipStreamSynchronize(ctx->eng->stream);
    hipFree(p);
}

void engine_copy(kite_ctx *ctx, void *dst, const void *src) {
    hipMemcpyAsync(dst, src, (size_t)ctx->eng->H * sizeof(double2), hipMemcpyDeviceToDevice, ctx->eng->stream);
}

static void launch_rows(Engine *E, double2 *x, const double2 *ysrc, int mode) {
    SlotDev s = E->slot;
    s.x = x;
    s.y_src = ysrc;
    hipLaunchKernelGGL(k_rows, dim3(E->nrowjobs), dim3(kNT), E->lds_rows, E->stream, (const SlotDev *)nullptr, s,
                       (const int2 *)nullptr, mode, 0);
}

static void launch_cols(Engine *E, double2 *x, int mode) {
    SlotDev s = E->slot;
    s.x = x;
    hipLaunchKernelGGL(k_cols_inv, dim3(E->ncoljobs), dim3(kNT), E->lds_cols, E->stream, (const SlotDev *)nullptr,
                       s, (const int2 *)nullptr, mode, 0);
    hipLaunchKernelGGL(k_cols_fwd, dim3(E->ncoljobs), dim3(kNT), E->lds_cols, E->stream, (const SlotDev *)nullptr,
                       s, (const int2 *)nullptr, mode, 0);
}

void engine_square(kite_ctx *ctx, void *x, uint64_t count) {
    Engine *E = ctx->eng;
    for (uint64_t i = 0; i < count; i++) {
        launch_rows(E, (double2 *)x, nullptr, kModeSquare);
        launch_cols(E, (double2 *)x, kModeSquare);
    }
}

void engine_mul(kite_ctx *ctx, void *x, const void *y) {
    Engine *E = ctx->eng;
    launch_rows(E, (double2 *)x, (const double2 *)y, kModeRowFft);
    launch_rows(E, (double2 *)x, nullptr, kModeMul);
    launch_cols(E, (double2 *)x, kModeMul);
}

void engine_addsub(kite_ctx *ctx, void *x, const void *y, int sign) {
    Engine *E = ctx->eng;
    const int n = (int)E->H;
    hipLaunchKernelGGL(k_addsub, dim3((n + 255) / 256), dim3(256), 0, E->stream, (double2 *)x, (const double2 *)y, n,
                       (double)sign);
}

// digit j = 2 (r C + g G) + i  <->  blocked index (g R + r) 2G + i
void engine_from_digits(kite_ctx *ctx, void *x, const int64_t *dig) {
    Engine *E = ctx->eng;
    const ShapeDev &sh = E->shape->host;
    const uint32_t R = sh.R, C = sh.C, G = sh.G, NG = sh.NG;
    for (uint32_t g = 0; g < NG; g++)
        for (uint32_t r = 0; r < R; r++)
            for (uint32_t i = 0; i < 2 * G; i++)
                E->hdig[((size_t)g * R + r) * 2 * G + i] = (int32_t)dig[2 * ((size_t)r * C + g * G) + i];
    hipMemcpyAsync(E->dig, E->hdig.data(), (size_t)ctx->L * 4, hipMemcpyHostToDevice, E->stream);
    SlotDev s = E->slot;
    s.x = (double2 *)x;
    hipLaunchKernelGGL(k_cols_fwd, dim3(E->ncoljobs), dim3(kNT), E->lds_cols, E->stream, (const SlotDev *)nullptr,
                       s, (const int2 *)nullptr, (int)kModeDigits, 0);
    hipStreamSynchronize(E->stream);  // hdig is reused
}

void engine_to_digits(kite_ctx *ctx, const void *x, int64_t *dig) {
    Engine *E = ctx->eng;
    const ShapeDev &sh = E->shape->host;
    SlotDev s = E->slot;
    s.x = (double2 *)const_cast<void *>(x);
    // the conversion's own rounding error is not an operation's round-off: count it aside
    s.maxerr = E->maxerr + 1;
    hipLaunchKernelGGL(k_cols_inv, dim3(E->ncoljobs), dim3(kNT), E->lds_cols, E->stream, (const SlotDev *)nullptr,
                       s, (const int2 *)nullptr, (int)kModeDigits, 0);
    hipMemcpyAsync(E->hdig.data(), E->dig, (size_t)ctx->L * 4, hipMemcpyDeviceToHost, E->stream);
    hipStreamSynchronize(E->stream);
    const uint32_t R = sh.R, C = sh.C, G = sh.G, NG = sh.NG;
    for (uint32_t g = 0; g < NG; g++)
        for (uint32_t r = 0; r < R; r++)
            for (uint32_t i = 0; i < 2 * G; i++)
                dig[2 * ((size_t)r * C + g * G) + i] = E->hdig[((size_t)g * R + r) * 2 * G + i];
}

double engine_maxerr(kite_ctx *ctx) {
    unsigned long long v = 0;
    hipMemcpyAsync(&v, ctx->eng->maxerr, 8, hipMemcpyDeviceToHost, ctx->eng->stream);
    hipStreamSynchronize(ctx->eng->stream);
    double d;
    memcpy(&d, &v, 8);
    return d;
}

void engine_reset_maxerr(kite_ctx *ctx) {
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, ctx->eng->stream, ctx->eng->maxerr, 0ull, 1);
}

void engine_sync(kite_ctx *ctx) { hipStreamSynchronize(ctx->eng->stream); }

}  // namespace kite

// ================================================================================================
// Batch: many contexts advanced by shared launches (include/kite_batch.h)
// ================================================================================================

namespace kite {

struct BatchSlot {
    kite_ctx *ctx = nullptr;
    std::vector<kite_num *> nums;
};

}  // namespace kite

struct kite_batch {
    std::vector<kite::BatchSlot> slots;
    hipStream_t stream = nullptr;
    kite::SlotDev *d_slots = nullptr, *h_slots = nullptr;  // h_slots is pinned
    size_t slot_cap = 0;
    int2 *d_rows = nullptr, *d_cols = nullptr;
    int nrows = 0, ncols = 0;
    size_t jobs_cap = 0;
    size_t lds_rows = 0, lds_cols = 0;
    unsigned long long *d_maxerr = nullptr;
    bool dirty = true;
    bool timing = false;
    std::vector<hipEvent_t> ev;
    double ms[3] = {0, 0, 0};
    uint64_t launches[3] = {0, 0, 0};
};

using namespace kite;

static int batch_refresh(kite_batch *b) {
    if (!b->dirty) return 0;
    const size_t cap = b->slots.size();
    if (cap > b->slot_cap) {
        const size_t nc = std::max<size_t>(cap, 2 * b->slot_cap);
        if (b->d_slots) hipFree(b->d_slots);
        if (b->h_slots) hipHostFree(b->h_slots);
        unsigned long long *old = b->d_maxerr;
        HIP_OK(hipMalloc(&b->d_slots, nc * sizeof(SlotDev)));
        HIP_OK(hipHostMalloc((void **)&b->h_slots, nc * sizeof(SlotDev), hipHostMallocDefault));
        HIP_OK(hipMalloc(&b->d_maxerr, nc * 8));
        HIP_OK(hipMemsetAsync(b->d_maxerr, 0, nc * 8, b->stream));
        if (old) {
            HIP_OK(hipMemcpyAsync(b->d_maxerr, old, b->slot_cap * 8, hipMemcpyDeviceToDevice, b->stream));
            HIP_OK(hipStreamSynchronize(b->stream));
            hipFree(old);
        }
        b->slot_cap = nc;
    }
    std::vector<int2> rows, cols;
    size_t lr = 0, lc = 0;
    for (size_t s = 0; s < cap; s++) {
        kite_ctx *ctx = b->slots[s].ctx;
        if (!ctx) continue;
        Engine *E = ctx->eng;
        for (int u = 0; u < E->nrowjobs; u++) rows.push_back(make_int2((int)s, u));
        for (int u = 0; u < E->ncoljobs; u++) cols.push_back(make_int2((int)s, u));
        lr = std::max(lr, E->lds_rows);
        lc = std::max(lc, E->lds_cols);
    }
    const size_t need = std::max(rows.size(), cols.size());
    if (need > b->jobs_cap) {
        if (b->d_rows) hipFree(b->d_rows);
        if (b->d_cols) hipFree(b->d_cols);
        const size_t nc = std::max(need, 2 * b->jobs_cap);
        HIP_OK(hipMalloc(&b->d_rows, nc * sizeof(int2)));
        HIP_OK(hipMalloc(&b->d_cols, nc * sizeof(int2)));
        b->jobs_cap = nc;
    }
    if (!rows.empty()) HIP_OK(hipMemcpy(b->d_rows, rows.data(), rows.size() * sizeof(int2), hipMemcpyHostToDevice));
    if (!cols.empty()) HIP_OK(hipMemcpy(b->d_cols, cols.data(), cols.size() * sizeof(int2), hipMemcpyHostToDevice));
    b->nrows = (int)rows.size();
    b->ncols = (int)cols.size();
    b->lds_rows = lr;
    b->lds_cols = lc;
    b->dirty = false;
    return 0;
}

// Fills the slot table: residue `x` (and operand `y`) per slot, step count per slot (0 = idle).
static int batch_upload(kite_batch *b, const int32_t *xi, const int32_t *yi, const uint32_t *count) {
    if (batch_refresh(b)) return -1;
    HIP_OK(hipStreamSynchronize(b->stream));  // the previous table may still be read
    for (size_t s = 0; s < b->slots.size(); s++) {
        SlotDev &d = b->h_slots[s];
        kite_ctx *ctx = b->slots[s].ctx;
        if (!ctx) {
            memset(&d, 0, sizeof d);
            continue;
        }
        d = ctx->eng->slot;
        d.maxerr = b->d_maxerr + s;
        const bool on = xi[s] >= 0 && (size_t)xi[s] < b->slots[s].nums.size() && count[s] > 0;
        d.count = on ? count[s] : 0;
        d.x = on ? (double2 *)b->slots[s].nums[xi[s]]->dev : nullptr;
        d.y_src = (on && yi && yi[s] >= 0) ? (const double2 *)b->slots[s].nums[yi[s]]->dev : nullptr;
    }
    HIP_OK(hipMemcpyAsync(b->d_slots, b->h_slots, b->slots.size() * sizeof(SlotDev), hipMemcpyHostToDevice,
                          b->stream));
    return 0;
}

static void batch_launch(kite_batch *b, int pass, int mode, int step) {
    const bool rows = pass == 0;
    const int grid = rows ? b->nrows : b->ncols;
    if (!grid) return;
    SlotDev none{};
    const size_t lds = rows ? b->lds_rows : b->lds_cols;
    if (pass == 0)
        hipLaunchKernelGGL(k_rows, dim3(grid), dim3(kNT), lds, b->stream, b->d_slots, none, b->d_rows, mode, step);
    else if (pass == 1)
        hipLaunchKernelGGL(k_cols_inv, dim3(grid), dim3(kNT), lds, b->stream, b->d_slots, none, b->d_cols, mode, step);
    else
        hipLaunchKernelGGL(k_cols_fwd, dim3(grid), dim3(kNT), lds, b->stream, b->d_slots, none, b->d_cols, mode, step);
}

extern "C" kite_batch *kite_batch_new(void) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return nullptr;
    set_lds_limits();
    kite_batch *b = new kite_batch();
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return nullptr;
    }
    return b;
}

extern "C" void kite_batch_free(kite_batch *b) {
    if (!b) return;
    hipStreamSynchronize(b->stream);
    for (size_t s = 0; s < b->slots.size(); s++) kite_batch_remove(b, (int)s);
    for (hipEvent_t e : b->ev) hipEventDestroy(e);
    if (b->d_slots) hipFree(b->d_slots);
    if (b->h_slots) hipHostFree(b->h_slots);
    if (b->d_rows) hipFree(b->d_rows);
    if (b->d_cols) hipFree(b->d_cols);
    if (b->d_maxerr) hipFree(b->d_maxerr);
    hipStreamDestroy(b->stream);
    delete b;
}

extern "C" int kite_batch_add(kite_batch *b, uint64_t k, uint64_t n, uint32_t fft_len, uint32_t nnums) {
    kite_opts o{};
    o.backend = KITE_BACKEND_HIP;
    o.fft_len = fft_len;
    kite_ctx *ctx = kite_ctx_new(k, n, -1, &o);
    if (!ctx) return -1;
    BatchSlot bs;
    bs.ctx = ctx;
    for (uint32_t i = 0; i < nnums; i++) {
        kite_num *x = kite_num_new(ctx);
        if (!x) {
            for (kite_num *y : bs.nums) kite_num_free(ctx, y);
            kite_ctx_free(ctx);
            return -1;
        }
        bs.nums.push_back(x);
    }
    hipStreamSynchronize(ctx->eng->stream);
    size_t s = 0;
    while (s < b->slots.size() && b->slots[s].ctx) s++;
    if (s == b->slots.size()) b->slots.emplace_back();
    b->slots[s] = std::move(bs);
    b->dirty = true;
    if (batch_refresh(b)) return -1;
    // clear this slot's round-off accumulator
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, b->stream, b->d_maxerr + s, 0ull, 1);
    hipStreamSynchronize(b->stream);
    return (int)s;
}

extern "C" void kite_batch_remove(kite_batch *b, int slot) {
    if (slot < 0 || (size_t)slot >= b->slots.size() || !b->slots[slot].ctx) return;
    hipStreamSynchronize(b->stream);
    BatchSlot &bs = b->slots[slot];
    for (kite_num *x : bs.nums) kite_num_free(bs.ctx, x);
    kite_ctx_free(bs.ctx);
    bs = BatchSlot();
    b->dirty = true;
}

extern "C" int kite_batch_capacity(const kite_batch *b) { return (int)b->slots.size(); }

extern "C" kite_ctx *kite_batch_ctx(kite_batch *b, int slot) {
    return slot >= 0 && (size_t)slot < b->slots.size() ? b->slots[slot].ctx : nullptr;
}

extern "C" kite_num *kite_batch_num(kite_batch *b, int slot, uint32_t i) {
    if (slot < 0 || (size_t)slot >= b->slots.size() || i >= b->slots[slot].nums.size()) return nullptr;
    return b->slots[slot].nums[i];
}

extern "C" int kite_batch_sync(kite_batch *b) { return hipStreamSynchronize(b->stream) == hipSuccess ? 0 : -1; }

static void timed_launch(kite_batch *b, int pass, int mode, int step, size_t evi) {
    if (b->timing) hipEventRecord(b->ev[evi], b->stream);
    batch_launch(b, pass, mode, step);
    if (b->timing) hipEventRecord(b->ev[evi + 1], b->stream);
}

extern "C" int kite_batch_square(kite_batch *b, const int32_t *which, const uint32_t *count) {
    uint32_t maxc = 0;
    for (size_t s = 0; s < b->slots.size(); s++)
        if (b->slots[s].ctx && which[s] >= 0) maxc = std::max(maxc, count[s]);
    if (!maxc) return 0;
    if (batch_upload(b, which, nullptr, count)) return -1;
    if (b->timing && b->ev.size() < 6 * (size_t)maxc) {
        size_t old = b->ev.size();
        b->ev.resize(6 * (size_t)maxc);
        for (size_t i = old; i < b->ev.size(); i++) hipEventCreate(&b->ev[i]);
    }
    for (uint32_t st = 0; st < maxc; st++) {
        timed_launch(b, 0, kModeSquare, (int)st, 6 * st);
        timed_launch(b, 1, kModeSquare, (int)st, 6 * st + 2);
        timed_launch(b, 2, kModeSquare, (int)st, 6 * st + 4);
    }
    if (b->timing) {
        hipStreamSynchronize(b->stream);
        for (int p = 0; p < 3; p++) b->ms[p] = 0;
        for (uint32_t st = 0; st < maxc; st++)
            for (int p = 0; p < 3; p++) {
                float ms = 0;
                hipEventElapsedTime(&ms, b->ev[6 * st + 2 * p], b->ev[6 * st + 2 * p + 1]);
                b->ms[p] += ms;
            }
        for (int p = 0; p < 3; p++) b->launches[p] = maxc;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int kite_batch_mul(kite_batch *b, const int32_t *dst, const int32_t *src) {
    std::vector<uint32_t> one(b->slots.size(), 1);
    bool any = false;
    for (size_t s = 0; s < b->slots.size(); s++) any |= b->slots[s].ctx && dst[s] >= 0;
    if (!any) return 0;
    if (batch_upload(b, dst, src, one.data())) return -1;
    batch_launch(b, 0, kModeRowFft, 0);
    batch_launch(b, 0, kModeMul, 0);
    batch_launch(b, 1, kModeMul, 0);
    batch_launch(b, 2, kModeMul, 0);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int kite_batch_copy(kite_batch *b, const int32_t *dst, const int32_t *src) {
    for (size_t s = 0; s < b->slots.size(); s++) {
        BatchSlot &bs = b->slots[s];
        if (!bs.ctx || dst[s] < 0 || dst[s] == src[s]) continue;
        hipMemcpyAsync(bs.nums[dst[s]]->dev, bs.nums[src[s]]->dev, (size_t)bs.ctx->eng->H * sizeof(double2),
                       hipMemcpyDeviceToDevice, b->stream);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" int kite_batch_maxerr(kite_batch *b, double *out, int reset) {
    if (batch_refresh(b)) return -1;
    const size_t cap = b->slots.size();
    if (!cap) return 0;
    std::vector<unsigned long long> v(cap);
    HIP_OK(hipMemcpyAsync(v.data(), b->d_maxerr, cap * 8, hipMemcpyDeviceToHost, b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    for (size_t s = 0; s < cap; s++) memcpy(&out[s], &v[s], 8);
    if (reset) {
        hipLaunchKernelGGL(k_set_u64, dim3((unsigned)((cap + 63) / 64)), dim3(64), 0, b->stream, b->d_maxerr, 0ull,
                           (int)cap);
        HIP_OK(hipStreamSynchronize(b->stream));
    }
    return 0;
}

extern "C" int kite_batch_set_timing(kite_batch *b, int on) {
    b->timing = on != 0;
    return 0;
}

extern "C" int kite_batch_pass_times(kite_batch *b, double *out3, uint64_t *launches3) {
    for (int p = 0; p < 3; p++) {
        out3[p] = b->ms[p];
        if (launches3) launches3[p] = b->launches[p];
    }
    return 0;
}
