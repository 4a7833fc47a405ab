// CPU oracle for squaring mod N = k*2^n - 1. TEST INFRASTRUCTURE ONLY (see kite_oracle.h).
//
// Each function names the reference file:line whose algorithm it restates. The code is a plain
// scalar restatement written for checking, not for speed (the two-pass backend doubles as the
// CPU baseline the benchmark times beside the GPU, like the reference's `--backend c2p`).
#include "kite_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdbool.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;
typedef __int128 i128;
typedef double complex cplx;

#define GRP 8  // columns per pass-1 group in the two-pass layout (src/fast2p.c:26)

// ---------------------------------------------------------------------------------------------
// Accurate roots of unity: cos/sin(2*pi*j/n) with the fraction j/n reduced to [0, 1/8] in exact
// integer arithmetic, then evaluated in long double. The reference uses Apple's __cospi/__sinpi
// (src/ref_fft.c:61, src/fast2p.c:108-109), which this replaces.
void or_sincos2pi(uint64_t j, uint64_t n, double *c, double *s) {
    j %= n;
    // fraction f = j/n in [0,1); use octant symmetries on the integer pair (j, n)
    int neg_s = 0, neg_c = 0, swap = 0;
    uint64_t a = j;                   // angle 2*pi*a/n
    if (2 * a > n) { a = n - a; neg_s = 1; }            // sin(2pi - x) = -sin x
    if (4 * a > n) { a = n - 2 * a; neg_c = 1; }        // now use x -> pi - x on 2a/(2n)
    // after the second step the angle is 2*pi*a/(2n) when neg_c is set
    uint64_t den = neg_c ? 2 * n : n;
    if (8 * a > den) { a = den - 4 * a; den = 4 * den; swap = 1; }  // x -> pi/2 - x
    long double x = 2.0L * 3.14159265358979323846264338327950288L * (long double)a / (long double)den;
    long double cc = cosl(x), ss = sinl(x);
    if (swap) { long double t = cc; cc = ss; ss = t; }
    if (neg_c) cc = -cc;
    if (neg_s) ss = -ss;
    *c = (double)cc;
    *s = (double)ss;
}

// ---------------------------------------------------------------------------------------------
// FFT length rules: src/ctx.c:17-24 (7-smooth), :26-33 (safe bits/word), :94-148 (choice).

int or_fft_len_ok(uint32_t L) {
    if (L < 16 || (L & 1)) return 0;
    static const uint32_t pr[4] = {2, 3, 5, 7};
    for (int i = 0; i < 4; i++)
        while (L % pr[i] == 0) L /= pr[i];
    return L == 1;
}

double or_safe_bpw(uint32_t L, uint64_t k) {
    return 21.44 - 0.305 * (log2((double)L) - 12) - 0.50 * log2((double)k) - 0.2;
}

static double radix_cost(uint32_t L) {  // src/ctx.c:96-103
    double c = log2((double)L);
    uint32_t m = L;
    while (m % 3 == 0) { m /= 3; c += 0.6; }
    while (m % 5 == 0) { m /= 5; c += 1.0; }
    while (m % 7 == 0) { m /= 7; c += 1.4; }
    return (double)L * c;
}

static int usable(uint32_t L) { return L % 64 == 0 && or_fft_len_ok(L); }

uint32_t or_choose_fft_len(uint64_t k, uint64_t n) {
    if (n < 2048) {
        for (uint32_t L = 16; L <= n; L *= 2)
            if ((double)n / L <= or_safe_bpw(L, k)) return L;
        return 0;
    }
    uint32_t lo = 0;
    for (uint32_t L = 64; L < (1u << 28); L += 64) {
        if (usable(L) && (double)n / L <= or_safe_bpw(L, k)) { lo = L; break; }
    }
    if (!lo) return 0;
    uint32_t best = lo;
    double bc = 1e300;
    for (uint32_t L = lo; L <= lo + lo / 10 * 3; L += 64) {
        if (!usable(L)) continue;
        double c = radix_cost(L);  // no tuning file in the oracle: the cost model decides
        if (c < bc) { bc = c; best = L; }
    }
    return best;
}

uint32_t or_next_fft_len(uint32_t L) {  // src/ctx.c:143-148
    for (uint32_t M = (L / 64 + 1) * 64; M < (1u << 28); M += 64)
        if (or_fft_len_ok(M)) return M;
    return 0;
}

// ---------------------------------------------------------------------------------------------
// Context

typedef struct {  // two-pass state (src/fast2p.c:28-52)
    uint32_t n, nst;
    uint8_t rad[32];
    double *twr, *twi;  // per-stage twiddles
    uint32_t *rev, *pos;
} tp_plan;

typedef struct {
    uint32_t H, R, C, NG;
    tp_plan pr, pc;
    double *t4r, *t4i;            // w_H^{c*k_r} at [row position][c]
    double *rsr, *rsi, *csr, *csi;
    uint32_t *partner;
    double *sre, *sim, *g0;       // group scratch
    int64_t *carry;
    double *ar, *ai, *br, *bi;    // row scratch
    double *yr, *yi;              // transformed mul operand rows
    double roots_r[9][8], roots_i[9][8];  // cos/sin(2*pi*q/p) for each radix p <= 8
} tp_state;

typedef struct {
    cplx *tw, *a, *b;
} full_state;

struct or_ctx {
    uint64_t k, n;
    uint32_t L;
    int backend;
    uint8_t *bits;
    uint64_t lam, sL, top_st;
    uint32_t top_t, top_m;
    double *w, *oscale;
    size_t nlimbs;
    uint64_t *Nl;
    double maxerr;
    int64_t *dig;
    full_state fs;
    tp_state ts;
};

// log2(k) - floor(log2 k) as a 64-bit binary fraction, one bit per squaring of a fixed-point
// mantissa with 62 fractional bits (src/ctx.c:150-167).
static uint64_t frac_log2(uint64_t k) {
    int e = 63 - __builtin_clzll(k);
    u128 m = (u128)k << (62 - e);  // in [2^62, 2^63): the mantissa in [1, 2)
    uint64_t f = 0;
    for (int i = 0; i < 64; i++) {
        m = (m * m) >> 62;
        int bit = (m >> 63) != 0;  // mantissa^2 >= 2
        f = (f << 1) | (uint64_t)bit;
        if (bit) m >>= 1;
    }
    return f;
}

static uint64_t ilog2(uint64_t k) { return (uint64_t)(63 - __builtin_clzll(k)); }

// s_j = ceil(j*T'/L) with T' = n + floor(log2 k) + lam/2^64, exactly (src/ctx.c:169-183).
uint64_t or_s_at(const or_ctx *c, uint64_t j) {
    u128 whole = (u128)j * (c->n + ilog2(c->k));
    u128 q = whole / c->L, r = whole % c->L;
    u128 num = (r << 64) + (u128)j * c->lam;
    u128 den = (u128)c->L << 64;
    return (uint64_t)(q + (num + den - 1) / den);
}

// w_j = 2^(s_j - j*T'/L) from the exact fractional exponent (src/ctx.c:185-196).
double or_weight(const or_ctx *c, uint64_t j, uint64_t sj) {
    u128 whole = (u128)j * (c->n + ilog2(c->k));
    u128 q = whole / c->L, r = whole % c->L;
    u128 num = (((u128)(sj - q) * c->L) << 64) - ((r << 64) + (u128)j * c->lam);
    double f = (double)(num / c->L) / 18446744073709551616.0;
    return exp2(f);
}

static int layout(or_ctx *c) {  // src/ctx.c:198-221
    c->lam = frac_log2(c->k);
    uint64_t s = 0;
    c->top_t = c->L;
    for (uint32_t j = 0; j < c->L; j++) {
        uint64_t s1 = or_s_at(c, j + 1);
        if (s1 < s || s1 - s > 40) return -1;
        c->bits[j] = (uint8_t)(s1 - s);
        if (s1 > c->n && c->top_t == c->L) {
            c->top_t = j;
            c->top_st = s;
        }
        c->w[j] = or_weight(c, j, s);
        c->oscale[j] = (double)c->k / c->w[j];
        s = s1;
    }
    c->sL = s;
    if (c->top_t == c->L) c->top_st = s;
    c->top_m = (uint32_t)(c->n - c->top_st);
    if (c->sL - c->top_st > 100 || c->top_t == 0) return -1;
    return 0;
}

static int full_init(or_ctx *c);
static int tp_init(or_ctx *c);
static void tp_free(or_ctx *c);

or_ctx *or_ctx_new(uint64_t k, uint64_t n, uint32_t L, int backend) {
    if (k == 0 || !(k & 1) || n < 8 || k >= (1ull << 30)) return NULL;
    or_ctx *c = calloc(1, sizeof *c);
    c->k = k;
    c->n = n;
    c->L = L ? L : or_choose_fft_len(k, n);
    c->backend = backend;
    if (!or_fft_len_ok(c->L) || c->L > n) { free(c); return NULL; }
    c->bits = malloc(c->L);
    c->w = malloc(c->L * sizeof(double));
    c->oscale = malloc(c->L * sizeof(double));
    c->dig = malloc(c->L * sizeof(int64_t));
    c->nlimbs = (n + 127) / 64;
    c->Nl = calloc(c->nlimbs, 8);
    if (layout(c)) goto fail;
    // N = k*2^n - 1 (src/ctx.c:246-254)
    c->Nl[n / 64] = k << (n % 64);
    if (n % 64 && n / 64 + 1 < c->nlimbs) c->Nl[n / 64 + 1] = k >> (64 - n % 64);
    for (size_t i = 0; i < c->nlimbs; i++)
        if (c->Nl[i]--) break;
    if (backend == OR_BACKEND_REF ? full_init(c) : tp_init(c)) goto fail;
    return c;
fail:
    or_ctx_free(c);
    return NULL;
}

void or_ctx_free(or_ctx *c) {
    if (!c) return;
    free(c->fs.tw);
    free(c->fs.a);
    free(c->fs.b);
    tp_free(c);
    free(c->bits);
    free(c->w);
    free(c->oscale);
    free(c->dig);
    free(c->Nl);
    free(c);
}

uint32_t or_fft_len(const or_ctx *c) { return c->L; }
size_t or_nlimbs(const or_ctx *c) { return c->nlimbs; }
size_t or_num_doubles(const or_ctx *c) { return c->L; }
double or_maxerr(const or_ctx *c) { return c->maxerr; }
void or_reset_maxerr(or_ctx *c) { c->maxerr = 0; }

void or_layout(const or_ctx *c, uint8_t *bits, double *w, uint64_t *info) {
    if (bits) memcpy(bits, c->bits, c->L);
    if (w) memcpy(w, c->w, c->L * sizeof(double));
    if (info) {
        info[0] = c->top_t;
        info[1] = c->top_st;
        info[2] = c->top_m;
        info[3] = c->sL;
        info[4] = c->lam;
    }
}

// ---------------------------------------------------------------------------------------------
// Carry and the top-region fold (src/ctx.c:345-407).

static inline int64_t bal_split(int64_t v, unsigned b, int64_t *digit) {
    int64_t q = (v + ((int64_t)1 << (b - 1))) >> b;  // digit in [-2^(b-1), 2^(b-1))
    *digit = v - (q << b);
    return q;
}

// Top region t..L-1 plus carry at 2^{s_L}: H = q*k + r -> q into digit 0, r*2^m rebalanced.
static int64_t fold(const or_ctx *c, int64_t *dig, int64_t carry) {
    i128 H = (i128)carry << (c->sL - c->top_st);
    unsigned sh = 0;
    for (uint32_t j = c->top_t; j < c->L; j++) {
        H += (i128)dig[j] << sh;
        sh += c->bits[j];
    }
    unsigned m = c->top_m;
    i128 lo = H & (((i128)1 << m) - 1), hi = H >> m;
    i128 K = (i128)c->k;
    i128 q = hi >= 0 ? hi / K : -((-hi + K - 1) / K);  // floor(hi / k)
    i128 top = lo + ((hi - q * K) << m);
    if (2 * top >= (K << m)) {
        top -= K << m;
        q += 1;
    }
    for (uint32_t j = c->top_t; j + 1 < c->L; j++) {
        unsigned b = c->bits[j];
        i128 qq = (top + ((i128)1 << (b - 1))) >> b;
        dig[j] = (int64_t)(top - (qq << b));
        top = qq;
    }
    if (c->top_t < c->L) dig[c->L - 1] = (int64_t)top;
    return (int64_t)q;
}

void or_carry_digits(const or_ctx *c, int64_t *dig) {
    int64_t cy = 0;
    for (uint32_t j = 0; j < c->L; j++) cy = bal_split(dig[j] + cy, c->bits[j], &dig[j]);
    for (int pass = 0; pass < 64; pass++) {
        cy = fold(c, dig, cy);
        for (uint32_t j = 0; j < c->top_t && cy; j++) cy = bal_split(dig[j] + cy, c->bits[j], &dig[j]);
        if (!cy) return;
        if (c->top_t < c->L) {
            dig[c->top_t] += cy;
            cy = 0;
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Limbs (src/ctx.c:409-551).

static void add_at(uint64_t *a, size_t na, uint64_t pos, uint64_t v) {
    size_t w = pos / 64;
    unsigned sh = pos % 64;
    if (w >= na) return;
    u128 t = (u128)v << sh;
    u128 s = (u128)a[w] + (uint64_t)t;
    a[w] = (uint64_t)s;
    uint64_t cy = (uint64_t)(s >> 64) + (uint64_t)(t >> 64);
    for (size_t i = w + 1; cy && i < na; i++) {
        s = (u128)a[i] + cy;
        a[i] = (uint64_t)s;
        cy = (uint64_t)(s >> 64);
    }
}

static int cmp_limbs(const uint64_t *a, const uint64_t *b, size_t n) {
    for (size_t i = n; i-- > 0;)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}

static void sub_limbs(uint64_t *a, const uint64_t *b, size_t n) {
    uint64_t br = 0;
    for (size_t i = 0; i < n; i++) {
        u128 d = (u128)a[i] - b[i] - br;
        a[i] = (uint64_t)d;
        br = (uint64_t)(d >> 64) & 1;
    }
}

static uint64_t bits_at(const uint64_t *a, size_t na, uint64_t pos, unsigned cnt) {
    size_t w = pos / 64;
    unsigned sh = pos % 64;
    if (w >= na) return 0;
    u128 v = a[w];
    if (w + 1 < na) v |= (u128)a[w + 1] << 64;
    v >>= sh;
    return cnt >= 64 ? (uint64_t)v : (uint64_t)v & ((1ull << cnt) - 1);
}

// a mod N using k*2^n == 1: a = lo + hi*2^n, hi = q*k + r -> lo + q + r*2^n.
static void reduce(const or_ctx *c, uint64_t *a, size_t na) {
    for (int pass = 0; pass < 4; pass++) {
        size_t w = c->n / 64;
        unsigned sh = c->n % 64;
        u128 h = 0;
        for (size_t i = na; i-- > w;) h = (h << 64) | a[i];
        h >>= sh;
        if (!h) break;
        a[w] &= sh ? (1ull << sh) - 1 : 0;
        for (size_t i = w + 1; i < na; i++) a[i] = 0;
        u128 q = h / c->k;
        add_at(a, na, c->n, (uint64_t)(h % c->k));
        add_at(a, na, 0, (uint64_t)q);
        add_at(a, na, 64, (uint64_t)(q >> 64));
    }
    while (cmp_limbs(a, c->Nl, c->nlimbs) >= 0) sub_limbs(a, c->Nl, c->nlimbs);
}

static void part_mod_N(const or_ctx *c, const int64_t *dig, int sign, uint64_t *a, size_t na) {
    memset(a, 0, na * 8);
    uint64_t s = 0;
    for (uint32_t j = 0; j < c->L; j++) {
        int64_t d = dig[j];
        if (sign > 0 && d > 0) add_at(a, na, s, (uint64_t)d);
        if (sign < 0 && d < 0) add_at(a, na, s, (uint64_t)-d);
        s += c->bits[j];
    }
    u128 cy = 0;  // the stored value is x/k, so x = k*V
    for (size_t i = 0; i < na; i++) {
        cy += (u128)a[i] * c->k;
        a[i] = (uint64_t)cy;
        cy >>= 64;
    }
    reduce(c, a, na);
}

void or_digits_to_limbs(or_ctx *c, const int64_t *dig, uint64_t *limbs) {
    size_t na = (c->sL + 64) / 64 + 3;
    uint64_t *p = calloc(na, 8), *m = calloc(na, 8);
    part_mod_N(c, dig, 1, p, na);
    part_mod_N(c, dig, -1, m, na);
    if (cmp_limbs(p, m, c->nlimbs) < 0) {
        uint64_t cy = 0;
        for (size_t i = 0; i < c->nlimbs; i++) {
            u128 t = (u128)p[i] + c->Nl[i] + cy;
            p[i] = (uint64_t)t;
            cy = (uint64_t)(t >> 64);
        }
    }
    sub_limbs(p, m, c->nlimbs);
    memcpy(limbs, p, c->nlimbs * 8);
    free(p);
    free(m);
}

void or_get_limbs(or_ctx *c, const double *x, uint64_t *limbs) {
    or_to_digits(c, x, c->dig);
    or_digits_to_limbs(c, c->dig, limbs);
}

void or_set_limbs(or_ctx *c, double *x, const uint64_t *limbs) {
    size_t nl = c->nlimbs + 1;
    uint64_t *a = calloc(nl, 8);
    memcpy(a, limbs, c->nlimbs * 8);
    // V = x*2^n mod N: x = q*k + r -> q + r*2^n
    u128 r = 0;
    for (size_t i = nl; i-- > 0;) {
        r = (r << 64) | a[i];
        a[i] = (uint64_t)(r / c->k);
        r %= c->k;
    }
    add_at(a, nl, c->n, (uint64_t)r);
    if (cmp_limbs(a, c->Nl, c->nlimbs) >= 0 && !a[nl - 1]) sub_limbs(a, c->Nl, c->nlimbs);
    uint64_t s = 0;
    for (uint32_t j = 0; j + 1 < c->L; j++) {
        c->dig[j] = (int64_t)bits_at(a, nl, s, c->bits[j]);
        s += c->bits[j];
    }
    c->dig[c->L - 1] = (int64_t)bits_at(a, nl, s, 64);
    free(a);
    or_carry_digits(c, c->dig);
    or_from_digits(c, x, c->dig);
}

void or_set_u64(or_ctx *c, double *x, uint64_t v) {
    uint64_t *t = calloc(c->nlimbs, 8);
    t[0] = v;
    or_set_limbs(c, x, t);
    free(t);
}

void or_add(or_ctx *c, double *x, const double *y) {
    for (uint32_t i = 0; i < c->L; i++) x[i] += y[i];
}
void or_sub(or_ctx *c, double *x, const double *y) {
    for (uint32_t i = 0; i < c->L; i++) x[i] -= y[i];
}

// ---------------------------------------------------------------------------------------------
// Full-length backend (src/ref_fft.c): complex FFT of length L on the real weighted digits,
// pointwise square, conjugate-FFT inverse, scale by k/(L*w), round, carry.

static uint32_t least_factor(uint32_t n) {
    for (uint32_t p = 2; p * p <= n; p++)
        if (n % p == 0) return p;
    return n;
}

// out[0..n) = DFT of in[0], in[s], in[2s], ... (decimation in time by the least prime factor).
static void rec_fft(const cplx *tw, uint32_t L, const cplx *in, size_t s, cplx *out, uint32_t n) {
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    uint32_t p = least_factor(n), m = n / p;
    for (uint32_t r = 0; r < p; r++) rec_fft(tw, L, in + r * s, s * p, out + r * m, m);
    cplx t[16], u[16];
    cplx *tp = p <= 16 ? t : malloc(2 * p * sizeof(cplx)), *up = p <= 16 ? u : tp + p;
    for (uint32_t q = 0; q < m; q++) {
        for (uint32_t r = 0; r < p; r++) tp[r] = out[r * m + q] * tw[(uint64_t)r * q * (L / n) % L];
        for (uint32_t v = 0; v < p; v++) {
            cplx acc = 0;
            for (uint32_t r = 0; r < p; r++) acc += tp[r] * tw[(uint64_t)(r * v % p) * (L / p)];
            up[v] = acc;
        }
        for (uint32_t v = 0; v < p; v++) out[v * m + q] = up[v];
    }
    if (p > 16) free(tp);
}

static int full_init(or_ctx *c) {
    uint32_t L = c->L;
    c->fs.tw = malloc(L * sizeof(cplx));
    c->fs.a = malloc(L * sizeof(cplx));
    c->fs.b = malloc(L * sizeof(cplx));
    for (uint32_t j = 0; j < L; j++) {
        double co, si;
        or_sincos2pi(j, L, &co, &si);
        c->fs.tw[j] = co - I * si;
    }
    return 0;
}

static void full_forward(or_ctx *c, const double *x, cplx *dst) {
    cplx *in = dst == c->fs.a ? c->fs.b : c->fs.a;
    for (uint32_t j = 0; j < c->L; j++) in[j] = x[j] * c->w[j];
    rec_fft(c->fs.tw, c->L, in, 1, dst, c->L);
}

static void full_inverse_carry(or_ctx *c, cplx *prod, double *x) {
    cplx *out = prod == c->fs.a ? c->fs.b : c->fs.a;
    for (uint32_t j = 0; j < c->L; j++) prod[j] = conj(prod[j]);
    rec_fft(c->fs.tw, c->L, prod, 1, out, c->L);
    double sc = 1.0 / c->L, me = c->maxerr;
    for (uint32_t j = 0; j < c->L; j++) {
        double v = creal(out[j]) * sc * c->oscale[j], r = rint(v), e = fabs(v - r);
        if (e > me) me = e;
        c->dig[j] = (int64_t)r;
    }
    c->maxerr = me;
    or_carry_digits(c, c->dig);
    for (uint32_t j = 0; j < c->L; j++) x[j] = (double)c->dig[j];
}

// ---------------------------------------------------------------------------------------------
// Two-pass backend (src/fast2p.c): H = L/2 = R x C, element e = r*C + c at row r, column c,
// complex pairs stored [re c, re c+1, im c, im c+1]; pass 1 = column FFTs over groups of 8
// columns with the four-step twiddle and the carry; pass 2 = row FFTs with the real-split square.

static int radices(uint32_t n, uint8_t *rad) {  // src/fast2p.c:75-83
    int m = 0;
    while (n % 8 == 0 && n > 8) { rad[m++] = 8; n /= 8; }
    static const uint8_t pr[6] = {8, 4, 2, 3, 5, 7};
    for (int i = 0; i < 6; i++)
        while (n % pr[i] == 0) { rad[m++] = pr[i]; n /= pr[i]; }
    return n == 1 ? m : -1;
}

static int plan_make(tp_plan *pl, uint32_t n) {  // src/fast2p.c:85-128
    pl->n = n;
    int m = radices(n, pl->rad);
    if (m < 0) return -1;
    pl->nst = (uint32_t)m;
    size_t tot = 0;
    for (uint32_t st = 0, sz = n; st < pl->nst; sz /= pl->rad[st], st++) tot += sz;
    pl->twr = malloc(tot * sizeof(double));
    pl->twi = malloc(tot * sizeof(double));
    pl->rev = malloc(n * sizeof(uint32_t));
    pl->pos = malloc(n * sizeof(uint32_t));
    size_t off = 0;
    for (uint32_t st = 0, sz = n; st < pl->nst; st++) {
        uint32_t p = pl->rad[st], s = sz / p;
        for (uint32_t t = 0; t < p; t++)
            for (uint32_t j = 0; j < s; j++) {
                double co, si;
                or_sincos2pi((uint64_t)t * j % sz, sz, &co, &si);
                pl->twr[off + t * s + j] = co;
                pl->twi[off + t * s + j] = -si;
            }
        off += sz;
        sz = s;
    }
    // position of frequency k after the in-place DIF: mixed-radix digit reversal
    for (uint32_t k = 0; k < n; k++) {
        uint32_t kk = k, p0 = 0, size = n;
        for (uint32_t st = 0; st < pl->nst; st++) {
            size /= pl->rad[st];
            p0 += (kk % pl->rad[st]) * size;
            kk /= pl->rad[st];
        }
        pl->pos[k] = p0;
        pl->rev[p0] = k;
    }
    return 0;
}

static void plan_drop(tp_plan *pl) {
    free(pl->twr);
    free(pl->twi);
    free(pl->rev);
    free(pl->pos);
}

// radix-p DFT of (xr, xi) by the O(p^2) sum (src/fast2p.c:56-71), roots from a per-radix table
static void small_dft(const tp_state *st, uint32_t p, double *xr, double *xi, int inv) {
    double yr[8], yi[8];
    for (uint32_t t = 0; t < p; t++) {
        double sr = 0, si = 0;
        for (uint32_t u = 0; u < p; u++) {
            uint32_t q = u * t % p;
            double cr = st->roots_r[p][q], ci = inv ? st->roots_i[p][q] : -st->roots_i[p][q];
            sr += xr[u] * cr - xi[u] * ci;
            si += xr[u] * ci + xi[u] * cr;
        }
        yr[t] = sr;
        yi[t] = si;
    }
    memcpy(xr, yr, p * sizeof(double));
    memcpy(xi, yi, p * sizeof(double));
}

// In-place DIF over `lanes` interleaved sequences (src/fast2p.c:137-166).
static void dif(const tp_state *st, const tp_plan *pl, double *re, double *im, uint32_t lanes) {
    uint32_t n = pl->n;
    size_t off = 0;
    double xr[8], xi[8];
    for (uint32_t s_ = 0, sz = n; s_ < pl->nst; s_++) {
        uint32_t p = pl->rad[s_], s = sz / p;
        for (uint32_t b = 0; b < n; b += sz)
            for (uint32_t j = 0; j < s; j++)
                for (uint32_t l = 0; l < lanes; l++) {
                    for (uint32_t t = 0; t < p; t++) {
                        size_t i = (size_t)(b + j + t * s) * lanes + l;
                        xr[t] = re[i];
                        xi[t] = im[i];
                    }
                    small_dft(st, p, xr, xi, 0);
                    for (uint32_t t = 0; t < p; t++) {
                        double wr = pl->twr[off + t * s + j], wi = pl->twi[off + t * s + j];
                        size_t i = (size_t)(b + j + t * s) * lanes + l;
                        re[i] = xr[t] * wr - xi[t] * wi;
                        im[i] = xr[t] * wi + xi[t] * wr;
                    }
                }
        off += sz;
        sz = s;
    }
}

// In-place inverse DIT, stages in reverse (src/fast2p.c:168-206).
static void dit_inv(const tp_state *st, const tp_plan *pl, double *re, double *im, uint32_t lanes) {
    uint32_t n = pl->n;
    size_t offs[32];
    uint32_t sizes[32];
    size_t off = 0;
    for (uint32_t s_ = 0, sz = n; s_ < pl->nst; sz /= pl->rad[s_], s_++) {
        offs[s_] = off;
        sizes[s_] = sz;
        off += sz;
    }
    double xr[8], xi[8];
    for (uint32_t s_ = pl->nst; s_-- > 0;) {
        uint32_t p = pl->rad[s_], sz = sizes[s_], s = sz / p;
        for (uint32_t b = 0; b < n; b += sz)
            for (uint32_t j = 0; j < s; j++)
                for (uint32_t l = 0; l < lanes; l++) {
                    for (uint32_t t = 0; t < p; t++) {
                        double wr = pl->twr[offs[s_] + t * s + j], wi = -pl->twi[offs[s_] + t * s + j];
                        size_t i = (size_t)(b + j + t * s) * lanes + l;
                        xr[t] = re[i] * wr - im[i] * wi;
                        xi[t] = re[i] * wi + im[i] * wr;
                    }
                    small_dft(st, p, xr, xi, 1);
                    for (uint32_t t = 0; t < p; t++) {
                        size_t i = (size_t)(b + j + t * s) * lanes + l;
                        re[i] = xr[t];
                        im[i] = xi[t];
                    }
                }
    }
}

int or_c2p_shape(uint32_t L, uint32_t *R, uint32_t *C) {  // src/fast2p.c:216-238
    uint32_t H = L / 2, br = 0, bc = 0;
    double best = 1e30;
    uint8_t tmp[32];
    for (uint32_t c = 16; c <= H / 2; c += 8) {
        if (H % c) continue;
        uint32_t r = H / c;
        if (radices(r, tmp) < 0 || radices(c, tmp) < 0) continue;
        double cost = fabs(log2((double)r) - log2((double)c) + 2);
        if (r > 512) cost += 10;
        if (c > 4096) cost += 10;
        if (cost < best) {
            best = cost;
            br = r;
            bc = c;
        }
    }
    if (!br) return -1;
    *R = br;
    *C = bc;
    return 0;
}

static inline size_t re_at(const tp_state *st, uint32_t r, uint32_t c) {
    size_t e = (size_t)r * st->C + c;
    return (e >> 1) * 4 + (e & 1);
}

static int tp_init(or_ctx *c) {
    tp_state *st = &c->ts;
    if (c->L % 32) return -1;
    st->H = c->L / 2;
    if (or_c2p_shape(c->L, &st->R, &st->C)) return -1;
    if (2 * GRP < c->L - c->top_t) return -1;  // top region must fit the last group's row
    st->NG = st->C / GRP;
    for (uint32_t p = 2; p <= 8; p++)
        for (uint32_t q = 0; q < p; q++) or_sincos2pi(q, p, &st->roots_r[p][q], &st->roots_i[p][q]);
    if (plan_make(&st->pr, st->R) || plan_make(&st->pc, st->C)) return -1;
    uint32_t R = st->R, C = st->C, H = st->H;
    st->t4r = malloc((size_t)H * sizeof(double));
    st->t4i = malloc((size_t)H * sizeof(double));
    for (uint32_t p = 0; p < R; p++)
        for (uint32_t cc = 0; cc < C; cc++) {
            double co, si;
            or_sincos2pi((uint64_t)cc * st->pr.rev[p] % H, H, &co, &si);
            st->t4r[(size_t)p * C + cc] = co;
            st->t4i[(size_t)p * C + cc] = -si;
        }
    st->rsr = malloc(R * sizeof(double));
    st->rsi = malloc(R * sizeof(double));
    st->partner = malloc(R * sizeof(uint32_t));
    for (uint32_t p = 0; p < R; p++) {
        double co, si;
        uint32_t kr = st->pr.rev[p];
        or_sincos2pi(kr, H, &co, &si);
        st->rsr[p] = co;
        st->rsi[p] = -si;
        st->partner[p] = st->pr.pos[(R - kr) % R];
    }
    st->csr = malloc(C * sizeof(double));
    st->csi = malloc(C * sizeof(double));
    for (uint32_t q = 0; q < C; q++) {
        double co, si;
        or_sincos2pi(st->pc.rev[q], C, &co, &si);
        st->csr[q] = co;
        st->csi[q] = -si;
    }
    st->sre = malloc((size_t)R * GRP * sizeof(double));
    st->sim = malloc((size_t)R * GRP * sizeof(double));
    st->g0 = malloc((size_t)R * 2 * GRP * sizeof(double));
    st->carry = malloc(R * sizeof(int64_t));
    st->ar = malloc(C * sizeof(double));
    st->ai = malloc(C * sizeof(double));
    st->br = malloc(C * sizeof(double));
    st->bi = malloc(C * sizeof(double));
    st->yr = malloc((size_t)H * sizeof(double));
    st->yi = malloc((size_t)H * sizeof(double));
    return 0;
}

static void tp_free(or_ctx *c) {
    tp_state *st = &c->ts;
    if (!st->H) return;
    plan_drop(&st->pr);
    plan_drop(&st->pc);
    void *ps[] = {st->t4r, st->t4i, st->rsr, st->rsi, st->csr, st->csi, st->partner, st->sre,
                  st->sim, st->g0, st->carry, st->ar, st->ai, st->br, st->bi, st->yr, st->yi};
    for (size_t i = 0; i < sizeof ps / sizeof *ps; i++) free(ps[i]);
    memset(st, 0, sizeof *st);
}

static void grp_load(tp_state *st, const double *d, uint32_t g) {
    for (uint32_t r = 0; r < st->R; r++)
        for (uint32_t l = 0; l < GRP; l++) {
            size_t i = re_at(st, r, g * GRP + l);
            st->sre[r * GRP + l] = d[i];
            st->sim[r * GRP + l] = d[i + 2];
        }
}

static void grp_store(tp_state *st, double *d, uint32_t g) {
    for (uint32_t r = 0; r < st->R; r++)
        for (uint32_t l = 0; l < GRP; l++) {
            size_t i = re_at(st, r, g * GRP + l);
            d[i] = st->sre[r * GRP + l];
            d[i + 2] = st->sim[r * GRP + l];
        }
}

static void grp_twiddle(tp_state *st, uint32_t g, int conj_) {
    for (uint32_t p = 0; p < st->R; p++)
        for (uint32_t l = 0; l < GRP; l++) {
            size_t t = (size_t)p * st->C + g * GRP + l;
            double wr = st->t4r[t], wi = conj_ ? -st->t4i[t] : st->t4i[t];
            double xr = st->sre[p * GRP + l], xi = st->sim[p * GRP + l];
            st->sre[p * GRP + l] = xr * wr - xi * wi;
            st->sim[p * GRP + l] = xr * wi + xi * wr;
        }
}

static void grp_inverse(tp_state *st, uint32_t g) {
    grp_twiddle(st, g, 1);
    dit_inv(st, &st->pr, st->sre, st->sim, GRP);
}

static void grp_forward(tp_state *st, uint32_t g) {
    dif(st, &st->pr, st->sre, st->sim, GRP);
    grp_twiddle(st, g, 0);
}

static inline size_t dig_of(const tp_state *st, uint32_t r, uint32_t c, uint32_t h) {
    return 2 * ((size_t)r * st->C + c) + h;
}

// Round the scaled row r of group g into 2*GRP digits with a carry chain (src/fast2p.c:377-394).
static int64_t row_round(or_ctx *c, uint32_t g, uint32_t r, double scale, int64_t cy,
                         int64_t *dg, double *me) {
    tp_state *st = &c->ts;
    for (uint32_t l = 0; l < GRP; l++)
        for (uint32_t h = 0; h < 2; h++) {
            size_t di = dig_of(st, r, g * GRP + l, h);
            double v = (h ? st->sim : st->sre)[r * GRP + l] * scale * c->oscale[di];
            double rv = rint(v), e = fabs(v - rv);
            if (e > *me) *me = e;
            cy = bal_split((int64_t)rv + cy, c->bits[di], &dg[2 * l + h]);
        }
    return cy;
}

static void row_weight(or_ctx *c, uint32_t g, uint32_t r, const int64_t *dg) {
    tp_state *st = &c->ts;
    for (uint32_t l = 0; l < GRP; l++) {
        st->sre[r * GRP + l] = (double)dg[2 * l] * c->w[dig_of(st, r, g * GRP + l, 0)];
        st->sim[r * GRP + l] = (double)dg[2 * l + 1] * c->w[dig_of(st, r, g * GRP + l, 1)];
    }
}

// Add a carry into the group's row; the last digit of the row keeps any remainder
// (src/fast2p.c:404-419).
static void absorb(const or_ctx *c, uint32_t r, uint32_t g, int64_t *dg, int64_t cy) {
    const tp_state *st = &c->ts;
    for (uint32_t i = 0; i < 2 * GRP && cy; i++) {
        if (i == 2 * GRP - 1) {
            dg[i] += cy;
            return;
        }
        cy = bal_split(dg[i] + cy, c->bits[dig_of(st, r, g * GRP + i / 2, i & 1)], &dg[i]);
    }
}

// The fold on the last group's last row, digits L-2*GRP .. L-1 (src/fast2p.c:421-451).
static int64_t fold_row(const or_ctx *c, int64_t *dg, int64_t cy) {
    uint32_t base = c->L - 2 * GRP;
    int64_t *full = c->dig;  // ctx scratch of L digits; only the top window is read/written
    for (uint32_t i = 0; i < 2 * GRP; i++) full[base + i] = dg[i];
    int64_t q = fold(c, full, cy);
    for (uint32_t i = 0; i < 2 * GRP; i++) dg[i] = full[base + i];
    return q;
}

static void tp_pass1(or_ctx *c, double *d) {  // src/fast2p.c:453-491
    tp_state *st = &c->ts;
    uint32_t R = st->R, NG = st->NG;
    double sc = 1.0 / (4.0 * st->H), me = c->maxerr;
    int64_t dg[2 * GRP];
    grp_load(st, d, 0);
    grp_inverse(st, 0);
    for (uint32_t r = 0; r < R; r++) {
        st->carry[r] = row_round(c, 0, r, sc, 0, dg, &me);
        for (int i = 0; i < 2 * GRP; i++) st->g0[r * 2 * GRP + i] = (double)dg[i];
    }
    int64_t wrapq = 0;
    for (uint32_t g = 1; g < NG; g++) {
        grp_load(st, d, g);
        grp_inverse(st, g);
        for (uint32_t r = 0; r < R; r++) {
            int64_t cy = row_round(c, g, r, sc, st->carry[r], dg, &me);
            if (g == NG - 1 && r == R - 1) {
                wrapq = fold_row(c, dg, cy);
                cy = 0;
            }
            st->carry[r] = cy;
            row_weight(c, g, r, dg);
        }
        grp_forward(st, g);
        grp_store(st, d, g);
    }
    for (uint32_t r = 0; r < R; r++) {
        for (int i = 0; i < 2 * GRP; i++) dg[i] = (int64_t)st->g0[r * 2 * GRP + i];
        absorb(c, r, 0, dg, r == 0 ? wrapq : st->carry[r - 1]);
        row_weight(c, 0, r, dg);
    }
    grp_forward(st, 0);
    grp_store(st, d, 0);
    c->maxerr = me;
}

static void row_load(const tp_state *st, const double *d, uint32_t p, double *re, double *im) {
    for (uint32_t q = 0; q < st->C; q++) {
        size_t i = re_at(st, p, q);
        re[q] = d[i];
        im[q] = d[i + 2];
    }
}

static void row_store(const tp_state *st, double *d, uint32_t p, const double *re, const double *im) {
    for (uint32_t q = 0; q < st->C; q++) {
        size_t i = re_at(st, p, q);
        d[i] = re[q];
        d[i + 2] = im[q];
    }
}

// One real-split pair (a = U_k, b = U_{H-k}) with t = w_H^k: square (y == NULL) or multiply by the
// operand's pair y = {Y_k, Y_{H-k}} (src/fast2p.c:512-543). With e = a + conj(b), o = -i(a - conj(b)):
//   square: X = e^2 + t o^2, Y = 2 e o;  multiply: X = e ey + t o oy, Y = e oy + o ey
//   new a = X + iY, new b = conj(X) + i conj(Y)   (both 4x the packed product's DFT)
static void split_pair(double *ar, double *ai, double *br, double *bi, double tr, double ti,
                       const double *y) {
    double er = *ar + *br, ei = *ai - *bi;
    double orr = *ai + *bi, oi = -(*ar - *br);
    double Xr, Xi, Yr, Yi;
    if (!y) {
        double e2r = er * er - ei * ei, e2i = 2 * er * ei;
        double o2r = orr * orr - oi * oi, o2i = 2 * orr * oi;
        Xr = e2r + (tr * o2r - ti * o2i);
        Xi = e2i + (tr * o2i + ti * o2r);
        Yr = 2 * (er * orr - ei * oi);
        Yi = 2 * (er * oi + ei * orr);
    } else {
        double fr = y[0] + y[2], fi = y[1] - y[3];        // e of the operand
        double gr = y[1] + y[3], gi = -(y[0] - y[2]);     // o of the operand
        double pr = orr * gr - oi * gi, pi = orr * gi + oi * gr;
        Xr = (er * fr - ei * fi) + (tr * pr - ti * pi);
        Xi = (er * fi + ei * fr) + (tr * pi + ti * pr);
        Yr = (er * gr - ei * gi) + (orr * fr - oi * fi);
        Yi = (er * gi + ei * gr) + (orr * fi + oi * fr);
    }
    *ar = Xr - Yi;
    *ai = Xi + Yr;
    *br = Xr + Yi;
    *bi = -Xi + Yr;
}

static void tp_pass2(or_ctx *c, double *d, int mul) {  // src/fast2p.c:545-617
    tp_state *st = &c->ts;
    uint32_t R = st->R, C = st->C;
    for (uint32_t p = 0; p < R; p++) {
        uint32_t pp = st->partner[p];
        if (pp < p) continue;
        row_load(st, d, p, st->ar, st->ai);
        dif(st, &st->pc, st->ar, st->ai, 1);
        if (pp != p) {
            row_load(st, d, pp, st->br, st->bi);
            dif(st, &st->pc, st->br, st->bi, 1);
        }
        uint32_t kr = st->pr.rev[p];
        const double *yar = st->yr + (size_t)p * C, *yai = st->yi + (size_t)p * C;
        const double *ybr = st->yr + (size_t)pp * C, *ybi = st->yi + (size_t)pp * C;
        for (uint32_t q = 0; q < C; q++) {
            uint32_t kc = st->pc.rev[q], qb;
            double *xr2, *xi2;
            const double *pbr = ybr, *pbi = ybi;
            if (kr == 0) {
                qb = st->pc.pos[(C - kc) % C];
                if (qb < q) continue;
                xr2 = &st->ar[qb];
                xi2 = &st->ai[qb];
                pbr = yar;
                pbi = yai;
            } else if (pp == p) {
                qb = st->pc.pos[C - 1 - kc];
                if (qb < q) continue;
                xr2 = &st->ar[qb];
                xi2 = &st->ai[qb];
                pbr = yar;
                pbi = yai;
            } else {
                qb = st->pc.pos[C - 1 - kc];
                xr2 = &st->br[qb];
                xi2 = &st->bi[qb];
            }
            // w_H^k with k = k_r + R*k_c
            double tr = st->rsr[p] * st->csr[q] - st->rsi[p] * st->csi[q];
            double ti = st->rsr[p] * st->csi[q] + st->rsi[p] * st->csr[q];
            double ar = st->ar[q], ai = st->ai[q], br = *xr2, bi = *xi2;
            double y4[4];
            const double *yp = NULL;
            if (mul) {
                y4[0] = yar[q];
                y4[1] = yai[q];
                y4[2] = pbr[qb];
                y4[3] = pbi[qb];
                yp = y4;
            }
            split_pair(&ar, &ai, &br, &bi, tr, ti, yp);
            st->ar[q] = ar;
            st->ai[q] = ai;
            if (!(qb == q && pp == p)) {
                *xr2 = br;
                *xi2 = bi;
            }
        }
        dit_inv(st, &st->pc, st->ar, st->ai, 1);
        row_store(st, d, p, st->ar, st->ai);
        if (pp != p) {
            dit_inv(st, &st->pc, st->br, st->bi, 1);
            row_store(st, d, pp, st->br, st->bi);
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Backend dispatch

void or_from_digits(or_ctx *c, double *x, const int64_t *dig) {
    if (c->backend == OR_BACKEND_REF) {
        for (uint32_t j = 0; j < c->L; j++) x[j] = (double)dig[j];
        return;
    }
    tp_state *st = &c->ts;  // weight, forward column FFT per group (src/fast2p.c:621-634)
    for (uint32_t g = 0; g < st->NG; g++) {
        for (uint32_t r = 0; r < st->R; r++)
            for (uint32_t l = 0; l < GRP; l++) {
                size_t d0 = dig_of(st, r, g * GRP + l, 0);
                st->sre[r * GRP + l] = (double)dig[d0] * c->w[d0];
                st->sim[r * GRP + l] = (double)dig[d0 + 1] * c->w[d0 + 1];
            }
        grp_forward(st, g);
        grp_store(st, x, g);
    }
}

void or_to_digits(or_ctx *c, const double *x, int64_t *dig) {
    if (c->backend == OR_BACKEND_REF) {
        for (uint32_t j = 0; j < c->L; j++) dig[j] = (int64_t)x[j];
        return;
    }
    tp_state *st = &c->ts;  // src/fast2p.c:636-649
    for (uint32_t g = 0; g < st->NG; g++) {
        grp_load(st, x, g);
        grp_inverse(st, g);
        for (uint32_t r = 0; r < st->R; r++)
            for (uint32_t l = 0; l < GRP; l++) {
                size_t d0 = dig_of(st, r, g * GRP + l, 0);
                dig[d0] = (int64_t)rint(st->sre[r * GRP + l] / (st->R * c->w[d0]));
                dig[d0 + 1] = (int64_t)rint(st->sim[r * GRP + l] / (st->R * c->w[d0 + 1]));
            }
    }
}

void or_square(or_ctx *c, double *x) {
    if (c->backend == OR_BACKEND_REF) {
        full_forward(c, x, c->fs.a);
        for (uint32_t j = 0; j < c->L; j++) c->fs.a[j] *= c->fs.a[j];
        full_inverse_carry(c, c->fs.a, x);
        return;
    }
    tp_pass2(c, x, 0);
    tp_pass1(c, x);
}

void or_square_n(or_ctx *c, double *x, uint64_t count) {
    for (uint64_t i = 0; i < count; i++) or_square(c, x);
}

void or_mul(or_ctx *c, double *x, const double *y) {
    if (c->backend == OR_BACKEND_REF) {
        cplx *fy = malloc(c->L * sizeof(cplx));
        full_forward(c, y, c->fs.a);
        memcpy(fy, c->fs.a, c->L * sizeof(cplx));
        full_forward(c, x, c->fs.a);
        for (uint32_t j = 0; j < c->L; j++) c->fs.a[j] *= fy[j];
        free(fy);
        full_inverse_carry(c, c->fs.a, x);
        return;
    }
    tp_state *st = &c->ts;
    for (uint32_t p = 0; p < st->R; p++) {
        double *yr = st->yr + (size_t)p * st->C, *yi = st->yi + (size_t)p * st->C;
        row_load(st, y, p, yr, yi);
        dif(st, &st->pc, yr, yi, 1);
    }
    tp_pass2(c, x, 1);
    tp_pass1(c, x);
}

void or_mul_small(or_ctx *c, double *x, int64_t m) {
    or_to_digits(c, x, c->dig);
    for (uint32_t j = 0; j < c->L; j++) c->dig[j] *= m;
    or_carry_digits(c, c->dig);
    or_from_digits(c, x, c->dig);
}

void or_dft(const double *re, const double *im, double *ore, double *oim, uint32_t n, int inv) {
    for (uint32_t k = 0; k < n; k++) {
        long double sr = 0, si = 0;
        for (uint32_t j = 0; j < n; j++) {
            double co, s;
            or_sincos2pi((uint64_t)j * k % n, n, &co, &s);
            if (!inv) s = -s;
            sr += (long double)re[j] * co - (long double)im[j] * s;
            si += (long double)re[j] * s + (long double)im[j] * co;
        }
        ore[k] = (double)sr;
        oim[k] = (double)si;
    }
}

// ---------------------------------------------------------------------------------------------
// base^(N+1) -> base^(N-1) mod N, with the CRT when gcd(base, N) > 1 (src/prp.c:36-144).

static uint64_t mod_small(const uint64_t *a, size_t n, uint64_t m) {
    u128 r = 0;
    for (size_t i = n; i-- > 0;) r = ((r << 64) | a[i]) % m;
    return (uint64_t)r;
}

static void divexact_small(uint64_t *a, size_t n, uint64_t d, u128 top) {
    u128 r = top;
    for (size_t i = n; i-- > 0;) {
        r = (r << 64) | a[i];
        a[i] = (uint64_t)(r / d);
        r %= d;
    }
}

// a <- a/d mod M (a < M, gcd(d, M) = 1): add t*M so that d divides, then divide exactly.
static void div_mod(uint64_t *a, const uint64_t *M, size_t n, uint64_t d) {
    uint64_t am = mod_small(a, n, d), mm = mod_small(M, n, d), t = 0;
    while ((am + (u128)t * mm) % d) t++;
    u128 cy = 0;
    for (size_t i = 0; i < n; i++) {
        cy += (u128)a[i] + (u128)M[i] * t;
        a[i] = (uint64_t)cy;
        cy >>= 64;
    }
    divexact_small(a, n, d, cy);
}

int or_residue_nm1(uint64_t *y, const uint64_t *N, size_t n, uint64_t base) {
    uint64_t Q = 1, b = base;
    uint64_t *M = calloc(n + 1, 8), *t = calloc(n + 1, 8);
    memcpy(M, N, n * 8);
    for (uint64_t p = 2; b > 1; p++) {
        if (b % p) continue;
        while (b % p == 0) b /= p;
        while (mod_small(M, n, p) == 0) {
            if (Q > (1ull << 40)) {
                free(M);
                free(t);
                return -1;
            }
            Q *= p;
            divexact_small(M, n, p, 0);
        }
    }
    if (Q == 1) {
        div_mod(y, N, n, base);
        div_mod(y, N, n, base);
        free(M);
        free(t);
        return 0;
    }
    // r = y mod M, computed as (y*Q mod N) / Q
    u128 cy = 0;
    for (size_t i = 0; i < n; i++) {
        cy += (u128)y[i] * Q;
        t[i] = (uint64_t)cy;
        cy >>= 64;
    }
    t[n] = (uint64_t)cy;
    for (;;) {  // t -= N while t >= N
        int ge = t[n] != 0;
        if (!ge) ge = cmp_limbs(t, N, n) >= 0;
        if (!ge) break;
        uint64_t br = 0;
        for (size_t i = 0; i <= n; i++) {
            u128 d = (u128)t[i] - (i < n ? N[i] : 0) - br;
            t[i] = (uint64_t)d;
            br = (uint64_t)(d >> 64) & 1;
        }
    }
    divexact_small(t, n + 1, Q, 0);
    div_mod(t, M, n, base);
    div_mod(t, M, n, base);
    // x = t + M*u with x == 0 (mod Q)
    uint64_t rm = mod_small(t, n, Q), mm = mod_small(M, n, Q), u = 0;
    while ((rm + (u128)mm * u) % Q) u++;
    cy = 0;
    for (size_t i = 0; i < n; i++) {
        cy += (u128)t[i] + (u128)M[i] * u;
        y[i] = (uint64_t)cy;
        cy >>= 64;
    }
    free(M);
    free(t);
    return 0;
}

// ---------------------------------------------------------------------------------------------
// Fermat PRP driver: Gerbicz blocks, rollback, round-off retry, doubled tail (src/prp.c:298-583).

static void power_start(or_ctx *c, double *x0, uint64_t k, uint32_t base, uint64_t *l) {
    or_set_u64(c, x0, 1);
    for (int b = 63; b >= 0; b--) {
        if (k >> b == 0) continue;
        or_square(c, x0);
        if ((k >> b) & 1) or_mul_small(c, x0, base);
    }
    if (k * log2((double)base) + 2 < (double)c->n) {  // exact power while below N
        memset(l, 0, c->nlimbs * 8);
        l[0] = 1;
        size_t used = 1;
        for (uint64_t i = 0; i < k; i++) {
            u128 cy = 0;
            for (size_t j = 0; j < used; j++) {
                cy += (u128)l[j] * base;
                l[j] = (uint64_t)cy;
                cy >>= 64;
            }
            if (cy) l[used++] = (uint64_t)cy;
        }
        or_set_limbs(c, x0, l);
    }
}

static void corrupt(or_ctx *c, double *x) {  // src/prp.c:26-34
    int64_t *d = malloc(c->L * sizeof(int64_t));
    or_to_digits(c, x, d);
    d[c->L / 3] += 1;
    or_from_digits(c, x, d);
    free(d);
}

int or_prp(uint64_t k, uint64_t n, uint32_t base, uint32_t fft_len, uint32_t block,
           uint64_t fault_at, int backend, or_prp_result *res) {
    memset(res, 0, sizeof *res);
    uint32_t B = block;
    if (!B) {
        B = (uint32_t)sqrt((double)n);
        if (B < 32) B = 32;
        if (B > 2000) B = 2000;
    }
    uint32_t M = B / 4 > 8 ? B / 4 : 8;
    or_ctx *c = or_ctx_new(k, n, fft_len, backend);
    if (!c) return -1;
    size_t nl = c->nlimbs, nd = c->L;
    double *x = malloc(nd * 8), *d = malloc(nd * 8), *x0 = malloc(nd * 8), *t = malloc(nd * 8);
    uint64_t *la = calloc(nl, 8), *lb = calloc(nl, 8), *vx = calloc(nl, 8), *vd = calloc(nl, 8);
    uint64_t vi = 0, full = n / B * B;
    int rc = 0, nfail = 0;
    power_start(c, x0, k, base, la);
    or_get_limbs(c, x0, vx);
    memcpy(vd, vx, nl * 8);
    for (;;) {
        or_set_limbs(c, x, vx);
        or_set_limbs(c, d, vd);
        or_reset_maxerr(c);
        uint64_t i = vi, end = vi + (uint64_t)B * M;
        if (end > full) end = full;
        int failed = 0, rfail = 0;
        while (i < end) {
            for (uint32_t j = 0; j < B; j++) {
                or_square(c, x);
                i++;
                if (fault_at && i == fault_at) {
                    corrupt(c, x);
                    fault_at = 0;
                }
            }
            memcpy(t, d, nd * 8);
            or_mul(c, d, x);
            if (c->maxerr > 0.40) {
                rfail = 1;
                break;
            }
        }
        if (!rfail && i > vi) {  // d == x0 * t^(2^B)
            or_square_n(c, t, B);
            or_mul(c, t, x0);
            if (c->maxerr > 0.40) rfail = 1;
            else {
                or_get_limbs(c, t, la);
                or_get_limbs(c, d, lb);
                failed = memcmp(la, lb, nl * 8) != 0;
            }
        }
        if (c->maxerr > res->maxerr) res->maxerr = c->maxerr;
        if (rfail) {
            uint32_t L = or_fft_len(c), L2 = or_next_fft_len(L + L / 25);
            or_ctx_free(c);
            c = L2 ? or_ctx_new(k, n, L2, backend) : NULL;
            if (!c) {
                rc = -2;
                goto out;
            }
            free(x); free(d); free(x0); free(t);
            nd = c->L;
            x = malloc(nd * 8); d = malloc(nd * 8); x0 = malloc(nd * 8); t = malloc(nd * 8);
            power_start(c, x0, k, base, la);
            res->roundoff_retries++;
            continue;
        }
        if (failed) {
            res->gerbicz_errors++;
            if (++nfail > 5) {
                rc = -3;
                goto out;
            }
            continue;
        }
        nfail = 0;
        or_get_limbs(c, x, vx);
        or_get_limbs(c, d, vd);
        vi = i;
        if (vi >= full) break;
    }
    for (int attempt = 0;; attempt++) {  // tail of n - full squarings, twice
        or_set_limbs(c, x, vx);
        or_set_limbs(c, t, vx);
        or_square_n(c, x, n - full);
        or_square_n(c, t, n - full);
        or_get_limbs(c, x, la);
        or_get_limbs(c, t, lb);
        if (!memcmp(la, lb, nl * 8)) break;
        if (attempt > 5) {
            rc = -3;
            goto out;
        }
        res->gerbicz_errors++;
    }
    {
        int prp = la[0] == (uint64_t)base * base;
        for (size_t j = 1; j < nl; j++) prp &= la[j] == 0;
        if (or_residue_nm1(la, c->Nl, nl, base)) rc = -4;
        res->res64 = la[0];
        res->is_prp = prp;
        res->fft_len = or_fft_len(c);
        if (c->maxerr > res->maxerr) res->maxerr = c->maxerr;
    }
out:
    free(x); free(d); free(x0); free(t);
    free(la); free(lb); free(vx); free(vd);
    or_ctx_free(c);
    return rc;
}

// ---------------------------------------------------------------------------------------------
// LLR (src/llr.c:29-305): P choice by Jacobi symbols, Lucas ladder for V_k, n-1 doublings.

static uint64_t mulm(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((u128)a * b % m); }

static uint64_t N_mod(uint64_t k, uint64_t n, uint64_t m) {  // (k*2^n - 1) mod m
    uint64_t r = 1 % m, b = 2 % m;
    for (uint64_t e = n; e; e >>= 1, b = mulm(b, b, m))
        if (e & 1) r = mulm(r, b, m);
    return (mulm(k % m, r, m) + m - 1 % m) % m;
}

static int jac(uint64_t a, uint64_t m) {  // Jacobi (a | m), m odd
    int s = 1;
    a %= m;
    while (a) {
        while (!(a & 1)) {
            a >>= 1;
            if ((m & 7) == 3 || (m & 7) == 5) s = -s;
        }
        uint64_t t = a;
        a = m;
        m = t;
        if ((a & 3) == 3 && (m & 3) == 3) s = -s;
        a %= m;
    }
    return m == 1 ? s : 0;
}

static int jac_N(uint64_t a, uint64_t k, uint64_t n) {  // Jacobi (a | N) by reciprocity
    int s = 1;
    uint64_t n8 = N_mod(k, n, 8);
    while (a && !(a & 1)) {
        a >>= 1;
        if (n8 == 3 || n8 == 5) s = -s;
    }
    if (a == 1) return s;
    if ((a & 3) == 3 && (n8 & 3) == 3) s = -s;
    return s * jac(N_mod(k, n, a), a);
}

static uint64_t gcd64(uint64_t a, uint64_t b) {
    while (b) {
        uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

int or_llr_choose_P(uint64_t k, uint64_t n, uint32_t *P, uint64_t *factor) {
    uint32_t p = 1;
    while (jac_N((uint64_t)p * p + 4, k, n) == 1)
        if (++p > 1000000) return -1;
    *P = p;
    uint64_t d = 4ull * p * ((uint64_t)p * p + 4), g = gcd64(d, N_mod(k, n, d));
    *factor = g > 1 ? g : 0;
    return 0;
}

int or_llr(uint64_t k, uint64_t n, uint32_t fft_len, int backend, or_llr_result *res) {
    memset(res, 0, sizeof *res);
    if (n < 3 || !(k & 1)) return -1;
    if (or_llr_choose_P(k, n, &res->P, &res->factor)) return -1;
    if (res->factor) return 0;
    or_ctx *c = or_ctx_new(k, n, fft_len, backend);
    if (!c) return -1;
    size_t nd = c->L, nl = c->nlimbs;
    double *x = malloc(nd * 8), *v1 = malloc(nd * 8), *cP = malloc(nd * 8), *c2 = malloc(nd * 8),
           *t = malloc(nd * 8);
    or_set_u64(c, cP, res->P);
    or_set_u64(c, c2, 2);
    // (V_m, V_{m+1}) from m = 0 over the bits of k, Q = -1 (src/llr.c:150-184)
    or_set_u64(c, x, 2);
    memcpy(v1, cP, nd * 8);
    uint64_t m = 0;
    int top = 63;
    while (top > 0 && !((k >> top) & 1)) top--;
    for (int b = top; b >= 0; b--) {
        int odd = (int)(m & 1);
        memcpy(t, x, nd * 8);
        or_mul(c, t, v1);  // V_{2m+1} = V_m V_{m+1} - P Q^m
        if (odd) or_add(c, t, cP);
        else or_sub(c, t, cP);
        if ((k >> b) & 1) {
            or_square(c, v1);  // V_{2m+2} = V_{m+1}^2 - 2 Q^{m+1}
            if (odd) or_sub(c, v1, c2);
            else or_add(c, v1, c2);
            memcpy(x, t, nd * 8);
            m = 2 * m + 1;
        } else {
            or_square(c, x);  // V_{2m} = V_m^2 - 2 Q^m
            if (odd) or_add(c, x, c2);
            else or_sub(c, x, c2);
            memcpy(v1, t, nd * 8);
            m = 2 * m;
        }
    }
    for (uint64_t i = 0; i < n - 1; i++) {  // V_2k = V_k^2 + 2, then V -> V^2 - 2
        or_square(c, x);
        if (i == 0) or_add(c, x, c2);
        else or_sub(c, x, c2);
    }
    uint64_t *l = calloc(nl, 8);
    or_get_limbs(c, x, l);
    int zero = 1;
    for (size_t i = 0; i < nl; i++) zero &= l[i] == 0;
    res->is_prime = zero;
    res->res64 = l[0];
    res->fft_len = c->L;
    res->maxerr = c->maxerr;
    free(l);
    free(x); free(v1); free(cP); free(c2); free(t);
    or_ctx_free(c);
    return 0;
}
