// The arithmetic around an LLR (Morrison) test of N = k*2^n - 1: N modulo a small number, the
// Jacobi symbol (D | N), the choice of the V-sequence parameter P (reference src/llr.c:29-90),
// and V_k(P, -1) by a ladder of GPU operations (reference src/llr.c:150-184).
#include <cstdint>

#include "kite_internal.h"

namespace kite {
namespace {

typedef unsigned __int128 u128;

uint64_t mul_mod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((u128)a * b % m); }

// Jacobi symbol (a | m) for odd m > 0, by the binary method: factors of 2 come out of a with
// (2 | m) = -1 exactly when m = 3 or 5 (mod 8), and an odd a trades places with m, the sign
// flipping when both are 3 (mod 4).
int jacobi_odd(uint64_t a, uint64_t m) {
    int sign = 1;
    a %= m;
    while (a != 0) {
        const int twos = __builtin_ctzll(a);
        a >>= twos;
        const uint64_t m8 = m & 7;
        if ((twos & 1) && (m8 == 3 || m8 == 5)) sign = -sign;
        if ((a & 3) == 3 && (m & 3) == 3) sign = -sign;
        const uint64_t r = m % a;
        m = a;
        a = r;
    }
    return m == 1 ? sign : 0;
}

// (a | N) for N = k 2^n - 1 (odd, too large for a machine word): with a = 2^e b, b odd,
// (2 | N) follows from N mod 8, and reciprocity turns (b | N) into (N mod b | b) with a sign
// that is -1 when b and N are both 3 (mod 4).
int jacobi_riesel(uint64_t a, uint64_t k, uint64_t n) {
    const uint64_t N8 = n_mod(k, n, 8);
    int sign = 1;
    const int e = a ? __builtin_ctzll(a) : 0;
    const uint64_t b = a >> e;
    if ((e & 1) && (N8 == 3 || N8 == 5)) sign = -sign;
    if (b == 1) return sign;
    if ((b & 3) == 3 && (N8 & 3) == 3) sign = -sign;
    return sign * jacobi_odd(n_mod(k, n, b), b);
}

uint64_t gcd_u64(uint64_t a, uint64_t b) {
    while (b) {
        a %= b;
        const uint64_t t = a;
        a = b;
        b = t;
    }
    return a;
}

}  // namespace

// (k 2^n - 1) mod m, by square-and-multiply of 2 modulo m.
uint64_t n_mod(uint64_t k, uint64_t n, uint64_t m) {
    if (m == 1) return 0;
    uint64_t p = 1, sq = 2 % m;
    for (uint64_t e = n; e; e >>= 1) {
        if (e & 1) p = mul_mod(p, sq, m);
        sq = mul_mod(sq, sq, m);
    }
    return (mul_mod(k % m, p, m) + m - 1) % m;
}

// V_k(P, -1) into x. The ladder keeps (V_j, V_{j+1}) and reads k from its top bit: a 0 bit makes
// j -> 2j, a 1 bit j -> 2j + 1, through V_{2j} = V_j^2 - 2 Q^j and V_{2j+1} = V_j V_{j+1} - P Q^j
// with Q^j = (-1)^j. Numbers: lo = V_j, hi = V_{j+1}; cP and c2 hold P and 2; mid is scratch.
void llr_lucas_vk(kite_ctx *ctx, uint64_t k, kite_num *lo, kite_num *hi, kite_num *cP, kite_num *c2, kite_num *mid) {
    kite_set_u64(ctx, lo, 2);  // V_0
    kite_copy(ctx, hi, cP);    // V_1
    bool j_odd = false;        // parity of j, so Q^j = -1 when set
    for (int bit = 63 - __builtin_clzll(k); bit >= 0; bit--) {
        // mid = V_{2j+1} = V_j V_{j+1} - P Q^j
        kite_copy(ctx, mid, lo);
        kite_mul(ctx, mid, hi);
        (j_odd ? kite_add : kite_sub)(ctx, mid, cP);
        if ((k >> bit) & 1) {
            // j -> 2j+1: (V_{2j+1}, V_{2j+2}), V_{2j+2} = V_{j+1}^2 - 2 Q^{j+1}
            kite_square(ctx, hi);
            (j_odd ? kite_sub : kite_add)(ctx, hi, c2);
            kite_copy(ctx, lo, mid);
            j_odd = true;
        } else {
            // j -> 2j: (V_{2j}, V_{2j+1})
            kite_square(ctx, lo);
            (j_odd ? kite_add : kite_sub)(ctx, lo, c2);
            kite_copy(ctx, hi, mid);
            j_odd = false;
        }
    }
}

}  // namespace kite

// The smallest P >= 1 with (P^2 + 4 | N) != 1; a common factor of 4P(P^2 + 4) and N is reported
// in *factor (the test is then not needed). -1 if no P up to 10^6 qualifies.
extern "C" int kite_llr_choose_P(uint64_t k, uint64_t n, uint32_t *P, uint64_t *factor) {
    for (uint64_t p = 1; p <= 1000000; p++) {
        const uint64_t D = p * p + 4;
        if (kite::jacobi_riesel(D, k, n) == 1) continue;
        *P = (uint32_t)p;
        const uint64_t q = 4 * p * D;
        const uint64_t g = kite::gcd_u64(q, kite::n_mod(k, n, q));
        *factor = g > 1 ? g : 0;
        return 0;
    }
    return -1;
}
