// CPU oracle for the modular-squaring engine mod N = k*2^n - 1.
//
// TEST INFRASTRUCTURE ONLY. This restates, in plain C, the algorithm of the reference's
// portable path so that tests can check the HIP engine against it:
//   layout, weights, carry/fold, limb conversions   src/ctx.c:17-559
//   full-length FFT squaring ("ref" backend)          src/ref_fft.c:1-150
//   two-pass layout squaring ("c2p" backend)          src/fast2p.c:1-684
//   Fermat PRP driver with Gerbicz checking           src/prp.c:36-583
//   LLR (Morrison, Lucas V) driver                    src/llr.c:29-384
// Nothing in the product library links, loads or calls this code; only tests/, the smoke test
// and the benchmark's CPU-baseline leg do.
//
// Pinning (see tests/test_oracle.py): every arithmetic entry point is checked against GMP
// vectors committed under tests/golden/ (random residues: set/get round trip, 20 squarings,
// mul, mul_small, add, sub, the u^2-2 chain, 3^(2^30)); full PRP runs against the reference's
// known_residues.csv rows; layout invariants (w in [1,2), bits in {b,b+1}, s_L = n+ceil(log2 k))
// against closed forms; each FFT against a brute-force O(L^2) DFT on small inputs.
#ifndef KITE_ORACLE_H
#define KITE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_BACKEND_REF = 0, OR_BACKEND_C2P = 1 };

typedef struct or_ctx or_ctx;

// FFT-length rules (src/ctx.c:17-148)
int or_fft_len_ok(uint32_t L);
double or_safe_bpw(uint32_t L, uint64_t k);
uint32_t or_choose_fft_len(uint64_t k, uint64_t n);
uint32_t or_next_fft_len(uint32_t L);

// Context (src/ctx.c:223-296). L = 0 chooses the length. Returns NULL on bad input.
or_ctx *or_ctx_new(uint64_t k, uint64_t n, uint32_t L, int backend);
void or_ctx_free(or_ctx *c);
uint32_t or_fft_len(const or_ctx *c);
size_t or_nlimbs(const or_ctx *c);
size_t or_num_doubles(const or_ctx *c);
// Layout export: bits[L], w[L]; info = {top_t, top_st, top_m, sL, lam}
void or_layout(const or_ctx *c, uint8_t *bits, double *w, uint64_t *info);
uint64_t or_s_at(const or_ctx *c, uint64_t j);
double or_weight(const or_ctx *c, uint64_t j, uint64_t sj);

// Numbers are arrays of or_num_doubles(c) doubles owned by the caller.
void or_set_limbs(or_ctx *c, double *x, const uint64_t *limbs);
void or_get_limbs(or_ctx *c, const double *x, uint64_t *limbs);
void or_set_u64(or_ctx *c, double *x, uint64_t v);
void or_square(or_ctx *c, double *x);
void or_square_n(or_ctx *c, double *x, uint64_t count);
void or_mul(or_ctx *c, double *x, const double *y);
void or_mul_small(or_ctx *c, double *x, int64_t m);
void or_add(or_ctx *c, double *x, const double *y);
void or_sub(or_ctx *c, double *x, const double *y);
double or_maxerr(const or_ctx *c);
void or_reset_maxerr(or_ctx *c);
void or_to_digits(or_ctx *c, const double *x, int64_t *dig);
void or_from_digits(or_ctx *c, double *x, const int64_t *dig);
void or_carry_digits(const or_ctx *c, int64_t *dig);
// Digits (balanced, any magnitude < 2^62) -> limbs of the value mod N they represent.
void or_digits_to_limbs(or_ctx *c, const int64_t *dig, uint64_t *limbs);

// y = base^(N+1) -> base^(N-1) mod N (src/prp.c:63-144)
int or_residue_nm1(uint64_t *y, const uint64_t *N, size_t n, uint64_t base);

typedef struct {
    uint64_t res64;
    int is_prp;
    uint32_t fft_len;
    double maxerr;
    uint32_t gerbicz_errors;
    uint32_t roundoff_retries;
} or_prp_result;

// Fermat PRP driver (src/prp.c:329-583) without checkpoints or proof points.
int or_prp(uint64_t k, uint64_t n, uint32_t base, uint32_t fft_len, uint32_t block,
           uint64_t fault_at, int backend, or_prp_result *res);

typedef struct {
    int is_prime;
    uint64_t factor;
    uint64_t res64;
    uint32_t P;
    uint32_t fft_len;
    double maxerr;
} or_llr_result;

int or_llr_choose_P(uint64_t k, uint64_t n, uint32_t *P, uint64_t *factor);
// One LLR run (src/llr.c:192-305 without retries/checkpoints): V_{k*2^(n-1)}(P,-1) mod N.
int or_llr(uint64_t k, uint64_t n, uint32_t fft_len, int backend, or_llr_result *res);

// Two-pass shape chosen by the c2p restatement (src/fast2p.c:216-238).
int or_c2p_shape(uint32_t L, uint32_t *R, uint32_t *C);

// Brute-force complex DFT (sign -1 forward), for checking FFT code on tiny inputs.
void or_dft(const double *re, const double *im, double *ore, double *oim, uint32_t n, int inv);
// cos/sin(2*pi*j/n) with exact argument reduction (the twiddle convention of the oracle).
void or_sincos2pi(uint64_t j, uint64_t n, double *c, double *s);

#ifdef __cplusplus
}
#endif
#endif
