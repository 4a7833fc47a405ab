// Internal definitions of the MI355X kite library (host side).
//
// Representation (the k-scaled IBDWT of the reference, src/internal.h:3-17): for N = k*2^n - 1
// and L real words, T' = n + floor(log2 k) + lam/2^64 (lam = frac(log2 k) to 64 bits),
// s_j = ceil(j*T'/L), digit j has s_{j+1} - s_j bits, weight w_j = 2^(s_j - j*T'/L) in [1, 2).
// The stored value is V = x/k (mod N) as sum d_j 2^{s_j}; squaring the weighted convolution
// and scaling by k/w_j maps V to k*V^2, which is V(x^2). The part at or above 2^n (top digits
// t..L-1 plus the final carry) is folded with k*2^n == 1 (mod N).
//
// Numbers live in HBM as H = L/2 complex doubles in the "after forward column pass" state of
// the four-step transform (the engine in engine.hip owns that layout).
#ifndef KITE_INTERNAL_H
#define KITE_INTERNAL_H

#include <cstddef>
#include <cstdint>
#include <vector>

#include "kite.h"
#include "kite_batch.h"

namespace kite {
struct Engine;  // device state of one context (engine.hip)
}

struct kite_num {
    void *dev;  // H complex doubles in device memory
};

struct kite_ctx {
    uint64_t k = 0, n = 0;
    int32_t c = -1;
    uint32_t L = 0;
    uint32_t plan_order = 0;   // shape variant (0 = tuned/default)
    std::vector<uint8_t> bits; // bits[j] = s_{j+1} - s_j
    uint64_t lam = 0;          // frac(log2 k) * 2^64
    uint64_t sL = 0;           // s_L
    uint32_t top_t = 0;        // first digit with s_{t+1} > n (L if none)
    uint64_t top_st = 0;       // s_t
    uint32_t top_m = 0;        // n - s_t
    uint32_t b0 = 0;           // floor(T'/L): every digit has b0 or b0+1 bits
    uint32_t min_sub = 1;      // carry sub-chunk the engine's shape must have at least (elements)
    size_t nlimbs = 0;
    std::vector<uint64_t> Nl;  // N as limbs
    uint32_t err_every = 1;
    kite::Engine *eng = nullptr;
    std::vector<int64_t> tmpdig;
};

namespace kite {

// ---- layout (layout.cpp) --------------------------------------------------------------------
uint64_t frac_log2_64(uint64_t k);
int build_layout(kite_ctx *ctx);
// Expected peak magnitude, in bits, of a product before its carry: log2 k + 2 (b0 + 1) + log2(L)/2.
double peak_product_bits(const kite_ctx *ctx);
// A length whose products would pass this many bits is refused: past 2^53 an FP64 product has no
// fraction bits left, so its round-off can no longer be measured, and the digits would outgrow int32.
constexpr double kMaxProductBits = 53.0;
// The smallest carry sub-chunk (elements of 2 digits, a power of two) whose carry the next sub-chunk
// absorbs with a leftover below 2^(b0 - 1); 0 if none up to 64 elements. Above 1 only for short digits
// with a large k.
uint32_t carry_sub_needed(const kite_ctx *ctx);
// Exact fraction numerator of j*T'/L: (j * T'fix) mod (L*2^64), as a 128-bit value.
unsigned __int128 phase_num(const kite_ctx *ctx, uint64_t j);
unsigned __int128 tfix(const kite_ctx *ctx);  // T' * 2^64
// 2^(-num/(L*2^64)) rounded to double, num < L*2^64.
double exp2_neg_frac(const kite_ctx *ctx, unsigned __int128 num);

// ---- digits and limbs (layout.cpp) -----------------------------------------------------------
void digits_to_limbs(kite_ctx *ctx, const int64_t *dig, uint64_t *limbs);
void limbs_to_digits(kite_ctx *ctx, const uint64_t *limbs, int64_t *dig);

// ---- FFT length rules (layout.cpp) -----------------------------------------------------------
bool fft_len_ok(uint32_t L);

// ---- engine (engine.hip) ---------------------------------------------------------------------
int engine_init(kite_ctx *ctx);
void engine_free(kite_ctx *ctx);
// A shape for L whose top region (top_digits, 0 if unknown) fits the last carry sub-chunk and whose
// carry sub-chunks hold at least min_sub elements (2 digits each).
bool engine_shape_ok(uint32_t L, uint32_t top_digits, uint32_t min_sub = 1);
bool engine_shape_fast(uint32_t L);  // a shape with compile-time kernels exists
void *engine_num_alloc(kite_ctx *ctx);
void engine_num_free(kite_ctx *ctx, void *p);
void engine_copy(kite_ctx *ctx, void *dst, const void *src);
void engine_from_digits(kite_ctx *ctx, void *x, const int64_t *dig);
void engine_to_digits(kite_ctx *ctx, const void *x, int64_t *dig);
void engine_square(kite_ctx *ctx, void *x, uint64_t count, int32_t xadd = 0);  // x -> x^2 + xadd, count times
void engine_mul(kite_ctx *ctx, void *x, const void *y);
void engine_addsub(kite_ctx *ctx, void *x, const void *y, int sign);
double engine_maxerr(kite_ctx *ctx);
void engine_reset_maxerr(kite_ctx *ctx);
void engine_sync(kite_ctx *ctx);
int engine_error(const kite_ctx *ctx);  // 0, or -1 once a HIP call of this context has failed
// count squarings of x with no allocation (static device scratch): 0, -1 (failure), -2 (too long)
int engine_square_static(kite_ctx *ctx, void *x, uint64_t count);
const char *engine_last_error();

// ---- LLR arithmetic (llr_math.cpp) -----------------------------------------------------------
// x = V_k(P, -1) by the Lucas ladder (cP holds P, c2 holds 2; v1, t are scratch).
void llr_lucas_vk(kite_ctx *ctx, uint64_t k, kite_num *x, kite_num *v1, kite_num *cP, kite_num *c2, kite_num *t);
uint64_t n_mod(uint64_t k, uint64_t n, uint64_t m);  // (k 2^n - 1) mod m

// ---- batch slots with context options (engine.hip) --------------------------------------------
// kite_batch_add with the context options of a kite.h caller (backend, length, shape variant).
int batch_add_opts(kite_batch *b, uint64_t k, uint64_t n, const kite_opts *o, uint32_t nnums);

// ---- checkpoint files (ckpt.cpp) -------------------------------------------------------------
inline size_t limbs_for_n(uint64_t n) { return (n + 64 + 63) / 64; }  // limbs of a residue, k < 2^64
// KITECKP1: Fermat state (x, d) at iteration iter, written at length L; 0 on success. The reader
// checks k, n, base, the limb count and the checksum.
int prp_checkpoint_write(const char *path, uint64_t k, uint64_t n, uint32_t base, uint64_t iter, uint32_t L,
                         const uint64_t *x, const uint64_t *d, size_t nl);
int prp_checkpoint_read(const char *path, uint64_t k, uint64_t n, uint32_t base, uint64_t *iter, uint32_t *L,
                        uint64_t *x, uint64_t *d, size_t nl);
// KITELLR1: LLR run `run`'s x at step iter (length L) and the low words of the earlier runs' results.
struct LlrCheckpoint {
    uint64_t iter = 0;
    uint32_t P = 0, L = 0;
    int run = 0;
    uint64_t prev[2] = {0, 0};
};
int llr_checkpoint_write(const char *path, uint64_t k, uint64_t n, const LlrCheckpoint &c, const uint64_t *x,
                         size_t nl);
int llr_checkpoint_read(const char *path, uint64_t k, uint64_t n, LlrCheckpoint *c, uint64_t *x, size_t nl);

}  // namespace kite

#endif
