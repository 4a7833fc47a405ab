// kite (MI355X build): modular squaring mod N = k*2^n - 1 on one AMD Instinct MI355X (gfx950),
// for Fermat PRP and LLR (Lucas V) tests.
//
// Drop-in replacement for the reference library's public API: every declaration below keeps the
// reference's name, signature and meaning (reference include/kite.h:1-174). Numbers live in
// device memory (HBM) in the engine's weighted, transformed representation; all arithmetic is
// modulo N. The library is HIP-only: KITE_BACKEND_AUTO and KITE_BACKEND_HIP select the gfx950
// engine, and the CPU/AArch64 backends of the reference (REF, NEON, SME, C2P) are not built, so
// kite_ctx_new returns NULL for them (as the reference does for a backend that cannot start).
#ifndef KITE_H
#define KITE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KITE_VERSION "0.1.8-mi355x"

typedef struct kite_ctx kite_ctx;
typedef struct kite_num kite_num;

// reference include/kite.h:20-26, plus the HIP engine
typedef enum {
    KITE_BACKEND_AUTO = 0,
    KITE_BACKEND_REF = 1,   // not built here (CPU)
    KITE_BACKEND_NEON = 2,  // not built here (AArch64)
    KITE_BACKEND_SME = 3,   // not built here (AArch64)
    KITE_BACKEND_C2P = 4,   // not built here (CPU)
    KITE_BACKEND_HIP = 5,   // MI355X engine
} kite_backend;

// reference include/kite.h:28-35
typedef struct {
    kite_backend backend;
    uint32_t fft_len;      // 0 = choose automatically
    uint32_t threads;      // accepted for compatibility; the GPU engine ignores it
    // Shape order: 0 = from the tuning file, else 1 + variant index (see `kite tune`)
    uint32_t plan_order;
} kite_opts;

// reference include/kite.h:37-69
kite_ctx *kite_ctx_new(uint64_t k, uint64_t n, int32_t c, const kite_opts *o);
void kite_ctx_free(kite_ctx *ctx);
uint32_t kite_fft_len(const kite_ctx *ctx);
const char *kite_backend_name(const kite_ctx *ctx);

kite_num *kite_num_new(kite_ctx *ctx);
void kite_num_free(kite_ctx *ctx, kite_num *x);
void kite_copy(kite_ctx *ctx, kite_num *dst, const kite_num *src);
void kite_set_u64(kite_ctx *ctx, kite_num *x, uint64_t v);

void kite_square(kite_ctx *ctx, kite_num *x);
void kite_square_n(kite_ctx *ctx, kite_num *x, uint64_t count);
void kite_mul(kite_ctx *ctx, kite_num *x, const kite_num *y);   // x = x*y
void kite_mul_small(kite_ctx *ctx, kite_num *x, int64_t m);   // x = x*m, |m| < 2^20
// x = x + y and x = x - y, element-wise on the (linear) stored form; at most one add or sub
// between multiplications.
void kite_add(kite_ctx *ctx, kite_num *x, const kite_num *y);
void kite_sub(kite_ctx *ctx, kite_num *x, const kite_num *y);

// Largest |rounded - exact| seen since the last reset (0..0.5).
double kite_max_roundoff(const kite_ctx *ctx);
void kite_reset_roundoff(kite_ctx *ctx);
// Round-off is measured on every operation by the GPU engine; `every` is kept for the API.
void kite_set_roundoff_check(kite_ctx *ctx, uint32_t every);

size_t kite_limbs(const kite_ctx *ctx);
void kite_get_limbs(kite_ctx *ctx, const kite_num *x, uint64_t *limbs);
void kite_set_limbs(kite_ctx *ctx, kite_num *x, const uint64_t *limbs);
uint64_t kite_res64(kite_ctx *ctx, const kite_num *x);

// ---- Fermat PRP test (reference include/kite.h:71-118) ----------------------------------

typedef struct {
    kite_opts fft;
    uint32_t gerbicz_block;    // 0 = auto (~sqrt(n), in [32, 2000])
    uint32_t gerbicz_blocks;   // blocks between checks; 0 = auto (block/4, at least 8)
    double max_roundoff;       // 0 = 0.40
    const char *checkpoint;    // verified state file (NULL = none)
    double checkpoint_secs;    // 0 = 600
    const char *checkpoint_unverified;  // unverified state file (NULL = none)
    double unverified_secs;             // 0 = 60
    void (*progress)(void *arg, uint64_t iter, uint64_t n, double maxerr);
    void *progress_arg;
    uint64_t inject_fault_at;  // testing: corrupt the state at this iteration (0 = never)
    const uint32_t *points;
    uint32_t npoints;
    int (*on_point)(void *arg, uint32_t index, uint64_t iter, const uint64_t *limbs);
    void *point_arg;
    volatile int *abort_flag;
} kite_prp_opts;

enum { KITE_PRP_ABORTED = -5, KITE_PRP_POINT_FAILED = -6 };

typedef struct {
    uint64_t res64;            // low 64 bits of base^(N-1) mod N
    int is_prp;
    uint32_t fft_len;
    double maxerr;
    uint32_t gerbicz_errors;
    uint32_t roundoff_retries;
    int resumed;
    uint64_t resumed_at;
    double seconds;
} kite_prp_result;

int kite_prp(uint64_t k, uint64_t n, uint32_t base, const kite_prp_opts *o, kite_prp_result *res);

// ---- LLR (Morrison) test (reference include/kite.h:120-166) -----------------------------

typedef struct {
    kite_opts fft;
    double max_roundoff;       // 0 = 0.40
    const char *checkpoint;
    double checkpoint_secs;    // 0 = 600
    void (*progress)(void *arg, uint64_t iter, uint64_t total, double maxerr);
    void (*on_confirm)(void *arg, int run);
    void *progress_arg;
    volatile int *abort_flag;
    uint64_t inject_fault_at;
    uint32_t inject_fault_runs;
    uint64_t inject_roundoff_at;
    int inject_roundoff_ladder;
} kite_llr_opts;

typedef struct {
    int is_prime;
    uint64_t factor;
    uint64_t res64;
    uint32_t P;
    uint32_t fft_len;
    double maxerr;
    uint32_t roundoff_retries;
    uint32_t confirm_runs;
    int resumed;
    uint64_t resumed_at;
    double seconds;
} kite_llr_result;

enum { KITE_LLR_ABORTED = -5 };

int kite_llr(uint64_t k, uint64_t n, const kite_llr_opts *o, kite_llr_result *res);
int kite_llr_choose_P(uint64_t k, uint64_t n, uint32_t *P, uint64_t *factor);
int kite_llr_checkpoint_iter(const char *path, uint64_t k, uint64_t n, uint64_t *iter, int *run);
int kite_prp_checkpoint_iter(const char *path, uint64_t k, uint64_t n, uint32_t base,
                            uint64_t *iter);

// Testing hook: add 1 to one digit (a valid representation of a wrong value).
void kite_debug_corrupt(kite_ctx *ctx, kite_num *x);

// ---- Internal symbols the reference exports to its tools and tests (src/internal.h:74-93) --

uint32_t kite_choose_fft_len(uint64_t k, uint64_t n);
uint32_t kite_next_fft_len(uint32_t L);
double kite_safe_bpw(uint32_t L, uint64_t k);
int kite_residue_nm1(uint64_t *y, const uint64_t *N, size_t n, uint64_t base);
uint64_t kite_s_at(const kite_ctx *ctx, uint64_t j);
double kite_weight(const kite_ctx *ctx, uint64_t j, uint64_t sj);
void kite_carry_digits(const kite_ctx *ctx, int64_t *dig);

#ifdef __cplusplus
}
#endif
#endif
