// Batched and serving extensions of the MI355X kite library.
//
// These entry points have no counterpart in the reference, which runs one test per CPU core
// (tools/kite.c:54-105); they exist for callers that run many independent tests on one GPU. A
// batch holds number slots for independent moduli N = k*2^n - 1 of any lengths; each call below
// advances every slot with one launch per pass of the squaring (rows, column inverse + carry,
// column forward), so the GPU works on all tests at once.
#ifndef KITE_BATCH_H
#define KITE_BATCH_H

#include <stddef.h>
#include <stdint.h>

#include "kite.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kite_batch kite_batch;

kite_batch *kite_batch_new(void);
void kite_batch_free(kite_batch *b);
// New slot for k*2^n-1 holding `nnums` residues (all 0); fft_len 0 chooses. Returns the slot id
// (>= 0) or -1. Slot ids are reused after kite_batch_remove.
int kite_batch_add(kite_batch *b, uint64_t k, uint64_t n, uint32_t fft_len, uint32_t nnums);
void kite_batch_remove(kite_batch *b, int slot);
int kite_batch_capacity(const kite_batch *b);  // highest slot id + 1
// The slot's context and residues, for the kite.h conversions (kite_set_limbs/kite_get_limbs).
// Call kite_batch_sync before converting a residue the batch has just worked on.
kite_ctx *kite_batch_ctx(kite_batch *b, int slot);
kite_num *kite_batch_num(kite_batch *b, int slot, uint32_t i);
// For each slot s < capacity with which[s] >= 0: residue which[s] = residue^(2^count[s]).
int kite_batch_square(kite_batch *b, const int32_t *which, const uint32_t *count);
// For each slot s with dst[s] >= 0: residue dst[s] = residue dst[s] * residue src[s].
int kite_batch_mul(kite_batch *b, const int32_t *dst, const int32_t *src);
// For each slot s with dst[s] >= 0: residue dst[s] = residue src[s].
int kite_batch_copy(kite_batch *b, const int32_t *dst, const int32_t *src);
// out[s] = max round-off of slot s since its last reset (capacity entries); reset if asked.
int kite_batch_maxerr(kite_batch *b, double *out, int reset);
int kite_batch_sync(kite_batch *b);
// Device time of each pass over the last kite_batch_square call timed with HIP events (ms):
// out[0] rows, out[1] column inverse + carry, out[2] column forward. With timing every N > 0,
// each launch of every N-th step is bracketed by events on the batch stream (0 turns it off);
// launches3 counts the timed launches.
int kite_batch_set_timing(kite_batch *b, int every);
int kite_batch_pass_times(kite_batch *b, double *out3, uint64_t *launches3);
// Per-kernel totals over the same timed call: kernel i's name, summed device time (ms), number of
// launches and summed algorithmic bytes (16 L per test for a row launch, 8 L per test for each
// column launch). Returns the number of kernels (i < 0 only queries it), -1 for i out of range.
int kite_batch_kernel_times(kite_batch *b, int i, char *name, size_t name_len, double *ms,
                            uint64_t *launches, double *bytes);

// ---- batched Fermat PRP tests ---------------------------------------------------------------

typedef struct {
    uint64_t k, n;
    uint32_t base;         // 0 = 3
    uint32_t fft_len;      // 0 = automatic
} kite_batch_test;

typedef struct {
    uint32_t slots;          // tests in flight (0 = all)
    uint32_t gerbicz_block;  // common block size B (0 = 1000)
    uint32_t gerbicz_blocks; // blocks between checks (0 = B/4)
    double max_roundoff;     // 0 = 0.40
    uint64_t inject_fault_test;  // testing: with inject_fault_at, corrupt test (index) at iteration
    uint64_t inject_fault_at;
    // called as each test finishes: its index, status (0 or kite_prp's negative codes), result
    void (*on_done)(void *arg, size_t index, int status, const kite_prp_result *r);
    void *arg;
} kite_batch_prp_opts;

// Runs every test to completion (Gerbicz-checked, round-off retry at a longer FFT per test).
// res[i] and status[i] receive test i's result and status (0, or kite_prp's -1/-2/-3/-4);
// returns 0 if every test succeeded, else the first failing status.
int kite_batch_prp(const kite_batch_test *tests, size_t ntests, const kite_batch_prp_opts *o,
                   kite_prp_result *res, int *status);

#ifdef __cplusplus
}
#endif
#endif
