// Batched and serving extensions of the MI355X kite library.
//
// These entry points have no counterpart in the reference, which runs one test per CPU core
// (tools/kite.c:54-105); they exist for callers that run many independent tests on one GPU. A
// batch holds number slots for independent moduli N = k*2^n - 1 of any lengths; each call below
// advances every slot with shared launches per pass of the squaring (rows, then the column pass:
// one fused kernel for squarings at compile-time lengths, else column inverse + carry and column
// forward), so the GPU works on all tests at once.
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
// For each slot s < capacity with which[s] >= 0: residue which[s] squared count[s] times
// (x -> x^2 + xadd each time, see kite_batch_set_xadd).
int kite_batch_square(kite_batch *b, const int32_t *which, const uint32_t *count);
// The slots are dealt to `parts` streams (1..4, default 2) whose launches run concurrently, so
// kernels of different parts overlap their memory and compute phases. Returns the count in use.
int kite_batch_set_parts(kite_batch *b, int parts);
// Column jobs are laid out for `xcds` XCDs (work-group i of a launch runs on XCD i mod xcds; 8 on
// MI355X, the default): each test's column groups run on one XCD. 1 keeps plain slot order.
// Returns the value in use.
int kite_batch_set_xcds(kite_batch *b, int xcds);
// Shape parts (off by default): the slots of the most common shape (R x C split) are dealt to the
// `parts` streams as usual, and each other shape's slots to a stream of its own while streams are
// left (4 in all): a few tests of another FFT length then run their short launches beside the
// main ones instead of after them in the same stream, and each length keeps its own column kernel
// instead of sharing a two-length one. Returns the setting in use.
int kite_batch_set_shape_parts(kite_batch *b, int on);
// Persistent launches (off by default): the square-mode row and fused column passes run as
// persistent kernels where those exist, on at most the device's resident capacity divided among
// the parts, each work-group running a fixed share of the jobs, instead of one work-group per job.
// Two batches must not run persistent launches at the same time on one device: the launches
// would exceed what the device holds at once, and a fused column job would then wait for one
// that cannot start (its wait is bounded: the test then reports a round-off failure). Returns the
// setting in use.
int kite_batch_set_persistent(kite_batch *b, int on);
// Square calls of more than one step are captured as a HIP graph and replayed while the batch's
// slots are unchanged (on by default; the launches are the same either way). Off, each step's
// kernels are launched directly, on streams that each have a hardware queue of their own (created
// at the first switch; the batch's work so far is waited for). Returns 0, -1 on a HIP failure.
int kite_batch_set_graphs(kite_batch *b, int on);
// Squarings of this slot become x -> x^2 + xadd (|xadd| < 2^20; 0 by default, -2 for LLR steps).
int kite_batch_set_xadd(kite_batch *b, int slot, int32_t xadd);
// For each slot s with dst[s] >= 0: residue dst[s] = residue dst[s] * residue src[s].
int kite_batch_mul(kite_batch *b, const int32_t *dst, const int32_t *src);
// For each slot s with dst[s] >= 0: residue dst[s] = residue src[s].
int kite_batch_copy(kite_batch *b, const int32_t *dst, const int32_t *src);
// For each slot s with a[s] >= 0 and c[s] >= 0: eq[s] = 1 if residues a[s] and c[s] have the same
// digits (so the same value mod N), 0 if their digits differ (the values may still agree: digits
// are balanced, but not canonical in every position; compare limbs to decide); -1 elsewhere. The
// comparison runs on the device; only the answers come back.
int kite_batch_equal(kite_batch *b, const int32_t *a, const int32_t *c, int32_t *eq);
// out[s] = max round-off of slot s since its last reset (capacity entries); reset if asked.
int kite_batch_maxerr(kite_batch *b, double *out, int reset);
int kite_batch_sync(kite_batch *b);
// Device time of each pass over the last kite_batch_square call timed with HIP events (ms):
// out[0] rows, out[1] column inverse + carry, out[2] column forward. With timing every N > 0,
// each launch of every N-th step carries a start/stop event pair stamped from its own dispatch
// (hipExtLaunchKernel), on the batch's first stream (0 turns it off); launches3 counts the timed
// launches.
int kite_batch_set_timing(kite_batch *b, int every);
int kite_batch_pass_times(kite_batch *b, double *out3, uint64_t *launches3);
// Per-kernel totals over the same timed call: kernel i's name, summed device time (ms), number of
// launches and summed algorithmic bytes (16 L per test for a row launch, 8 L per test for each
// column launch). Returns the number of kernels (i < 0 only queries it), -1 for i out of range.
int kite_batch_kernel_times(kite_batch *b, int i, char *name, size_t name_len, double *ms,
                            uint64_t *launches, double *bytes);

// kite_square_n without allocation. No counterpart in the reference: it is for callers that cannot
// allocate device memory (code under HIP stream capture, tight serving loops, harnesses that forbid
// allocation). x is squared count times, like kite_square_n, but with direct launches (no graph)
// and the squaring's device scratch (its slot record and the fused column pass's carry flags) in
// a static 2 MiB array of the library, one per GPU, that holds carry flags for FFT lengths up to
// about 2M. Returns 0 on success, -2 if the context's length needs more flags than the scratch
// holds (nothing is done), -1 on a failure (the context's error is latched). Waits for the
// squarings before it returns. Every call on one GPU shares that GPU's scratch: calls that overlap
// on one GPU (from several threads) race; calls in sequence do not.
int kite_square_n_static(kite_ctx *ctx, kite_num *x, uint64_t count);

// Testing hook: the n-th HIP call the library checks from now on (in any context or batch) is
// reported as failed (n <= 0: off). Nothing fails on the device; the library takes its error path:
// the context's round-off reads +inf, batch calls return -1 and the drivers a negative code.
void kite_debug_fail_hip(long n);
// Testing hook (no device needed): cos and sin(2 pi j / n) as the library computes them for its
// twiddle tables (exact integer reduction of j/n to [0, 1/8], long double evaluation).
void kite_debug_root_of_unity(uint64_t j, uint64_t n, double *c, double *s);

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

// ---- mixed Fermat / LLR jobs with arrival times (continuous batching, serving) ---------------

enum { KITE_JOB_PRP = 0, KITE_JOB_LLR = 1 };
enum { KITE_JOB_PENDING = 1 };  // status of a job the horizon stopped before it finished
enum { KITE_ORDER_ARRIVAL = 0, KITE_ORDER_LEAST_WORK = 1 };  // kite_run_opts.order
enum { KITE_SHARE_SQUARINGS = 0, KITE_SHARE_TIME = 1 };      // kite_run_opts.share
enum { KITE_LAUNCH_GRAPH = 0, KITE_LAUNCH_DIRECT = 1 };       // kite_run_opts.launch

typedef struct {
    uint64_t k, n;
    uint32_t kind;      // KITE_JOB_PRP (Fermat, Gerbicz-checked) or KITE_JOB_LLR
    uint32_t base;      // Fermat base (0 = 3)
    uint32_t fft_len;   // 0 = automatic
    double arrival;     // seconds after the start of kite_batch_run before which it may not start
} kite_job;

typedef struct {
    int status;         // 0 done; kite_prp / kite_llr negative codes; KITE_JOB_PENDING
    int is_prime;       // Fermat: probable prime; LLR: prime
    uint64_t res64;     // Fermat: low 64 bits of base^(N-1); LLR: of the final V (0 if prime)
    uint64_t factor;    // LLR: a factor found while choosing P (the test is then not run)
    uint32_t P;         // LLR: the V-sequence parameter
    uint32_t fft_len;   // length of the final run
    double maxerr;
    uint32_t gerbicz_errors, roundoff_retries, confirm_runs;
    double start, finish;  // seconds after the start of kite_batch_run: start, when the job's setup
                           // (x_0 or the ladder, or its checkpoint) was done; finish, result known
    uint64_t resumed_at;   // iteration restored from a checkpoint (0: none, or an LLR run's start)
    uint32_t resumed;      // 1 if the job started from a checkpoint
    uint32_t run;          // LLR: the run in progress or last done (0, or a confirmation run)
    uint64_t iter;         // squarings done in the current run (Fermat: of the chain x; LLR: doublings)
    uint64_t verified;     // Fermat: iteration of the last state that passed a Gerbicz check;
                           // LLR: step of the last block that passed its round-off check
} kite_job_result;

typedef struct {
    uint32_t slots;          // jobs in flight at most (0 = no limit)
    uint64_t hbm_budget;     // device bytes for the jobs in flight at most (0 = no limit)
    uint32_t block;          // squarings per round, the Fermat Gerbicz block B (0 = 1000)
    uint32_t gerbicz_blocks; // Fermat: blocks between checks (0 = B/4)
    double max_roundoff;     // 0 = 0.40
    double horizon;          // seconds; jobs unfinished by then end KITE_JOB_PENDING (0 = none)
    uint64_t inject_fault_job;  // testing: with inject_fault_at, corrupt job (index) at iteration
    uint64_t inject_fault_at;   // (LLR: in its first run, so the vote has to outvote it)
    void (*on_done)(void *arg, size_t index, const kite_job_result *r);
    void *arg;
    // Checkpoints (NULL: none). The directory must exist and be writable, else kite_batch_run
    // returns -1 at once with every status -1. A job resumes from its file there when it starts, and
    // writes it after a verified block at most every checkpoint_secs (0 = 600) and for every job in
    // flight when the horizon stops the run; the file is removed when the job finishes. Fermat
    // jobs use kite_prp's KITECKP1 file "kite_<fp>.<base>.ckpt", LLR jobs kite_llr's KITELLR1
    // file "kite_<fp>.llr.ckpt" (fp = N mod 3417905339), the names kite-prst uses, so a test
    // stopped in one of them can be finished by the others.
    const char *checkpoint_dir;
    double checkpoint_secs;
    // Admission order among the jobs that have arrived: KITE_ORDER_ARRIVAL (first come, first
    // served) or KITE_ORDER_LEAST_WORK (least work first, work = n times the FFT length, twice
    // for LLR jobs, whose composite answers are confirmed by a second run: with a slot limit,
    // short tests finish while long ones wait). A job that has waited more than 60 s goes first
    // whatever its work, so a steady stream of short jobs cannot hold a long one back for ever.
    uint32_t order;
    // How the jobs in flight share the rounds: KITE_SHARE_SQUARINGS (every job squares one block
    // per round: the most work per launch, the highest throughput) or KITE_SHARE_TIME (a job of
    // FFT length L takes part in a fraction L_min / L of the rounds, L_min the shortest length in
    // flight: every job gets about the same device time per round, so a short test is not held to
    // the pace of the longest ones in flight, which lowers serving latency).
    uint32_t share;
    // The batch's launch settings (0 = its defaults): streams (kite_batch_set_parts), shape parts
    // (kite_batch_set_shape_parts: 1 on) and the XCD layout (kite_batch_set_xcds), as a batched
    // tuning run (python -m ahnum_amd.tune_batch) found best for the workload.
    uint32_t parts;
    uint32_t shape_parts;
    uint32_t xcds;
    // How a round's squarings reach the device: KITE_LAUNCH_GRAPH (replayed from a captured HIP
    // graph, as kite_batch_set_graphs(b, 1)) or KITE_LAUNCH_DIRECT (each step's kernels launched
    // on the batch's streams, as kite_batch_set_graphs(b, 0)). Which runs faster depends on the
    // workload's launches; the batched tuner measures both.
    uint32_t launch;
} kite_run_opts;

// Device bytes a job of this kind holds at length L while in flight (the admission estimate).
uint64_t kite_job_bytes(uint32_t kind, uint32_t fft_len);

// Runs the jobs through one batch: each job starts once its arrival time has passed and the slot
// and HBM limits allow (in the admission order of o->order); every round advances all jobs in flight by up
// to `block` squarings with shared launches, then checks each (Fermat: Gerbicz every M blocks with
// rollback, doubled tail; LLR: round-off after every block with a verified copy, confirmation
// runs at the next FFT length and a vote as kite_llr). A finished job's slot is refilled at
// once. Returns 0 if every job finished with status 0, else the first other status.
int kite_batch_run(const kite_job *jobs, size_t njobs, const kite_run_opts *o, kite_job_result *res);

#ifdef __cplusplus
}
#endif
#endif
