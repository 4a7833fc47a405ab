// Every Fermat PRP and LLR test the library runs goes through one scheduler: kite_batch_run (many
// jobs with arrival times), kite_batch_prp (Fermat jobs that all arrive at once), and the
// reference's single-test drivers kite_prp and kite_llr (one job each, with their hooks).
//
// The scheduler works in rounds. A round gives every job in flight one span of squarings through a
// shared kite_batch_square call (spans differ per job: a Gerbicz block, the rest of a block up to a
// proof point or an injected fault, a tail, an LLR block), then the per-job follow-up operations of
// the round in shared calls (copies, Gerbicz multiplies), one read-back of every job's round-off,
// and finally each job's state update on the host. A job that finishes or fails frees its slot for
// the next arrived job at once (continuous batching).
//
// Fermat (reference src/prp.c:329-583): x_0 = base^k, x_{i+1} = x_i^2, so x_n = base^(N+1). The
// Gerbicz product d (d_0 = x_0, d *= x after every block of B squarings) satisfies
// d_{t+1} = x_0 d_t^(2^B); every M blocks t (d before its last update) is squared B times and
// multiplied by x_0 and must equal d, else the job rolls back to its last verified (i, x, d).
// Round-off above the limit restarts the job from that state at a longer FFT. The last n mod B
// squarings are done twice and compared. The result is base^(N-1) = x_n / base^2.
//
// LLR (reference src/llr.c:192-384): V_k(P, -1) by the Lucas ladder, then n - 1 doublings: the
// first V -> V^2 + 2 (V_2k, k odd), the others V -> V^2 - 2, with the addend folded into the
// squaring's carry pass (kite_batch_set_xadd). N is prime iff the result is 0. Round-off is read
// after every block and a failing block is redone from a device copy of the last good state at a
// longer FFT; a composite answer is confirmed at the next FFT length, and a third run breaks a tie.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <deque>
#include <string>
#include <thread>
#include <vector>

#include <sys/stat.h>
#include <unistd.h>

#include "kite_batch.h"
#include "kite_internal.h"

using namespace kite;

namespace {

double now_s() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

// The reference's Gerbicz block for a single test: ~sqrt(n), clamped to [32, 2000], and its
// blocks per check, B/4 but at least 8 (src/prp.c:334-340).
uint32_t default_block(uint64_t n) {
    const uint32_t b = (uint32_t)std::sqrt((double)n);
    return b < 32 ? 32 : b > 2000 ? 2000 : b;
}
uint32_t default_blocks(uint32_t B) { return std::max<uint32_t>(B / 4, 8); }

enum Phase { kBlocks, kCheck, kTailX, kTailT, kLlrFirst, kLlrSteps, kDone };
// Fermat: x, the product d, x_0, t (d before its last update, or the second tail), and the last
// verified (x, d), kept on the device between checks
enum FermatNum { kX = 0, kD = 1, kX0 = 2, kT = 3, kVX = 4, kVD = 5 };
enum LlrNum { kV = 1 };  // LLR: x (kX) and its last verified copy

// What a caller asks of one test.
struct Spec {
    uint64_t k = 0, n = 0;
    uint32_t kind = KITE_JOB_PRP, base = 3, fft_len = 0, plan_order = 0;
    double arrival = 0;
    uint32_t B = 0, M = 0;         // Fermat block and blocks per check
    std::string ckpt, uckpt;       // verified and unverified state files ("" = none)
    double ck_secs = 600, u_secs = 60;
    bool save_at_end = false;      // Fermat: write the verified file when the blocks are done
    // Fermat: a file's Gerbicz product continues with this job's block size (kite_prp, as the
    // reference's readers do); otherwise only when the block is the reference's default for n
    bool product_any_block = false;
    bool drop_ckpt_when_done = true;
    // kite_prp / kite_llr hooks
    void (*progress)(void *, uint64_t, uint64_t, double) = nullptr;
    void *progress_arg = nullptr;
    const uint32_t *points = nullptr;
    uint32_t npoints = 0;
    int (*on_point)(void *, uint32_t, uint64_t, const uint64_t *) = nullptr;
    void *point_arg = nullptr;
    void (*on_confirm)(void *, int) = nullptr;
    volatile int *abort_flag = nullptr;
    uint64_t fault_at = 0;         // corrupt x once, after squaring number fault_at
    uint32_t fault_runs = 1;       // LLR: bit r: corrupt in run r
    uint64_t roundoff_at = 0;      // LLR run 0: report the block reaching this step as over the limit
    bool roundoff_ladder = false;  // LLR run 0: report the ladder as over the limit
};

// A proof point captured since the last verified state (reported once it is verified).
struct Point {
    uint32_t index;
    std::vector<uint64_t> limbs;
};

struct Job {
    Spec sp;
    size_t index = 0;
    int slot = -1;
    uint32_t L = 0;
    Phase phase = kBlocks;
    uint64_t bytes = 0;  // admission estimate while in flight
    bool started = false, stopping = false;
    double last_ck = 0, last_u = 0;
    kite_job_result res{};
    // Fermat: i = squarings done; blocks of B count from x0_at (0, or a restart point)
    uint64_t i = 0, full = 0, check_at = 0, blk = 0;
    uint64_t vi = 0, ui = 0;                 // verified and unverified (from file) positions
    std::vector<uint64_t> vx, vd, ux, ud;    // their x and d (vx, vd: host copies of kVX, kVD)
    bool host_stale = false;                 // vx, vd are older than kVX, kVD
    std::vector<uint64_t> x0;                // x_0, when it is not base^k (a restart point)
    uint64_t x0_at = 0;
    uint32_t fails = 0, tail_tries = 0;
    uint32_t pi = 0, vp = 0, up = 0;         // next proof point: now, at vi, at ui
    std::vector<Point> pending;
    bool fault_pending = false;
    // LLR
    uint32_t P = 0;
    int run = 0;
    uint64_t lres[3] = {0, 0, 0};
    int lzero[3] = {0, 0, 0};
    bool copy_pending = false;  // x passed its block: copy it to the verified number next round
    bool roundoff_pending = false, ladder_roundoff_pending = false;
    double credit = 0;  // KITE_SHARE_TIME: rounds owed to this job (it takes part at >= 1)
    std::vector<std::pair<int, int>> copies;  // device copies (dst, src) due before the next round
};

// x_0 = base^k mod N into x0 and its limbs into l: the exact power, computed on the host, while it
// is below N (always, for k below 2^30 and the n this engine runs); else left-to-right binary
// powering on the device.
void power_start(kite_ctx *ctx, kite_num *x0, uint64_t k, uint32_t base, std::vector<uint64_t> &l) {
    l.assign(ctx->nlimbs, 0);
    if ((double)k * std::log2((double)base) + 2 < (double)ctx->n) {
        l[0] = 1;
        size_t used = 1;
        for (uint64_t e = 0; e < k; e++) {
            unsigned __int128 c = 0;
            for (size_t j = 0; j < used; j++) {
                c += (unsigned __int128)l[j] * base;
                l[j] = (uint64_t)c;
                c >>= 64;
            }
            if (c) l[used++] = (uint64_t)c;
        }
        kite_set_limbs(ctx, x0, l.data());
        return;
    }
    kite_set_u64(ctx, x0, 1);
    for (int b = 63 - __builtin_clzll(k); b >= 0; b--) {
        kite_square(ctx, x0);
        if ((k >> b) & 1) kite_mul_small(ctx, x0, base);
    }
    kite_get_limbs(ctx, x0, l.data());
}

std::string ckpt_name(const Spec &s) {
    char name[64];
    const unsigned long long fp = n_mod(s.k, s.n, 3417905339ull);
    if (s.kind == KITE_JOB_LLR) snprintf(name, sizeof name, "kite_%llu.llr.ckpt", fp);
    else snprintf(name, sizeof name, "kite_%llu.%u.ckpt", fp, s.base);
    return name;
}

bool all_zero(const std::vector<uint64_t> &v) {
    for (uint64_t w : v)
        if (w) return false;
    return true;
}

// Status and results of one run of the scheduler.
struct Scheduler {
    kite_batch *b = nullptr;
    size_t max_slots = (size_t)-1;
    uint64_t hbm_budget = 0;
    uint32_t order = KITE_ORDER_ARRIVAL, share = KITE_SHARE_SQUARINGS;
    double maxround = 0.40, horizon = 0, t0 = 0;
    void (*on_done)(void *, size_t, const kite_job_result *) = nullptr;
    void *done_arg = nullptr;
    std::vector<Job> jobs;
    std::vector<size_t> active;
    std::deque<size_t> waiting;
    uint64_t bytes_in_flight = 0;
    bool engine_failed = false;

    kite_ctx *ctx(const Job &t) { return kite_batch_ctx(b, t.slot); }
    kite_num *num(const Job &t, int i) { return kite_batch_num(b, t.slot, (uint32_t)i); }

    // --- slots -------------------------------------------------------------------------------

    int add_slot(Job &t, uint32_t L, uint32_t nnums) {
        kite_opts o{};
        o.backend = KITE_BACKEND_HIP;
        o.fft_len = L;
        o.plan_order = t.sp.plan_order;
        t.slot = batch_add_opts(b, t.sp.k, t.sp.n, &o, nnums);
        if (t.slot < 0) return -1;
        t.L = kite_fft_len(ctx(t));
        return 0;
    }

    void drop_slot(Job &t) {
        if (t.slot >= 0) kite_batch_remove(b, t.slot);
        t.slot = -1;
    }

    bool engine_ok(const Job &t) { return t.slot < 0 || engine_error(ctx(t)) == 0; }

    // --- Fermat state --------------------------------------------------------------------------

    // (Re)starts Fermat job t in a fresh slot at length L from its verified state (or from the
    // unverified one, if newer).
    int start_fermat(Job &t, uint32_t L) {
        if (add_slot(t, L, 6)) return -1;
        kite_ctx *c = ctx(t);
        kite_num *x0 = num(t, kX0);
        std::vector<uint64_t> l0;
        if (t.x0.empty()) {
            power_start(c, x0, t.sp.k, t.sp.base, l0);
        } else {
            l0 = t.x0;
            kite_set_limbs(c, x0, l0.data());
        }
        if (t.vx.empty()) {  // a fresh start: x = d = x_0
            t.vx = l0;
            t.vd = l0;
            t.vi = 0;
        }
        const bool from_u = t.ui > t.vi;
        // every residue equal to x_0 is a device copy of it; the others come from their limbs
        auto load = [&](int i, const std::vector<uint64_t> &v) {
            if (v == l0) kite_copy(c, num(t, i), x0);
            else kite_set_limbs(c, num(t, i), v.data());
        };
        load(kVX, t.vx);
        load(kVD, t.vd);
        load(kX, from_u ? t.ux : t.vx);
        load(kD, from_u ? t.ud : t.vd);
        engine_sync(c);  // the copies ran on the context's stream; the batch's launches come next
        t.host_stale = false;
        t.copies.clear();
        t.i = from_u ? t.ui : t.vi;
        t.pi = from_u ? t.up : t.vp;
        t.blk = 0;
        t.pending.clear();
        t.full = t.x0_at + (t.sp.n - t.x0_at) / t.sp.B * t.sp.B;
        t.check_at = std::min(t.i + (uint64_t)t.sp.B * t.sp.M, t.full);
        t.phase = t.i >= t.full ? kTailX : kBlocks;
        if (t.phase == kTailX) {
            kite_set_limbs(c, num(t, kX), t.vx.data());
            kite_set_limbs(c, num(t, kT), t.vx.data());
            begin_tail(t);
        }
        return engine_ok(t) ? 0 : -1;
    }

    // The tail from the verified state at `full` (x and t hold it, or get it by the copies due).
    void begin_tail(Job &t) {
        t.i = t.full;
        t.pi = t.vp;
        t.pending.clear();
        t.phase = kTailX;
    }

    // x and d back to the verified state, by device copies before the next round
    void roll_back(Job &t) {
        t.copies.push_back({kX, kVX});
        t.copies.push_back({kD, kVD});
        t.i = t.vi;
        t.pi = t.vp;
        t.blk = 0;
        t.pending.clear();
        t.check_at = std::min(t.vi + (uint64_t)t.sp.B * t.sp.M, t.full);
        t.phase = kBlocks;
    }

    // host copies of the verified state from residues ix, id (after the batch is synchronised)
    void pull(Job &t, int ix, int id) {
        const size_t nl = ctx(t)->nlimbs;
        t.vx.resize(nl);
        t.vd.resize(nl);
        kite_get_limbs(ctx(t), num(t, ix), t.vx.data());
        kite_get_limbs(ctx(t), num(t, id), t.vd.data());
        t.host_stale = false;
    }

    // Reports the points captured since the last verified state; false if the caller refused one.
    bool flush_points(Job &t) {
        for (const Point &p : t.pending)
            if (t.sp.on_point(t.sp.point_arg, p.index, t.sp.points[p.index], p.limbs.data())) {
                t.pending.clear();
                return false;
            }
        t.pending.clear();
        return true;
    }

    void capture_point(Job &t) {
        Point p;
        p.index = t.pi++;
        p.limbs.resize(ctx(t)->nlimbs);
        kite_get_limbs(ctx(t), num(t, kX), p.limbs.data());
        t.pending.push_back(std::move(p));
    }

    // The next iteration after t.i at which the host has to look at x (a proof point or the
    // injected fault), or ~0.
    uint64_t next_event(const Job &t) const {
        uint64_t e = ~0ull;
        if (t.sp.on_point && t.pi < t.sp.npoints && t.sp.points[t.pi] > t.i) e = t.sp.points[t.pi];
        if (t.fault_pending && t.sp.fault_at > t.i) e = std::min<uint64_t>(e, t.sp.fault_at);
        return e;
    }

    // Writes the verified Fermat state. The file holds x and d; a reader with the same block
    // size continues the Gerbicz product from base^k when d != x, and restarts it at x when
    // d == x. The product continues only where the reference's reader can continue it too: from
    // base^k (no restart since) with the reference's own block for this n. Otherwise the file is
    // written with d = x, and the product here restarts at x as a reader's would.
    // (With host copies of the verified state that are current.)
    void save_fermat(Job &t) {
        if (t.sp.ckpt.empty() || t.slot < 0) return;
        const bool continues = t.x0_at == 0 && (t.sp.product_any_block || t.sp.B == default_block(t.sp.n));
        const bool restart = !continues && t.vi > 0 && t.vd != t.vx;
        if (restart && !t.sp.uckpt.empty()) {
            // an unverified file written after this one chains from what this one says, so the
            // product in flight restarts at the verified x as well (blocks stay aligned)
            t.x0 = t.vx;
            t.vd = t.vx;
            for (int i : {kX0, kX, kD, kVX, kVD}) kite_set_limbs(ctx(t), num(t, i), t.vx.data());
            t.copies.clear();
            if (t.phase == kBlocks || t.phase == kCheck) {
                t.i = t.vi;
                t.pi = t.vp;
                t.blk = 0;
                t.pending.clear();
                t.check_at = std::min(t.vi + (uint64_t)t.sp.B * t.sp.M, t.full);
                t.phase = t.vi >= t.full ? kTailX : kBlocks;
            }
            remove(t.sp.uckpt.c_str());
        }
        prp_checkpoint_write(t.sp.ckpt.c_str(), t.sp.k, t.sp.n, t.sp.base, t.vi, t.L, t.vx.data(),
                             restart ? t.vx.data() : t.vd.data(), t.vx.size());
        t.last_ck = now_s();
    }

    // Restores job t from its checkpoint files (before its first start).
    void resume(Job &t) {
        const size_t nl = limbs_for_n(t.sp.n);
        std::vector<uint64_t> x(nl), d(nl);
        uint32_t L = 0;
        if (t.sp.kind == KITE_JOB_PRP) {
            uint64_t it = 0;
            if (!t.sp.ckpt.empty() &&
                !prp_checkpoint_read(t.sp.ckpt.c_str(), t.sp.k, t.sp.n, t.sp.base, &it, &L, x.data(), d.data(), nl) &&
                it <= t.sp.n) {
                t.vi = it;
                t.vx = x;
                t.vd = d;
                // a product this job cannot continue (another block size, or a restart file)
                // restarts at the verified x
                if (it > 0 && (d == x || !(t.sp.product_any_block || t.sp.B == default_block(t.sp.n)))) {
                    t.x0 = x;
                    t.x0_at = it;
                    t.vd = x;
                }
                t.res.resumed = 1;
            }
            uint64_t uit = 0;
            uint32_t uL = 0;
            if (!t.sp.uckpt.empty() &&
                !prp_checkpoint_read(t.sp.uckpt.c_str(), t.sp.k, t.sp.n, t.sp.base, &uit, &uL, x.data(), d.data(), nl) &&
                uit > t.vi && uit <= t.sp.n) {
                t.ui = uit;
                t.ux = x;
                t.ud = d;
                L = std::max(L, uL);
                t.res.resumed = 1;
            }
            // proof points after the restored positions (the ones before them were reported)
            if (t.sp.on_point) {
                while (t.vp < t.sp.npoints && t.sp.points[t.vp] < t.vi + (t.vi > 0)) t.vp++;
                t.up = t.vp;
                while (t.up < t.sp.npoints && t.sp.points[t.up] < t.ui + 1) t.up++;
            }
        } else {
            LlrCheckpoint c;
            if (t.sp.ckpt.empty() || llr_checkpoint_read(t.sp.ckpt.c_str(), t.sp.k, t.sp.n, &c, x.data(), nl) ||
                c.P != t.P)
                return;
            t.run = c.run;
            for (int r = 0; r < 2; r++) {
                t.lres[r] = c.prev[r];
                t.lzero[r] = r < c.run && c.prev[r] == 0;
            }
            t.res.confirm_runs = (uint32_t)c.run;
            t.vi = c.iter;
            t.vx.swap(x);
            L = c.L;
            if (c.run > 0) t.L = L;  // a confirmation run keeps its length
            t.res.resumed = 1;
        }
        // a round-off retry had lengthened the FFT
        if (!t.L && L > kite_choose_fft_len(t.sp.k, t.sp.n)) t.L = L;
        if (t.res.resumed) t.res.resumed_at = std::max(t.vi, t.ui);
    }

    // --- LLR state -----------------------------------------------------------------------------

    // (Re)starts LLR job t's current run in a fresh slot at length L: from its verified state, or
    // from V_k by the ladder. Returns 1 if the ladder's round-off was over the limit.
    int start_llr(Job &t, uint32_t L) {
        if (add_slot(t, L, 2)) return -1;
        kite_ctx *c = ctx(t);
        kite_num *x = num(t, kX);
        if (t.vx.empty()) {
            kite_num *hi = kite_num_new(c), *cP = kite_num_new(c), *c2 = kite_num_new(c), *mid = kite_num_new(c);
            if (!hi || !cP || !c2 || !mid) {
                for (kite_num *y : {hi, cP, c2, mid}) kite_num_free(c, y);
                return -1;
            }
            kite_reset_roundoff(c);
            kite_set_u64(c, cP, t.P);
            kite_set_u64(c, c2, 2);
            llr_lucas_vk(c, t.sp.k, x, hi, cP, c2, mid);
            double e = kite_max_roundoff(c);
            t.vx.assign(c->nlimbs, 0);
            kite_get_limbs(c, x, t.vx.data());
            for (kite_num *y : {hi, cP, c2, mid}) kite_num_free(c, y);
            if (!engine_ok(t)) return -1;
            t.res.maxerr = std::max(t.res.maxerr, e);
            if (t.ladder_roundoff_pending && t.run == 0) {
                t.ladder_roundoff_pending = false;
                e = INFINITY;
            }
            if (e > maxround) {
                t.vx.clear();
                return 1;
            }
            t.vi = 0;
        } else {
            kite_set_limbs(c, x, t.vx.data());
        }
        kite_set_limbs(c, num(t, kV), t.vx.data());
        t.i = t.vi;
        t.phase = t.vi == 0 ? kLlrFirst : kLlrSteps;
        kite_batch_set_xadd(b, t.slot, t.phase == kLlrFirst ? 2 : -2);
        t.copy_pending = false;
        t.fault_pending = t.sp.fault_at && ((t.sp.fault_runs >> t.run) & 1) && t.sp.fault_at > t.vi;
        return engine_ok(t) ? 0 : -1;
    }

    void save_llr(Job &t) {
        if (t.sp.ckpt.empty()) return;
        LlrCheckpoint c;
        c.iter = t.vi;
        c.P = t.P;
        c.L = t.L;
        c.run = t.run;
        c.prev[0] = t.lres[0];
        c.prev[1] = t.lres[1];
        llr_checkpoint_write(t.sp.ckpt.c_str(), t.sp.k, t.sp.n, c, t.vx.data(), t.vx.size());
        t.last_ck = now_s();
    }

    // x after a good block (the slot's x, once the batch is synchronised) becomes the verified state
    void llr_take_state(Job &t) {
        t.vx.resize(ctx(t)->nlimbs);
        kite_get_limbs(ctx(t), num(t, kX), t.vx.data());
    }

    // --- starting, retrying, finishing ---------------------------------------------------------

    // Starts (or restarts) job t at length L; 0 on success, else the job's failure status.
    int start(Job &t, uint32_t L) {
        if (t.sp.kind == KITE_JOB_PRP) return start_fermat(t, L) ? -1 : 0;
        for (;;) {
            const int r = start_llr(t, L);
            if (r < 0) return -1;
            if (r == 0) return 0;
            // the ladder's round-off was over the limit: a longer FFT
            const uint32_t L2 = kite_next_fft_len(t.L + t.L / 25);
            drop_slot(t);
            t.res.roundoff_retries++;
            if (!L2 || t.res.roundoff_retries > 8) return -2;
            L = L2;
        }
    }

    // Restarts job t at a longer FFT after a round-off failure: 0, -2 if no longer length is left,
    // -1 if the restart failed.
    int retry_longer(Job &t) {
        if (t.sp.kind == KITE_JOB_LLR && t.vi > 0) {  // the verified copy holds step vi
            t.vx.resize(ctx(t)->nlimbs);
            kite_get_limbs(ctx(t), num(t, kV), t.vx.data());
            save_llr(t);
        }
        if (t.sp.kind == KITE_JOB_PRP) {
            if (t.host_stale) pull(t, kVX, kVD);  // the verified state moves to the new length by limbs
            t.ui = 0;  // the unverified state is abandoned with the failed attempt
            if (!t.sp.uckpt.empty()) remove(t.sp.uckpt.c_str());
        }
        const uint32_t L2 = kite_next_fft_len(t.L + t.L / 25);
        drop_slot(t);
        t.res.roundoff_retries++;
        if (!L2 || t.res.roundoff_retries > 8) return -2;
        return start(t, L2);
    }

    // the job's position into its result record
    static void note_progress(Job &t) {
        t.res.run = (uint32_t)t.run;
        t.res.iter = t.i;
        t.res.verified = t.vi;
    }

    void done(Job &t, int rc) {
        note_progress(t);
        const bool aborted = rc == KITE_PRP_ABORTED;
        if (t.sp.drop_ckpt_when_done && !aborted && !t.sp.ckpt.empty()) remove(t.sp.ckpt.c_str());
        drop_slot(t);
        t.phase = kDone;
        t.res.status = rc;
        t.res.fft_len = t.L;
        t.res.finish = now_s() - t0;
        bytes_in_flight -= t.bytes;
        if (on_done) on_done(done_arg, t.index, &t.res);
    }

    void finish_fermat(Job &t, std::vector<uint64_t> xn, int rc) {
        if (rc == 0) {
            const uint64_t bb = (uint64_t)t.sp.base * t.sp.base;
            bool prp = xn[0] == bb;
            for (size_t j = 1; j < xn.size(); j++) prp &= xn[j] == 0;
            if (kite_residue_nm1(xn.data(), ctx(t)->Nl.data(), xn.size(), t.sp.base)) rc = -4;
            t.res.res64 = xn[0];
            t.res.is_prime = prp;
        }
        done(t, rc);
    }

    // The end of an LLR run: record it and decide (prime, confirmed, vote) or start the next run.
    void llr_run_done(Job &t) {
        std::vector<uint64_t> out(ctx(t)->nlimbs);
        kite_get_limbs(ctx(t), num(t, kX), out.data());
        if (!engine_ok(t)) {
            done(t, -1);
            return;
        }
        const int r = t.run;
        t.lres[r] = out[0];
        t.lzero[r] = all_zero(out);
        auto same = [&](int a, int c) { return t.lres[a] == t.lres[c] && t.lzero[a] == t.lzero[c]; };
        int answer = -1;  // the run whose result stands
        if (r == 0 && t.lzero[0]) answer = 0;
        else if (r == 1 && same(1, 0)) answer = 1;
        else if (r == 2) answer = same(2, 1) ? 1 : same(2, 0) ? 0 : -2;
        if (answer == -2) {
            done(t, -3);
            return;
        }
        if (answer >= 0) {
            t.res.is_prime = t.lzero[answer];
            t.res.res64 = t.lres[answer];
            done(t, 0);
            return;
        }
        // a composite answer is confirmed at the next FFT length (src/llr.c:350-360)
        const uint32_t L2 = kite_next_fft_len(t.L + 1);
        drop_slot(t);
        t.run = r + 1;
        t.res.confirm_runs = (uint32_t)t.run;
        if (t.sp.on_confirm) t.sp.on_confirm(t.sp.progress_arg, t.run);
        t.vx.clear();
        t.vi = 0;
        if (!L2 || start(t, L2)) done(t, -2);
    }

    // --- admission -----------------------------------------------------------------------------

    void admit() {
        const double now = now_s() - t0;
        while (!waiting.empty() && active.size() < max_slots) {
            auto pick = waiting.begin();
            if (order == KITE_ORDER_LEAST_WORK) pick = least_work(now);
            Job &t = jobs[*pick];
            if (t.sp.arrival > now) break;
            const uint32_t Lest = t.sp.fft_len ? t.sp.fft_len : kite_choose_fft_len(t.sp.k, t.sp.n);
            const uint64_t need = kite_job_bytes(t.sp.kind, Lest);
            if (hbm_budget && !active.empty() && bytes_in_flight + need > hbm_budget) break;
            waiting.erase(pick);
            t.started = true;
            t.res.start = now;
            t.bytes = need;
            bytes_in_flight += need;
            t.L = t.sp.fft_len;
            if (t.sp.kind == KITE_JOB_LLR) {
                if (t.sp.n < 3 || !(t.sp.k & 1) || kite_llr_choose_P(t.sp.k, t.sp.n, &t.P, &t.res.factor)) {
                    done(t, -1);
                    continue;
                }
                t.res.P = t.P;
                if (t.res.factor) {
                    done(t, 0);
                    continue;
                }
            }
            resume(t);
            if (t.sp.kind == KITE_JOB_LLR && t.run > 0 && t.sp.on_confirm) t.sp.on_confirm(t.sp.progress_arg, t.run);
            t.last_ck = t.last_u = now_s();
            const int r = start(t, t.L);
            t.res.start = now_s() - t0;  // in the rounds from now on
            if (r) {
                done(t, r);
                continue;
            }
            if (t.sp.kind == KITE_JOB_PRP && t.sp.on_point && t.vi == 0 && t.ui == 0 && t.vp < t.sp.npoints &&
                t.sp.points[t.vp] == 0) {
                // point 0 is x_0 itself, reported at once
                t.pi = t.vp;
                capture_point(t);
                t.vp = t.pi;
                if (!flush_points(t)) {
                    done(t, KITE_PRP_POINT_FAILED);
                    continue;
                }
            }
            active.push_back(t.index);
        }
    }

    // Among the arrived jobs, the one with the least work: squarings times FFT length (n L, L from
    // the length chooser), twice for an LLR test, whose composite answers run again. A job that
    // has waited longer than the horizon's share (or 60 s) goes first, so long tests do not starve.
    std::deque<size_t>::iterator least_work(double now) {
        auto pick = waiting.begin();
        double best = 0;
        const double max_wait = 60;
        for (auto it = waiting.begin(); it != waiting.end() && jobs[*it].sp.arrival <= now; ++it) {
            const Spec &s = jobs[*it].sp;
            if (now - s.arrival > max_wait) return it;  // arrival order for jobs that waited too long
            const uint32_t L = s.fft_len ? s.fft_len : kite_choose_fft_len(s.k, s.n);
            const double w = (double)s.n * (double)L * (s.kind == KITE_JOB_LLR ? 2.0 : 1.0);
            if (it == waiting.begin() || w < best) {
                best = w;
                pick = it;
            }
        }
        return pick;
    }

    // --- one round -----------------------------------------------------------------------------

    // The squaring span of job t this round (0: nothing to square), and the numbers involved.
    uint32_t plan(Job &t, int32_t *which, int32_t *cdst, int32_t *csrc, int32_t *mdst, int32_t *msrc) {
        const uint64_t n = t.sp.n;
        uint64_t cnt = 0;
        switch (t.phase) {
        case kBlocks:
            cnt = std::min<uint64_t>(t.sp.B - t.blk, next_event(t) - t.i);
            *which = kX;
            if (t.blk + cnt == t.sp.B) {  // the block ends: t = d, d *= x
                *cdst = kT;
                *csrc = kD;
                *mdst = kD;
                *msrc = kX;
            }
            break;
        case kCheck:
            cnt = t.sp.B;
            *which = kT;
            *mdst = kT;
            *msrc = kX0;
            break;
        case kTailX:
            cnt = std::min<uint64_t>(n - t.i, next_event(t) - t.i);
            *which = kX;
            break;
        case kTailT:
            cnt = n - t.full;
            *which = kT;
            break;
        case kLlrFirst:
            cnt = 1;
            *which = kX;
            break;
        case kLlrSteps: {
            uint64_t stop = std::min<uint64_t>(t.i + t.sp.B, n - 1);
            if (t.fault_pending && t.sp.fault_at > t.i) stop = std::min(stop, t.sp.fault_at);
            cnt = stop - t.i;
            *which = kX;
            break;
        }
        default:
            break;
        }
        return (uint32_t)cnt;
    }

    // After the squarings: the injected fault and the proof point at the job's new position.
    void events(Job &t, uint32_t cnt) {
        const uint64_t at = t.i + cnt;
        if (t.fault_pending && at == t.sp.fault_at &&
            (t.phase == kBlocks || t.phase == kLlrSteps || t.phase == kLlrFirst)) {
            kite_debug_corrupt(ctx(t), num(t, kX));
            t.fault_pending = false;
        }
        if ((t.phase == kBlocks || t.phase == kTailX) && t.sp.on_point && t.pi < t.sp.npoints &&
            t.sp.points[t.pi] == at)
            capture_point(t);
    }

    bool aborting(const Job &t) const { return t.sp.abort_flag && *t.sp.abort_flag; }

    // The Fermat state update after a round; false when the job has ended.
    // Residues a and c of job t hold the same value: the device's digit comparison (eq = 1), or,
    // when the digits differ (eq = 0: equal values can still differ in a digit at a chunk's end),
    // the exact comparison of their limbs.
    bool same_value(Job &t, int a, int c, int eq) {
        if (eq == 1) return true;
        std::vector<uint64_t> la(ctx(t)->nlimbs), lc(ctx(t)->nlimbs);
        kite_get_limbs(ctx(t), num(t, a), la.data());
        kite_get_limbs(ctx(t), num(t, c), lc.data());
        return la == lc;
    }

    bool step_fermat(Job &t, uint32_t cnt, double e, int eq) {
        switch (t.phase) {
        case kBlocks: {
            t.i += cnt;
            t.blk += cnt;
            if (t.blk < t.sp.B) return true;  // the block goes on past a point or the fault
            t.blk = 0;
            if (t.sp.progress) t.sp.progress(t.sp.progress_arg, t.i, t.sp.n, e);
            if (aborting(t)) t.stopping = true;
            if (t.i >= t.check_at || t.stopping) {
                t.phase = kCheck;
            } else if (!t.sp.uckpt.empty() && now_s() - t.last_u > t.sp.u_secs) {
                // the unverified state (and the points so far) between checks
                if (!t.pending.empty() && !flush_points(t)) {
                    done(t, KITE_PRP_POINT_FAILED);
                    return false;
                }
                std::vector<uint64_t> lx(ctx(t)->nlimbs), ld(ctx(t)->nlimbs);
                kite_get_limbs(ctx(t), num(t, kX), lx.data());
                kite_get_limbs(ctx(t), num(t, kD), ld.data());
                prp_checkpoint_write(t.sp.uckpt.c_str(), t.sp.k, t.sp.n, t.sp.base, t.i, t.L, lx.data(), ld.data(),
                                     lx.size());
                t.last_u = now_s();
            }
            return true;
        }
        case kCheck: {
            if (!same_value(t, kT, kD, eq)) {  // Gerbicz failure: back to the verified state
                t.res.gerbicz_errors++;
                if (!t.sp.uckpt.empty()) remove(t.sp.uckpt.c_str());
                if (++t.fails > 5) {
                    done(t, -3);
                    return false;
                }
                t.ui = 0;
                roll_back(t);
                if (t.stopping) {
                    if (t.host_stale) pull(t, kVX, kVD);
                    save_fermat(t);
                    done(t, KITE_PRP_ABORTED);
                    return false;
                }
                return true;
            }
            if (!engine_ok(t)) {
                done(t, -1);
                return false;
            }
            // verified: (x, d) become the state a later failure rolls back to
            t.fails = 0;
            t.ui = 0;
            t.vi = t.i;
            t.vp = t.pi;
            t.copies.push_back({kVX, kX});
            t.copies.push_back({kVD, kD});
            t.host_stale = true;
            if (!t.pending.empty() && !flush_points(t)) {
                done(t, KITE_PRP_POINT_FAILED);
                return false;
            }
            t.phase = kBlocks;
            t.check_at = std::min(t.vi + (uint64_t)t.sp.B * t.sp.M, t.full);
            const bool at_full = t.vi >= t.full;
            if (now_s() - t.last_ck >= t.sp.ck_secs || t.stopping || (t.sp.save_at_end && at_full) ||
                (at_full && t.full == t.sp.n)) {
                pull(t, kX, kD);  // x and d are the verified state now
                if (now_s() - t.last_ck >= t.sp.ck_secs || t.stopping || (t.sp.save_at_end && at_full))
                    save_fermat(t);
            }
            if (t.stopping) {
                done(t, KITE_PRP_ABORTED);
                return false;
            }
            if (at_full) {
                if (t.full == t.sp.n) {
                    finish_fermat(t, t.vx, 0);
                    return false;
                }
                t.copies.push_back({kT, kX});  // the second tail starts from the same x
                begin_tail(t);
            }
            return true;
        }
        case kTailX:
            t.i += cnt;
            if (t.i >= t.sp.n) t.phase = kTailT;
            return true;
        case kTailT: {
            const bool same = same_value(t, kX, kT, eq);
            if (!engine_ok(t)) {
                done(t, -1);
                return false;
            }
            if (same) {
                if (!t.pending.empty() && !flush_points(t)) {
                    done(t, KITE_PRP_POINT_FAILED);
                    return false;
                }
                std::vector<uint64_t> lx(ctx(t)->nlimbs);
                kite_get_limbs(ctx(t), num(t, kX), lx.data());
                finish_fermat(t, lx, 0);
                return false;
            }
            if (++t.tail_tries > 6) {
                std::vector<uint64_t> lx(ctx(t)->nlimbs);
                kite_get_limbs(ctx(t), num(t, kX), lx.data());
                finish_fermat(t, lx, -3);
                return false;
            }
            t.res.gerbicz_errors++;
            t.copies.push_back({kX, kVX});  // both tails again from the verified x at `full`
            t.copies.push_back({kT, kVX});
            begin_tail(t);
            return true;
        }
        default:
            return true;
        }
    }

    // The LLR state update after a round; false when the job has ended.
    bool step_llr(Job &t, uint32_t cnt, double e) {
        t.i += cnt;
        if (t.phase == kLlrFirst) {
            t.phase = kLlrSteps;
            kite_batch_set_xadd(b, t.slot, -2);
        }
        t.vi = t.i;
        t.copy_pending = true;
        if (t.sp.progress) t.sp.progress(t.sp.progress_arg, t.vi, t.sp.n - 1, e);
        if (t.i >= t.sp.n - 1) {
            llr_run_done(t);
            return t.phase != kDone;
        }
        const bool stop = aborting(t);
        if (!t.sp.ckpt.empty() && (stop || now_s() - t.last_ck >= t.sp.ck_secs)) {
            llr_take_state(t);
            save_llr(t);
        }
        if (stop) {
            done(t, KITE_LLR_ABORTED);
            return false;
        }
        return true;
    }

    // Runs rounds until every job has finished (or the horizon); returns 0 or the first failure.
    void run() {
        admit();
        for (;;) {
            const double now = now_s() - t0;
            if (horizon > 0 && now >= horizon) break;
            if (active.empty()) {
                if (waiting.empty()) break;
                double wait = jobs[waiting.front()].sp.arrival - now;
                if (horizon > 0) wait = std::min(wait, horizon - now);
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.05)));
                admit();
                continue;
            }
            if (!round()) {
                engine_failed = true;
                for (size_t ti : active) done(jobs[ti], -1);
                active.clear();
                break;
            }
            admit();
        }
    }

    bool round() {
        const int cap = kite_batch_capacity(b);
        std::vector<int32_t> which(cap, -1), cdst(cap, -1), csrc(cap, -1), mdst(cap, -1), msrc(cap, -1);
        std::vector<int32_t> vdst(cap, -1), vsrc(cap, -1);
        std::vector<uint32_t> cnt(cap, 0);
        // LLR jobs whose last block passed: their verified copy before the next block
        for (size_t ti : active) {
            Job &t = jobs[ti];
            if (t.sp.kind == KITE_JOB_LLR && t.copy_pending) {
                vdst[t.slot] = kV;
                vsrc[t.slot] = kX;
                t.copy_pending = false;
            }
        }
        if (kite_batch_copy(b, vdst.data(), vsrc.data())) return false;
        // which jobs take part: all of them, or (time share) those whose credit has reached a round
        std::vector<char> due(cap, 1);
        if (share == KITE_SHARE_TIME) {
            uint32_t Lmin = ~0u;
            for (size_t ti : active) Lmin = std::min(Lmin, jobs[ti].L);
            for (size_t ti : active) {
                Job &t = jobs[ti];
                t.credit += (double)Lmin / (double)t.L;
                due[t.slot] = t.credit >= 1.0 - 1e-9;
                if (due[t.slot]) t.credit -= 1.0;
            }
        }
        for (size_t ti : active) {
            Job &t = jobs[ti];
            if (due[t.slot])
                cnt[t.slot] = plan(t, &which[t.slot], &cdst[t.slot], &csrc[t.slot], &mdst[t.slot], &msrc[t.slot]);
        }
        if (kite_batch_square(b, which.data(), cnt.data())) return false;
        bool any_event = false;
        for (size_t ti : active) {
            const Job &t = jobs[ti];
            if (!due[t.slot]) continue;
            const uint64_t at = t.i + cnt[t.slot];
            any_event |= (t.fault_pending && at == t.sp.fault_at) ||
                         (t.sp.on_point && t.pi < t.sp.npoints && t.sp.points[t.pi] == at);
        }
        if (any_event) {
            if (kite_batch_sync(b)) return false;
            for (size_t ti : active)
                if (due[jobs[ti].slot]) events(jobs[ti], cnt[jobs[ti].slot]);
        }
        if (kite_batch_copy(b, cdst.data(), csrc.data()) || kite_batch_mul(b, mdst.data(), msrc.data())) return false;
        std::vector<double> err(cap, 0);
        if (kite_batch_maxerr(b, err.data(), 1)) return false;
        // the checks of this round, compared on the device: t against d after a check's squarings,
        // the two tails against each other
        std::vector<int32_t> ea(cap, -1), ec(cap, -1), eq(cap, -1);
        bool compare = false;
        for (size_t ti : active) {
            const Job &t = jobs[ti];
            if (!due[t.slot] || err[t.slot] > maxround || t.sp.kind != KITE_JOB_PRP) continue;
            if (t.phase == kCheck) {
                ea[t.slot] = kT;
                ec[t.slot] = kD;
            } else if (t.phase == kTailT) {
                ea[t.slot] = kX;
                ec[t.slot] = kT;
            }
            compare |= ea[t.slot] >= 0;
        }
        if (compare && kite_batch_equal(b, ea.data(), ec.data(), eq.data())) return false;
        std::vector<size_t> still;
        for (size_t ti : active) {
            Job &t = jobs[ti];
            if (!due[t.slot]) {  // sat this round out
                still.push_back(ti);
                continue;
            }
            double e = err[t.slot];
            t.res.maxerr = std::max(t.res.maxerr, e);
            if (t.roundoff_pending && t.run == 0 && t.phase == kLlrSteps && t.i + cnt[t.slot] >= t.sp.roundoff_at) {
                t.roundoff_pending = false;
                e = INFINITY;
            }
            if (!engine_ok(t)) {
                done(t, -1);
                continue;
            }
            if (e > maxround) {  // redo from the verified state at a longer FFT
                const int r = retry_longer(t);
                if (r) {
                    done(t, r);
                    continue;
                }
                still.push_back(ti);
                continue;
            }
            const bool going = t.sp.kind == KITE_JOB_PRP ? step_fermat(t, cnt[t.slot], e, eq[t.slot])
                                                         : step_llr(t, cnt[t.slot], e);
            if (going) still.push_back(ti);
        }
        active.swap(still);
        return run_copies();
    }

    // The device copies the jobs' updates asked for, in order, one per job and call.
    bool run_copies() {
        const int cap = kite_batch_capacity(b);
        for (size_t k = 0;; k++) {
            std::vector<int32_t> dst(cap, -1), src(cap, -1);
            bool any = false;
            for (size_t ti : active) {
                Job &t = jobs[ti];
                if (k < t.copies.size()) {
                    dst[t.slot] = t.copies[k].first;
                    src[t.slot] = t.copies[k].second;
                    any = true;
                }
            }
            if (!any) break;
            if (kite_batch_copy(b, dst.data(), src.data())) return false;
        }
        for (size_t ti : active) jobs[ti].copies.clear();
        return true;
    }

    // Jobs stopped by the horizon keep their state on file.
    void save_in_flight() {
        if (engine_failed || kite_batch_sync(b)) return;
        for (size_t ti : active) {
            Job &t = jobs[ti];
            if (t.phase == kDone || t.slot < 0) continue;
            if (t.sp.kind == KITE_JOB_LLR) {
                if (t.copy_pending) llr_take_state(t);  // x is the state of the last good block
                else {
                    t.vx.resize(ctx(t)->nlimbs);
                    kite_get_limbs(ctx(t), num(t, kV), t.vx.data());
                }
                save_llr(t);
            } else if (t.vi > 0) {
                if (t.host_stale) pull(t, kVX, kVD);
                save_fermat(t);
            }
        }
    }
};

bool writable_dir(const char *dir) {
    struct stat st;
    return stat(dir, &st) == 0 && S_ISDIR(st.st_mode) && access(dir, W_OK | X_OK) == 0;
}

}  // namespace

extern "C" uint64_t kite_job_bytes(uint32_t kind, uint32_t fft_len) {
    // residues (8 L each: 4 for Fermat, x and its verified copy for LLR; the LLR ladder briefly
    // holds 4 more) plus the per-test tables and scratch (operand spectra 8 L, digits 4 L, codes)
    const uint64_t L = fft_len;
    return (kind == KITE_JOB_LLR ? 6 : 4) * 8 * L + 13 * L + (1u << 16);
}

extern "C" int kite_batch_run(const kite_job *jobs, size_t njobs, const kite_run_opts *opts, kite_job_result *res) {
    const kite_run_opts o = opts ? *opts : kite_run_opts{};
    const char *dir = o.checkpoint_dir && o.checkpoint_dir[0] ? o.checkpoint_dir : nullptr;
    if (dir && !writable_dir(dir)) {  // refused before any work
        for (size_t i = 0; i < njobs; i++) {
            res[i] = kite_job_result{};
            res[i].status = -1;
        }
        return -1;
    }
    Scheduler S;
    S.b = kite_batch_new();
    if (!S.b) return -1;
    if (o.parts) kite_batch_set_parts(S.b, (int)o.parts);
    if (o.shape_parts) kite_batch_set_shape_parts(S.b, 1);
    if (o.xcds) kite_batch_set_xcds(S.b, (int)o.xcds);
    // a failure here is latched in the batch: the first round then fails and the run reports it
    if (o.launch == KITE_LAUNCH_DIRECT) (void)kite_batch_set_graphs(S.b, 0);
    S.max_slots = o.slots ? o.slots : (size_t)-1;
    S.hbm_budget = o.hbm_budget;
    S.order = o.order;
    S.share = o.share;
    S.maxround = o.max_roundoff > 0 ? o.max_roundoff : 0.40;
    S.horizon = o.horizon;
    S.on_done = o.on_done;
    S.done_arg = o.arg;
    const uint32_t B = o.block ? o.block : 1000;
    const uint32_t M = o.gerbicz_blocks ? o.gerbicz_blocks : default_blocks(B);
    S.jobs.resize(njobs);
    std::vector<size_t> order(njobs);
    for (size_t i = 0; i < njobs; i++) {
        Job &t = S.jobs[i];
        Spec &s = t.sp;
        t.index = i;
        s.k = jobs[i].k;
        s.n = jobs[i].n;
        s.kind = jobs[i].kind;
        s.base = jobs[i].base ? jobs[i].base : 3;
        s.fft_len = jobs[i].fft_len;
        s.arrival = jobs[i].arrival;
        s.B = B;
        s.M = M;
        if (dir) s.ckpt = std::string(dir) + "/" + ckpt_name(s);
        s.ck_secs = o.checkpoint_secs > 0 ? o.checkpoint_secs : 600;
        if (o.inject_fault_at && o.inject_fault_job == i) s.fault_at = o.inject_fault_at;  // LLR: run 0
        t.fault_pending = s.fault_at > 0;
        t.res.status = KITE_JOB_PENDING;
        order[i] = i;
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t c) { return S.jobs[a].sp.arrival < S.jobs[c].sp.arrival; });
    S.waiting.assign(order.begin(), order.end());
    S.t0 = now_s();
    S.run();
    if (dir) S.save_in_flight();
    int rc = 0;
    for (size_t i = 0; i < njobs; i++) {
        Job &t = S.jobs[i];
        if (t.phase != kDone || !t.started) {  // stopped by the horizon
            t.res.status = KITE_JOB_PENDING;
            if (t.started) {
                t.res.fft_len = t.L;
                Scheduler::note_progress(t);
            }
        }
        res[i] = t.res;
        if (t.res.status && !rc) rc = t.res.status;
    }
    kite_batch_free(S.b);
    return rc;
}

// ---- kite_batch_prp: Fermat jobs that all arrive at once ------------------------------------

namespace {

struct PrpAdapter {
    const kite_batch_prp_opts *o;
};

kite_prp_result to_prp_result(const kite_job_result &r) {
    kite_prp_result p{};
    p.res64 = r.res64;
    p.is_prp = r.is_prime;
    p.fft_len = r.fft_len;
    p.maxerr = r.maxerr;
    p.gerbicz_errors = r.gerbicz_errors;
    p.roundoff_retries = r.roundoff_retries;
    p.resumed = (int)r.resumed;
    p.resumed_at = r.resumed_at;
    p.seconds = r.finish - r.start;
    return p;
}

void prp_done(void *arg, size_t index, const kite_job_result *r) {
    const kite_batch_prp_opts *o = static_cast<PrpAdapter *>(arg)->o;
    const kite_prp_result pr = to_prp_result(*r);
    o->on_done(o->arg, index, r->status, &pr);
}

}  // namespace

extern "C" int kite_batch_prp(const kite_batch_test *tests, size_t ntests, const kite_batch_prp_opts *opts,
                              kite_prp_result *res, int *status) {
    const kite_batch_prp_opts o = opts ? *opts : kite_batch_prp_opts{};
    std::vector<kite_job> jobs(ntests);
    for (size_t i = 0; i < ntests; i++) {
        jobs[i] = kite_job{};
        jobs[i].k = tests[i].k;
        jobs[i].n = tests[i].n;
        jobs[i].kind = KITE_JOB_PRP;
        jobs[i].base = tests[i].base;
        jobs[i].fft_len = tests[i].fft_len;
    }
    PrpAdapter ad{&o};
    kite_run_opts ro{};
    ro.slots = o.slots;
    ro.block = o.gerbicz_block;
    ro.gerbicz_blocks = o.gerbicz_blocks;
    ro.max_roundoff = o.max_roundoff;
    ro.inject_fault_job = o.inject_fault_test;
    ro.inject_fault_at = o.inject_fault_at;
    if (o.on_done) {
        ro.on_done = prp_done;
        ro.arg = &ad;
    }
    std::vector<kite_job_result> jr(ntests);
    const int rc = kite_batch_run(jobs.data(), ntests, &ro, jr.data());
    for (size_t i = 0; i < ntests; i++) {
        res[i] = to_prp_result(jr[i]);
        if (status) status[i] = jr[i].status == KITE_JOB_PENDING ? -1 : jr[i].status;
    }
    return rc;
}

// ---- the reference's single-test drivers: one job through the same scheduler ---------------

namespace {

// One job alone; its result record and the scheduler's return value.
int run_one(Job &&job, double maxround, kite_job_result *out) {
    Scheduler S;
    S.b = kite_batch_new();
    if (!S.b) return -1;
    S.maxround = maxround;
    job.index = 0;
    job.res.status = KITE_JOB_PENDING;
    job.fault_pending = job.sp.fault_at > 0;
    job.roundoff_pending = job.sp.roundoff_at > 0;
    job.ladder_roundoff_pending = job.sp.roundoff_ladder;
    S.jobs.push_back(std::move(job));
    S.waiting.push_back(0);
    S.t0 = now_s();
    S.run();
    *out = S.jobs[0].res;
    kite_batch_free(S.b);
    return out->status == KITE_JOB_PENDING ? -1 : out->status;
}

}  // namespace

extern "C" int kite_prp(uint64_t k, uint64_t n, uint32_t base, const kite_prp_opts *opts, kite_prp_result *res) {
    const kite_prp_opts o = opts ? *opts : kite_prp_opts{};
    memset(res, 0, sizeof *res);
    const double t_start = now_s();
    if (o.fft.backend != KITE_BACKEND_AUTO && o.fft.backend != KITE_BACKEND_HIP) return -1;
    Job t;
    Spec &s = t.sp;
    s.k = k;
    s.n = n;
    s.kind = KITE_JOB_PRP;
    s.base = base;
    s.fft_len = o.fft.fft_len;
    s.plan_order = o.fft.plan_order;
    s.B = o.gerbicz_block ? o.gerbicz_block : default_block(n);
    s.M = o.gerbicz_blocks ? o.gerbicz_blocks : default_blocks(s.B);
    if (o.checkpoint) s.ckpt = o.checkpoint;
    if (o.checkpoint_unverified) s.uckpt = o.checkpoint_unverified;
    s.ck_secs = o.checkpoint_secs > 0 ? o.checkpoint_secs : 600;
    s.u_secs = o.unverified_secs > 0 ? o.unverified_secs : 60;
    s.save_at_end = true;
    s.drop_ckpt_when_done = false;
    s.product_any_block = true;
    s.progress = o.progress;
    s.progress_arg = o.progress_arg;
    if (o.on_point && o.points) {
        s.points = o.points;
        s.npoints = o.npoints;
        s.on_point = o.on_point;
        s.point_arg = o.point_arg;
    }
    s.abort_flag = o.abort_flag;
    s.fault_at = o.inject_fault_at;
    kite_job_result jr{};
    const int rc = run_one(std::move(t), o.max_roundoff > 0 ? o.max_roundoff : 0.40, &jr);
    *res = to_prp_result(jr);
    res->seconds = now_s() - t_start;
    return rc;
}

extern "C" int kite_llr(uint64_t k, uint64_t n, const kite_llr_opts *opts, kite_llr_result *res) {
    const kite_llr_opts o = opts ? *opts : kite_llr_opts{};
    memset(res, 0, sizeof *res);
    const double t_start = now_s();
    if (n < 3 || !(k & 1)) return -1;
    if (o.fft.backend != KITE_BACKEND_AUTO && o.fft.backend != KITE_BACKEND_HIP) return -1;
    Job t;
    Spec &s = t.sp;
    s.k = k;
    s.n = n;
    s.kind = KITE_JOB_LLR;
    s.fft_len = o.fft.fft_len;
    s.plan_order = o.fft.plan_order;
    s.B = 1000;  // doublings between round-off reads (src/llr.c:245)
    if (o.checkpoint) s.ckpt = o.checkpoint;
    s.ck_secs = o.checkpoint_secs > 0 ? o.checkpoint_secs : 600;
    s.drop_ckpt_when_done = true;  // kept only when the test is aborted
    s.progress = o.progress;
    s.progress_arg = o.progress_arg;
    s.on_confirm = o.on_confirm;
    s.abort_flag = o.abort_flag;
    s.fault_at = o.inject_fault_at;
    s.fault_runs = o.inject_fault_runs ? o.inject_fault_runs : 1;
    s.roundoff_at = o.inject_roundoff_at;
    s.roundoff_ladder = o.inject_roundoff_ladder != 0;
    kite_job_result jr{};
    const int rc = run_one(std::move(t), o.max_roundoff > 0 ? o.max_roundoff : 0.40, &jr);
    res->is_prime = jr.is_prime;
    res->factor = jr.factor;
    res->res64 = jr.res64;
    res->P = jr.P;
    res->fft_len = jr.fft_len;
    res->maxerr = jr.maxerr;
    res->roundoff_retries = jr.roundoff_retries;
    res->confirm_runs = jr.confirm_runs;
    res->resumed = (int)jr.resumed;
    res->resumed_at = jr.resumed_at;
    res->seconds = now_s() - t_start;
    return rc;
}
