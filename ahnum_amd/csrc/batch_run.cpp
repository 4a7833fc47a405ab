// Batched Fermat PRP and LLR tests: one scheduler for many independent jobs on one GPU.
//
// Every round advances each job in flight by up to B squarings through shared launches
// (kite_batch_square), then runs its checks. A Fermat job keeps the reference driver's state
// machine (src/prp.c:329-583): blocks with d_{t+1} = d_t * x, a check d == x0 * t^(2^B) every M
// blocks with rollback to the last verified (i, x, d), a round-off retry at a >= 4% longer FFT, and
// the doubled tail. An LLR job keeps kite_llr's (src/llr.c:192-384): V_k by the Lucas ladder,
// then n - 1 doublings x -> x^2 - 2 (the first + 2) done by the squaring kernels themselves
// (kite_batch_set_xadd), a round-off check after every block with a device copy of the last
// verified state, and a composite answer confirmed at the next FFT length (a third run breaks a
// tie). Jobs start once they have arrived and the slot and HBM limits allow, by arrival or least
// work first; a finished job's
// slot goes to the next one at once (continuous batching). With a checkpoint directory a job
// resumes from its verified state on file and writes it back as it goes (src/prp.c:156-212 and
// src/llr.c:92-146 for the two file formats).
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

enum Phase { kBlocks, kCheck, kTailX, kTailT, kLlrFirst, kLlrSteps, kDone };
enum FermatNum { kX = 0, kD = 1, kX0 = 2, kT = 3 };
enum LlrNum { kV = 1 };  // LLR: x = 0 and its last verified copy

double now_s() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

struct Job {
    size_t index = 0;
    uint64_t k = 0, n = 0;
    uint32_t kind = KITE_JOB_PRP, base = 3, L = 0;
    double arrival = 0;
    int slot = -1;
    Phase phase = kBlocks;
    uint64_t bytes = 0;  // admission estimate while in flight
    // Fermat
    uint64_t full = 0, i = 0, vi = 0, check_at = 0;
    std::vector<uint64_t> vx, vd;
    // start of the Gerbicz product (x0 = base^k at iteration 0, or a verified x restored from a
    // checkpoint at iteration x0_at); blocks of B squarings count from there
    std::vector<uint64_t> x0;
    uint64_t x0_at = 0;
    uint32_t consecutive_fail = 0, tail_attempt = 0;
    // LLR
    uint32_t P = 0;
    int run = 0;
    uint64_t lres[3] = {0, 0, 0};
    int lzero[3] = {0, 0, 0};
    bool copy_pending = false;  // x passed its block: copy it to the verified slot next round
    bool fault_pending = false;
    kite_job_result res{};
    int rc = 0;
    bool started = false;
    double last_ck = 0;  // when its checkpoint was last written
};

void power_start(kite_ctx *ctx, kite_num *x0, uint64_t k, uint32_t base) {
    kite_set_u64(ctx, x0, 1);
    for (int b = 63; b >= 0; b--) {
        if (k >> b == 0) continue;
        kite_square(ctx, x0);
        if ((k >> b) & 1) kite_mul_small(ctx, x0, base);
    }
    if ((double)k * std::log2((double)base) + 2 < (double)ctx->n) {  // exact power below N
        std::vector<uint64_t> l(ctx->nlimbs, 0);
        l[0] = 1;
        size_t used = 1;
        for (uint64_t i = 0; i < k; i++) {
            unsigned __int128 cy = 0;
            for (size_t j = 0; j < used; j++) {
                cy += (unsigned __int128)l[j] * base;
                l[j] = (uint64_t)cy;
                cy >>= 64;
            }
            if (cy) l[used++] = (uint64_t)cy;
        }
        kite_set_limbs(ctx, x0, l.data());
    }
}

// (Re)starts Fermat job t in a fresh slot at length L from its verified state (or x0 if none).
int start_fermat(kite_batch *b, Job &t, uint32_t L, uint32_t B, uint32_t M) {
    t.slot = kite_batch_add(b, t.k, t.n, L, 4);
    if (t.slot < 0) return -1;
    kite_ctx *ctx = kite_batch_ctx(b, t.slot);
    t.L = kite_fft_len(ctx);
    kite_num *x0 = kite_batch_num(b, t.slot, kX0);
    if (t.x0.empty()) power_start(ctx, x0, t.k, t.base);
    else kite_set_limbs(ctx, x0, t.x0.data());
    if (t.vx.empty()) {
        t.vx.assign(ctx->nlimbs, 0);
        kite_get_limbs(ctx, x0, t.vx.data());
        t.vd = t.vx;
        t.vi = 0;
    }
    kite_set_limbs(ctx, kite_batch_num(b, t.slot, kX), t.vx.data());
    kite_set_limbs(ctx, kite_batch_num(b, t.slot, kD), t.vd.data());
    t.i = t.vi;
    t.full = t.x0_at + (t.n - t.x0_at) / B * B;
    t.check_at = std::min(t.vi + (uint64_t)B * M, t.full);
    t.phase = t.vi >= t.full ? kTailX : kBlocks;
    if (t.phase == kTailX) {
        kite_set_limbs(ctx, kite_batch_num(b, t.slot, kT), t.vx.data());
        if (t.full == t.n) t.phase = kDone;
    }
    return 0;
}

// (Re)starts LLR job t's current run in a fresh slot at length L: from its verified state, or
// from V_k by the Lucas ladder. Returns 1 if the ladder's round-off was over the limit.
int start_llr(kite_batch *b, Job &t, uint32_t L, double maxround) {
    t.slot = kite_batch_add(b, t.k, t.n, L, 2);
    if (t.slot < 0) return -1;
    kite_ctx *ctx = kite_batch_ctx(b, t.slot);
    t.L = kite_fft_len(ctx);
    kite_num *x = kite_batch_num(b, t.slot, kX);
    if (t.vx.empty()) {
        kite_num *v1 = kite_num_new(ctx), *cP = kite_num_new(ctx), *c2 = kite_num_new(ctx), *tmp = kite_num_new(ctx);
        if (!v1 || !cP || !c2 || !tmp) return -1;
        kite_set_roundoff_check(ctx, 1);
        kite_reset_roundoff(ctx);
        kite_set_u64(ctx, cP, t.P);
        kite_set_u64(ctx, c2, 2);
        llr_lucas_vk(ctx, t.k, x, v1, cP, c2, tmp);
        const double e = kite_max_roundoff(ctx);
        t.vx.assign(ctx->nlimbs, 0);
        kite_get_limbs(ctx, x, t.vx.data());
        for (kite_num *y : {v1, cP, c2, tmp}) kite_num_free(ctx, y);
        if (e > t.res.maxerr) t.res.maxerr = e;
        if (e > maxround) {
            t.vx.clear();
            return 1;
        }
        t.vi = 0;
    } else {
        kite_set_limbs(ctx, x, t.vx.data());
    }
    kite_set_limbs(ctx, kite_batch_num(b, t.slot, kV), t.vx.data());
    t.i = t.vi;
    t.phase = t.vi == 0 ? kLlrFirst : kLlrSteps;
    kite_batch_set_xadd(b, t.slot, t.phase == kLlrFirst ? 2 : -2);
    t.copy_pending = false;
    return 0;
}

bool nums_equal(kite_batch *b, int slot, int a, int c, std::vector<uint64_t> &la, std::vector<uint64_t> &lb) {
    kite_ctx *ctx = kite_batch_ctx(b, slot);
    la.resize(ctx->nlimbs);
    lb.resize(ctx->nlimbs);
    kite_get_limbs(ctx, kite_batch_num(b, slot, a), la.data());
    kite_get_limbs(ctx, kite_batch_num(b, slot, c), lb.data());
    return la == lb;
}

// Name of job t's checkpoint file: kite_<N mod 3417905339>.<base>.ckpt (Fermat) or
// kite_<N mod 3417905339>.llr.ckpt (LLR), as kite-prst names them (tools/kite_prst.c).
std::string ckpt_name(const Job &t) {
    char name[64];
    const unsigned long long fp = n_mod(t.k, t.n, 3417905339ull);
    if (t.kind == KITE_JOB_LLR) snprintf(name, sizeof name, "kite_%llu.llr.ckpt", fp);
    else snprintf(name, sizeof name, "kite_%llu.%u.ckpt", fp, t.base);
    return name;
}

struct Runner {
    kite_batch *b = nullptr;
    kite_run_opts o{};
    uint32_t B = 1000, M = 250;
    double maxround = 0.40, t0 = 0, ck_secs = 600;
    std::string dir;  // checkpoint directory ("" = no checkpoints)
    std::vector<Job> all;
    uint64_t bytes_in_flight = 0;
    int rc_all = 0;

    std::string ckpt_path(const Job &t) const { return dir + "/" + ckpt_name(t); }

    // Restores job t from its checkpoint file, if a valid one exists. A Fermat state is trusted
    // for its x only (verified when written): the Gerbicz product restarts at x, so files from
    // writers with other block sizes resume too. An LLR file restores the run, its step and the
    // earlier runs' results; a confirmation run keeps the length it was written at.
    void resume(Job &t) {
        const size_t nl = limbs_for_n(t.n);
        const std::string path = ckpt_path(t);
        std::vector<uint64_t> x(nl), d(nl);
        uint32_t L = 0;
        if (t.kind == KITE_JOB_PRP) {
            uint64_t it = 0;
            if (prp_checkpoint_read(path.c_str(), t.k, t.n, t.base, &it, &L, x.data(), d.data(), nl) || it > t.n)
                return;
            t.vi = t.x0_at = it;
            t.vd = t.x0 = x;
            t.vx.swap(x);
        } else {
            LlrCheckpoint c;
            if (llr_checkpoint_read(path.c_str(), t.k, t.n, &c, x.data(), nl) || c.P != t.P) return;
            t.run = c.run;
            for (int r = 0; r < 2; r++) {
                t.lres[r] = c.prev[r];
                t.lzero[r] = r < c.run && c.prev[r] == 0;
            }
            t.res.confirm_runs = (uint32_t)c.run;
            t.vi = c.iter;
            t.vx.swap(x);
            L = c.L;
            if (c.run > 0) t.L = L;
        }
        if (!t.L && L > kite_choose_fft_len(t.k, t.n)) t.L = L;  // a round-off retry had lengthened it
        t.res.resumed = 1;
        t.res.resumed_at = t.vi;
    }

    // Writes Fermat job t's verified state: x at iteration vi, with d = x, which tells a reader
    // to restart the Gerbicz product at x.
    void save_fermat(Job &t) {
        if (dir.empty() || t.vi == 0) return;
        prp_checkpoint_write(ckpt_path(t).c_str(), t.k, t.n, t.base, t.vi, t.L, t.vx.data(), t.vx.data(), t.vx.size());
        t.last_ck = now_s();
    }

    // Writes LLR job t's state: x (in its slot, at step vi of the current run) after a sync.
    void save_llr(Job &t, std::vector<uint64_t> &buf) {
        if (dir.empty()) return;
        kite_ctx *ctx = kite_batch_ctx(b, t.slot);
        buf.resize(ctx->nlimbs);
        kite_batch_sync(b);
        kite_get_limbs(ctx, kite_batch_num(b, t.slot, kX), buf.data());
        LlrCheckpoint c;
        c.iter = t.vi;
        c.P = t.P;
        c.L = t.L;
        c.run = t.run;
        c.prev[0] = t.lres[0];
        c.prev[1] = t.lres[1];
        llr_checkpoint_write(ckpt_path(t).c_str(), t.k, t.n, c, buf.data(), buf.size());
        t.last_ck = now_s();
    }

    void done(Job &t, int rc) {
        if (!dir.empty()) remove(ckpt_path(t).c_str());
        if (t.slot >= 0) kite_batch_remove(b, t.slot);
        t.slot = -1;
        t.phase = kDone;
        t.rc = rc;
        t.res.status = rc;
        t.res.fft_len = t.L;
        t.res.finish = now_s() - t0;
        bytes_in_flight -= t.bytes;
        if (rc && !rc_all) rc_all = rc;
        if (o.on_done) o.on_done(o.arg, t.index, &t.res);
    }

    void finish_fermat(Job &t, const std::vector<uint64_t> &xn, int rc) {
        if (rc == 0) {
            kite_ctx *ctx = kite_batch_ctx(b, t.slot);
            std::vector<uint64_t> la = xn;
            bool prp = la[0] == (uint64_t)t.base * t.base;
            for (size_t j = 1; j < la.size(); j++) prp &= la[j] == 0;
            if (kite_residue_nm1(la.data(), ctx->Nl.data(), la.size(), t.base)) rc = -4;
            t.res.res64 = la[0];
            t.res.is_prime = prp;
        }
        done(t, rc);
    }

    // Restarts job t at a longer FFT after a round-off failure; false if none is left.
    bool retry_longer(Job &t) {
        kite_ctx *ctx = kite_batch_ctx(b, t.slot);
        if (t.kind == KITE_JOB_LLR && t.vi > 0) {  // the verified copy holds step vi
            kite_batch_sync(b);
            kite_get_limbs(ctx, kite_batch_num(b, t.slot, kV), t.vx.data());
        }
        const uint32_t L2 = kite_next_fft_len(t.L + t.L / 25);
        kite_batch_remove(b, t.slot);
        t.slot = -1;
        t.res.roundoff_retries++;
        if (!L2 || t.res.roundoff_retries > 8) return false;
        return start(t, L2) == 0;
    }

    // Starts (or restarts) job t at length L; 0 on success.
    int start(Job &t, uint32_t L) {
        if (t.kind == KITE_JOB_PRP) return start_fermat(b, t, L, B, M);
        for (;;) {
            const int r = start_llr(b, t, L, maxround);
            if (r <= 0) return r;
            // the ladder's round-off was over the limit: a longer FFT
            const uint32_t L2 = kite_next_fft_len(t.L + t.L / 25);
            kite_batch_remove(b, t.slot);
            t.slot = -1;
            t.res.roundoff_retries++;
            if (!L2 || t.res.roundoff_retries > 8) return -2;
            L = L2;
        }
    }

    // The end of an LLR run: record it and decide (prime, confirmed, vote) or start the next run.
    void llr_run_done(Job &t) {
        kite_ctx *ctx = kite_batch_ctx(b, t.slot);
        std::vector<uint64_t> out(ctx->nlimbs);
        kite_batch_sync(b);
        kite_get_limbs(ctx, kite_batch_num(b, t.slot, kX), out.data());
        bool zero = true;
        for (uint64_t w : out) zero &= w == 0;
        const int r = t.run;
        t.lres[r] = out[0];
        t.lzero[r] = zero;
        auto agree = [&](int a, int c) { return t.lres[a] == t.lres[c] && t.lzero[a] == t.lzero[c]; };
        int fin = -1;
        if (r == 0 && zero) fin = 0;
        else if (r == 1 && agree(1, 0)) fin = 1;
        else if (r == 2) fin = agree(2, 1) ? 1 : agree(2, 0) ? 0 : -2;
        if (fin == -2) {
            done(t, -3);
            return;
        }
        if (fin >= 0) {
            t.res.is_prime = t.lzero[fin];
            t.res.res64 = t.lres[fin];
            done(t, 0);
            return;
        }
        // confirm at the next FFT length (src/llr.c:350-360)
        const uint32_t L2 = kite_next_fft_len(t.L + 1);
        kite_batch_remove(b, t.slot);
        t.slot = -1;
        t.run = r + 1;
        t.res.confirm_runs = (uint32_t)t.run;
        t.vx.clear();
        t.vi = 0;
        if (!L2 || start(t, L2)) done(t, -2);
    }
};

}  // namespace

extern "C" uint64_t kite_job_bytes(uint32_t kind, uint32_t fft_len) {
    // residues (8 L each: 4 for Fermat, x and its verified copy for LLR; the LLR ladder briefly
    // holds 4 more) plus the per-test tables and scratch (operand spectra 8 L, digits 4 L, codes)
    const uint64_t L = fft_len;
    return (kind == KITE_JOB_LLR ? 6 : 4) * 8 * L + 13 * L + (1u << 16);
}

extern "C" int kite_batch_run(const kite_job *jobs, size_t njobs, const kite_run_opts *opts, kite_job_result *res) {
    Runner R;
    R.o = opts ? *opts : kite_run_opts{};
    R.B = R.o.block ? R.o.block : 1000;
    R.M = R.o.gerbicz_blocks ? R.o.gerbicz_blocks : std::max<uint32_t>(R.B / 4, 8);
    R.maxround = R.o.max_roundoff > 0 ? R.o.max_roundoff : 0.40;
    if (R.o.checkpoint_dir && R.o.checkpoint_dir[0]) {
        R.dir = R.o.checkpoint_dir;
        struct stat st;
        if (stat(R.dir.c_str(), &st) != 0 || !S_ISDIR(st.st_mode) || access(R.dir.c_str(), W_OK | X_OK) != 0) {
            // not a writable directory: refuse before any work
            for (size_t i = 0; i < njobs; i++) {
                res[i] = kite_job_result{};
                res[i].status = -1;
            }
            return -1;
        }
    }
    R.ck_secs = R.o.checkpoint_secs > 0 ? R.o.checkpoint_secs : 600;
    const size_t max_slots = R.o.slots ? R.o.slots : (size_t)-1;
    R.b = kite_batch_new();
    if (!R.b) return -1;
    kite_batch *b = R.b;
    const uint32_t B = R.B;
    R.all.resize(njobs);
    std::vector<size_t> order(njobs);
    for (size_t i = 0; i < njobs; i++) {
        Job &t = R.all[i];
        t.index = i;
        t.k = jobs[i].k;
        t.n = jobs[i].n;
        t.kind = jobs[i].kind;
        t.base = jobs[i].base ? jobs[i].base : 3;
        t.L = jobs[i].fft_len;
        t.arrival = jobs[i].arrival;
        t.fault_pending = R.o.inject_fault_at && R.o.inject_fault_job == i;
        t.res.status = KITE_JOB_PENDING;
        order[i] = i;
    }
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t c) { return R.all[a].arrival < R.all[c].arrival; });
    std::deque<size_t> waiting(order.begin(), order.end());
    std::vector<size_t> active;
    std::vector<uint64_t> la, lb;
    bool engine_failed = false;
    R.t0 = now_s();
    auto admit = [&]() {
        const double now = now_s() - R.t0;
        while (!waiting.empty() && active.size() < max_slots) {
            auto pick = waiting.begin();
            if (R.o.order == KITE_ORDER_LEAST_WORK) {  // the arrived job with the least work: n^2
                double best = 0;                        // (n squarings at a length proportional to n)
                for (auto it = waiting.begin(); it != waiting.end() && R.all[*it].arrival <= now; ++it) {
                    const double w = (double)R.all[*it].n * (double)R.all[*it].n;
                    if (it == waiting.begin() || w < best) {
                        best = w;
                        pick = it;
                    }
                }
            }
            Job &t = R.all[*pick];
            if (t.arrival > now) break;
            const uint32_t Lest = t.L ? t.L : kite_choose_fft_len(t.k, t.n);
            const uint64_t need = kite_job_bytes(t.kind, Lest);
            if (R.o.hbm_budget && !active.empty() && R.bytes_in_flight + need > R.o.hbm_budget) break;
            waiting.erase(pick);
            t.started = true;
            t.res.start = now;
            t.bytes = need;
            R.bytes_in_flight += need;
            if (t.kind == KITE_JOB_LLR) {
                if (t.n < 3 || !(t.k & 1) || kite_llr_choose_P(t.k, t.n, &t.P, &t.res.factor)) {
                    R.done(t, -1);
                    continue;
                }
                t.res.P = t.P;
                if (t.res.factor) {
                    R.done(t, 0);
                    continue;
                }
            }
            if (!R.dir.empty()) R.resume(t);
            t.last_ck = now_s();
            const int r = R.start(t, t.L);
            if (r) {
                R.done(t, r < 0 ? r : -1);
                continue;
            }
            if (t.phase == kDone) {  // Fermat: n below one block and the tail already covered
                R.finish_fermat(t, t.vx, 0);
                continue;
            }
            active.push_back(t.index);
        }
    };
    admit();
    for (;;) {
        const double now = now_s() - R.t0;
        if (R.o.horizon > 0 && now >= R.o.horizon) break;
        if (active.empty()) {
            if (waiting.empty()) break;
            const double next = R.all[waiting.front()].arrival;
            double wait = next - now;
            if (R.o.horizon > 0) wait = std::min(wait, R.o.horizon - now);
            if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.05)));
            admit();
            continue;
        }
        const int cap = kite_batch_capacity(b);
        std::vector<int32_t> which(cap, -1), dst(cap, -1), src(cap, -1), cdst(cap, -1), csrc(cap, -1);
        std::vector<uint32_t> cnt(cap, 0);
        // LLR jobs whose last block passed: their verified copy before the next block
        for (size_t ti : active) {
            Job &t = R.all[ti];
            if (t.kind == KITE_JOB_LLR && t.copy_pending) {
                cdst[t.slot] = kV;
                csrc[t.slot] = kX;
                t.copy_pending = false;
            }
        }
        kite_batch_copy(b, cdst.data(), csrc.data());
        std::fill(cdst.begin(), cdst.end(), -1);
        std::fill(csrc.begin(), csrc.end(), -1);
        for (size_t ti : active) {
            Job &t = R.all[ti];
            const uint32_t tail = (uint32_t)(t.n - t.full);
            switch (t.phase) {
            case kBlocks:
                which[t.slot] = kX;
                cnt[t.slot] = B;
                cdst[t.slot] = kT;  // t = d before the update
                csrc[t.slot] = kD;
                dst[t.slot] = kD;
                src[t.slot] = kX;
                break;
            case kCheck:
                which[t.slot] = kT;
                cnt[t.slot] = B;
                dst[t.slot] = kT;
                src[t.slot] = kX0;
                break;
            case kTailX:
                which[t.slot] = kX;
                cnt[t.slot] = tail;
                break;
            case kTailT:
                which[t.slot] = kT;
                cnt[t.slot] = tail;
                break;
            case kLlrFirst:
                which[t.slot] = kX;
                cnt[t.slot] = 1;
                break;
            case kLlrSteps:
                which[t.slot] = kX;
                cnt[t.slot] = (uint32_t)std::min<uint64_t>(B, t.n - 1 - t.i);
                break;
            default:
                break;
            }
        }
        if (kite_batch_square(b, which.data(), cnt.data())) {
            if (!R.rc_all) R.rc_all = -1;
            engine_failed = true;
            break;
        }
        // testing hook: a +1 digit corruption at the end of the block holding inject_fault_at
        for (size_t ti : active) {
            Job &t = R.all[ti];
            const bool in_block = (t.phase == kBlocks || t.phase == kLlrSteps) && t.i + cnt[t.slot] >= R.o.inject_fault_at;
            if (t.fault_pending && in_block && (t.kind == KITE_JOB_PRP || t.run == 0)) {
                kite_batch_sync(b);
                kite_debug_corrupt(kite_batch_ctx(b, t.slot), kite_batch_num(b, t.slot, kX));
                t.fault_pending = false;
            }
        }
        kite_batch_copy(b, cdst.data(), csrc.data());
        kite_batch_mul(b, dst.data(), src.data());
        std::vector<double> err(cap, 0);
        kite_batch_maxerr(b, err.data(), 1);
        std::vector<size_t> still;
        for (size_t ti : active) {
            Job &t = R.all[ti];
            const double e = err[t.slot];
            if (e > t.res.maxerr) t.res.maxerr = e;
            if (e > R.maxround) {  // redo from the verified state at a longer FFT
                if (!R.retry_longer(t)) {
                    R.done(t, -2);
                    continue;
                }
                still.push_back(ti);
                continue;
            }
            switch (t.phase) {
            case kLlrFirst:
            case kLlrSteps:
                t.i += cnt[t.slot];
                if (t.phase == kLlrFirst) {
                    t.phase = kLlrSteps;
                    kite_batch_set_xadd(b, t.slot, -2);
                }
                t.vi = t.i;
                t.copy_pending = true;
                if (t.i >= t.n - 1) {
                    R.llr_run_done(t);
                    if (t.phase != kDone) still.push_back(ti);
                } else {
                    if (now_s() - t.last_ck >= R.ck_secs) R.save_llr(t, la);
                    still.push_back(ti);
                }
                break;
            case kBlocks:
                t.i += B;
                if (t.i >= t.check_at) t.phase = kCheck;
                still.push_back(ti);
                break;
            case kCheck: {
                if (nums_equal(b, t.slot, kT, kD, la, lb)) {
                    kite_ctx *ctx = kite_batch_ctx(b, t.slot);
                    kite_get_limbs(ctx, kite_batch_num(b, t.slot, kX), t.vx.data());
                    t.vd = lb;
                    t.vi = t.i;
                    t.consecutive_fail = 0;
                    if (now_s() - t.last_ck >= R.ck_secs && t.vi < t.n) R.save_fermat(t);
                    if (t.vi >= t.full) {
                        if (t.full == t.n) {
                            R.finish_fermat(t, t.vx, 0);
                            break;
                        }
                        kite_set_limbs(ctx, kite_batch_num(b, t.slot, kT), t.vx.data());
                        t.phase = kTailX;
                    } else {
                        t.check_at = std::min(t.vi + (uint64_t)B * R.M, t.full);
                        t.phase = kBlocks;
                    }
                } else {
                    t.res.gerbicz_errors++;
                    if (++t.consecutive_fail > 5) {
                        R.finish_fermat(t, t.vx, -3);
                        break;
                    }
                    kite_ctx *ctx = kite_batch_ctx(b, t.slot);
                    kite_set_limbs(ctx, kite_batch_num(b, t.slot, kX), t.vx.data());
                    kite_set_limbs(ctx, kite_batch_num(b, t.slot, kD), t.vd.data());
                    t.i = t.vi;
                    t.check_at = std::min(t.vi + (uint64_t)B * R.M, t.full);
                    t.phase = kBlocks;
                }
                if (t.phase != kDone) still.push_back(ti);
                break;
            }
            case kTailX:
                t.phase = kTailT;
                still.push_back(ti);
                break;
            case kTailT:
                if (nums_equal(b, t.slot, kX, kT, la, lb)) {
                    R.finish_fermat(t, la, 0);
                } else if (++t.tail_attempt > 5) {
                    R.finish_fermat(t, la, -3);
                } else {
                    t.res.gerbicz_errors++;
                    kite_ctx *ctx = kite_batch_ctx(b, t.slot);
                    kite_set_limbs(ctx, kite_batch_num(b, t.slot, kX), t.vx.data());
                    kite_set_limbs(ctx, kite_batch_num(b, t.slot, kT), t.vx.data());
                    t.phase = kTailX;
                    still.push_back(ti);
                }
                break;
            default:
                break;
            }
        }
        active.swap(still);
        admit();
    }
    if (!R.dir.empty() && !engine_failed) {  // stopped by the horizon: save the jobs in flight
        for (size_t ti : active) {
            Job &t = R.all[ti];
            if (t.phase == kDone || t.slot < 0) continue;
            if (t.kind == KITE_JOB_LLR) R.save_llr(t, la);
            else R.save_fermat(t);
        }
    }
    for (size_t i = 0; i < njobs; i++) {
        Job &t = R.all[i];
        if (t.phase != kDone || !t.started) {  // stopped by the horizon
            t.res.status = KITE_JOB_PENDING;
            if (t.started) t.res.fft_len = t.L;
        }
        res[i] = t.res;
        if (t.res.status && !R.rc_all) R.rc_all = t.res.status;
    }
    kite_batch_free(b);
    return R.rc_all;
}

// ---- kite_batch_prp: Fermat jobs that all arrive at once ------------------------------------

namespace {

struct PrpAdapter {
    const kite_batch_prp_opts *o;
};

void prp_done(void *arg, size_t index, const kite_job_result *r) {
    const kite_batch_prp_opts *o = static_cast<PrpAdapter *>(arg)->o;
    kite_prp_result pr{};
    pr.res64 = r->res64;
    pr.is_prp = r->is_prime;
    pr.fft_len = r->fft_len;
    pr.maxerr = r->maxerr;
    pr.gerbicz_errors = r->gerbicz_errors;
    pr.roundoff_retries = r->roundoff_retries;
    pr.seconds = r->finish - r->start;
    o->on_done(o->arg, index, r->status, &pr);
}

}  // namespace

extern "C" int kite_batch_prp(const kite_batch_test *tests, size_t ntests, const kite_batch_prp_opts *opts,
                              kite_prp_result *res, int *status) {
    kite_batch_prp_opts o = opts ? *opts : kite_batch_prp_opts{};
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
        res[i] = kite_prp_result{};
        res[i].res64 = jr[i].res64;
        res[i].is_prp = jr[i].is_prime;
        res[i].fft_len = jr[i].fft_len;
        res[i].maxerr = jr[i].maxerr;
        res[i].gerbicz_errors = jr[i].gerbicz_errors;
        res[i].roundoff_retries = jr[i].roundoff_retries;
        res[i].seconds = jr[i].finish - jr[i].start;
        if (status) status[i] = jr[i].status == KITE_JOB_PENDING ? -1 : jr[i].status;
    }
    return rc;
}
