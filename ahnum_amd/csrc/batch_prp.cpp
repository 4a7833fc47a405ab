// Batched Fermat PRP tests with Gerbicz checking: every test in flight advances one block of B
// squarings per round through shared launches (kite_batch_square), then its d (or check) multiply
// (kite_batch_mul). Each test keeps the reference driver's state machine (src/prp.c:329-583):
// blocks with d_{t+1} = d_t * x, a check d == x0 * t^(2^B) every M blocks with rollback to the
// last verified (i, x, d), a round-off retry at a >= 4% longer FFT, and the doubled tail. A
// finished test's slot is given to the next waiting test (continuous batching). The block size
// is common to the batch so the tests stay in step.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <deque>
#include <vector>

#include "kite_batch.h"
#include "kite_internal.h"

using namespace kite;

namespace {

enum Phase { kBlocks, kCheck, kTailX, kTailT, kDone };
enum Num { kX = 0, kD = 1, kX0 = 2, kT = 3 };

double now_s() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

struct Test {
    size_t index = 0;
    uint64_t k = 0, n = 0;
    uint32_t base = 3, L = 0;
    int slot = -1;
    uint64_t full = 0, i = 0, vi = 0, check_at = 0;
    std::vector<uint64_t> vx, vd;
    Phase phase = kBlocks;
    uint32_t consecutive_fail = 0, tail_attempt = 0;
    bool fault_pending = false;
    kite_prp_result res{};
    int rc = 0;
    double t0 = 0;
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

// (Re)starts test t in a fresh slot at length L from its verified state (or x0 if none).
int start_slot(kite_batch *b, Test &t, uint32_t L, uint32_t B, uint32_t M) {
    t.slot = kite_batch_add(b, t.k, t.n, L, 4);
    if (t.slot < 0) return -1;
    kite_ctx *ctx = kite_batch_ctx(b, t.slot);
    t.L = kite_fft_len(ctx);
    kite_num *x0 = kite_batch_num(b, t.slot, kX0);
    power_start(ctx, x0, t.k, t.base);
    if (t.vx.empty()) {
        t.vx.assign(ctx->nlimbs, 0);
        kite_get_limbs(ctx, x0, t.vx.data());
        t.vd = t.vx;
        t.vi = 0;
    }
    kite_set_limbs(ctx, kite_batch_num(b, t.slot, kX), t.vx.data());
    kite_set_limbs(ctx, kite_batch_num(b, t.slot, kD), t.vd.data());
    t.i = t.vi;
    t.full = t.n / B * B;
    t.check_at = std::min(t.vi + (uint64_t)B * M, t.full);
    t.phase = t.vi >= t.full ? kTailX : kBlocks;
    if (t.phase == kTailX) {
        kite_set_limbs(ctx, kite_batch_num(b, t.slot, kT), t.vx.data());
        if (t.full == t.n) t.phase = kDone;
    }
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

void finish(kite_batch *b, Test &t, const std::vector<uint64_t> &xn, const kite_batch_prp_opts &o, int rc) {
    kite_ctx *ctx = kite_batch_ctx(b, t.slot);
    t.res.fft_len = t.L;
    if (rc == 0) {
        std::vector<uint64_t> la = xn;
        bool prp = la[0] == (uint64_t)t.base * t.base;
        for (size_t j = 1; j < la.size(); j++) prp &= la[j] == 0;
        if (kite_residue_nm1(la.data(), ctx->Nl.data(), la.size(), t.base)) rc = -4;
        t.res.res64 = la[0];
        t.res.is_prp = prp;
    }
    t.res.seconds = now_s() - t.t0;
    kite_batch_remove(b, t.slot);
    t.slot = -1;
    t.phase = kDone;
    t.rc = rc;
    if (o.on_done) o.on_done(o.arg, t.index, rc, &t.res);
}

}  // namespace

extern "C" int kite_batch_prp(const kite_batch_test *tests, size_t ntests, const kite_batch_prp_opts *opts,
                              kite_prp_result *res, int *status) {
    kite_batch_prp_opts o = opts ? *opts : kite_batch_prp_opts{};
    const uint32_t B = o.gerbicz_block ? o.gerbicz_block : 1000;
    const uint32_t M = o.gerbicz_blocks ? o.gerbicz_blocks : std::max<uint32_t>(B / 4, 8);
    const double maxround = o.max_roundoff > 0 ? o.max_roundoff : 0.40;
    const size_t inflight = o.slots ? o.slots : ntests;
    kite_batch *b = kite_batch_new();
    if (!b) return -1;
    std::vector<Test> all(ntests);
    std::deque<size_t> waiting;
    for (size_t i = 0; i < ntests; i++) {
        Test &t = all[i];
        t.index = i;
        t.k = tests[i].k;
        t.n = tests[i].n;
        t.base = tests[i].base ? tests[i].base : 3;
        t.L = tests[i].fft_len;
        t.fault_pending = o.inject_fault_at && o.inject_fault_test == i;
        waiting.push_back(i);
    }
    std::vector<size_t> active;
    std::vector<uint64_t> la, lb;
    int rc_all = 0;
    auto admit = [&]() {
        while (active.size() < inflight && !waiting.empty()) {
            Test &t = all[waiting.front()];
            waiting.pop_front();
            t.t0 = now_s();
            if (start_slot(b, t, t.L, B, M)) {
                t.phase = kDone;
                t.rc = -1;
                if (!rc_all) rc_all = -1;
                if (o.on_done) o.on_done(o.arg, t.index, t.rc, &t.res);
                continue;
            }
            if (t.phase == kDone) {
                finish(b, t, t.vx, o, 0);
                continue;
            }
            active.push_back(t.index);
        }
    };
    admit();
    while (!active.empty()) {
        const int cap = kite_batch_capacity(b);
        std::vector<int32_t> which(cap, -1), dst(cap, -1), src(cap, -1), cdst(cap, -1), csrc(cap, -1);
        std::vector<uint32_t> cnt(cap, 0);
        for (size_t ti : active) {
            Test &t = all[ti];
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
            default:
                break;
            }
        }
        if (kite_batch_square(b, which.data(), cnt.data())) {
            if (!rc_all) rc_all = -1;
            break;
        }
        // testing hook: a +1 digit corruption of x at the end of the block holding inject_fault_at
        for (size_t ti : active) {
            Test &t = all[ti];
            if (t.fault_pending && t.phase == kBlocks && t.i + B >= o.inject_fault_at) {
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
            Test &t = all[ti];
            const double e = err[t.slot];
            if (e > t.res.maxerr) t.res.maxerr = e;
            if (e > maxround) {  // redo from the verified state at a longer FFT
                const uint32_t L2 = kite_next_fft_len(t.L + t.L / 25);
                kite_batch_remove(b, t.slot);
                t.res.roundoff_retries++;
                if (!L2 || start_slot(b, t, L2, B, M)) {
                    t.slot = -1;
                    t.phase = kDone;
                    t.rc = -2;
                    if (!rc_all) rc_all = -2;
                    if (o.on_done) o.on_done(o.arg, t.index, t.rc, &t.res);
                    continue;
                }
                still.push_back(ti);
                continue;
            }
            switch (t.phase) {
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
                    if (t.vi >= t.full) {
                        if (t.full == t.n) {
                            finish(b, t, t.vx, o, 0);
                            break;
                        }
                        kite_set_limbs(ctx, kite_batch_num(b, t.slot, kT), t.vx.data());
                        t.phase = kTailX;
                    } else {
                        t.check_at = std::min(t.vi + (uint64_t)B * M, t.full);
                        t.phase = kBlocks;
                    }
                } else {
                    t.res.gerbicz_errors++;
                    if (++t.consecutive_fail > 5) {
                        finish(b, t, t.vx, o, -3);
                        if (!rc_all) rc_all = -3;
                        break;
                    }
                    kite_ctx *ctx = kite_batch_ctx(b, t.slot);
                    kite_set_limbs(ctx, kite_batch_num(b, t.slot, kX), t.vx.data());
                    kite_set_limbs(ctx, kite_batch_num(b, t.slot, kD), t.vd.data());
                    t.i = t.vi;
                    t.check_at = std::min(t.vi + (uint64_t)B * M, t.full);
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
                    finish(b, t, la, o, 0);
                } else if (++t.tail_attempt > 5) {
                    finish(b, t, la, o, -3);
                    if (!rc_all) rc_all = -3;
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
    for (size_t i = 0; i < ntests; i++) {
        res[i] = all[i].res;
        if (status) status[i] = all[i].phase == kDone ? all[i].rc : -1;
        if (all[i].rc && !rc_all) rc_all = all[i].rc;
    }
    kite_batch_free(b);
    return rc_all;
}
