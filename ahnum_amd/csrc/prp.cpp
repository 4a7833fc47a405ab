// Fermat PRP test of N = k*2^n - 1 with Gerbicz checking (reference src/prp.c:1-583).
//
// Chain x_0 = base^k, x_{i+1} = x_i^2, so x_n = base^(N+1); the residue is x_n / base^2.
// Gerbicz: d_t = prod_{j<=t} x_{jB} satisfies d_{t+1} = x_0 * d_t^(2^B). d is updated every B
// squarings; every M blocks the identity is checked with B more squarings of t (d before the
// last update); a failure rolls back to the last verified (i, x, d). Round-off above the limit
// restarts from the verified state at a >= 4% longer FFT. The tail (n mod B squarings) is done
// twice and compared.
//
// The B squarings of a block are issued as one device-resident run (engine_square with a count),
// so the host touches the GPU once per block: the multiply, then a round-off read-back.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "kite_internal.h"

using namespace kite;

namespace {

double now_s() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

uint64_t fnv1a(const void *p, size_t len, uint64_t h) {
    const uint8_t *b = static_cast<const uint8_t *>(p);
    for (size_t i = 0; i < len; i++) h = (h ^ b[i]) * 0x100000001b3ull;
    return h;
}

// KITECKP1 checkpoint: {magic, k, n, base, iter, nlimbs, fft_len}, x limbs, d limbs, FNV-1a-64 of
// all before it (src/prp.c:156-212). "AHNCKPT1" is read too.
struct CkptHdr {
    char magic[8];
    uint64_t k, n, base, iter, nlimbs, fft_len;
};
const char kMagic[8] = {'K', 'I', 'T', 'E', 'C', 'K', 'P', '1'};
const char kMagicOld[8] = {'A', 'H', 'N', 'C', 'K', 'P', 'T', '1'};

bool magic_ok(const char m[8]) { return !memcmp(m, kMagic, 8) || !memcmp(m, kMagicOld, 8); }

int ckpt_write(const char *path, uint64_t k, uint64_t n, uint32_t base, uint64_t iter, uint32_t L,
               const uint64_t *x, const uint64_t *d, size_t nl) {
    std::string tmp = std::string(path) + ".tmp";
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) return -1;
    CkptHdr h{};
    memcpy(h.magic, kMagic, 8);
    h.k = k;
    h.n = n;
    h.base = base;
    h.iter = iter;
    h.nlimbs = nl;
    h.fft_len = L;
    uint64_t sum = fnv1a(&h, sizeof h, 0xcbf29ce484222325ull);
    sum = fnv1a(x, nl * 8, sum);
    sum = fnv1a(d, nl * 8, sum);
    bool ok = fwrite(&h, sizeof h, 1, f) == 1 && fwrite(x, 8, nl, f) == nl && fwrite(d, 8, nl, f) == nl &&
              fwrite(&sum, 8, 1, f) == 1;
    ok &= fclose(f) == 0;
    if (!ok || rename(tmp.c_str(), path)) {
        remove(tmp.c_str());
        return -1;
    }
    return 0;
}

int ckpt_read(const char *path, uint64_t k, uint64_t n, uint32_t base, uint64_t *iter, uint32_t *L,
              uint64_t *x, uint64_t *d, size_t nl) {
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    CkptHdr h;
    uint64_t sum = 0;
    bool ok = fread(&h, sizeof h, 1, f) == 1 && magic_ok(h.magic) && h.k == k && h.n == n && h.base == base &&
              h.nlimbs == nl && fread(x, 8, nl, f) == nl && fread(d, 8, nl, f) == nl && fread(&sum, 8, 1, f) == 1;
    fclose(f);
    if (!ok) return -1;
    uint64_t s = fnv1a(&h, sizeof h, 0xcbf29ce484222325ull);
    s = fnv1a(x, nl * 8, s);
    s = fnv1a(d, nl * 8, s);
    if (s != sum) return -1;
    *iter = h.iter;
    *L = (uint32_t)h.fft_len;
    return 0;
}

// proof points captured since the last verified state, reported once verified
struct Pending {
    std::vector<uint32_t> index;
    std::vector<std::vector<uint64_t>> limbs;
    void clear() {
        index.clear();
        limbs.clear();
    }
};

void pending_add(Pending &p, kite_ctx *ctx, uint32_t index, const kite_num *x) {
    p.index.push_back(index);
    p.limbs.emplace_back(kite_limbs(ctx));
    kite_get_limbs(ctx, x, p.limbs.back().data());
}

int pending_flush(Pending &p, const kite_prp_opts &o) {
    int rc = 0;
    for (size_t i = 0; i < p.index.size() && !rc; i++)
        rc = o.on_point(o.point_arg, p.index[i], o.points[p.index[i]], p.limbs[i].data());
    p.clear();
    return rc;
}

struct State {
    kite_ctx *ctx = nullptr;
    kite_num *x = nullptr, *d = nullptr, *x0 = nullptr, *t = nullptr;
    void drop() {
        if (!ctx) return;
        kite_num_free(ctx, x);
        kite_num_free(ctx, d);
        kite_num_free(ctx, x0);
        kite_num_free(ctx, t);
        kite_ctx_free(ctx);
        *this = State();
    }
    int make(uint64_t k, uint64_t n, kite_opts o) {
        ctx = kite_ctx_new(k, n, -1, &o);
        if (!ctx) return -1;
        kite_set_roundoff_check(ctx, 4);
        x = kite_num_new(ctx);
        d = kite_num_new(ctx);
        x0 = kite_num_new(ctx);
        t = kite_num_new(ctx);
        return x && d && x0 && t ? 0 : -1;
    }
};

bool equal_nums(kite_ctx *ctx, const kite_num *a, const kite_num *b, uint64_t *la, uint64_t *lb) {
    kite_get_limbs(ctx, a, la);
    kite_get_limbs(ctx, b, lb);
    return !memcmp(la, lb, kite_limbs(ctx) * 8);
}

// x0 = base^k mod N: left-to-right binary powering, replaced by the exact power while it is
// below N (src/prp.c:298-327).
void power_start(State &s, uint64_t k, uint32_t base, uint64_t *l) {
    kite_ctx *ctx = s.ctx;
    const size_t nl = kite_limbs(ctx);
    kite_set_u64(ctx, s.x0, 1);
    for (int b = 63; b >= 0; b--) {
        if (k >> b == 0) continue;
        kite_square(ctx, s.x0);
        if ((k >> b) & 1) kite_mul_small(ctx, s.x0, base);
    }
    if ((double)k * std::log2((double)base) + 2 < (double)ctx->n) {
        memset(l, 0, nl * 8);
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
        kite_set_limbs(ctx, s.x0, l);
    }
}

}  // namespace

namespace kite {

int prp_checkpoint_write(const char *path, uint64_t k, uint64_t n, uint32_t base, uint64_t iter, uint32_t L,
                         const uint64_t *x, const uint64_t *d, size_t nl) {
    return ckpt_write(path, k, n, base, iter, L, x, d, nl);
}

int prp_checkpoint_read(const char *path, uint64_t k, uint64_t n, uint32_t base, uint64_t *iter, uint32_t *L,
                        uint64_t *x, uint64_t *d, size_t nl) {
    return ckpt_read(path, k, n, base, iter, L, x, d, nl);
}

}  // namespace kite

extern "C" int kite_prp_checkpoint_iter(const char *path, uint64_t k, uint64_t n, uint32_t base, uint64_t *iter) {
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    CkptHdr h;
    bool ok = fread(&h, sizeof h, 1, f) == 1 && magic_ok(h.magic) && h.k == k && h.n == n && h.base == base;
    fclose(f);
    if (!ok) return -1;
    *iter = h.iter;
    return 0;
}

extern "C" int kite_prp(uint64_t k, uint64_t n, uint32_t base, const kite_prp_opts *opts, kite_prp_result *res) {
    kite_prp_opts o = opts ? *opts : kite_prp_opts{};
    memset(res, 0, sizeof *res);
    const double t_start = now_s();
    uint32_t B = o.gerbicz_block;
    if (!B) {
        B = (uint32_t)std::sqrt((double)n);
        B = B < 32 ? 32 : B > 2000 ? 2000 : B;
    }
    const uint32_t M = o.gerbicz_blocks ? o.gerbicz_blocks : (B / 4 > 8 ? B / 4 : 8);
    const double maxround = o.max_roundoff > 0 ? o.max_roundoff : 0.40;
    const double ck_secs = o.checkpoint_secs > 0 ? o.checkpoint_secs : 600;
    const double u_secs = o.unverified_secs > 0 ? o.unverified_secs : 60;

    State s;
    if (s.make(k, n, o.fft)) {
        s.drop();
        return -1;
    }
    const size_t nl = kite_limbs(s.ctx);
    std::vector<uint64_t> la(nl), lb(nl), vx(nl), vd(nl), ux, ud;
    uint64_t vi = 0, ui = 0;
    int rc = 0;
    Pending pend;

    power_start(s, k, base, la.data());
    kite_get_limbs(s.ctx, s.x0, vx.data());
    vd = vx;
    {
        uint32_t L = 0, uL = 0;
        if (o.checkpoint && !ckpt_read(o.checkpoint, k, n, base, &vi, &L, vx.data(), vd.data(), nl)) res->resumed = 1;
        if (o.checkpoint_unverified) {
            ux.assign(nl, 0);
            ud.assign(nl, 0);
            if (ckpt_read(o.checkpoint_unverified, k, n, base, &ui, &uL, ux.data(), ud.data(), nl) || ui <= vi) ui = 0;
            else {
                res->resumed = 1;
                if (uL > L) L = uL;
            }
        }
        if (L > kite_fft_len(s.ctx)) {
            kite_opts fo = o.fft;
            fo.fft_len = L;
            s.drop();
            if (s.make(k, n, fo)) {
                s.drop();
                return -1;
            }
            power_start(s, k, base, la.data());
        }
    }
    // A verified checkpoint with d == x (kite_batch_run writes them so; a product from base^k
    // has d != x past iteration 0) restarts the Gerbicz product at that x: it becomes x0, and
    // blocks count from its iteration. Every verified write then restarts the product again, so
    // the file stays in that form and an unverified file chains from it.
    std::vector<uint64_t> x0l;
    uint64_t x0_at = 0;
    if (vi > 0 && vx == vd) {
        x0l = vx;
        x0_at = vi;
        kite_set_limbs(s.ctx, s.x0, x0l.data());
    }
    auto write_verified = [&]() {
        if (!x0l.empty()) {
            x0l = vx;
            vd = vx;
            kite_set_limbs(s.ctx, s.x0, x0l.data());
            if (o.checkpoint_unverified) remove(o.checkpoint_unverified);
        }
        ckpt_write(o.checkpoint, k, n, base, vi, kite_fft_len(s.ctx), vx.data(), vd.data(), nl);
    };
    const uint64_t full = x0_at + (n - x0_at) / B * B;
    double last_ck = now_s(), last_u = last_ck;
    int consecutive_fail = 0;
    res->resumed_at = ui ? ui : vi;
    uint32_t vp = 0, up = 0;
    if (o.on_point) {
        while (vp < o.npoints && o.points[vp] < vi + (vi > 0)) vp++;
        if (vp < o.npoints && o.points[vp] == 0) {  // x_0 = base^k
            kite_set_limbs(s.ctx, s.x, vx.data());
            pending_add(pend, s.ctx, vp++, s.x);
            if (pending_flush(pend, o)) {
                s.drop();
                return KITE_PRP_POINT_FAILED;
            }
        }
        up = vp;
        while (up < o.npoints && o.points[up] < ui + 1) up++;
    }
    for (;;) {
        const bool from_u = ui > vi;
        kite_set_limbs(s.ctx, s.x, from_u ? ux.data() : vx.data());
        kite_set_limbs(s.ctx, s.d, from_u ? ud.data() : vd.data());
        kite_reset_roundoff(s.ctx);
        const uint64_t start = from_u ? ui : vi;
        uint64_t i = start;
        uint32_t pi = from_u ? up : vp;
        ui = 0;
        bool failed = false, round_fail = false;
        pend.clear();
        if (o.abort_flag && *o.abort_flag) {
            if (o.checkpoint) write_verified();
            rc = KITE_PRP_ABORTED;
            goto out;
        }
        {
            uint64_t check_at = start + (uint64_t)B * M;
            if (check_at > full) check_at = full;
            while (i < check_at) {
                if (o.abort_flag && *o.abort_flag) break;
                // one block of B squarings, device-resident unless a fault or a point falls in it
                const uint64_t end = i + B;
                const bool fault_here = o.inject_fault_at && o.inject_fault_at > i && o.inject_fault_at <= end;
                const bool point_here = o.on_point && pi < o.npoints && o.points[pi] > i && o.points[pi] <= end;
                if (!fault_here && !point_here) {
                    kite_square_n(s.ctx, s.x, B);
                    i = end;
                } else {
                    while (i < end) {
                        kite_square(s.ctx, s.x);
                        i++;
                        if (o.inject_fault_at && i == o.inject_fault_at) {
                            kite_debug_corrupt(s.ctx, s.x);
                            o.inject_fault_at = 0;
                        }
                        if (pi < o.npoints && o.points[pi] == i && o.on_point) pending_add(pend, s.ctx, pi++, s.x);
                    }
                }
                kite_copy(s.ctx, s.t, s.d);
                kite_mul(s.ctx, s.d, s.x);
                const double e = kite_max_roundoff(s.ctx);
                if (e > maxround) {
                    round_fail = true;
                    break;
                }
                if (o.progress) o.progress(o.progress_arg, i, n, e);
                if (o.checkpoint_unverified && now_s() - last_u > u_secs && i < check_at) {
                    if (!pend.index.empty() && pending_flush(pend, o)) {
                        rc = KITE_PRP_POINT_FAILED;
                        goto out;
                    }
                    kite_get_limbs(s.ctx, s.x, la.data());
                    kite_get_limbs(s.ctx, s.d, lb.data());
                    ckpt_write(o.checkpoint_unverified, k, n, base, i, kite_fft_len(s.ctx), la.data(), lb.data(), nl);
                    last_u = now_s();
                }
            }
            const bool stopping = o.abort_flag && *o.abort_flag;
            if (stopping && i == start) continue;
            if (!round_fail && i > start) {  // d == x0 * t^(2^B)
                kite_square_n(s.ctx, s.t, B);
                kite_mul(s.ctx, s.t, s.x0);
                if (kite_max_roundoff(s.ctx) > maxround) round_fail = true;
                else failed = !equal_nums(s.ctx, s.t, s.d, la.data(), lb.data());
            }
            const double e = kite_max_roundoff(s.ctx);
            if (e > res->maxerr) res->maxerr = e;
            if ((round_fail || failed) && o.checkpoint_unverified) remove(o.checkpoint_unverified);
            if (round_fail) {
                const uint32_t L = kite_fft_len(s.ctx);
                kite_opts fo = o.fft;
                fo.fft_len = kite_next_fft_len(L + L / 25);
                s.drop();
                if (!fo.fft_len || s.make(k, n, fo)) {
                    rc = -2;
                    goto out;
                }
                power_start(s, k, base, la.data());
                if (!x0l.empty()) kite_set_limbs(s.ctx, s.x0, x0l.data());
                res->roundoff_retries++;
                continue;
            }
            if (failed) {
                res->gerbicz_errors++;
                if (++consecutive_fail > 5) {
                    rc = -3;
                    goto out;
                }
                continue;
            }
            consecutive_fail = 0;
            kite_get_limbs(s.ctx, s.x, vx.data());
            kite_get_limbs(s.ctx, s.d, vd.data());
            vi = i;
            vp = pi;
            if (!pend.index.empty() && pending_flush(pend, o)) {
                rc = KITE_PRP_POINT_FAILED;
                goto out;
            }
            if (o.checkpoint && (now_s() - last_ck > ck_secs || vi == full || stopping)) {
                write_verified();
                last_ck = now_s();
            }
            if (stopping) {
                rc = KITE_PRP_ABORTED;
                goto out;
            }
            if (vi >= full) break;
        }
    }
    // unprotected tail (< B squarings): twice, compared
    for (int attempt = 0;; attempt++) {
        kite_set_limbs(s.ctx, s.x, vx.data());
        kite_set_limbs(s.ctx, s.t, vx.data());
        uint64_t i = full;
        uint32_t pi = vp;
        pend.clear();
        while (i < n) {
            uint64_t next = n;
            if (o.on_point && pi < o.npoints && o.points[pi] > i && o.points[pi] < next) next = o.points[pi];
            kite_square_n(s.ctx, s.x, next - i);
            kite_square_n(s.ctx, s.t, next - i);
            i = next;
            if (o.on_point && pi < o.npoints && o.points[pi] == i) pending_add(pend, s.ctx, pi++, s.x);
        }
        if (equal_nums(s.ctx, s.x, s.t, la.data(), lb.data())) {
            if (!pend.index.empty() && pending_flush(pend, o)) {
                rc = KITE_PRP_POINT_FAILED;
                goto out;
            }
            break;
        }
        if (attempt > 5) {
            rc = -3;
            goto out;
        }
        res->gerbicz_errors++;
    }
    {
        bool prp = la[0] == (uint64_t)base * base;
        for (size_t j = 1; j < nl; j++) prp &= la[j] == 0;
        if (kite_residue_nm1(la.data(), s.ctx->Nl.data(), nl, base)) rc = -4;
        res->res64 = la[0];
        res->is_prp = prp;
    }
    res->fft_len = kite_fft_len(s.ctx);
    {
        const double e = kite_max_roundoff(s.ctx);
        if (e > res->maxerr) res->maxerr = e;
    }
out:
    res->seconds = now_s() - t_start;
    s.drop();
    return rc;
}
