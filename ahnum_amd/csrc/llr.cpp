// LLR (Morrison) test of N = k*2^n - 1 as PRST computes it (reference src/llr.c:1-384).
//
// P is the smallest P >= 1 with Jacobi(P^2 + 4 | N) != 1 (a common factor of 4P(P^2+4) and N is
// reported as a factor). V_k(P, -1) by a Lucas ladder, V_2k = V_k^2 + 2 (k odd), then n - 2
// doublings V -> V^2 - 2 give V_{k 2^(n-1)}: N is prime iff it is 0, else RES64 = low 64 bits.
// Round-off is checked after every block of 1000 doublings (each op measures it on the GPU); a
// block over the limit is redone from the last verified state at a longer FFT. A composite result
// is confirmed by a second run at the next FFT length, and a third if they disagree.
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

uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((unsigned __int128)a * b % m); }


int jacobi(uint64_t a, uint64_t m) {  // m odd
    int r = 1;
    a %= m;
    while (a) {
        while (!(a & 1)) {
            a >>= 1;
            if ((m & 7) == 3 || (m & 7) == 5) r = -r;
        }
        std::swap(a, m);
        if ((a & 3) == 3 && (m & 3) == 3) r = -r;
        a %= m;
    }
    return m == 1 ? r : 0;
}

int jacobi_N(uint64_t a, uint64_t k, uint64_t n) {  // (a | N) by reciprocity against N mod a
    int r = 1;
    const uint64_t n8 = n_mod(k, n, 8);
    while (a && !(a & 1)) {
        a >>= 1;
        if (n8 == 3 || n8 == 5) r = -r;
    }
    if (a == 1) return r;
    if ((a & 3) == 3 && (n8 & 3) == 3) r = -r;
    return r * jacobi(n_mod(k, n, a), a);
}

uint64_t gcd64(uint64_t a, uint64_t b) {
    while (b) {
        uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

// KITELLR1 checkpoint (src/llr.c:92-146)
struct LlrHdr {
    char magic[8];
    uint64_t k, n, P, iter, nlimbs, fft_len, run;
    uint64_t prev[2];
};

uint64_t fnv1a(const void *p, size_t len, uint64_t h) {
    const uint8_t *b = static_cast<const uint8_t *>(p);
    for (size_t i = 0; i < len; i++) h = (h ^ b[i]) * 0x100000001b3ull;
    return h;
}

int ckpt_write(const char *path, const LlrHdr &h, const uint64_t *x) {
    std::string tmp = std::string(path) + ".tmp";
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) return -1;
    uint64_t s = fnv1a(&h, sizeof h, 0xcbf29ce484222325ull);
    s = fnv1a(x, h.nlimbs * 8, s);
    bool ok = fwrite(&h, sizeof h, 1, f) == 1 && fwrite(x, 8, h.nlimbs, f) == h.nlimbs && fwrite(&s, 8, 1, f) == 1;
    ok &= fclose(f) == 0;
    if (!ok || rename(tmp.c_str(), path)) {
        remove(tmp.c_str());
        return -1;
    }
    return 0;
}

int ckpt_read(const char *path, LlrHdr *h, uint64_t *x, size_t nl) {
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    uint64_t s = 0;
    bool ok = fread(h, sizeof *h, 1, f) == 1 && !memcmp(h->magic, "KITELLR1", 8) && h->nlimbs == nl &&
              fread(x, 8, nl, f) == nl && fread(&s, 8, 1, f) == 1;
    fclose(f);
    if (!ok) return -1;
    return fnv1a(x, nl * 8, fnv1a(h, sizeof *h, 0xcbf29ce484222325ull)) == s ? 0 : -1;
}

LlrHdr make_hdr(uint64_t k, uint64_t n, uint32_t P, uint64_t iter, size_t nl, uint32_t L, int run,
                const uint64_t *prev) {
    LlrHdr h{};
    memcpy(h.magic, "KITELLR1", 8);
    h.k = k;
    h.n = n;
    h.P = P;
    h.iter = iter;
    h.nlimbs = nl;
    h.fft_len = L;
    h.run = (uint64_t)run;
    h.prev[0] = prev[0];
    h.prev[1] = prev[1];
    return h;
}

bool all_zero(const uint64_t *l, size_t nl) {
    for (size_t i = 0; i < nl; i++)
        if (l[i]) return false;
    return true;
}

// One complete run from the checkpoint (if any) to the end (src/llr.c:192-305).
int llr_run(uint64_t k, uint64_t n, uint32_t P, const kite_llr_opts *o, uint32_t fft_len, int run,
            const uint64_t *prev, uint64_t *out, kite_llr_result *res) {
    kite_opts fo = o->fft;
    if (fft_len) fo.fft_len = fft_len;
    const double maxround = o->max_roundoff > 0 ? o->max_roundoff : 0.40;
    const double ck_secs = o->checkpoint_secs > 0 ? o->checkpoint_secs : 600;
    const uint64_t total = n - 1;
    bool fault_due = o->inject_fault_at && (((o->inject_fault_runs ? o->inject_fault_runs : 1u) >> run) & 1);
    bool roundoff_due = run == 0 && (o->inject_roundoff_at || o->inject_roundoff_ladder);
    for (int attempt = 0;; attempt++) {
        kite_ctx *ctx = kite_ctx_new(k, n, -1, &fo);
        if (!ctx) return -1;
        kite_set_roundoff_check(ctx, 1);
        const size_t nl = kite_limbs(ctx);
        std::vector<uint64_t> vx(nl);
        kite_num *x = kite_num_new(ctx), *v1 = kite_num_new(ctx), *cP = kite_num_new(ctx), *c2 = kite_num_new(ctx),
                 *t = kite_num_new(ctx);
        kite_set_u64(ctx, cP, P);
        kite_set_u64(ctx, c2, 2);
        uint64_t vi = 0;
        LlrHdr h;
        int rc = 0;
        bool round_fail = false, have_state = false;
        if (o->checkpoint && ckpt_read(o->checkpoint, &h, vx.data(), nl) == 0 && h.k == k && h.n == n && h.P == P &&
            h.run == (uint64_t)run && h.iter <= total) {
            vi = h.iter;
            if (!res->resumed) {
                res->resumed = 1;
                res->resumed_at = vi;
            }
            kite_set_limbs(ctx, x, vx.data());
            have_state = true;
        } else {
            kite_reset_roundoff(ctx);
            llr_lucas_vk(ctx, k, x, v1, cP, c2, t);
            const double e = kite_max_roundoff(ctx);
            if (e > res->maxerr) res->maxerr = e;
            if (e > maxround || (roundoff_due && o->inject_roundoff_ladder)) {
                roundoff_due = false;
                round_fail = true;
            } else {
                kite_get_limbs(ctx, x, vx.data());
                have_state = true;
            }
        }
        double last_ck = now_s();
        uint64_t i = vi;
        const uint64_t B = 1000;
        while (!round_fail && i < total) {
            kite_reset_roundoff(ctx);
            const uint64_t stop = i + B < total ? i + B : total;
            while (i < stop) {  // V_2k = V_k^2 + 2 (k odd), then x -> x^2 - 2, both inside the squaring
                uint64_t end = stop;
                if (fault_due && o->inject_fault_at > i && o->inject_fault_at < end) end = o->inject_fault_at;
                if (i == 0) end = 1;
                square_add(ctx, x, end - i, i == 0 ? 2 : -2);
                i = end;
                if (fault_due && i == o->inject_fault_at) {
                    kite_debug_corrupt(ctx, x);
                    fault_due = false;
                }
            }
            const double e = kite_max_roundoff(ctx);
            if (e > res->maxerr) res->maxerr = e;
            if (e > maxround || (roundoff_due && o->inject_roundoff_at && stop >= o->inject_roundoff_at)) {
                roundoff_due = false;
                round_fail = true;
                break;
            }
            kite_get_limbs(ctx, x, vx.data());
            vi = i;
            if (o->progress) o->progress(o->progress_arg, vi, total, e);
            const bool stop_now = o->abort_flag && *o->abort_flag;
            if (o->checkpoint && (stop_now || now_s() - last_ck > ck_secs)) {
                ckpt_write(o->checkpoint, make_hdr(k, n, P, vi, nl, kite_fft_len(ctx), run, prev), vx.data());
                last_ck = now_s();
            }
            if (stop_now) {
                rc = KITE_LLR_ABORTED;
                break;
            }
        }
        const uint32_t L = kite_fft_len(ctx);
        if (!round_fail && rc == 0) {
            memcpy(out, vx.data(), nl * 8);
            res->fft_len = L;
        }
        kite_num_free(ctx, x);
        kite_num_free(ctx, v1);
        kite_num_free(ctx, cP);
        kite_num_free(ctx, c2);
        kite_num_free(ctx, t);
        kite_ctx_free(ctx);
        if (!round_fail) return rc;
        if (o->checkpoint && have_state)
            ckpt_write(o->checkpoint, make_hdr(k, n, P, vi, nl, L, run, prev), vx.data());
        fo.fft_len = kite_next_fft_len(L + L / 25);
        res->roundoff_retries++;
        if (!fo.fft_len || attempt > 8) return -2;
    }
}

}  // namespace

namespace kite {

uint64_t n_mod(uint64_t k, uint64_t n, uint64_t m) {
    uint64_t r = 1 % m, b = 2 % m;
    for (uint64_t e = n; e; e >>= 1, b = mulmod(b, b, m))
        if (e & 1) r = mulmod(r, b, m);
    return (mulmod(k % m, r, m) + m - 1 % m) % m;
}

int llr_checkpoint_write(const char *path, uint64_t k, uint64_t n, const LlrCheckpoint &c, const uint64_t *x,
                         size_t nl) {
    return ckpt_write(path, make_hdr(k, n, c.P, c.iter, nl, c.L, c.run, c.prev), x);
}

int llr_checkpoint_read(const char *path, uint64_t k, uint64_t n, LlrCheckpoint *c, uint64_t *x, size_t nl) {
    LlrHdr h;
    if (ckpt_read(path, &h, x, nl) || h.k != k || h.n != n || h.run > 2 || h.iter > n - 1) return -1;
    c->iter = h.iter;
    c->P = (uint32_t)h.P;
    c->L = (uint32_t)h.fft_len;
    c->run = (int)h.run;
    c->prev[0] = h.prev[0];
    c->prev[1] = h.prev[1];
    return 0;
}

// V_k(P, -1) by the ladder (V_m, V_{m+1}) over the bits of k (src/llr.c:150-184).
void llr_lucas_vk(kite_ctx *ctx, uint64_t k, kite_num *x, kite_num *v1, kite_num *cP, kite_num *c2, kite_num *t) {
    kite_set_u64(ctx, x, 2);
    kite_copy(ctx, v1, cP);
    uint64_t m = 0;
    int top = 63;
    while (top > 0 && !((k >> top) & 1)) top--;
    for (int b = top; b >= 0; b--) {
        const bool odd = m & 1;
        kite_copy(ctx, t, x);
        kite_mul(ctx, t, v1);  // V_{2m+1} = V_m V_{m+1} - P Q^m
        if (odd) kite_add(ctx, t, cP);
        else kite_sub(ctx, t, cP);
        if ((k >> b) & 1) {
            kite_square(ctx, v1);  // V_{2m+2} = V_{m+1}^2 - 2 Q^{m+1}
            if (odd) kite_sub(ctx, v1, c2);
            else kite_add(ctx, v1, c2);
            kite_copy(ctx, x, t);
            m = 2 * m + 1;
        } else {
            kite_square(ctx, x);  // V_{2m} = V_m^2 - 2 Q^m
            if (odd) kite_add(ctx, x, c2);
            else kite_sub(ctx, x, c2);
            kite_copy(ctx, v1, t);
            m = 2 * m;
        }
    }
}

}  // namespace kite

extern "C" int kite_llr_choose_P(uint64_t k, uint64_t n, uint32_t *P, uint64_t *factor) {
    uint32_t p = 1;
    while (jacobi_N((uint64_t)p * p + 4, k, n) == 1)
        if (++p > 1000000) return -1;
    *P = p;
    const uint64_t d = 4ull * p * ((uint64_t)p * p + 4);
    const uint64_t g = gcd64(d, n_mod(k, n, d));
    *factor = g > 1 ? g : 0;
    return 0;
}

extern "C" int kite_llr_checkpoint_iter(const char *path, uint64_t k, uint64_t n, uint64_t *iter, int *run) {
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    LlrHdr h;
    const bool ok = fread(&h, sizeof h, 1, f) == 1 && !memcmp(h.magic, "KITELLR1", 8) && h.k == k && h.n == n;
    fclose(f);
    if (!ok) return -1;
    *iter = h.iter;
    if (run) *run = (int)h.run;
    return 0;
}

extern "C" int kite_llr(uint64_t k, uint64_t n, const kite_llr_opts *opts, kite_llr_result *res) {
    kite_llr_opts o = opts ? *opts : kite_llr_opts{};
    memset(res, 0, sizeof *res);
    const double t0 = now_s();
    if (n < 3 || !(k & 1)) return -1;
    uint32_t P;
    uint64_t factor;
    if (kite_llr_choose_P(k, n, &P, &factor)) return -1;
    res->P = P;
    if (factor) {
        res->factor = factor;
        res->seconds = now_s() - t0;
        return 0;
    }
    kite_ctx *probe = kite_ctx_new(k, n, -1, &o.fft);
    if (!probe) return -1;
    const size_t nl = kite_limbs(probe);
    const uint32_t L0 = kite_fft_len(probe);
    kite_ctx_free(probe);
    uint64_t res64[3] = {0, 0, 0};
    int zero[3] = {0, 0, 0}, rc = 0, run = 0, fin = 0;
    uint32_t L = 0;
    {
        LlrHdr h;
        std::vector<uint64_t> tmp(nl);
        if (o.checkpoint && ckpt_read(o.checkpoint, &h, tmp.data(), nl) == 0 && h.k == k && h.n == n && h.P == P &&
            h.run < 3) {
            run = (int)h.run;
            res64[0] = h.prev[0];
            res64[1] = h.prev[1];
            zero[0] = run > 0 && res64[0] == 0;
            zero[1] = run > 1 && res64[1] == 0;
            if (run > 0) L = (uint32_t)h.fft_len;
        }
    }
    std::vector<uint64_t> out(nl);
    auto agree = [&](int a, int b) { return res64[a] == res64[b] && zero[a] == zero[b]; };
    for (; run < 3; run++) {
        if (run > 0) {
            res->confirm_runs = (uint32_t)run;
            if (o.on_confirm) o.on_confirm(o.progress_arg, run);
        }
        rc = llr_run(k, n, P, &o, L, run, res64, out.data(), res);
        if (rc) break;
        res64[run] = out[0];
        zero[run] = all_zero(out.data(), nl);
        if (run == 0 && zero[0]) break;  // prime
        if (run == 1 && agree(1, 0)) break;
        if (run == 2) {
            if (agree(2, 1)) fin = 1;
            else if (agree(2, 0)) fin = 0;
            else rc = -3;
            break;
        }
        const uint32_t base = res->fft_len ? res->fft_len : L0;
        L = kite_next_fft_len(base + 1);
        if (!L) {
            rc = -2;
            break;
        }
    }
    if (rc == 0) {
        res->is_prime = zero[fin];
        res->res64 = res64[fin];
    }
    if (o.checkpoint && rc != KITE_LLR_ABORTED) remove(o.checkpoint);
    res->seconds = now_s() - t0;
    return rc;
}
