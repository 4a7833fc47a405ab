// kite: Fermat PRP / LLR tests, timing and tuning for Riesel numbers k*2^n-1 on the MI355X
// engine. Same commands, options and output lines as the reference CLI (tools/kite.c:1-330):
//
//   kite prp  K N | "K*2^N-1"  [--expect HEX] [--base B] [--ckpt FILE] [--ckpt-secs S]
//                                 [--uckpt FILE] [--uckpt-secs S] [--block B] [--fault ITER]
//   kite llr  K N [--expect HEX|prime] [--ckpt FILE]
//   kite time K N [--iters N]
//   kite throughput K N [--configs 10x1,5x2,...] [--secs S] [--batch 1]
//   kite tune LMIN LMAX [--out FILE]
// Common options: --backend hip --fft L --threads T (ignored) --order 0..3 (shape variant)
//
// `throughput` runs T concurrent tests, each with its own context and HIP stream on the GPU,
// from T host threads (the second number of each config, threads per test, is ignored); with
// --batch 1 the T tests form one batch instead (include/kite_batch.h: shared launches per pass).
#include <atomic>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <thread>
#include <vector>

#include "kite.h"
#include "kite_batch.h"

static double now_s() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

static kite_backend parse_backend(const char *s) {
    if (!strcmp(s, "hip")) return KITE_BACKEND_HIP;
    if (!strcmp(s, "ref")) return KITE_BACKEND_REF;
    if (!strcmp(s, "c2p")) return KITE_BACKEND_C2P;
    if (!strcmp(s, "neon")) return KITE_BACKEND_NEON;
    if (!strcmp(s, "sme")) return KITE_BACKEND_SME;
    return KITE_BACKEND_AUTO;
}

static int parse_number(int argc, char **argv, int i, uint64_t *k, uint64_t *n) {
    if (i < argc && strchr(argv[i], '*')) {
        unsigned long long kk, nn;
        if (sscanf(argv[i], "%llu*2^%llu-1", &kk, &nn) != 2) return 0;
        *k = kk;
        *n = nn;
        return 1;
    }
    if (i + 1 >= argc) return 0;
    *k = strtoull(argv[i], nullptr, 10);
    *n = strtoull(argv[i + 1], nullptr, 10);
    return 2;
}

static void progress(void *, uint64_t i, uint64_t n, double maxerr) {
    static double last;
    double t = now_s();
    if (t - last < 5) return;
    last = t;
    fprintf(stderr, "\r%5.1f%%  maxerr %.4f", 100.0 * i / n, maxerr);
}

static void usage(const char *p) {
    fprintf(stderr,
            "usage: %s prp K N | \"K*2^N-1\" [--expect HEX] [--base B] [--ckpt FILE]\n"
            "       %s llr K N | \"K*2^N-1\" [--expect HEX|prime] [--ckpt FILE]\n"
            "       %s time K N [--iters N]\n"
            "       %s throughput K N [--configs 10x1,5x2,...] [--secs S] [--batch 1]\n"
            "       %s tune LMIN LMAX [--out FILE]\n"
            "options: --backend hip --fft L --threads T\n"
            "         --order 0..3 (shape variant, default from the tuning file)\n",
            p, p, p, p, p);
}

// T concurrent tests for `secs`; returns squarings/s summed over tests.
static double throughput(uint64_t k, uint64_t n, kite_opts o, int tests, double secs) {
    std::atomic<int> ready{0};
    std::atomic<bool> go{false};
    std::vector<uint64_t> iters(tests, 0);
    std::vector<double> took(tests, 1);
    std::vector<std::thread> th;
    for (int t = 0; t < tests; t++) {
        th.emplace_back([&, t] {
            kite_ctx *ctx = kite_ctx_new(k, n, -1, &o);
            kite_num *x = ctx ? kite_num_new(ctx) : nullptr;
            if (x) {
                kite_set_u64(ctx, x, 3);
                kite_square_n(ctx, x, 5);
                kite_max_roundoff(ctx);  // waits for the warm-up
            }
            ready++;
            while (!go.load()) std::this_thread::yield();
            if (x) {
                const double t0 = now_s(), dl = t0 + secs;
                uint64_t it = 0;
                while (now_s() < dl) {
                    kite_square_n(ctx, x, 64);
                    kite_max_roundoff(ctx);
                    it += 64;
                }
                took[t] = now_s() - t0;
                iters[t] = it;
                kite_num_free(ctx, x);
            }
            kite_ctx_free(ctx);
        });
    }
    while (ready.load() < tests) std::this_thread::yield();
    go = true;
    for (auto &x : th) x.join();
    double tot = 0;
    for (int t = 0; t < tests; t++) tot += iters[t] / took[t];
    return tot;
}

// T tests of the number as one batch for `secs`, in square calls of 1000 steps (a Gerbicz block,
// replayed from the batch's captured graph); returns squarings/s summed over tests, 0 on failure.
static double throughput_batched(uint64_t k, uint64_t n, uint32_t L, int tests, double secs) {
    kite_batch *b = kite_batch_new();
    if (!b) return 0;
    for (int t = 0; t < tests; t++) {
        const int s = kite_batch_add(b, k, n, L, 1);
        if (s < 0) {
            kite_batch_free(b);
            return 0;
        }
        kite_set_u64(kite_batch_ctx(b, s), kite_batch_num(b, s, 0), 3 + (uint64_t)t);
    }
    const int cap = kite_batch_capacity(b);
    std::vector<int32_t> which(cap, 0);
    std::vector<uint32_t> cnt(cap, 1000);
    bool ok = kite_batch_square(b, which.data(), cnt.data()) == 0 && kite_batch_sync(b) == 0;  // warm-up
    uint64_t it = 0;
    const double t0 = now_s();
    while (ok && now_s() - t0 < secs) {
        ok = kite_batch_square(b, which.data(), cnt.data()) == 0 && kite_batch_sync(b) == 0;
        it += 1000;
    }
    const double dt = now_s() - t0;
    kite_batch_free(b);
    return ok ? (double)it * tests / dt : 0;
}

int main(int argc, char **argv) {
    if (argc < 3) {
        usage(argv[0]);
        return 2;
    }
    const char *cmd = argv[1];
    uint64_t k = 0, n = 0;
    const int used = parse_number(argc, argv, 2, &k, &n);
    if (!used) {
        usage(argv[0]);
        return 2;
    }
    kite_opts o{};
    kite_prp_opts po{};
    const char *expect = nullptr, *out = nullptr, *cfg = "1x1,2x1,4x1,8x1,16x1";
    uint64_t iters = 1000;
    uint32_t base = 3;
    double secs = 3;
    bool batched = false;
    for (int i = 2 + used; i + 1 < argc; i += 2) {
        const char *a = argv[i], *v = argv[i + 1];
        if (!strcmp(a, "--backend")) o.backend = parse_backend(v);
        else if (!strcmp(a, "--fft")) o.fft_len = (uint32_t)atoi(v);
        else if (!strcmp(a, "--threads")) o.threads = (uint32_t)atoi(v);
        else if (!strcmp(a, "--order")) o.plan_order = 1 + (uint32_t)atoi(v);
        else if (!strcmp(a, "--iters")) iters = strtoull(v, nullptr, 10);
        else if (!strcmp(a, "--expect")) expect = v;
        else if (!strcmp(a, "--base")) base = (uint32_t)atoi(v);
        else if (!strcmp(a, "--ckpt")) po.checkpoint = v;
        else if (!strcmp(a, "--ckpt-secs")) po.checkpoint_secs = atof(v);
        else if (!strcmp(a, "--uckpt")) po.checkpoint_unverified = v;
        else if (!strcmp(a, "--uckpt-secs")) po.unverified_secs = atof(v);
        else if (!strcmp(a, "--block")) po.gerbicz_block = (uint32_t)atoi(v);
        else if (!strcmp(a, "--fault")) po.inject_fault_at = strtoull(v, nullptr, 10);
        else if (!strcmp(a, "--configs")) cfg = v;
        else if (!strcmp(a, "--secs")) secs = atof(v);
        else if (!strcmp(a, "--batch")) batched = atoi(v) != 0;
        else if (!strcmp(a, "--out")) out = v;
        else {
            fprintf(stderr, "unknown option %s\n", a);
            return 2;
        }
    }

    if (!strcmp(cmd, "prp")) {
        po.fft = o;
        po.progress = progress;
        kite_prp_result r;
        const int rc = kite_prp(k, n, base, &po, &r);
        fprintf(stderr, "\r                                        \r");
        if (rc) {
            printf("%" PRIu64 "*2^%" PRIu64 "-1  ERROR %d\n", k, n, rc);
            return 1;
        }
        printf("%" PRIu64 "*2^%" PRIu64 "-1 is %s  RES64 %016" PRIX64
               "  FFT %u  %.1f s (%.1f us/iter)  maxerr %.4f  gerbicz_errors %u  fft_retries %u%s\n",
               k, n, r.is_prp ? "a probable prime!" : "composite", r.res64, r.fft_len, r.seconds,
               r.seconds * 1e6 / n, r.maxerr, r.gerbicz_errors, r.roundoff_retries, r.resumed ? "  (resumed)" : "");
        if (expect) {
            const bool ok = strtoull(expect, nullptr, 16) == r.res64;
            printf("%s expected %s\n", ok ? "MATCH" : "MISMATCH", expect);
            return !ok;
        }
        return 0;
    }

    if (!strcmp(cmd, "llr")) {
        kite_llr_opts lo{};
        lo.fft = o;
        lo.checkpoint = po.checkpoint;
        lo.checkpoint_secs = po.checkpoint_secs;
        lo.progress = progress;
        kite_llr_result r;
        const int rc = kite_llr(k, n, &lo, &r);
        fprintf(stderr, "\r                                        \r");
        if (rc) {
            printf("%" PRIu64 "*2^%" PRIu64 "-1  ERROR %d\n", k, n, rc);
            return 1;
        }
        char what[64];
        if (r.factor) snprintf(what, sizeof what, "divisible by %" PRIu64, r.factor);
        else if (r.is_prime) snprintf(what, sizeof what, "prime!");
        else snprintf(what, sizeof what, "composite  RES64 %016" PRIX64, r.res64);
        printf("%" PRIu64 "*2^%" PRIu64 "-1 is %s  P = %u  FFT %u  %.1f s  maxerr %.4f  confirm_runs %u  fft_retries %u%s\n",
               k, n, what, r.P, r.fft_len, r.seconds, r.maxerr, r.confirm_runs, r.roundoff_retries,
               r.resumed ? "  (resumed)" : "");
        if (expect) {
            const bool ok = !strcmp(expect, "prime") ? r.is_prime && !r.factor
                                                     : !r.is_prime && !r.factor && strtoull(expect, nullptr, 16) == r.res64;
            printf("%s expected %s\n", ok ? "MATCH" : "MISMATCH", expect);
            return !ok;
        }
        return 0;
    }

    if (!strcmp(cmd, "time")) {
        kite_ctx *ctx = kite_ctx_new(k, n, -1, &o);
        if (!ctx) {
            fprintf(stderr, "cannot create context\n");
            return 1;
        }
        kite_num *x = kite_num_new(ctx);
        printf("%" PRIu64 "*2^%" PRIu64 "-1  FFT %u (%.2f bits/word)  backend %s\n", k, n, kite_fft_len(ctx),
               (double)n / kite_fft_len(ctx), kite_backend_name(ctx));
        kite_set_u64(ctx, x, 3);
        kite_square_n(ctx, x, 10);
        kite_max_roundoff(ctx);
        const double t0 = now_s();
        kite_square_n(ctx, x, iters);
        const double e = kite_max_roundoff(ctx);  // waits for the squarings
        const double dt = now_s() - t0;
        printf("%.3f ms/iter  (%.1f us)  maxerr %.4f\n", dt * 1e3 / iters, dt * 1e6 / iters, e);
        kite_num_free(ctx, x);
        kite_ctx_free(ctx);
        return 0;
    }

    if (!strcmp(cmd, "throughput")) {
        kite_ctx *c0 = kite_ctx_new(k, n, -1, &o);
        if (!c0) {
            fprintf(stderr, "cannot create context\n");
            return 1;
        }
        o.fft_len = kite_fft_len(c0);
        kite_ctx_free(c0);
        printf("%" PRIu64 "*2^%" PRIu64 "-1  FFT %u\n", k, n, o.fft_len);
        char buf[256];
        snprintf(buf, sizeof buf, "%s", cfg);
        double best = 0;
        char bestc[32] = "";
        for (char *tok = strtok(buf, ","); tok; tok = strtok(nullptr, ",")) {
            int T = 1, S = 1;
            if (sscanf(tok, "%dx%d", &T, &S) < 1 || T < 1 || T > (batched ? 4096 : 256)) continue;
            const double r = batched ? throughput_batched(k, n, o.fft_len, T, secs) : throughput(k, n, o, T, secs);
            printf("%-6s %10.0f sq/s  %8.1f tests/hour  (%.1f us/iter per test)\n", tok, r, r * 3600.0 / n, 1e6 * T / r);
            fflush(stdout);
            if (r > best) {
                best = r;
                snprintf(bestc, sizeof bestc, "%s", tok);
            }
        }
        printf("best: %s  %.1f tests/hour\n", bestc, best * 3600.0 / n);
        return 0;
    }

    if (!strcmp(cmd, "tune")) {
        // every usable length in [LMIN, LMAX] in each of four shape variants, interleaved; the
        // fastest variant is written as "L us order" (the reference's tuning-file format)
        FILE *f = out ? fopen(out, "a") : nullptr;
        for (uint32_t L = ((uint32_t)k + 63) / 64 * 64; L <= n; L += 64) {
            kite_ctx *c[4] = {};
            kite_num *x[4] = {};
            double best[4] = {1e30, 1e30, 1e30, 1e30};
            bool any = false;
            for (int ord = 0; ord < 4; ord++) {
                kite_opts lo = o;
                lo.fft_len = L;
                lo.plan_order = 1 + ord;
                c[ord] = kite_ctx_new(1, (uint64_t)L * 12, -1, &lo);
                if (!c[ord]) continue;
                x[ord] = kite_num_new(c[ord]);
                kite_set_u64(c[ord], x[ord], 3);
                kite_square_n(c[ord], x[ord], 3);
                any = true;
            }
            for (int rep = 0; any && rep < 3; rep++)
                for (int ord = 0; ord < 4; ord++) {
                    if (!c[ord]) continue;
                    uint64_t it = 0;
                    double t0 = now_s(), t1;
                    do {
                        kite_square_n(c[ord], x[ord], 16);
                        kite_max_roundoff(c[ord]);
                        it += 16;
                    } while ((t1 = now_s()) - t0 < 0.05);
                    const double us = (t1 - t0) * 1e6 / it;
                    if (us < best[ord]) best[ord] = us;
                }
            if (any) {
                int bo = 0;
                for (int ord = 1; ord < 4; ord++)
                    if (best[ord] < best[bo]) bo = ord;
                printf("%u %.2f %d\n", L, best[bo], bo);
                if (f) fprintf(f, "%u %.2f %d\n", L, best[bo], bo);
                fflush(stdout);
            }
            for (int ord = 0; ord < 4; ord++) {
                if (x[ord]) kite_num_free(c[ord], x[ord]);
                if (c[ord]) kite_ctx_free(c[ord]);
            }
        }
        if (f) fclose(f);
        return 0;
    }

    usage(argv[0]);
    return 2;
}
