// kite: Fermat PRP / LLR tests, timing, throughput, round-off sweeps and tuning for Riesel numbers
// k*2^n-1 on the MI355X engine. Commands, options and result lines are the reference CLI's
// (tools/kite.c: result lines :180-184, :210-213, :238-239, :262-263, tuning lines :315-316;
// sweep lines tools/sweep_err.c:33-34), so scripts that parse it keep working:
//
//   kite prp  K N | "K*2^N-1"  [--expect HEX] [--base B] [--ckpt FILE] [--ckpt-secs S]
//                                 [--uckpt FILE] [--uckpt-secs S] [--block B] [--fault ITER]
//   kite llr  K N [--expect HEX|prime] [--ckpt FILE]
//   kite time K N [--iters N]
//   kite throughput K N [--configs 10x1,5x2,...] [--secs S] [--batch 1]
//   kite sweep L K BPW_LO BPW_HI STEP [--iters N]
//   kite tune LMIN LMAX [--out FILE]
// Common options: --backend hip --fft L --threads T (ignored) --order 0..3 (shape variant)
//
// On the GPU a "thread" of the reference's throughput configs has no meaning: a config TxS runs T
// tests, all driven from this one host thread (each on its own context and stream, or with
// --batch 1 as one batch advanced by shared launches).
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <functional>
#include <string>
#include <vector>

#include "kite.h"
#include "kite_batch.h"

namespace {

double seconds_now() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

// Everything a command can be given.
struct Invocation {
    std::string command;
    uint64_t k = 0, n = 0;       // the number (or, for tune, LMIN and LMAX)
    std::vector<std::string> rest;  // further positional words (sweep)
    kite_opts engine{};
    kite_prp_opts prp{};
    std::string expect, out, configs = "1x1,2x1,4x1,8x1,16x1";
    uint64_t iters = 0;  // 0: the command's default
    uint32_t base = 3;
    double secs = 3;
    bool batched = false;
};

kite_backend backend_named(const std::string &s) {
    struct {
        const char *name;
        kite_backend b;
    } const table[] = {{"hip", KITE_BACKEND_HIP}, {"ref", KITE_BACKEND_REF}, {"c2p", KITE_BACKEND_C2P},
                       {"neon", KITE_BACKEND_NEON}, {"sme", KITE_BACKEND_SME}};
    for (const auto &e : table)
        if (s == e.name) return e.b;
    return KITE_BACKEND_AUTO;
}

// Option name -> what it sets (every option takes one value).
using Setter = std::function<void(Invocation &, const char *)>;
struct OptionSpec {
    const char *name;
    Setter set;
};

const std::vector<OptionSpec> &options() {
    static const std::vector<OptionSpec> table = {
        {"--backend", [](Invocation &v, const char *s) { v.engine.backend = backend_named(s); }},
        {"--fft", [](Invocation &v, const char *s) { v.engine.fft_len = (uint32_t)atoi(s); }},
        {"--threads", [](Invocation &v, const char *s) { v.engine.threads = (uint32_t)atoi(s); }},
        {"--order", [](Invocation &v, const char *s) { v.engine.plan_order = 1 + (uint32_t)atoi(s); }},
        {"--iters", [](Invocation &v, const char *s) { v.iters = strtoull(s, nullptr, 10); }},
        {"--expect", [](Invocation &v, const char *s) { v.expect = s; }},
        {"--base", [](Invocation &v, const char *s) { v.base = (uint32_t)atoi(s); }},
        {"--ckpt", [](Invocation &v, const char *s) { v.prp.checkpoint = s; }},
        {"--ckpt-secs", [](Invocation &v, const char *s) { v.prp.checkpoint_secs = atof(s); }},
        {"--uckpt", [](Invocation &v, const char *s) { v.prp.checkpoint_unverified = s; }},
        {"--uckpt-secs", [](Invocation &v, const char *s) { v.prp.unverified_secs = atof(s); }},
        {"--block", [](Invocation &v, const char *s) { v.prp.gerbicz_block = (uint32_t)atoi(s); }},
        {"--fault", [](Invocation &v, const char *s) { v.prp.inject_fault_at = strtoull(s, nullptr, 10); }},
        {"--configs", [](Invocation &v, const char *s) { v.configs = s; }},
        {"--secs", [](Invocation &v, const char *s) { v.secs = atof(s); }},
        {"--batch", [](Invocation &v, const char *s) { v.batched = atoi(s) != 0; }},
        {"--out", [](Invocation &v, const char *s) { v.out = s; }},
    };
    return table;
}

// "K*2^N-1" in one word, or K and N in two; returns the words used (0: not a number).
int read_number(const std::vector<std::string> &w, size_t at, uint64_t *k, uint64_t *n) {
    if (at < w.size() && w[at].find('*') != std::string::npos) {
        unsigned long long kk = 0, nn = 0;
        if (sscanf(w[at].c_str(), "%llu*2^%llu-1", &kk, &nn) != 2) return 0;
        *k = kk;
        *n = nn;
        return 1;
    }
    if (at + 1 >= w.size()) return 0;
    *k = strtoull(w[at].c_str(), nullptr, 10);
    *n = strtoull(w[at + 1].c_str(), nullptr, 10);
    return 2;
}

// false (and a message) on a malformed command line
bool parse(int argc, char **argv, Invocation &v) {
    std::vector<std::string> words(argv + 1, argv + argc);
    if (words.size() < 2) return false;
    v.command = words[0];
    size_t at = 1;
    if (v.command == "sweep") {  // L K BPW_LO BPW_HI STEP
        while (at < words.size() && words[at].rfind("--", 0) != 0) v.rest.push_back(words[at++]);
        if (v.rest.size() != 5) return false;
    } else {
        const int used = read_number(words, 1, &v.k, &v.n);
        if (!used) return false;
        at += (size_t)used;
    }
    for (; at < words.size(); at += 2) {
        const OptionSpec *spec = nullptr;
        for (const OptionSpec &o : options())
            if (words[at] == o.name) spec = &o;
        if (!spec || at + 1 >= words.size()) {
            fprintf(stderr, "unknown option %s\n", words[at].c_str());
            return false;
        }
        spec->set(v, words[at + 1].c_str());
    }
    return true;
}

void usage(const char *p) {
    fprintf(stderr,
            "usage: %s prp K N | \"K*2^N-1\" [--expect HEX] [--base B] [--ckpt FILE]\n"
            "       %s llr K N | \"K*2^N-1\" [--expect HEX|prime] [--ckpt FILE]\n"
            "       %s time K N [--iters N]\n"
            "       %s throughput K N [--configs 10x1,5x2,...] [--secs S] [--batch 1]\n"
            "       %s sweep L K BPW_LO BPW_HI STEP [--iters N]\n"
            "       %s tune LMIN LMAX [--out FILE]\n"
            "options: --backend hip --fft L --threads T\n"
            "         --order 0..3 (shape variant, default from the tuning file)\n",
            p, p, p, p, p, p);
}

// A progress reading on stderr at most every 5 s, wiped before the result line.
struct Progress {
    double last = 0;
    static void report(void *arg, uint64_t i, uint64_t n, double maxerr) {
        Progress &p = *static_cast<Progress *>(arg);
        const double t = seconds_now();
        if (t - p.last < 5) return;
        p.last = t;
        fprintf(stderr, "\r%5.1f%%  maxerr %.4f", 100.0 * (double)i / (double)n, maxerr);
    }
    void wipe() const { fprintf(stderr, "\r                                        \r"); }
};

// MATCH / MISMATCH line of --expect; the exit code
int verdict(const std::string &want, bool ok) {
    printf("%s expected %s\n", ok ? "MATCH" : "MISMATCH", want.c_str());
    return ok ? 0 : 1;
}

int cmd_prp(Invocation &v) {
    Progress pr;
    kite_prp_opts o = v.prp;
    o.fft = v.engine;
    o.progress = Progress::report;
    o.progress_arg = &pr;
    kite_prp_result r;
    const int rc = kite_prp(v.k, v.n, v.base, &o, &r);
    pr.wipe();
    if (rc) {
        printf("%" PRIu64 "*2^%" PRIu64 "-1  ERROR %d\n", v.k, v.n, rc);
        return 1;
    }
    printf("%" PRIu64 "*2^%" PRIu64 "-1 is %s  RES64 %016" PRIX64
           "  FFT %u  %.1f s (%.1f us/iter)  maxerr %.4f  gerbicz_errors %u  fft_retries %u%s\n",
           v.k, v.n, r.is_prp ? "a probable prime!" : "composite", r.res64, r.fft_len, r.seconds,
           r.seconds * 1e6 / (double)v.n, r.maxerr, r.gerbicz_errors, r.roundoff_retries,
           r.resumed ? "  (resumed)" : "");
    return v.expect.empty() ? 0 : verdict(v.expect, strtoull(v.expect.c_str(), nullptr, 16) == r.res64);
}

int cmd_llr(Invocation &v) {
    Progress pr;
    kite_llr_opts o{};
    o.fft = v.engine;
    o.checkpoint = v.prp.checkpoint;
    o.checkpoint_secs = v.prp.checkpoint_secs;
    o.progress = Progress::report;
    o.progress_arg = &pr;
    kite_llr_result r;
    const int rc = kite_llr(v.k, v.n, &o, &r);
    pr.wipe();
    if (rc) {
        printf("%" PRIu64 "*2^%" PRIu64 "-1  ERROR %d\n", v.k, v.n, rc);
        return 1;
    }
    char what[64];
    if (r.factor) snprintf(what, sizeof what, "divisible by %" PRIu64, r.factor);
    else if (r.is_prime) snprintf(what, sizeof what, "prime!");
    else snprintf(what, sizeof what, "composite  RES64 %016" PRIX64, r.res64);
    printf("%" PRIu64 "*2^%" PRIu64 "-1 is %s  P = %u  FFT %u  %.1f s  maxerr %.4f  confirm_runs %u  fft_retries %u%s\n",
           v.k, v.n, what, r.P, r.fft_len, r.seconds, r.maxerr, r.confirm_runs, r.roundoff_retries,
           r.resumed ? "  (resumed)" : "");
    if (v.expect.empty()) return 0;
    const bool ok = v.expect == "prime" ? (r.is_prime && !r.factor)
                                        : (!r.is_prime && !r.factor && strtoull(v.expect.c_str(), nullptr, 16) == r.res64);
    return verdict(v.expect, ok);
}

// One context with one number (3, squared a few times), for the timing commands.
struct Bench {
    kite_ctx *ctx = nullptr;
    kite_num *x = nullptr;
    Bench(uint64_t k, uint64_t n, const kite_opts &o, uint64_t seed = 3) {
        ctx = kite_ctx_new(k, n, -1, &o);
        if (!ctx) return;
        x = kite_num_new(ctx);
        if (!x) return;
        kite_set_u64(ctx, x, seed);
        kite_square_n(ctx, x, 10);
        kite_max_roundoff(ctx);  // waits for the warm-up
    }
    ~Bench() {
        if (x) kite_num_free(ctx, x);
        if (ctx) kite_ctx_free(ctx);
    }
    bool ok() const { return x != nullptr; }
};

int cmd_time(Invocation &v) {
    const uint64_t iters = v.iters ? v.iters : 1000;
    Bench bn(v.k, v.n, v.engine);
    if (!bn.ok()) {
        fprintf(stderr, "cannot create context\n");
        return 1;
    }
    const uint32_t L = kite_fft_len(bn.ctx);
    printf("%" PRIu64 "*2^%" PRIu64 "-1  FFT %u (%.2f bits/word)  backend %s\n", v.k, v.n, L, (double)v.n / L,
           kite_backend_name(bn.ctx));
    kite_reset_roundoff(bn.ctx);
    const double t0 = seconds_now();
    kite_square_n(bn.ctx, bn.x, iters);
    const double e = kite_max_roundoff(bn.ctx);  // waits for the squarings
    const double dt = seconds_now() - t0;
    printf("%.3f ms/iter  (%.1f us)  maxerr %.4f\n", dt * 1e3 / (double)iters, dt * 1e6 / (double)iters, e);
    return 0;
}

// Squarings per second summed over `tests` independent tests of the number for `secs`: every
// test has its own context and stream, all driven from this thread (the squarings of one round are
// queued on every stream before any is waited for, so the tests overlap on the device).
double rate_contexts(uint64_t k, uint64_t n, const kite_opts &o, int tests, double secs) {
    std::vector<Bench *> all;
    for (int t = 0; t < tests; t++) {
        all.push_back(new Bench(k, n, o, 3 + (uint64_t)t));
        if (!all.back()->ok()) break;
    }
    bool ok = (int)all.size() == tests && all.back()->ok();
    const uint64_t per_round = 64;
    uint64_t rounds = 0;
    const double t0 = seconds_now();
    while (ok && seconds_now() - t0 < secs) {
        for (Bench *b : all) kite_square_n(b->ctx, b->x, per_round);
        for (Bench *b : all) ok &= kite_max_roundoff(b->ctx) < 0.5;
        rounds++;
    }
    const double dt = seconds_now() - t0;
    for (Bench *b : all) delete b;
    return ok ? (double)(rounds * per_round) * tests / dt : 0;
}

// The same with the tests as one batch, in square calls of 1000 steps (a Gerbicz block, replayed
// from the batch's captured graph).
double rate_batch(uint64_t k, uint64_t n, uint32_t L, int tests, double secs) {
    kite_batch *b = kite_batch_new();
    if (!b) return 0;
    bool ok = true;
    for (int t = 0; t < tests && ok; t++) {
        const int s = kite_batch_add(b, k, n, L, 1);
        ok = s >= 0;
        if (ok) kite_set_u64(kite_batch_ctx(b, s), kite_batch_num(b, s, 0), 3 + (uint64_t)t);
    }
    const int cap = kite_batch_capacity(b);
    std::vector<int32_t> which(cap, 0);
    std::vector<uint32_t> count(cap, 1000);
    auto call = [&]() { return kite_batch_square(b, which.data(), count.data()) == 0 && kite_batch_sync(b) == 0; };
    ok = ok && call() && call();  // capture, then one warm replay
    uint64_t steps = 0;
    const double t0 = seconds_now();
    while (ok && seconds_now() - t0 < secs) {
        ok = call();
        steps += 1000;
    }
    const double dt = seconds_now() - t0;
    kite_batch_free(b);
    return ok ? (double)steps * tests / dt : 0;
}

int cmd_throughput(Invocation &v) {
    uint32_t L = 0;
    {
        Bench probe(v.k, v.n, v.engine);
        if (!probe.ok()) {
            fprintf(stderr, "cannot create context\n");
            return 1;
        }
        L = kite_fft_len(probe.ctx);
    }
    kite_opts o = v.engine;
    o.fft_len = L;
    printf("%" PRIu64 "*2^%" PRIu64 "-1  FFT %u\n", v.k, v.n, L);
    std::string best_cfg;
    double best = 0;
    size_t from = 0;
    while (from <= v.configs.size()) {
        size_t to = v.configs.find(',', from);
        if (to == std::string::npos) to = v.configs.size();
        const std::string cfg = v.configs.substr(from, to - from);
        from = to + 1;
        const int tests = atoi(cfg.c_str());
        if (tests < 1 || tests > (v.batched ? 4096 : 256)) continue;
        const double r = v.batched ? rate_batch(v.k, v.n, L, tests, v.secs) : rate_contexts(v.k, v.n, o, tests, v.secs);
        printf("%-6s %10.0f sq/s  %8.1f tests/hour  (%.1f us/iter per test)\n", cfg.c_str(), r,
               r * 3600.0 / (double)v.n, r > 0 ? 1e6 * tests / r : 0.0);
        fflush(stdout);
        if (r > best) {
            best = r;
            best_cfg = cfg;
        }
    }
    printf("best: %s  %.1f tests/hour\n", best_cfg.c_str(), best * 3600.0 / (double)v.n);
    return 0;
}

// Max round-off against bits per word at one FFT length (the reference's tools/sweep_err.c):
// for each bpw in [lo, hi] by step, n = bpw L, `iters` squarings of 3 after 64 to fill the words.
int cmd_sweep(Invocation &v) {
    kite_opts o = v.engine;
    o.fft_len = (uint32_t)atoi(v.rest[0].c_str());
    const uint64_t k = strtoull(v.rest[1].c_str(), nullptr, 10);
    const double lo = atof(v.rest[2].c_str()), hi = atof(v.rest[3].c_str()), step = atof(v.rest[4].c_str());
    const uint64_t iters = v.iters ? v.iters : 200;
    if (step <= 0) return 2;
    for (int i = 0; lo + i * step <= hi + 1e-9; i++) {
        const double bpw = lo + i * step;
        const uint64_t n = (uint64_t)(bpw * o.fft_len);
        kite_ctx *ctx = kite_ctx_new(k, n, -1, &o);
        kite_num *x = ctx ? kite_num_new(ctx) : nullptr;
        if (!x) {
            printf("bpw=%.2f: no ctx\n", bpw);
            if (ctx) kite_ctx_free(ctx);
            continue;
        }
        kite_set_u64(ctx, x, 3);
        kite_square_n(ctx, x, 64);
        kite_reset_roundoff(ctx);
        kite_square_n(ctx, x, iters);
        printf("L=%u k=%llu bpw=%.2f n=%llu maxerr=%.4f\n", o.fft_len, (unsigned long long)k, bpw,
               (unsigned long long)n, kite_max_roundoff(ctx));
        fflush(stdout);
        kite_num_free(ctx, x);
        kite_ctx_free(ctx);
    }
    return 0;
}

// Times every usable length in [LMIN, LMAX] in each of four shape variants (interleaved, best of
// three windows of 50 ms each) and writes the fastest as "L us order", the reference's tuning-file
// format, to stdout and --out. The number tested has k = 1 and n = 12 L, so that every length
// holds it.
int cmd_tune(Invocation &v) {
    FILE *f = v.out.empty() ? nullptr : fopen(v.out.c_str(), "a");
    constexpr int kVariants = 4;
    for (uint32_t L = ((uint32_t)v.k + 63) / 64 * 64; L <= v.n; L += 64) {
        std::vector<Bench *> cand(kVariants, nullptr);
        for (int var = 0; var < kVariants; var++) {
            kite_opts o = v.engine;
            o.fft_len = L;
            o.plan_order = 1 + (uint32_t)var;
            cand[var] = new Bench(1, (uint64_t)L * 12, o);
        }
        std::vector<double> us(kVariants, 1e30);
        for (int pass = 0; pass < 3; pass++)
            for (int var = 0; var < kVariants; var++) {
                if (!cand[var]->ok()) continue;
                uint64_t done = 0;
                const double t0 = seconds_now();
                double t1;
                do {
                    kite_square_n(cand[var]->ctx, cand[var]->x, 16);
                    kite_max_roundoff(cand[var]->ctx);
                    done += 16;
                } while ((t1 = seconds_now()) - t0 < 0.05);
                us[var] = std::min(us[var], (t1 - t0) * 1e6 / (double)done);
            }
        int pick = -1;
        for (int var = 0; var < kVariants; var++)
            if (cand[var]->ok() && (pick < 0 || us[var] < us[pick])) pick = var;
        for (Bench *b : cand) delete b;
        if (pick < 0) continue;
        printf("%u %.2f %d\n", L, us[pick], pick);
        if (f) fprintf(f, "%u %.2f %d\n", L, us[pick], pick);
        fflush(stdout);
    }
    if (f) fclose(f);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    Invocation v;
    if (!parse(argc, argv, v)) {
        usage(argv[0]);
        return 2;
    }
    const struct {
        const char *name;
        int (*run)(Invocation &);
    } commands[] = {{"prp", cmd_prp},         {"llr", cmd_llr},     {"time", cmd_time},
                    {"throughput", cmd_throughput}, {"sweep", cmd_sweep}, {"tune", cmd_tune}};
    for (const auto &c : commands)
        if (v.command == c.name) return c.run(v);
    usage(argv[0]);
    return 2;
}
