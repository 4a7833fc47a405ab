// kite-prst: a PRST-compatible command line for Fermat tests (with PRST proofs) and LLR (Morrison)
// tests of k*2^n-1 on the MI355X engine, for callers that drive PRST, such as the primecrunch
// client (reference tools/kite_prst.c, docs/prst-compat.md).
//
// The contract with those callers is PRST's: the accepted arguments, the stdout lines and their
// order, the result.txt lines, the checkpoint and proof file names (PRST's fingerprint N mod
// 3417905339), SIGTERM/SIGINT stopping the run with its state saved, and the exit codes. Anything
// outside the slice run here (other number forms and options) is handed to PRST itself ($KITE_PRST,
// or `prst` beside this binary) by exec, before the GPU is touched.
#include <cinttypes>
#include <climits>
#include <cmath>
#include <csignal>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <string>
#include <vector>
#include <sys/time.h>
#include <unistd.h>

#include "kite.h"
#include "prst_proof.h"

namespace {

volatile int g_stop = 0;  // set by SIGTERM / SIGINT, polled by the drivers between blocks

void request_stop(int) { g_stop = 1; }

double wall_clock() {
    timeval tv;
    gettimeofday(&tv, nullptr);
    return (double)tv.tv_sec + 1e-6 * (double)tv.tv_usec;
}

// ---- what was asked --------------------------------------------------------------------------

enum class Proof { kNone, kSave, kBuild, kCert };

struct Request {
    std::string expr;  // "K*2^N-1" as given
    uint64_t k = 0, n = 0;
    uint32_t a = 3;    // Fermat base
    int threads = 1;
    double progress_secs = 10, write_secs = 60;  // -time progress / write
    bool fermat = false, llr = false, fft_info = false, version = false;
    Proof proof = Proof::kNone;
    int proof_count = 0;
    bool proof_keep = false;
    std::string cert_path;
};

enum class Outcome { kRun, kPrst, kUsage };

// The words of the command line, consumed front to back.
struct Cursor {
    std::vector<std::string> words;
    size_t at = 0;
    bool has(size_t ahead) const { return at + ahead < words.size(); }
    const std::string &peek(size_t ahead) const { return words[at + ahead]; }
    std::string take() { return at < words.size() ? words[at++] : std::string(); }
    // the next word is `w` and at least `extra` words follow it: consume it
    bool accept(const char *w, size_t extra = 0) {
        if (!has(extra) || peek(0) != w) return false;
        at++;
        return true;
    }
};

// K*2^N-1 with K odd and N >= 2
bool read_expression(const std::string &s, Request &q) {
    unsigned long long k = 0, n = 0;
    int end = 0;
    if (sscanf(s.c_str(), "%llu*2^%llu-1%n", &k, &n, &end) != 2 || (size_t)end != s.size()) return false;
    if (k == 0 || k % 2 == 0 || n < 2) return false;
    q.expr = s;
    q.k = k;
    q.n = n;
    return true;
}

bool power_of_two(int c) { return c >= 2 && (c & (c - 1)) == 0; }

// What each flag of the slice does with the words after it; any other flag goes to PRST.
using FlagRule = std::function<Outcome(Cursor &, Request &)>;

const std::map<std::string, FlagRule> &flag_rules() {
    static const std::map<std::string, FlagRule> rules = {
        {"-v", [](Cursor &, Request &q) { q.version = true; return Outcome::kRun; }},
        {"-d", [](Cursor &, Request &) { return Outcome::kRun; }},
        {"-q", [](Cursor &, Request &) { return Outcome::kRun; }},
        {"-t", [](Cursor &c, Request &q) { q.threads = atoi(c.take().c_str()); return Outcome::kRun; }},
        // PRST's instruction-set choice: the GPU engine has none to make
        {"-cpu", [](Cursor &c, Request &) { c.take(); return Outcome::kRun; }},
        {"-fermat", [](Cursor &c, Request &q) {
             q.fermat = true;
             if (c.accept("a", 1)) q.a = (uint32_t)strtoul(c.take().c_str(), nullptr, 10);
             return Outcome::kRun;
         }},
        // -time [write S] [progress S] [coarse]: at least one of them
        {"-time", [](Cursor &c, Request &q) {
             int seen = 0;
             for (;; seen++) {
                 if (c.accept("write", 1)) q.write_secs = atof(c.take().c_str());
                 else if (c.accept("progress", 1)) q.progress_secs = atof(c.take().c_str());
                 else if (!c.accept("coarse")) break;
             }
             return seen ? Outcome::kRun : Outcome::kPrst;
         }},
        {"-fft", [](Cursor &c, Request &q) {
             q.fft_info = c.take() == "info";
             return q.fft_info ? Outcome::kRun : Outcome::kPrst;
         }},
        // -proof save C [keep] | build C [keep] | cert FILE
        {"-proof", [](Cursor &c, Request &q) {
             const std::string op = c.take();
             if (op == "cert") {
                 q.proof = Proof::kCert;
                 q.cert_path = c.take();
                 return Outcome::kRun;
             }
             if (op != "save" && op != "build") return Outcome::kPrst;
             q.proof = op == "save" ? Proof::kSave : Proof::kBuild;
             q.proof_count = atoi(c.take().c_str());
             if (!power_of_two(q.proof_count)) {
                 fprintf(stderr, "proof count is not a power of 2.\n");
                 return Outcome::kUsage;
             }
             q.proof_keep = c.accept("keep");
             return Outcome::kRun;
         }},
    };
    return rules;
}

Outcome read_request(int argc, char **argv, Request &q) {
    Cursor c;
    c.words.assign(argv + 1, argv + argc);
    bool have_number = false;
    while (c.has(0)) {
        const std::string w = c.take();
        if (!w.empty() && w[0] == '-') {
            const auto rule = flag_rules().find(w);
            if (rule == flag_rules().end()) return Outcome::kPrst;
            const Outcome o = rule->second(c, q);
            if (o != Outcome::kRun) return o;
        } else if (have_number || !read_expression(w, q)) {
            return Outcome::kPrst;
        } else {
            have_number = true;
        }
    }
    if (q.version) return Outcome::kRun;
    if (!have_number) return Outcome::kUsage;
    if (q.proof == Proof::kCert) return Outcome::kRun;
    if (q.fermat) return q.a < 2 ? Outcome::kUsage : Outcome::kRun;
    // PRST's default test of k*2^n-1 is Morrison's (LLR); it has no -fft info or proofs here
    if (q.fft_info || q.proof != Proof::kNone) return Outcome::kPrst;
    q.llr = true;
    return Outcome::kRun;
}

[[noreturn]] void run_prst_instead(char **argv) {
    std::string path;
    if (const char *env = getenv("KITE_PRST"); env && *env) {
        path = env;
    } else {
        const std::string self = argv[0];
        const size_t slash = self.rfind('/');
        path = (slash == std::string::npos ? std::string(".") : self.substr(0, slash)) + "/prst";
    }
    argv[0] = const_cast<char *>(path.c_str());
    execv(path.c_str(), argv);
    fprintf(stderr, "kite-prst: cannot run PRST at %s for this test\n", path.c_str());
    exit(2);
}

// ---- what is printed -------------------------------------------------------------------------

// PRST's terminal: full lines, and progress readings that end in '\r' and are wiped (70 spaces
// and '\r') before the next line; result.txt in the working directory gets one line per result.
class Console {
public:
    void say(const char *fmt, ...) __attribute__((format(printf, 2, 3))) {
        wipe_reading();
        va_list ap;
        va_start(ap, fmt);
        vprintf(fmt, ap);
        va_end(ap);
        fflush(stdout);
    }
    void record(const char *fmt, ...) __attribute__((format(printf, 2, 3))) {
        FILE *f = fopen("result.txt", "a");
        if (!f) return;
        va_list ap;
        va_start(ap, fmt);
        vfprintf(f, fmt, ap);
        va_end(ap);
        fclose(f);
    }
    // "EXPR P% done, time per op: T ms." at most every `every` seconds
    void reading(const std::string &expr, double every, uint64_t iter, double pct) {
        const double t = wall_clock();
        if (t - last_ < every) return;
        const double ms = iter > last_iter_ ? (t - last_) * 1e3 / (double)(iter - last_iter_) : 0.0;
        wipe_reading();
        printf("%s %.1f%% done, time per op: %.3f ms. \r", expr.c_str(), pct, ms);
        fflush(stdout);
        shown_ = true;
        last_ = t;
        last_iter_ = iter;
    }
    void restart_readings() {
        last_ = wall_clock();
        last_iter_ = 0;
    }

private:
    void wipe_reading() {
        if (!shown_) return;
        fputs("                                                                      \r", stdout);
        shown_ = false;
    }
    bool shown_ = false;
    double last_ = 0;
    uint64_t last_iter_ = 0;
};

std::string fft_label(uint32_t L) {
    char b[32];
    if (L % (1u << 20) == 0) snprintf(b, sizeof b, "%uM", L >> 20);
    else if (L % 1024 == 0) snprintf(b, sizeof b, "%uK", L >> 10);
    else snprintf(b, sizeof b, "%u", L);
    return b;
}

kite_opts engine_options(const Request &q) {
    kite_opts o;
    memset(&o, 0, sizeof o);
    o.backend = KITE_BACKEND_HIP;
    o.threads = q.threads > 1 ? (uint32_t)q.threads : 1;
    return o;
}

// A context for the number, with "Using HIP FFT length X." announced; nullptr if none.
kite_ctx *open_context(const Request &q, Console &out) {
    kite_opts o = engine_options(q);
    kite_ctx *c = kite_ctx_new(q.k, q.n, -1, &o);
    if (c) out.say("Using HIP FFT length %s.\n", fft_label(kite_fft_len(c)).c_str());
    return c;
}

std::string fp_name(const Request &q) { return std::to_string(prst::fingerprint(q.k, q.n)); }

int depth_of(int count) {
    int d = 0;
    while ((1 << d) < count) d++;
    return d;
}

// ---- a Fermat test ---------------------------------------------------------------------------

// The record of a finished Fermat test kept while its proof is compressed, so that a restart
// after the test does not repeat it: "kite done k n a RES64 prp seconds".
struct DoneRecord {
    static bool write(const std::string &path, const Request &q, const kite_prp_result &r) {
        const std::string tmp = path + ".tmp";
        FILE *f = fopen(tmp.c_str(), "w");
        if (!f) return false;
        bool ok = fprintf(f, "kite done %" PRIu64 " %" PRIu64 " %u %016" PRIX64 " %d %.1f\n", q.k, q.n, q.a, r.res64,
                          r.is_prp, r.seconds) > 0;
        ok = fclose(f) == 0 && ok;
        if (!ok || rename(tmp.c_str(), path.c_str())) {
            remove(tmp.c_str());
            return false;
        }
        return true;
    }
    static bool read(const std::string &path, const Request &q, kite_prp_result *r) {
        FILE *f = fopen(path.c_str(), "r");
        if (!f) return false;
        uint64_t k = 0, n = 0, res = 0;
        unsigned a = 0;
        int prp = 0;
        double secs = 0;
        const int got =
            fscanf(f, "kite done %" SCNu64 " %" SCNu64 " %u %" SCNx64 " %d %lf", &k, &n, &a, &res, &prp, &secs);
        fclose(f);
        if (got != 6 || k != q.k || n != q.n || a != q.a) return false;
        r->res64 = res;
        r->is_prp = prp;
        r->seconds = secs;
        return true;
    }
};

class FermatSession {
public:
    FermatSession(const Request &q, Console &out) : q_(q), out_(out) {
        const std::string fp = fp_name(q);
        const std::string stem = "kite_" + fp + "." + std::to_string(q.a);
        ckpt_ = stem + ".ckpt";
        ckpt_u_ = stem + ".u.ckpt";
        done_ = stem + ".done";
        proof_prefix_ = "prst_" + fp;
        B_ = (uint32_t)std::sqrt((double)q.n);  // kite_prp's own block choice, in PRST's terms
        B_ = B_ < 32 ? 32 : B_ > 2000 ? 2000 : B_;
        M_ = B_ / 4 > 8 ? B_ / 4 : 8;
    }

    int run() {
        {
            kite_ctx *c = open_context(q_, out_);
            if (!c) {
                fprintf(stderr, "kite-prst: cannot set up arithmetic for %s\n", q_.expr.c_str());
                return 1;
            }
            nlimbs_ = kite_limbs(c);
            kite_ctx_free(c);
        }
        out_.say("Fermat probabilistic test of %s, a = %u, complexity = %" PRIu64 ".\n", q_.expr.c_str(), q_.a,
                 q_.n + q_.n / B_);
        if (q_.fft_info) return 0;
        if (base_divides_N()) {
            out_.say("%s is not prime, divisible by %u.\n", q_.expr.c_str(), q_.a);
            out_.record("%s is not prime, divisible by %u.\n", q_.expr.c_str(), q_.a);
            return 0;
        }
        out_.say("%s Gerbicz check enabled, L2 = %u*%u.\n", q_.expr.c_str(), B_, M_);
        const bool proving = q_.proof == Proof::kSave;
        if (proving && !plan_points()) return 1;
        kite_prp_result r;
        memset(&r, 0, sizeof r);
        // a test that finished while its proof was being compressed is not repeated: its result
        // is printed again and the compression resumes
        const bool already = proving && DoneRecord::read(done_, q_, &r);
        if (!already) {
            announce_resume();
            if (const int rc = test(&r)) return rc;
            if (proving && !DoneRecord::write(done_, q_, r)) {
                out_.say("Write to %s failed.\n", done_.c_str());
                return 1;
            }
            for (const std::string &f : {ckpt_, ckpt_u_, ckpt_ + ".md5"}) remove(f.c_str());
        }
        report(r, !already);
        return proving ? compress() : 0;
    }

private:
    bool base_divides_N() const {
        // N mod a = (k mod a) 2^n - 1 (mod a)
        uint64_t p = 1 % q_.a, sq = 2 % q_.a;
        for (uint64_t e = q_.n; e; e >>= 1) {
            if (e & 1) p = (uint64_t)((unsigned __int128)p * sq % q_.a);
            sq = (uint64_t)((unsigned __int128)sq * sq % q_.a);
        }
        return (unsigned __int128)(q_.k % q_.a) * p % q_.a == 1 % q_.a;
    }

    bool plan_points() {
        if (q_.n > INT32_MAX) {
            fprintf(stderr, "kite-prst: n too large for PRST proof files\n");
            return false;
        }
        proof_fp_ = prst::proof_fingerprint(q_.k, q_.n, q_.a);
        prst::points((uint32_t)q_.n, q_.proof_count, &points_);
        out_.say("%s Saving %d proof points.\n", q_.expr.c_str(), q_.proof_count);
        return true;
    }

    // "restarting at P%." from the later of the two state files, before the test starts
    void announce_resume() const {
        uint64_t verified = 0, unverified = 0;
        if (kite_prp_checkpoint_iter(ckpt_.c_str(), q_.k, q_.n, q_.a, &verified)) verified = 0;
        if (kite_prp_checkpoint_iter(ckpt_u_.c_str(), q_.k, q_.n, q_.a, &unverified)) unverified = 0;
        const uint64_t at = std::max(verified, unverified);
        if (at > 0) out_.say("%s restarting at %.1f%%.\n", q_.expr.c_str(), 100.0 * (double)at / (double)q_.n);
    }

    static void on_progress(void *arg, uint64_t iter, uint64_t n, double) {
        FermatSession &s = *static_cast<FermatSession *>(arg);
        s.out_.reading(s.q_.expr, s.q_.progress_secs, iter, 100.0 * (double)iter / (double)n);
    }

    // a proof point as the test reaches it: prefix.proof.I holding x_pos in PRST's giant form
    static int on_point(void *arg, uint32_t index, uint64_t iter, const uint64_t *limbs) {
        FermatSession &s = *static_cast<FermatSession *>(arg);
        prst::Giant g;
        g.w.resize(2 * s.nlimbs_);
        for (size_t i = 0; i < s.nlimbs_; i++) {
            g.w[2 * i] = (uint32_t)limbs[i];
            g.w[2 * i + 1] = (uint32_t)(limbs[i] >> 32);
        }
        g.trim();
        const std::string path = s.proof_prefix_ + ".proof." + std::to_string(index);
        const int rc = prst::write_state(path, prst::kTypeValue, s.proof_fp_, (int32_t)iter, {g});
        if (rc) s.out_.say("Write to %s failed.\n", path.c_str());
        return rc;
    }

    int test(kite_prp_result *r) {
        kite_prp_opts o;
        memset(&o, 0, sizeof o);
        o.fft = engine_options(q_);
        o.checkpoint = ckpt_.c_str();
        o.checkpoint_secs = 300;
        o.checkpoint_unverified = ckpt_u_.c_str();
        o.unverified_secs = q_.write_secs;
        o.progress = on_progress;
        o.progress_arg = this;
        o.abort_flag = &g_stop;
        if (!points_.empty()) {
            o.points = points_.data();
            o.npoints = (uint32_t)points_.size();
            o.on_point = on_point;
            o.point_arg = this;
        }
        out_.restart_readings();
        const int rc = kite_prp(q_.k, q_.n, q_.a, &o, r);
        if (rc == KITE_PRP_ABORTED) {
            out_.say("Terminating...\n");
            return 1;
        }
        if (rc) {
            out_.say("%s test failed (error %d).\n", q_.expr.c_str(), rc);
            return 1;
        }
        return 0;
    }

    void report(const kite_prp_result &r, bool to_file) {
        const char *e = q_.expr.c_str();
        if (r.is_prp) {
            out_.say("%s is a probable prime. Time: %.1f s.\n", e, r.seconds);
            if (to_file) out_.record("%s is a probable prime. Time: %d s.\n", e, (int)r.seconds);
        } else {
            out_.say("%s is not prime. RES64: %016" PRIX64 ", time: %.1f s.\n", e, r.res64, r.seconds);
            if (to_file) out_.record("%s is not prime. RES64: %016" PRIX64 ", time: %d s.\n", e, r.res64, (int)r.seconds);
        }
    }

    // the proof's save step: the points become products and the raw certificate
    int compress() {
        const double t0 = wall_clock();
        kite_opts o = engine_options(q_);
        kite_ctx *c = kite_ctx_new(q_.k, q_.n, -1, &o);
        uint64_t raw = 0;
        const int rc = c ? prst::save_products(c, q_.k, q_.n, q_.a, q_.proof_count, proof_prefix_, q_.proof_keep,
                                               &g_stop, &raw)
                         : -1;
        if (c) kite_ctx_free(c);
        if (rc == prst::kAborted) {
            out_.say("Terminating...\n");
            return 1;
        }
        if (rc) return 1;
        remove(done_.c_str());
        const double secs = wall_clock() - t0;
        out_.say("%s compressed %d points to %d products, time: %.1f s.\n", q_.expr.c_str(), q_.proof_count,
                 depth_of(q_.proof_count), secs);
        out_.say("%s raw certificate RES64: %016" PRIX64 ".\n", q_.expr.c_str(), raw);
        out_.record("%s raw certificate RES64: %016" PRIX64 ", time: %d s.\n", q_.expr.c_str(), raw, (int)secs);
        return 0;
    }

    const Request &q_;
    Console &out_;
    std::string ckpt_, ckpt_u_, done_, proof_prefix_;
    uint32_t B_ = 0, M_ = 0, proof_fp_ = 0;
    size_t nlimbs_ = 0;
    std::vector<uint32_t> points_;
};

// ---- an LLR test -----------------------------------------------------------------------------

class LlrSession {
public:
    LlrSession(const Request &q, Console &out)
        : q_(q), out_(out), ckpt_("kite_" + fp_name(q) + ".llr.ckpt") {}

    int run() {
        uint32_t P = 0;
        uint64_t factor = 0;
        if (kite_llr_choose_P(q_.k, q_.n, &P, &factor)) {
            fprintf(stderr, "kite-prst: no P found for %s\n", q_.expr.c_str());
            return 1;
        }
        kite_ctx *c = open_context(q_, out_);
        if (!c) {
            fprintf(stderr, "kite-prst: cannot set up arithmetic for %s\n", q_.expr.c_str());
            return 1;
        }
        kite_ctx_free(c);
        const char *e = q_.expr.c_str();
        if (factor) {  // a factor of 4P(P^2+4) divides N: reported before any test line
            out_.say("%s is not prime, divisible by %" PRIu64 ". Time: 0.0 s.\n", e, factor);
            out_.record("%s is not prime, divisible by %" PRIu64 ". Time: 0 s.\n", e, factor);
            return 0;
        }
        out_.say("Morrison (LLR) test of %s, P = %u, Q = -1, complexity = %" PRIu64 ".\n", e, P, q_.n);
        uint64_t at = 0;
        int at_run = 0;
        if (kite_llr_checkpoint_iter(ckpt_.c_str(), q_.k, q_.n, &at, &at_run) == 0 && (at > 0 || at_run > 0))
            out_.say("%s restarting at %.1f%%.\n", e, percent(at_run, at));
        kite_llr_opts o;
        memset(&o, 0, sizeof o);
        o.fft = engine_options(q_);
        o.checkpoint = ckpt_.c_str();
        o.checkpoint_secs = 300;
        o.progress = on_progress;
        o.on_confirm = on_confirm;
        o.progress_arg = this;
        o.abort_flag = &g_stop;
        kite_llr_result r;
        memset(&r, 0, sizeof r);
        out_.restart_readings();
        const int rc = kite_llr(q_.k, q_.n, &o, &r);
        if (rc == KITE_LLR_ABORTED) {
            out_.say("Terminating...\n");
            return 1;
        }
        if (rc) {
            out_.say("%s LLR test failed (error %d).\n", e, rc);
            return 1;
        }
        remove(ckpt_.c_str());
        if (r.is_prime) {
            out_.say("%s is prime! Time: %.1f s.\n", e, r.seconds);
            out_.record("%s is prime! Time: %d s.\n", e, (int)r.seconds);
        } else {
            out_.say("%s is not a probable prime. Have you run Fermat test first? RES64: %016" PRIX64
                     ", time: %.1f s.\n", e, r.res64, r.seconds);
            out_.record("%s is not a probable prime. Have you run Fermat test first? RES64: %016" PRIX64
                        ", time: %d s.\n", e, r.res64, (int)r.seconds);
        }
        return 0;
    }

private:
    // PRST's LLR progress only goes forward: the first run counts to 99%, a confirmation run (of a
    // composite result) holds there until the result line
    double percent(int run, uint64_t iter) const { return run ? 99.0 : 99.0 * (double)iter / (double)(q_.n - 1); }

    static void on_progress(void *arg, uint64_t iter, uint64_t, double) {
        LlrSession &s = *static_cast<LlrSession *>(arg);
        s.out_.reading(s.q_.expr, s.q_.progress_secs, iter, s.percent(s.run_, iter));
    }

    static void on_confirm(void *arg, int run) {
        LlrSession &s = *static_cast<LlrSession *>(arg);
        s.run_ = run;
        s.out_.restart_readings();
        fprintf(stderr, "%s result is composite; confirming with run %d at a longer FFT length.\n", s.q_.expr.c_str(),
                run + 1);
    }

    const Request &q_;
    Console &out_;
    std::string ckpt_;
    int run_ = 0;
};

// ---- the proof's later steps -----------------------------------------------------------------

// -proof build C [keep]: the certificate from point 0, point C and the products
int build_certificate(const Request &q, Console &out) {
    kite_ctx *c = open_context(q, out);
    if (!c) return 1;
    const double t0 = wall_clock();
    out.say("%s Building certificate from %d products.\n", q.expr.c_str(), depth_of(q.proof_count));
    uint64_t res = 0;
    const int rc = prst::build_cert(c, q.k, q.n, q.a, q.proof_count, "prst_" + fp_name(q), q.proof_keep, &res);
    kite_ctx_free(c);
    if (rc) return 1;
    const double secs = wall_clock() - t0;
    out.say("%s certificate RES64: %016" PRIX64 ", time: %.1f s.\n", q.expr.c_str(), res, secs);
    out.record("%s certificate RES64: %016" PRIX64 ", time: %d s.\n", q.expr.c_str(), res, (int)secs);
    return 0;
}

// -proof cert FILE: verify a certificate (FILE "default": prst_FP.cert)
int verify_certificate(const Request &q, Console &out) {
    kite_ctx *c = open_context(q, out);
    if (!c) return 1;
    const std::string path =
        q.cert_path.empty() || q.cert_path == "default" ? "prst_" + fp_name(q) + ".cert" : q.cert_path;
    const double t0 = wall_clock();
    uint64_t res = 0;
    uint32_t M = 0;
    const int rc = prst::verify_cert(c, q.k, q.n, path, &res, &M);
    kite_ctx_free(c);
    if (rc) return 1;
    const double secs = wall_clock() - t0;
    out.say("Verifying certificate of %s, complexity = %u.\n", q.expr.c_str(), M);
    out.say("%s certificate RES64: %016" PRIX64 ", time: %.1f s.\n", q.expr.c_str(), res, secs);
    out.record("%s certificate RES64: %016" PRIX64 ", time: %d s.\n", q.expr.c_str(), res, (int)secs);
    return 0;
}

void usage(const char *p) {
    fprintf(stderr,
            "usage: %s K*2^N-1 -d -fermat a 3 [-t T] [-time [write S] [progress S] [coarse]]\n"
            "              [-proof save|build C [keep]]\n"
            "       %s K*2^N-1 -d -proof cert FILE\n"
            "       %s K*2^N-1 -d [-t T]          (Morrison/LLR test)\n"
            "       %s K*2^N-1 -fermat a 3 -fft info\n"
            "       %s -v\n"
            "(other forms and options run PRST: $KITE_PRST or prst beside this binary)\n",
            p, p, p, p, p);
}

}  // namespace

int main(int argc, char **argv) {
    Request q;
    switch (read_request(argc, argv, q)) {
    case Outcome::kPrst:
        run_prst_instead(argv);
    case Outcome::kUsage:
        usage(argv[0]);
        return 1;
    case Outcome::kRun:
        break;
    }
    if (q.version) {
        printf("kite version %s\n", KITE_VERSION);
        return 0;
    }
    signal(SIGTERM, request_stop);
    signal(SIGINT, request_stop);
    Console out;
    if (q.proof == Proof::kBuild) return build_certificate(q, out);
    if (q.proof == Proof::kCert) return verify_certificate(q, out);
    if (q.llr) return LlrSession(q, out).run();
    return FermatSession(q, out).run();
}
