// kite-prst: the PRST-compatible command line for Fermat tests, Fermat proofs and LLR (Morrison)
// tests of k*2^n-1 on the MI355X engine (reference tools/kite_prst.c: parse :98-182, run_fermat
// :288-430, run_llr :451-511, run_build/run_cert :513-553, main :555-578). Its arguments, stdout
// lines, checkpoint and proof file names, result.txt lines, SIGTERM behaviour and exit codes are
// the ones the reference reproduces from PRST (docs/prst-compat.md), so a caller that drives PRST
// (the primecrunch client) can run this instead. Anything outside that slice -- other number
// forms and options -- is handed to PRST ($KITE_PRST, or `prst` beside this binary) by exec,
// before anything touches the GPU.
#include <cinttypes>
#include <climits>
#include <cmath>
#include <csignal>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <sys/time.h>
#include <unistd.h>

#include "kite.h"
#include "prst_proof.h"

namespace {

volatile int g_abort = 0;
bool g_reading_shown = false;  // a progress reading ending in '\r' is on the terminal

void on_signal(int) { g_abort = 1; }

double wall() {
    timeval tv;
    gettimeofday(&tv, nullptr);
    return (double)tv.tv_sec + 1e-6 * (double)tv.tv_usec;
}

// A progress reading is wiped with 70 spaces and '\r' before the next full line, as PRST does.
void clear_reading() {
    if (!g_reading_shown) return;
    fputs("                                                                      \r", stdout);
    g_reading_shown = false;
}

void line(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
void line(const char *fmt, ...) {
    clear_reading();
    va_list ap;
    va_start(ap, fmt);
    vprintf(fmt, ap);
    va_end(ap);
    fflush(stdout);
}

// One line appended to result.txt in the working directory per result (whole seconds).
void result_line(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
void result_line(const char *fmt, ...) {
    FILE *f = fopen("result.txt", "a");
    if (!f) return;
    va_list ap;
    va_start(ap, fmt);
    vfprintf(f, fmt, ap);
    va_end(ap);
    fclose(f);
}

struct Args {
    std::string expr;
    uint64_t k = 0, n = 0;
    uint32_t a = 3;
    int threads = 1;
    double progress_secs = 10;
    double write_secs = 60;  // how often the unverified checkpoint is written
    bool fermat = false, fft_info = false, version = false, llr = false;
    int proof_op = 0;  // 0 none, 1 save, 2 build, 3 cert
    int proof_count = 0;
    bool proof_keep = false;
    std::string cert_path;
};

enum class Parsed { kRun, kHandOff, kUsage };

// EXPR = K*2^N-1 with K odd and N >= 2
bool parse_expr(Args &A, const char *s) {
    unsigned long long k = 0, n = 0;
    int end = 0;
    if (sscanf(s, "%llu*2^%llu-1%n", &k, &n, &end) != 2 || s[end] != '\0') return false;
    if (k == 0 || k % 2 == 0 || n < 2) return false;
    A.expr = s;
    A.k = k;
    A.n = n;
    return true;
}

Parsed parse(Args &A, int argc, char **argv) {
    bool have_expr = false;
    for (int i = 1; i < argc; i++) {
        const std::string s = argv[i];
        auto next = [&]() -> std::string { return i + 1 < argc ? std::string(argv[++i]) : std::string(); };
        auto peek = [&](const char *w, int need) { return i + need < argc && !strcmp(argv[i + 1], w); };
        if (s == "-v") {
            A.version = true;
        } else if (s == "-d" || s == "-q") {
        } else if (s == "-t") {
            A.threads = atoi(next().c_str());
        } else if (s == "-cpu") {
            next();  // an instruction-set choice for PRST's x86 kernels: nothing to choose here
        } else if (s == "-fermat") {
            A.fermat = true;
            if (peek("a", 2)) {
                i++;
                A.a = (uint32_t)strtoul(next().c_str(), nullptr, 10);
            }
        } else if (s == "-time") {  // [write S] [progress S] [coarse], at least one
            bool any = false;
            for (;;) {
                if (peek("write", 2)) {
                    i++;
                    A.write_secs = atof(next().c_str());
                } else if (peek("progress", 2)) {
                    i++;
                    A.progress_secs = atof(next().c_str());
                } else if (peek("coarse", 1)) {
                    i++;
                } else {
                    break;
                }
                any = true;
            }
            if (!any) return Parsed::kHandOff;
        } else if (s == "-fft") {
            if (next() != "info") return Parsed::kHandOff;
            A.fft_info = true;
        } else if (s == "-proof") {
            const std::string op = next();
            if (op == "save" || op == "build") {
                A.proof_op = op == "save" ? 1 : 2;
                A.proof_count = atoi(next().c_str());
                if (A.proof_count < 2 || (A.proof_count & (A.proof_count - 1))) {
                    fprintf(stderr, "proof count is not a power of 2.\n");
                    return Parsed::kUsage;
                }
                if (peek("keep", 1)) {
                    i++;
                    A.proof_keep = true;
                }
            } else if (op == "cert") {
                A.proof_op = 3;
                A.cert_path = next();
            } else {
                return Parsed::kHandOff;
            }
        } else if (!s.empty() && s[0] == '-') {
            return Parsed::kHandOff;
        } else if (!have_expr) {
            if (!parse_expr(A, s.c_str())) return Parsed::kHandOff;
            have_expr = true;
        } else {
            return Parsed::kHandOff;
        }
    }
    if (A.version) return Parsed::kRun;
    if (!have_expr) return Parsed::kUsage;
    if (A.proof_op == 3) return Parsed::kRun;
    if (!A.fermat) {  // PRST's default test of k*2^n-1 is Morrison's (LLR)
        if (A.fft_info || A.proof_op) return Parsed::kHandOff;
        A.llr = true;
        return Parsed::kRun;
    }
    return A.a < 2 ? Parsed::kUsage : Parsed::kRun;
}

[[noreturn]] void hand_off(char **argv) {
    std::string path;
    const char *env = getenv("KITE_PRST");
    if (env && *env) {
        path = env;
    } else {
        path = argv[0];
        const size_t slash = path.rfind('/');
        path = (slash == std::string::npos ? std::string(".") : path.substr(0, slash)) + "/prst";
    }
    argv[0] = const_cast<char *>(path.c_str());
    execv(path.c_str(), argv);
    fprintf(stderr, "kite-prst: cannot run PRST at %s for this test\n", path.c_str());
    exit(2);
}

uint64_t pow_mod(uint64_t b, uint64_t e, uint64_t m) {
    unsigned __int128 r = 1 % m, x = b % m;
    for (; e; e >>= 1, x = x * x % m)
        if (e & 1) r = r * x % m;
    return (uint64_t)r;
}

// PRST's number fingerprint (N mod 3417905339) names its files and so the caller's
uint32_t fingerprint(uint64_t k, uint64_t n) { return prst::fingerprint(k, n); }

std::string fft_label(uint32_t L) {
    char b[32];
    if (L % (1u << 20) == 0) snprintf(b, sizeof b, "%uM", L >> 20);
    else if (L % 1024 == 0) snprintf(b, sizeof b, "%uK", L >> 10);
    else snprintf(b, sizeof b, "%u", L);
    return b;
}

kite_opts engine_opts(const Args &A) {
    kite_opts o;
    memset(&o, 0, sizeof o);
    o.backend = KITE_BACKEND_HIP;
    o.threads = A.threads > 1 ? (uint32_t)A.threads : 1;
    return o;
}

// The FFT length the test will use, announced the way PRST does; 0 if no context can be made.
uint32_t announce_fft(const Args &A) {
    kite_opts o = engine_opts(A);
    kite_ctx *c = kite_ctx_new(A.k, A.n, -1, &o);
    if (!c) return 0;
    const uint32_t L = kite_fft_len(c);
    kite_ctx_free(c);
    line("Using HIP FFT length %s.\n", fft_label(L).c_str());
    return L;
}

struct Progress {
    const Args *A = nullptr;
    double last = 0;
    uint64_t last_iter = 0;
    int run = 0;  // LLR: 0, or the confirmation run
};

void show_progress(Progress &P, uint64_t iter, double pct) {
    const double t = wall();
    if (t - P.last < P.A->progress_secs) return;
    const double ms = iter > P.last_iter ? (t - P.last) * 1e3 / (double)(iter - P.last_iter) : 0.0;
    clear_reading();
    printf("%s %.1f%% done, time per op: %.3f ms. \r", P.A->expr.c_str(), pct, ms);
    fflush(stdout);
    g_reading_shown = true;
    P.last = t;
    P.last_iter = iter;
}

void fermat_progress(void *arg, uint64_t iter, uint64_t n, double) {
    show_progress(*static_cast<Progress *>(arg), iter, 100.0 * (double)iter / (double)n);
}

// Proof points as the Fermat test reaches them: prefix.proof.I with the value x_pos, trimmed to
// PRST's giant form (reference tools/kite_prst.c:212-232).
struct PointWriter {
    std::string prefix;
    uint32_t fp = 0;  // proof fingerprint inside the files
    size_t nlimbs = 0;
};

int write_point(void *arg, uint32_t index, uint64_t iter, const uint64_t *limbs) {
    const PointWriter &W = *static_cast<PointWriter *>(arg);
    prst::Giant g;
    g.w.resize(2 * W.nlimbs);
    for (size_t i = 0; i < W.nlimbs; i++) {
        g.w[2 * i] = (uint32_t)limbs[i];
        g.w[2 * i + 1] = (uint32_t)(limbs[i] >> 32);
    }
    g.trim();
    const std::string path = W.prefix + ".proof." + std::to_string(index);
    const int rc = prst::write_state(path, prst::kTypeValue, W.fp, (int32_t)iter, {g});
    if (rc) line("Write to %s failed.\n", path.c_str());
    return rc;
}

// The record of a finished Fermat test kept while its proof is compressed: k, n, a, RES64, PRP,
// seconds (reference tools/kite_prst.c:236-272).
bool write_done(const char *path, const Args &A, const kite_prp_result &r) {
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = fopen(tmp.c_str(), "w");
    if (!f) return false;
    bool ok = fprintf(f, "kite done %" PRIu64 " %" PRIu64 " %u %016" PRIX64 " %d %.1f\n", A.k, A.n, A.a, r.res64, r.is_prp,
                      r.seconds) > 0;
    ok = (fclose(f) == 0) && ok;
    if (!ok || rename(tmp.c_str(), path)) {
        remove(tmp.c_str());
        return false;
    }
    return true;
}

bool read_done(const char *path, const Args &A, kite_prp_result *r) {
    FILE *f = fopen(path, "r");
    if (!f) return false;
    uint64_t k = 0, n = 0, res = 0;
    unsigned a = 0;
    int prp = 0;
    double secs = 0;
    const bool ok = fscanf(f, "kite done %" SCNu64 " %" SCNu64 " %u %" SCNx64 " %d %lf", &k, &n, &a, &res, &prp, &secs) == 6 &&
                    k == A.k && n == A.n && a == A.a;
    fclose(f);
    if (!ok) return false;
    r->res64 = res;
    r->is_prp = prp;
    r->seconds = secs;
    return true;
}

// The Fermat test through kite_prp with PRST's checkpoint files, progress readings, SIGTERM stop
// and (pos != nullptr) the proof points; 0, or the exit code.
int run_prp(const Args &A, const char *ckpt, const char *ckpt_u, const std::vector<uint32_t> *pos, PointWriter *W,
            kite_prp_result *r) {
    Progress P;
    P.A = &A;
    P.last = wall();
    kite_prp_opts po;
    memset(&po, 0, sizeof po);
    po.fft = engine_opts(A);
    po.checkpoint = ckpt;
    po.checkpoint_secs = 300;
    po.checkpoint_unverified = ckpt_u;
    po.unverified_secs = A.write_secs;
    po.progress = fermat_progress;
    po.progress_arg = &P;
    po.abort_flag = &g_abort;
    if (pos) {
        kite_ctx *c = kite_ctx_new(A.k, A.n, -1, &po.fft);
        if (!c) return 1;
        W->nlimbs = kite_limbs(c);
        kite_ctx_free(c);
        po.points = pos->data();
        po.npoints = (uint32_t)pos->size();
        po.on_point = write_point;
        po.point_arg = W;
    }
    const int rc = kite_prp(A.k, A.n, A.a, &po, r);
    if (rc == KITE_PRP_ABORTED) {
        line("Terminating...\n");
        return 1;
    }
    if (rc) {
        line("%s test failed (error %d).\n", A.expr.c_str(), rc);
        return 1;
    }
    return 0;
}

int run_fermat(const Args &A) {
    const uint32_t L = announce_fft(A);
    if (!L) {
        fprintf(stderr, "kite-prst: cannot set up arithmetic for %s\n", A.expr.c_str());
        return 1;
    }
    // the Gerbicz block kite_prp picks (~sqrt(n) in [32, 2000], checks every B/4 >= 8 blocks),
    // printed in PRST's terms
    uint32_t B = (uint32_t)sqrt((double)A.n);
    B = B < 32 ? 32 : B > 2000 ? 2000 : B;
    const uint32_t M = B / 4 > 8 ? B / 4 : 8;
    line("Fermat probabilistic test of %s, a = %u, complexity = %" PRIu64 ".\n", A.expr.c_str(), A.a,
         A.n + A.n / B);
    if (A.fft_info) return 0;
    if ((unsigned __int128)(A.k % A.a) * pow_mod(2, A.n, A.a) % A.a == 1 % A.a) {  // a divides N
        line("%s is not prime, divisible by %u.\n", A.expr.c_str(), A.a);
        result_line("%s is not prime, divisible by %u.\n", A.expr.c_str(), A.a);
        return 0;
    }
    line("%s Gerbicz check enabled, L2 = %u*%u.\n", A.expr.c_str(), B, M);

    const uint32_t fp = fingerprint(A.k, A.n);
    char ckpt[96], ckpt_u[96], done[96];
    snprintf(ckpt, sizeof ckpt, "kite_%u.%u.ckpt", fp, A.a);
    snprintf(ckpt_u, sizeof ckpt_u, "kite_%u.%u.u.ckpt", fp, A.a);
    snprintf(done, sizeof done, "kite_%u.%u.done", fp, A.a);
    PointWriter W;
    W.prefix = "prst_" + std::to_string(fp);
    std::vector<uint32_t> pos;
    if (A.proof_op == 1) {
        if (A.n > INT32_MAX) {
            fprintf(stderr, "kite-prst: n too large for PRST proof files\n");
            return 1;
        }
        W.fp = prst::proof_fingerprint(A.k, A.n, A.a);
        prst::points((uint32_t)A.n, A.proof_count, &pos);
        line("%s Saving %d proof points.\n", A.expr.c_str(), A.proof_count);
    }
    // a test that finished while its proof was being compressed: its result is printed again and
    // the compression finished, the test is not repeated
    kite_prp_result r;
    memset(&r, 0, sizeof r);
    const bool finished = A.proof_op == 1 && read_done(done, A, &r);
    if (!finished) {
        {  // a resume is announced before the test starts: the later of the two files
            uint64_t at = 0, u = 0;
            if (kite_prp_checkpoint_iter(ckpt, A.k, A.n, A.a, &at)) at = 0;
            if (!kite_prp_checkpoint_iter(ckpt_u, A.k, A.n, A.a, &u) && u > at) at = u;
            if (at > 0) line("%s restarting at %.1f%%.\n", A.expr.c_str(), 100.0 * (double)at / (double)A.n);
        }
        const int rc = run_prp(A, ckpt, ckpt_u, A.proof_op == 1 ? &pos : nullptr, &W, &r);
        if (rc) return rc;
        if (A.proof_op == 1 && !write_done(done, A, r)) {
            line("Write to %s failed.\n", done);
            return 1;
        }
        remove(ckpt);
        remove(ckpt_u);
        remove((std::string(ckpt) + ".md5").c_str());
    }
    if (r.is_prp) {
        line("%s is a probable prime. Time: %.1f s.\n", A.expr.c_str(), r.seconds);
        if (!finished) result_line("%s is a probable prime. Time: %d s.\n", A.expr.c_str(), (int)r.seconds);
    } else {
        line("%s is not prime. RES64: %016" PRIX64 ", time: %.1f s.\n", A.expr.c_str(), r.res64, r.seconds);
        if (!finished)
            result_line("%s is not prime. RES64: %016" PRIX64 ", time: %d s.\n", A.expr.c_str(), r.res64,
                        (int)r.seconds);
    }
    if (A.proof_op != 1) return 0;
    // compress the points into products (the proof's save step)
    const double t0 = wall();
    kite_opts o = engine_opts(A);
    kite_ctx *c = kite_ctx_new(A.k, A.n, -1, &o);
    uint64_t raw = 0;
    const int prc = c ? prst::save_products(c, A.k, A.n, A.a, A.proof_count, W.prefix, A.proof_keep, &g_abort, &raw)
                      : -1;
    if (c) kite_ctx_free(c);
    if (prc == prst::kAborted) {
        line("Terminating...\n");
        return 1;
    }
    if (prc) return 1;
    remove(done);
    int depth = 0;
    while ((1 << depth) < A.proof_count) depth++;
    line("%s compressed %d points to %d products, time: %.1f s.\n", A.expr.c_str(), A.proof_count, depth, wall() - t0);
    line("%s raw certificate RES64: %016" PRIX64 ".\n", A.expr.c_str(), raw);
    result_line("%s raw certificate RES64: %016" PRIX64 ", time: %d s.\n", A.expr.c_str(), raw, (int)(wall() - t0));
    return 0;
}


// PRST's LLR progress only goes forward: the first run counts to 99%, a confirmation run (of a
// composite result) holds there until the result line.
double llr_pct(int run, uint64_t iter, uint64_t total) {
    return run ? 99.0 : 99.0 * (double)iter / (double)total;
}

void llr_progress(void *arg, uint64_t iter, uint64_t total, double) {
    Progress &P = *static_cast<Progress *>(arg);
    show_progress(P, iter, llr_pct(P.run, iter, total));
}

void llr_confirm(void *arg, int run) {
    Progress &P = *static_cast<Progress *>(arg);
    P.run = run;
    P.last_iter = 0;
    fprintf(stderr, "%s result is composite; confirming with run %d at a longer FFT length.\n",
            P.A->expr.c_str(), run + 1);
}

int run_llr(const Args &A) {
    uint32_t Pv = 0;
    uint64_t factor = 0;
    if (kite_llr_choose_P(A.k, A.n, &Pv, &factor)) {
        fprintf(stderr, "kite-prst: no P found for %s\n", A.expr.c_str());
        return 1;
    }
    if (!announce_fft(A)) {
        fprintf(stderr, "kite-prst: cannot set up arithmetic for %s\n", A.expr.c_str());
        return 1;
    }
    if (factor) {  // a factor of 4P(P^2+4) divides N: reported before the test line
        line("%s is not prime, divisible by %" PRIu64 ". Time: 0.0 s.\n", A.expr.c_str(), factor);
        result_line("%s is not prime, divisible by %" PRIu64 ". Time: 0 s.\n", A.expr.c_str(), factor);
        return 0;
    }
    line("Morrison (LLR) test of %s, P = %u, Q = -1, complexity = %" PRIu64 ".\n", A.expr.c_str(), Pv, A.n);
    char ckpt[96];
    snprintf(ckpt, sizeof ckpt, "kite_%u.llr.ckpt", fingerprint(A.k, A.n));
    uint64_t at = 0;
    int at_run = 0;
    if (kite_llr_checkpoint_iter(ckpt, A.k, A.n, &at, &at_run) == 0 && (at > 0 || at_run > 0))
        line("%s restarting at %.1f%%.\n", A.expr.c_str(), llr_pct(at_run, at, A.n - 1));
    Progress P;
    P.A = &A;
    P.last = wall();
    kite_llr_opts lo;
    memset(&lo, 0, sizeof lo);
    lo.fft = engine_opts(A);
    lo.checkpoint = ckpt;
    lo.checkpoint_secs = 300;
    lo.progress = llr_progress;
    lo.on_confirm = llr_confirm;
    lo.progress_arg = &P;
    lo.abort_flag = &g_abort;
    kite_llr_result r;
    memset(&r, 0, sizeof r);
    const int rc = kite_llr(A.k, A.n, &lo, &r);
    if (rc == KITE_LLR_ABORTED) {
        line("Terminating...\n");
        return 1;
    }
    if (rc) {
        line("%s LLR test failed (error %d).\n", A.expr.c_str(), rc);
        return 1;
    }
    remove(ckpt);
    if (r.is_prime) {
        line("%s is prime! Time: %.1f s.\n", A.expr.c_str(), r.seconds);
        result_line("%s is prime! Time: %d s.\n", A.expr.c_str(), (int)r.seconds);
    } else {
        line("%s is not a probable prime. Have you run Fermat test first? RES64: %016" PRIX64 ", time: %.1f s.\n",
             A.expr.c_str(), r.res64, r.seconds);
        result_line("%s is not a probable prime. Have you run Fermat test first? RES64: %016" PRIX64
                    ", time: %d s.\n",
                    A.expr.c_str(), r.res64, (int)r.seconds);
    }
    return 0;
}

// -proof build C [keep]: the certificate from point 0, point C and the products
int run_build(const Args &A) {
    kite_opts o = engine_opts(A);
    kite_ctx *c = kite_ctx_new(A.k, A.n, -1, &o);
    if (!c) return 1;
    const std::string prefix = "prst_" + std::to_string(fingerprint(A.k, A.n));
    const double t0 = wall();
    line("Using HIP FFT length %s.\n", fft_label(kite_fft_len(c)).c_str());
    int depth = 0;
    while ((1 << depth) < A.proof_count) depth++;
    line("%s Building certificate from %d products.\n", A.expr.c_str(), depth);
    uint64_t res = 0;
    const int rc = prst::build_cert(c, A.k, A.n, A.a, A.proof_count, prefix, A.proof_keep, &res);
    kite_ctx_free(c);
    if (rc) return 1;
    line("%s certificate RES64: %016" PRIX64 ", time: %.1f s.\n", A.expr.c_str(), res, wall() - t0);
    result_line("%s certificate RES64: %016" PRIX64 ", time: %d s.\n", A.expr.c_str(), res, (int)(wall() - t0));
    return 0;
}

// -proof cert FILE: verify a certificate (FILE "default": prst_FP.cert)
int run_cert(const Args &A) {
    kite_opts o = engine_opts(A);
    kite_ctx *c = kite_ctx_new(A.k, A.n, -1, &o);
    if (!c) return 1;
    const std::string path = A.cert_path.empty() || A.cert_path == "default"
                                 ? "prst_" + std::to_string(fingerprint(A.k, A.n)) + ".cert"
                                 : A.cert_path;
    const double t0 = wall();
    line("Using HIP FFT length %s.\n", fft_label(kite_fft_len(c)).c_str());
    uint64_t res = 0;
    uint32_t M = 0;
    const int rc = prst::verify_cert(c, A.k, A.n, path, &res, &M);
    kite_ctx_free(c);
    if (rc) return 1;
    line("Verifying certificate of %s, complexity = %u.\n", A.expr.c_str(), M);
    line("%s certificate RES64: %016" PRIX64 ", time: %.1f s.\n", A.expr.c_str(), res, wall() - t0);
    result_line("%s certificate RES64: %016" PRIX64 ", time: %d s.\n", A.expr.c_str(), res, (int)(wall() - t0));
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    Args A;
    const Parsed p = parse(A, argc, argv);
    if (p == Parsed::kHandOff) hand_off(argv);
    if (p == Parsed::kUsage) {
        fprintf(stderr,
                "usage: %s K*2^N-1 -d -fermat a 3 [-t T] [-time [write S] [progress S] [coarse]]\n"
                "              [-proof save|build C [keep]]\n"
                "       %s K*2^N-1 -d -proof cert FILE\n"
                "       %s K*2^N-1 -d [-t T]          (Morrison/LLR test)\n"
                "       %s K*2^N-1 -fermat a 3 -fft info\n"
                "       %s -v\n"
                "(other forms and options run PRST: $KITE_PRST or prst beside this binary)\n",
                argv[0], argv[0], argv[0], argv[0], argv[0]);
        return 1;
    }
    if (A.version) {
        printf("kite version %s\n", KITE_VERSION);
        return 0;
    }
    signal(SIGTERM, on_signal);
    signal(SIGINT, on_signal);
    if (A.proof_op == 2) return run_build(A);
    if (A.proof_op == 3) return run_cert(A);
    return A.llr ? run_llr(A) : run_fermat(A);
}
