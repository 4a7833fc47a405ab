// PRST-compatible Fermat proofs of k*2^n-1 (reference src/prstproof.c, after PRST 14.1
// src/proof.cpp and arithmetic/file.cpp). File layouts, point positions, challenges and the order
// of every power and product must be PRST's for its certificates to verify here and ours there;
// the arithmetic itself is the GPU engine's, through kite.h.
#include "prst_proof.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>

namespace prst {

// ---- MD5 (RFC 1321) ----------------------------------------------------------------------------

namespace {
constexpr uint32_t kMd5Shift[64] = {7, 12, 17, 22, 7, 12, 17, 22, 7, 12, 17, 22, 7, 12, 17, 22,
                                    5, 9,  14, 20, 5, 9,  14, 20, 5, 9,  14, 20, 5, 9,  14, 20,
                                    4, 11, 16, 23, 4, 11, 16, 23, 4, 11, 16, 23, 4, 11, 16, 23,
                                    6, 10, 15, 21, 6, 10, 15, 21, 6, 10, 15, 21, 6, 10, 15, 21};

// K[i] = floor(|sin(i + 1)| 2^32), evaluated once in extended precision (the RFC's definition;
// the test vectors in the test suite pin the table)
struct Md5Table {
    uint32_t K[64];
    Md5Table() {
        for (int i = 0; i < 64; i++) K[i] = (uint32_t)std::floor(std::fabs(std::sin((long double)(i + 1))) * 4294967296.0L);
    }
};
const Md5Table &md5_table() {
    static const Md5Table t;
    return t;
}
inline uint32_t rotl(uint32_t x, uint32_t s) { return (x << s) | (x >> (32 - s)); }
}  // namespace

Md5::Md5() : h_{0x67452301u, 0xefcdab89u, 0x98badcfeu, 0x10325476u}, buf_{} {}

void Md5::block(const uint8_t *b) {
    const uint32_t *K = md5_table().K;
    uint32_t m[16];
    for (int i = 0; i < 16; i++) m[i] = b[4 * i] | b[4 * i + 1] << 8 | b[4 * i + 2] << 16 | (uint32_t)b[4 * i + 3] << 24;
    uint32_t a = h_[0], bb = h_[1], c = h_[2], d = h_[3];
    for (int i = 0; i < 64; i++) {
        uint32_t f;
        int g;
        if (i < 16) {
            f = (bb & c) | (~bb & d);
            g = i;
        } else if (i < 32) {
            f = (d & bb) | (~d & c);
            g = (5 * i + 1) % 16;
        } else if (i < 48) {
            f = bb ^ c ^ d;
            g = (3 * i + 5) % 16;
        } else {
            f = c ^ (bb | ~d);
            g = (7 * i) % 16;
        }
        const uint32_t t = d;
        d = c;
        c = bb;
        bb = bb + rotl(a + f + K[i] + m[g], kMd5Shift[i]);
        a = t;
    }
    h_[0] += a;
    h_[1] += bb;
    h_[2] += c;
    h_[3] += d;
}

void Md5::update(const void *p, size_t n) {
    const uint8_t *s = static_cast<const uint8_t *>(p);
    size_t fill = len_ % 64;
    len_ += n;
    while (n > 0) {
        const size_t take = std::min<size_t>(64 - fill, n);
        memcpy(buf_ + fill, s, take);
        fill += take;
        s += take;
        n -= take;
        if (fill == 64) {
            block(buf_);
            fill = 0;
        }
    }
}

void Md5::final(uint8_t out[16]) {
    const uint64_t bits = len_ * 8;
    const uint8_t one = 0x80, zero = 0;
    update(&one, 1);
    while (len_ % 64 != 56) update(&zero, 1);
    uint8_t lb[8];
    for (int i = 0; i < 8; i++) lb[i] = (uint8_t)(bits >> (8 * i));
    update(lb, 8);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) out[4 * i + j] = (uint8_t)(h_[i] >> (8 * j));
}

std::string Md5::hex(const void *p, size_t n) {
    Md5 m;
    m.update(p, n);
    uint8_t d[16];
    m.final(d);
    static const char *digits = "0123456789abcdef";
    std::string s(32, '0');
    for (int i = 0; i < 16; i++) {
        s[2 * i] = digits[d[i] >> 4];
        s[2 * i + 1] = digits[d[i] & 15];
    }
    return s;
}

// ---- fingerprints ------------------------------------------------------------------------------

uint32_t fingerprint(uint64_t k, uint64_t n) {
    const uint64_t p = 3417905339u;
    unsigned __int128 r = 1, b = 2;
    for (uint64_t e = n; e; e >>= 1, b = b * b % p)
        if (e & 1) r = r * b % p;
    return (uint32_t)(((unsigned __int128)(k % p) * r + p - 1) % p);
}

uint32_t unique_fingerprint(uint32_t fp, const std::string &id) {
    Md5 m;
    m.update(&fp, 4);
    m.update(id.data(), id.size());
    uint8_t d[16];
    m.final(d);
    return d[0] | d[1] << 8 | d[2] << 16 | (uint32_t)d[3] << 24;
}

uint32_t proof_fingerprint(uint64_t k, uint64_t n, uint32_t a) {
    return unique_fingerprint(fingerprint(k, n), std::to_string(a) + "." + std::to_string(n));
}

// ---- state files -------------------------------------------------------------------------------

namespace {
constexpr uint32_t kMagic = 0x9f2b3cd4u;
constexpr uint32_t kAppId = 4;  // PRST's application id; version 0

bool write_file(const std::string &path, const void *p, size_t n) {
    const std::string tmp = path + ".tmp";
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) return false;
    bool ok = fwrite(p, 1, n, f) == n;
    ok = (fclose(f) == 0) && ok;
    if (!ok || rename(tmp.c_str(), path.c_str())) {
        remove(tmp.c_str());
        return false;
    }
    return true;
}

bool read_file(const std::string &path, std::vector<uint8_t> *out) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return false;
    out->clear();
    uint8_t buf[1 << 16];
    size_t got;
    while ((got = fread(buf, 1, sizeof buf, f)) > 0) out->insert(out->end(), buf, buf + got);
    const bool ok = !ferror(f);
    fclose(f);
    return ok;
}

void put32(std::vector<uint8_t> *b, uint32_t v) {
    for (int i = 0; i < 4; i++) b->push_back((uint8_t)(v >> (8 * i)));
}
}  // namespace

int write_state(const std::string &path, int type, uint32_t fp, int32_t iteration, const std::vector<Giant> &g) {
    std::vector<uint8_t> b;
    put32(&b, kMagic);
    put32(&b, kAppId | (uint32_t)(uint8_t)type << 16);
    if (fp) put32(&b, fp);
    put32(&b, (uint32_t)iteration);
    for (const Giant &x : g) {
        put32(&b, (uint32_t)x.w.size());
        for (uint32_t w : x.w) put32(&b, w);
    }
    const std::string hex = Md5::hex(b.data(), b.size());
    return write_file(path, b.data(), b.size()) && write_file(path + ".md5", hex.data(), hex.size()) ? 0 : -1;
}

int read_state(const std::string &path, int type, uint32_t fp, int32_t *iteration, std::vector<Giant> *g, int maxg) {
    std::vector<uint8_t> b, m;
    if (!read_file(path, &b)) return -1;
    if (read_file(path + ".md5", &m)) {  // a sidecar that is present must match
        const std::string hex = Md5::hex(b.data(), b.size());
        if (m.size() < 32 || memcmp(hex.data(), m.data(), 32) != 0) return -1;
    }
    size_t off = 0;
    auto get = [&](uint32_t *v) {
        if (off + 4 > b.size()) return false;
        *v = b[off] | b[off + 1] << 8 | b[off + 2] << 16 | (uint32_t)b[off + 3] << 24;
        off += 4;
        return true;
    };
    uint32_t v;
    if (!get(&v) || v != kMagic) return -1;
    if (!get(&v) || (v & 0xff) != kAppId || ((v >> 16) & 0xff) != (uint32_t)type) return -1;
    if (fp && (!get(&v) || v != fp)) return -1;
    if (!get(&v)) return -1;
    *iteration = (int32_t)v;
    g->assign(maxg, Giant());
    for (int i = 0; i < maxg && off < b.size(); i++) {  // trailing giants may be absent
        uint32_t len;
        if (!get(&len) || (int32_t)len < 0 || off + 4 * (size_t)len > b.size()) return -1;
        (*g)[i].w.resize(len);
        for (uint32_t j = 0; j < len; j++) get(&(*g)[i].w[j]);
    }
    return 0;
}

void remove_state(const std::string &path) {
    remove(path.c_str());
    remove((path + ".md5").c_str());
}

// ---- point positions (proof.cpp:160) ----------------------------------------------------------
// Point i (0 < i < count) sits where the binary subdivision of the iterations leads: from the top
// level down, the interval length halves (rounding down) and the upper half is taken where i has
// its bit; every level visited adds one when the iteration count has that level's bit set. The
// walk stops below i's lowest set bit.
int points(uint32_t iterations, int count, std::vector<uint32_t> *pos) {
    pos->assign(count + 1, 0);
    int spacing = 0;
    for (int i = 1; i < count; i++) {
        spacing = (int)iterations;
        uint32_t at = 0;
        for (int bit = count / 2; bit > 0 && (i & (2 * bit - 1)) != 0; bit >>= 1) {
            spacing /= 2;
            if (i & bit) at += (uint32_t)spacing;
            if (iterations & (uint32_t)(count / bit / 2)) at++;
        }
        (*pos)[i] = at;
    }
    (*pos)[count] = iterations;
    return spacing;
}

// ---- numbers -----------------------------------------------------------------------------------

Giant from_num(kite_ctx *ctx, const kite_num *x) {
    const size_t nl = kite_limbs(ctx);
    std::vector<uint64_t> l(nl);
    kite_get_limbs(ctx, x, l.data());
    Giant g;
    g.w.resize(2 * nl);
    for (size_t i = 0; i < nl; i++) {
        g.w[2 * i] = (uint32_t)l[i];
        g.w[2 * i + 1] = (uint32_t)(l[i] >> 32);
    }
    g.trim();
    return g;
}

void to_num(kite_ctx *ctx, const Giant &g, kite_num *x) {
    const size_t nl = kite_limbs(ctx);
    std::vector<uint64_t> l(nl, 0);
    for (size_t i = 0; i < g.w.size() && i < 2 * nl; i++) l[i / 2] |= (uint64_t)g.w[i] << (32 * (i & 1));
    kite_set_limbs(ctx, x, l.data());
}

namespace {
bool is_small_prime(unsigned p) {
    if (p < 2) return false;
    for (unsigned q = 2; q * q <= p; q++)
        if (p % q == 0) return false;
    return true;
}

// X = X0^e, left to right (proof.cpp:538); X and X0 are different numbers
void power(kite_ctx *ctx, unsigned __int128 e, kite_num *X, const kite_num *X0) {
    int top = 0;
    while (top < 127 && (e >> (top + 1)) != 0) top++;
    kite_copy(ctx, X, X0);
    for (int bit = top - 1; bit >= 0; bit--) {
        kite_square(ctx, X);
        if ((e >> bit) & 1) kite_mul(ctx, X, X0);
    }
}

bool roundoff_ok(kite_ctx *ctx) { return kite_max_roundoff(ctx) < 0.40; }

// One number kept by the prover: kite_num with RAII release.
struct Num {
    kite_ctx *ctx;
    kite_num *p;
    explicit Num(kite_ctx *c) : ctx(c), p(kite_num_new(c)) {}
    ~Num() { kite_num_free(ctx, p); }
    Num(const Num &) = delete;
    Num &operator=(const Num &) = delete;
};

int depth_of(int count) {
    int t = 0;
    while ((1 << t) < count) t++;
    return t;
}

std::string file_of(const std::string &prefix, const char *kind, int index) {
    return prefix + "." + kind + "." + std::to_string(index);
}

// a point or product file with the expected iteration, into x
bool load(kite_ctx *ctx, const std::string &path, int type, uint32_t fp, int32_t expect, kite_num *x, Giant *keep = nullptr) {
    std::vector<Giant> g;
    int32_t it;
    if (read_state(path, type, fp, &it, &g, 1) != 0 || it != expect) {
        fprintf(stderr, "%s is missing or corrupt.\n", path.c_str());
        return false;
    }
    to_num(ctx, g[0], x);
    if (keep) *keep = g[0];
    return true;
}
}  // namespace

unsigned __int128 challenge(uint32_t fp, const Giant &Y, const Giant &D) {
    Md5 m;
    m.update(&fp, 4);
    if (!Y.w.empty()) m.update(Y.w.data(), 4 * Y.w.size());
    if (!D.w.empty()) m.update(D.w.data(), 4 * D.w.size());
    uint8_t d[16];
    m.final(d);
    uint32_t lo = d[0] | d[1] << 8 | d[2] << 16 | (uint32_t)d[3] << 24;
    uint32_t hi = d[4] | d[5] << 8 | d[6] << 16 | (uint32_t)d[7] << 24;
    lo |= 1;
    if (hi == 0) hi = 1;
    unsigned __int128 h = (unsigned __int128)hi << 32 | lo;
    for (;; h += 2) {  // proof.cpp:522: the next odd value without a prime factor below 1000
        bool smooth = false;
        for (unsigned p = 2; p < 1000 && !smooth; p++) smooth = is_small_prime(p) && h % p == 0;
        if (!smooth) return h;
    }
}

// ---- save: the products (ProofSave::execute, proof.cpp:549) -----------------------------------
// Level i combines the points at the odd multiples of count / 2^(i+1); each point is raised by
// the challenges of the levels above along the binary tree of its index.
int save_products(kite_ctx *ctx, uint64_t k, uint64_t n, uint32_t a, int count, const std::string &prefix,
                  bool keep, volatile int *abort_flag, uint64_t *raw_res64) {
    const uint32_t fpN = fingerprint(k, n), fpP = proof_fingerprint(k, n, a);
    const int t = depth_of(count);
    std::vector<uint32_t> pos;
    points((uint32_t)n, count, &pos);
    Num Y(ctx), D(ctx), T(ctx);
    std::vector<std::unique_ptr<Num>> tree;
    for (int i = 0; i < t; i++) tree.emplace_back(new Num(ctx));
    std::vector<unsigned __int128> h(t);
    kite_reset_roundoff(ctx);
    if (!load(ctx, file_of(prefix, "proof", count), kTypeValue, fpP, (int32_t)pos[count], Y.p)) return -1;

    // products of an interrupted run: Y and the challenges follow from them (one power each)
    int first = 0;
    for (; first < t; first++) {
        std::vector<Giant> g;
        int32_t it;
        if (read_state(file_of(prefix, "prod", first), kTypeProduct, fpP, &it, &g, 1) != 0 || it != first) break;
        to_num(ctx, g[0], D.p);
        h[first] = challenge(fpN, from_num(ctx, Y.p), g[0]);
        power(ctx, h[first], T.p, D.p);
        kite_mul(ctx, Y.p, T.p);
    }
    for (int i = first; i < t; i++) {
        if (abort_flag && *abort_flag) return kAborted;
        if (i == 0) {
            if (!load(ctx, file_of(prefix, "proof", count / 2), kTypeValue, fpP, (int32_t)pos[count / 2], D.p)) return -1;
        } else {
            for (int j = 0; j < (1 << i); j++) {
                if (abort_flag && *abort_flag) return kAborted;
                const int idx = (1 + 2 * j) << (t - i - 1);
                if (!load(ctx, file_of(prefix, "proof", idx), kTypeValue, fpP, (int32_t)pos[idx], D.p)) return -1;
                for (int up = 1; up <= i; up++) {
                    if ((j & (1 << (up - 1))) == 0) {  // left child: park it one level up
                        kite_copy(ctx, tree[i - up]->p, D.p);
                        break;
                    }
                    power(ctx, h[i - up], T.p, tree[i - up]->p);
                    kite_mul(ctx, D.p, T.p);
                }
            }
        }
        const Giant dg = from_num(ctx, D.p);
        const std::string path = file_of(prefix, "prod", i);
        if (write_state(path, kTypeProduct, fpP, i, {dg})) {
            fprintf(stderr, "Write to %s failed.\n", path.c_str());
            return -1;
        }
        h[i] = challenge(fpN, from_num(ctx, Y.p), dg);
        power(ctx, h[i], T.p, D.p);
        kite_mul(ctx, Y.p, T.p);
        if (!roundoff_ok(ctx)) {
            fprintf(stderr, "round-off error while compressing proof points.\n");
            return -1;
        }
    }
    *raw_res64 = from_num(ctx, Y.p).res64();
    if (!keep)
        for (int i = 1; i < count; i++) remove_state(file_of(prefix, "proof", i));
    return 0;
}

// ---- build: the certificate (ProofBuild::execute, proof.cpp:660) ------------------------------

namespace {
uint64_t mulmod(uint64_t x, uint64_t y, uint64_t m) { return (uint64_t)((unsigned __int128)x * y % m); }

// (N - 1) mod m = (k 2^n - 2) mod m
uint64_t nm1_mod(uint64_t k, uint64_t n, uint64_t m) {
    uint64_t r = 1 % m, b = 2 % m;
    for (uint64_t e = n; e; e >>= 1, b = mulmod(b, b, m))
        if (e & 1) r = mulmod(r, b, m);
    return (mulmod(k % m, r, m) + 2 * m - 2 % m) % m;
}

bool is_one(const std::vector<uint64_t> &l) {
    if (l.empty() || l[0] != 1) return false;
    for (size_t i = 1; i < l.size(); i++)
        if (l[i]) return false;
    return true;
}

// PRST's roots-of-unity check (proof.cpp:82, security 24): R = a^(N-1) from x_n; unless R = 1 (a
// probable prime), R raised to 2 and to the largest power below 2^31 of every odd prime below
// 2^12 dividing N-1 must not be 1.
bool roots_of_unity_ok(kite_ctx *ctx, uint64_t k, uint64_t n, uint32_t a, const Giant &xn) {
    const size_t nl = kite_limbs(ctx);
    std::vector<uint64_t> l(nl, 0), N(nl, 0);
    for (size_t i = 0; i < xn.w.size() && i < 2 * nl; i++) l[i / 2] |= (uint64_t)xn.w[i] << (32 * (i & 1));
    N[n / 64] = k << (n % 64);
    if (n % 64 && n / 64 + 1 < nl) N[n / 64 + 1] = k >> (64 - n % 64);
    for (size_t j = 0; j < nl; j++)  // N = k 2^n - 1
        if (N[j]--) break;
    if (kite_residue_nm1(l.data(), N.data(), nl, a) != 0 || is_one(l)) return true;
    Num R(ctx), X(ctx), T(ctx);
    kite_set_limbs(ctx, R.p, l.data());
    kite_copy(ctx, X.p, R.p);
    kite_square(ctx, X.p);  // 2 || N - 1 = 2 (k 2^(n-1) - 1)
    for (uint32_t p = 3; p < 4096; p += 2) {
        if (!is_small_prime(p) || nm1_mod(k, n, p) != 0) continue;
        uint64_t d = p;
        while (d * p < 2147483647u && nm1_mod(k, n, d * p) == 0) d *= p;
        kite_copy(ctx, T.p, X.p);
        int top = 0;
        while ((d >> (top + 1)) != 0) top++;
        for (int bit = top - 1; bit >= 0; bit--) {
            kite_square(ctx, X.p);
            if ((d >> bit) & 1) kite_mul(ctx, X.p, T.p);
        }
    }
    kite_get_limbs(ctx, X.p, l.data());
    return !is_one(l);
}
}  // namespace

int build_cert(kite_ctx *ctx, uint64_t k, uint64_t n, uint32_t a, int count, const std::string &prefix, bool keep,
               uint64_t *cert_res64) {
    const uint32_t fpN = fingerprint(k, n), fpP = proof_fingerprint(k, n, a);
    const int t = depth_of(count);
    std::vector<uint32_t> pos;
    points((uint32_t)n, count, &pos);
    Num X(ctx), Y(ctx), D(ctx), T(ctx);
    kite_reset_roundoff(ctx);
    Giant g0, gn;
    if (!load(ctx, file_of(prefix, "proof", 0), kTypeValue, fpP, 0, X.p, &g0)) return -1;
    {  // point 0 must be a^k (proof.cpp:219)
        kite_set_u64(ctx, T.p, a);
        kite_set_u64(ctx, D.p, 1);
        for (int b = 63; b >= 0; b--) {
            kite_square(ctx, D.p);
            if ((k >> b) & 1) kite_mul(ctx, D.p, T.p);
        }
        if (from_num(ctx, D.p).w != g0.w) {
            fprintf(stderr, "invalid a^k.\n");
            return -1;
        }
    }
    if (!load(ctx, file_of(prefix, "proof", count), kTypeValue, fpP, (int32_t)pos[count], Y.p, &gn)) return -1;
    if (!roots_of_unity_ok(ctx, k, n, a, gn)) {
        fprintf(stderr, "roots of unity check failed.\n");
        return -1;
    }
    uint32_t M = pos[count];
    for (int i = 0; i < t; i++, M >>= 1) {
        Giant dg;
        if (!load(ctx, file_of(prefix, "prod", i), kTypeProduct, fpP, i, D.p, &dg)) return -1;
        const unsigned __int128 h = challenge(fpN, from_num(ctx, Y.p), dg);
        power(ctx, h * (M % 2 ? 2 : 1), T.p, X.p);  // an odd level doubles the exponent (base 2)
        kite_copy(ctx, X.p, T.p);
        kite_mul(ctx, X.p, D.p);
        power(ctx, h, T.p, D.p);
        kite_mul(ctx, Y.p, T.p);
        if (!roundoff_ok(ctx)) {
            fprintf(stderr, "round-off error while building the certificate.\n");
            return -1;
        }
    }
    const Giant yg = from_num(ctx, Y.p);
    if (yg.w.empty()) {
        fprintf(stderr, "invalid proof, the certificate is zero.\n");
        return -1;
    }
    *cert_res64 = yg.res64();
    const std::string path = prefix + ".cert";
    if (write_state(path, kTypeCert, fpN, (int32_t)M, {from_num(ctx, X.p), Giant(), Giant()})) {
        fprintf(stderr, "Write to %s failed.\n", path.c_str());
        return -1;
    }
    if (!keep) {
        remove_state(file_of(prefix, "proof", 0));
        remove_state(file_of(prefix, "proof", count));
        for (int i = 0; i < t; i++) remove_state(file_of(prefix, "prod", i));
    }
    return 0;
}

// ---- verify: X^(2^M) under a Gerbicz check ----------------------------------------------------
// d_b = prod_{j<=b} x_{jB}; d_T = d_{T-1}^(2^B) x_0 is checked once after the blocks, the tail of
// M mod B squarings twice over.
int verify_cert(kite_ctx *ctx, uint64_t k, uint64_t n, const std::string &path, uint64_t *cert_res64, uint32_t *Mout) {
    std::vector<Giant> c;
    int32_t M;
    if (read_state(path, kTypeCert, fingerprint(k, n), &M, &c, 3) != 0 || M < 0) {
        fprintf(stderr, "Invalid certificate file.\n");
        return -1;
    }
    if (!c[1].w.empty()) {  // a Gerbicz-Li certificate (non-smooth tests)
        fprintf(stderr, "Gerbicz-Li certificates are not supported.\n");
        return -1;
    }
    if (Mout) *Mout = (uint32_t)M;
    Num x0(ctx), x(ctx), d(ctx), e(ctx);
    to_num(ctx, c[0], x0.p);
    uint32_t B = (uint32_t)std::sqrt((double)M);
    if (B < 1) B = 1;
    for (int attempt = 0; attempt < 3; attempt++) {
        kite_reset_roundoff(ctx);
        kite_copy(ctx, x.p, x0.p);
        kite_copy(ctx, d.p, x0.p);
        const uint32_t blocks = (uint32_t)M / B;
        for (uint32_t b = 0; b < blocks; b++) {
            if (b + 1 == blocks) kite_copy(ctx, e.p, d.p);
            kite_square_n(ctx, x.p, B);
            kite_mul(ctx, d.p, x.p);
        }
        if (blocks) {
            kite_square_n(ctx, e.p, B);
            kite_mul(ctx, e.p, x0.p);
            if (from_num(ctx, d.p).w != from_num(ctx, e.p).w) continue;
        }
        const uint32_t tail = (uint32_t)M - blocks * B;
        kite_copy(ctx, e.p, x.p);
        kite_square_n(ctx, x.p, tail);
        kite_square_n(ctx, e.p, tail);
        const Giant g1 = from_num(ctx, x.p);
        if (g1.w == from_num(ctx, e.p).w && roundoff_ok(ctx)) {
            *cert_res64 = g1.res64();
            return 0;
        }
    }
    return -1;
}

}  // namespace prst
