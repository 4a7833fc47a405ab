// Checkpoint files of the Fermat and LLR drivers.
//
// The formats are the reference's, byte for byte, so that a test stopped by one program can be
// finished by another (reference src/prp.c:146-212, src/llr.c:92-146):
//
//   KITECKP1  magic[8] k n base iter nlimbs fft_len | x[nlimbs] | d[nlimbs] | fnv
//   KITELLR1  magic[8] k n P iter nlimbs fft_len run prev0 prev1 | x[nlimbs] | fnv
//
// every field a little-endian u64 and fnv the FNV-1a-64 hash of all bytes before it ("AHNCKPT1"
// is accepted as a KITECKP1 magic). Here a file is built as one byte image (fields appended to a
// buffer, the hash taken over the buffer) and written to "<path>.tmp", then renamed over <path>,
// so a reader never sees a partial file; a reader loads the whole file, checks its length and
// hash, then decodes the fields.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "kite_internal.h"

namespace kite {
namespace {

constexpr uint64_t kFnvBasis = 0xcbf29ce484222325ull, kFnvPrime = 0x100000001b3ull;

uint64_t fnv1a64(const uint8_t *p, size_t n) {
    uint64_t h = kFnvBasis;
    for (size_t i = 0; i < n; i++) h = (h ^ p[i]) * kFnvPrime;
    return h;
}

// A file image under construction.
struct Image {
    std::vector<uint8_t> bytes;
    void tag(const char m[8]) { bytes.insert(bytes.end(), m, m + 8); }
    void u64(uint64_t v) {
        for (int i = 0; i < 8; i++) bytes.push_back((uint8_t)(v >> (8 * i)));
    }
    void limbs(const uint64_t *w, size_t n) {
        for (size_t i = 0; i < n; i++) u64(w[i]);
    }
    // appends the hash and replaces `path` atomically; 0 on success
    int commit(const char *path) {
        u64(fnv1a64(bytes.data(), bytes.size()));
        const std::string tmp = std::string(path) + ".tmp";
        FILE *f = fopen(tmp.c_str(), "wb");
        if (!f) return -1;
        const bool wrote = fwrite(bytes.data(), 1, bytes.size(), f) == bytes.size();
        const bool closed = fclose(f) == 0;
        if (!wrote || !closed || rename(tmp.c_str(), path) != 0) {
            remove(tmp.c_str());
            return -1;
        }
        return 0;
    }
};

// A whole file, verified against its trailing hash, read field by field.
struct Reader {
    std::vector<uint8_t> bytes;
    size_t at = 0;
    bool ok = false;
    // loads the whole of `path` (a caller decodes the header, then checks the length it implies
    // and the hash with sealed() before trusting the rest)
    explicit Reader(const char *path) {
        FILE *f = fopen(path, "rb");
        if (!f) return;
        uint8_t chunk[1 << 16];
        size_t got;
        while ((got = fread(chunk, 1, sizeof chunk, f)) > 0) bytes.insert(bytes.end(), chunk, chunk + got);
        fclose(f);
        ok = bytes.size() >= 16;
    }
    bool tag_is(const char *m) const { return ok && memcmp(bytes.data(), m, 8) == 0; }
    uint64_t u64() {
        if (at + 8 > bytes.size()) {
            ok = false;
            return 0;
        }
        uint64_t v = 0;
        for (int i = 7; i >= 0; i--) v = (v << 8) | bytes[at + i];
        at += 8;
        return v;
    }
    void limbs(uint64_t *w, size_t n) {
        for (size_t i = 0; i < n; i++) w[i] = u64();
    }
    // the file is exactly `body` bytes plus a matching hash
    bool sealed(size_t body) const {
        if (!ok || bytes.size() != body + 8) return false;
        uint64_t stored = 0;
        for (int i = 7; i >= 0; i--) stored = (stored << 8) | bytes[body + i];
        return stored == fnv1a64(bytes.data(), body);
    }
};

const char kPrpTag[8] = {'K', 'I', 'T', 'E', 'C', 'K', 'P', '1'};
const char kPrpTagOld[8] = {'A', 'H', 'N', 'C', 'K', 'P', 'T', '1'};
const char kLlrTag[8] = {'K', 'I', 'T', 'E', 'L', 'L', 'R', '1'};
constexpr size_t kPrpHeader = 8 + 6 * 8, kLlrHeader = 8 + 9 * 8;

struct PrpHeader {
    uint64_t k = 0, n = 0, base = 0, iter = 0, nlimbs = 0, fft_len = 0;
};

bool read_prp_header(Reader &r, PrpHeader *h) {
    if (!(r.tag_is(kPrpTag) || r.tag_is(kPrpTagOld))) return false;
    r.at = 8;
    h->k = r.u64();
    h->n = r.u64();
    h->base = r.u64();
    h->iter = r.u64();
    h->nlimbs = r.u64();
    h->fft_len = r.u64();
    return r.ok;
}

struct LlrHeader {
    uint64_t k = 0, n = 0, P = 0, iter = 0, nlimbs = 0, fft_len = 0, run = 0, prev0 = 0, prev1 = 0;
};

bool read_llr_header(Reader &r, LlrHeader *h) {
    if (!r.tag_is(kLlrTag)) return false;
    r.at = 8;
    h->k = r.u64();
    h->n = r.u64();
    h->P = r.u64();
    h->iter = r.u64();
    h->nlimbs = r.u64();
    h->fft_len = r.u64();
    h->run = r.u64();
    h->prev0 = r.u64();
    h->prev1 = r.u64();
    return r.ok;
}

}  // namespace

int prp_checkpoint_write(const char *path, uint64_t k, uint64_t n, uint32_t base, uint64_t iter, uint32_t L,
                         const uint64_t *x, const uint64_t *d, size_t nl) {
    Image im;
    im.bytes.reserve(kPrpHeader + 16 * nl + 8);
    im.tag(kPrpTag);
    for (uint64_t v : {k, n, (uint64_t)base, iter, (uint64_t)nl, (uint64_t)L}) im.u64(v);
    im.limbs(x, nl);
    im.limbs(d, nl);
    return im.commit(path);
}

int prp_checkpoint_read(const char *path, uint64_t k, uint64_t n, uint32_t base, uint64_t *iter, uint32_t *L,
                        uint64_t *x, uint64_t *d, size_t nl) {
    Reader r(path);
    PrpHeader h;
    if (!read_prp_header(r, &h) || h.k != k || h.n != n || h.base != base || h.nlimbs != nl) return -1;
    if (!r.sealed(kPrpHeader + 16 * nl)) return -1;
    r.limbs(x, nl);
    r.limbs(d, nl);
    *iter = h.iter;
    *L = (uint32_t)h.fft_len;
    return 0;
}

int llr_checkpoint_write(const char *path, uint64_t k, uint64_t n, const LlrCheckpoint &c, const uint64_t *x,
                         size_t nl) {
    Image im;
    im.bytes.reserve(kLlrHeader + 8 * nl + 8);
    im.tag(kLlrTag);
    for (uint64_t v : {k, n, (uint64_t)c.P, c.iter, (uint64_t)nl, (uint64_t)c.L, (uint64_t)c.run, c.prev[0], c.prev[1]})
        im.u64(v);
    im.limbs(x, nl);
    return im.commit(path);
}

int llr_checkpoint_read(const char *path, uint64_t k, uint64_t n, LlrCheckpoint *c, uint64_t *x, size_t nl) {
    Reader r(path);
    LlrHeader h;
    if (!read_llr_header(r, &h) || h.k != k || h.n != n || h.nlimbs != nl || h.run > 2 || h.iter > n - 1) return -1;
    if (!r.sealed(kLlrHeader + 8 * nl)) return -1;
    r.limbs(x, nl);
    c->iter = h.iter;
    c->P = (uint32_t)h.P;
    c->L = (uint32_t)h.fft_len;
    c->run = (int)h.run;
    c->prev[0] = h.prev0;
    c->prev[1] = h.prev1;
    return 0;
}

}  // namespace kite

// The iteration a file records, from its header alone (the reference's tools read it to announce
// a resume before a test starts).
extern "C" int kite_prp_checkpoint_iter(const char *path, uint64_t k, uint64_t n, uint32_t base, uint64_t *iter) {
    kite::Reader r(path);
    kite::PrpHeader h;
    if (!kite::read_prp_header(r, &h) || h.k != k || h.n != n || h.base != base) return -1;
    *iter = h.iter;
    return 0;
}

extern "C" int kite_llr_checkpoint_iter(const char *path, uint64_t k, uint64_t n, uint64_t *iter, int *run) {
    kite::Reader r(path);
    kite::LlrHeader h;
    if (!kite::read_llr_header(r, &h) || h.k != k || h.n != n) return -1;
    *iter = h.iter;
    if (run) *run = (int)h.run;
    return 0;
}
