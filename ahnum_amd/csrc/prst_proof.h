// PRST-compatible Fermat proofs of k*2^n-1 for kite-prst (the reference's src/prstproof.h:1-58,
// which follows PRST 14.1's src/proof.cpp and arithmetic/file.cpp): the proof point, product and
// certificate files PRST reads and writes, the compression of the points into products, the
// certificate build and its verification. Built on the kite.h number API only, so every power and
// product runs on the GPU engine. PRST's "smooth" path: x_0 = a^k, then n squarings.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "kite.h"

namespace prst {

// RFC 1321 MD5, incremental.
class Md5 {
public:
    Md5();
    void update(const void *p, size_t n);
    void final(uint8_t out[16]);
    static std::string hex(const void *p, size_t n);  // lowercase hex digest of a buffer

private:
    void block(const uint8_t *b);
    uint32_t h_[4];
    uint8_t buf_[64];
    uint64_t len_ = 0;
};

// A non-negative integer as PRST stores it: little-endian 32-bit words, no leading zero words.
struct Giant {
    std::vector<uint32_t> w;
    void trim() {
        while (!w.empty() && w.back() == 0) w.pop_back();
    }
    uint64_t res64() const { return (w.size() > 0 ? w[0] : 0) | (w.size() > 1 ? (uint64_t)w[1] << 32 : 0); }
};

enum : int { kTypeValue = 1, kTypeProduct = 3, kTypeCert = 4 };

uint32_t fingerprint(uint64_t k, uint64_t n);                      // N mod 3417905339
uint32_t unique_fingerprint(uint32_t fp, const std::string &id);   // first word of MD5(fp || id)
uint32_t proof_fingerprint(uint64_t k, uint64_t n, uint32_t a);    // unique(fp, "a.n")

// A state file (header, iteration, giants) and its .md5 sidecar. fp = 0: no fingerprint word.
int write_state(const std::string &path, int type, uint32_t fp, int32_t iteration, const std::vector<Giant> &g);
// Checks magic, type, fingerprint and the sidecar when present; reads up to maxg giants (missing
// trailing giants come back empty). 0 on success.
int read_state(const std::string &path, int type, uint32_t fp, int32_t *iteration, std::vector<Giant> *g, int maxg);
void remove_state(const std::string &path);

// Positions pos[0..count] of the proof points along `iterations` squarings (count a power of 2);
// returns the spacing PRST reports.
int points(uint32_t iterations, int count, std::vector<uint32_t> *pos);

Giant from_num(kite_ctx *ctx, const kite_num *x);
void to_num(kite_ctx *ctx, const Giant &g, kite_num *x);

// The 64-bit challenge of a product: MD5(fp || Y || D), its first two words, made odd and free of
// prime factors below 1000.
unsigned __int128 challenge(uint32_t fp, const Giant &Y, const Giant &D);

enum : int { kAborted = -2 };
// Save step: points 0..count of the test are on disk as prefix.proof.I; writes the products
// prefix.prod.i and returns the raw certificate RES64. Products already on disk (an interrupted
// run) are reused. Deletes points 1..count-1 unless keep. Stops between points when *abort_flag.
int save_products(kite_ctx *ctx, uint64_t k, uint64_t n, uint32_t a, int count, const std::string &prefix,
                  bool keep, volatile int *abort_flag, uint64_t *raw_res64);
// Build step: checks point 0 = a^k and that a^(N-1) is no small-order root of unity, writes
// prefix.cert, returns the certificate RES64. Deletes points 0, count and the products unless keep.
int build_cert(kite_ctx *ctx, uint64_t k, uint64_t n, uint32_t a, int count, const std::string &prefix, bool keep,
               uint64_t *cert_res64);
// Verify: squares the certificate's value M times under a Gerbicz check; the RES64 PRST prints.
int verify_cert(kite_ctx *ctx, uint64_t k, uint64_t n, const std::string &path, uint64_t *cert_res64, uint32_t *M);

}  // namespace prst
