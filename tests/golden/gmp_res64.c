// RES64 oracle by GMP: low 64 bits of base^(N-1) mod N, N = k*2^n - 1 (the semantics of the
// reference's tools/gmp_ref.c). Reads "k n [base]" lines on stdin, writes "k,n,base,RES64".
// Also: "k n base x i" lines write the low limb and an FNV-1a-64 hash of the limbs of
// base^(k*2^i) mod N (the Fermat chain's x_i, src/prp.c:4-6).
#include <gmp.h>
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t low64(const mpz_t x) {
    mpz_t t;
    mpz_init(t);
    mpz_fdiv_r_2exp(t, x, 64);
    uint64_t r = 0;
    size_t cnt = 0;
    mpz_export(&r, &cnt, -1, 8, 0, 0, t);
    mpz_clear(t);
    return r;
}

static uint64_t fnv_limbs(const mpz_t x, size_t nl) {
    uint64_t *l = calloc(nl, 8);
    size_t cnt = 0;
    mpz_export(l, &cnt, -1, 8, 0, 0, x);
    uint64_t h = 0xcbf29ce484222325ull;
    const uint8_t *b = (const uint8_t *)l;
    for (size_t i = 0; i < nl * 8; i++) h = (h ^ b[i]) * 0x100000001b3ull;
    free(l);
    return h;
}

int main(void) {
    char line[256];
    mpz_t N, e, b, r;
    mpz_inits(N, e, b, r, NULL);
    while (fgets(line, sizeof line, stdin)) {
        unsigned long k, n, base = 3, i = 0;
        char what[8] = "";
        int got = sscanf(line, "%lu %lu %lu %7s %lu", &k, &n, &base, what, &i);
        if (got < 2) continue;
        if (got < 3) base = 3;
        mpz_set_ui(N, k);
        mpz_mul_2exp(N, N, n);
        mpz_sub_ui(N, N, 1);
        mpz_set_ui(b, base);
        if (got == 5 && !strcmp(what, "x")) {
            mpz_set_ui(e, k);
            mpz_mul_2exp(e, e, i);
            mpz_powm(r, b, e, N);
            size_t nl = (n + 127) / 64;
            printf("%lu,%lu,%lu,%lu,%016" PRIX64 ",%016" PRIX64 "\n", k, n, base, i, low64(r),
                   fnv_limbs(r, nl));
        } else {
            mpz_sub_ui(e, N, 1);
            mpz_powm(r, b, e, N);
            printf("%lu,%lu,%lu,%016" PRIX64 "\n", k, n, base, low64(r));
        }
        fflush(stdout);
    }
    mpz_clears(N, e, b, r, NULL);
    return 0;
}
