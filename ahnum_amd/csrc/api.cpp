// kite.h number API on the gfx950 engine (reference src/ctx.c:223-343, src/prp.c:26-34).
#include <cstdlib>
#include <cstring>
#include <new>

#include "kite_internal.h"

using namespace kite;

uint32_t kite_tuned_order(uint32_t L);

extern "C" kite_ctx *kite_ctx_new(uint64_t k, uint64_t n, int32_t c, const kite_opts *o) {
    if (c != -1 || k == 0 || (k & 1) == 0 || n < 8 || k >= (1ull << 30)) return nullptr;
    const kite_backend b = o ? o->backend : KITE_BACKEND_AUTO;
    if (b != KITE_BACKEND_AUTO && b != KITE_BACKEND_HIP) return nullptr;  // CPU backends not built
    kite_ctx *ctx = new (std::nothrow) kite_ctx();
    if (!ctx) return nullptr;
    ctx->k = k;
    ctx->n = n;
    ctx->c = c;
    ctx->L = o && o->fft_len ? o->fft_len : kite_choose_fft_len(k, n);
    if (!fft_len_ok(ctx->L) || ctx->L > n) {
        delete ctx;
        return nullptr;
    }
    ctx->plan_order = o && o->plan_order ? o->plan_order - 1 : kite_tuned_order(ctx->L);
    ctx->nlimbs = limbs_for_n(n);
    ctx->Nl.assign(ctx->nlimbs, 0);
    ctx->tmpdig.assign(ctx->L, 0);
    if (build_layout(ctx) || peak_product_bits(ctx) > kMaxProductBits) {
        delete ctx;
        return nullptr;
    }
    ctx->min_sub = carry_sub_needed(ctx);
    if (!ctx->min_sub) {
        delete ctx;
        return nullptr;
    }
    // N = k*2^n - 1
    ctx->Nl[n / 64] = k << (n % 64);
    if (n % 64 && n / 64 + 1 < ctx->nlimbs) ctx->Nl[n / 64 + 1] = k >> (64 - n % 64);
    for (size_t i = 0; i < ctx->nlimbs; i++)
        if (ctx->Nl[i]--) break;
    if (engine_init(ctx)) {
        kite_ctx_free(ctx);
        return nullptr;
    }
    return ctx;
}

extern "C" void kite_ctx_free(kite_ctx *ctx) {
    if (!ctx) return;
    engine_free(ctx);
    delete ctx;
}

extern "C" uint32_t kite_fft_len(const kite_ctx *ctx) { return ctx->L; }
extern "C" const char *kite_backend_name(const kite_ctx *ctx) {
    (void)ctx;
    return "hip-gfx950";
}

extern "C" kite_num *kite_num_new(kite_ctx *ctx) {
    kite_num *x = new (std::nothrow) kite_num();
    if (!x) return nullptr;
    x->dev = engine_num_alloc(ctx);
    if (!x->dev) {
        delete x;
        return nullptr;
    }
    return x;
}

extern "C" void kite_num_free(kite_ctx *ctx, kite_num *x) {
    if (!x) return;
    engine_num_free(ctx, x->dev);
    delete x;
}

extern "C" void kite_copy(kite_ctx *ctx, kite_num *dst, const kite_num *src) {
    if (dst != src) engine_copy(ctx, dst->dev, src->dev);
}

extern "C" void kite_square(kite_ctx *ctx, kite_num *x) { engine_square(ctx, x->dev, 1); }
extern "C" void kite_square_n(kite_ctx *ctx, kite_num *x, uint64_t count) { engine_square(ctx, x->dev, count); }
extern "C" int kite_square_n_static(kite_ctx *ctx, kite_num *x, uint64_t count) {
    return engine_square_static(ctx, x->dev, count);
}

extern "C" void kite_mul(kite_ctx *ctx, kite_num *x, const kite_num *y) { engine_mul(ctx, x->dev, y->dev); }
extern "C" void kite_add(kite_ctx *ctx, kite_num *x, const kite_num *y) { engine_addsub(ctx, x->dev, y->dev, 1); }
extern "C" void kite_sub(kite_ctx *ctx, kite_num *x, const kite_num *y) { engine_addsub(ctx, x->dev, y->dev, -1); }

// Small multiplier on the integer digits, then carry (src/neon.c:1108-1114).
extern "C" void kite_mul_small(kite_ctx *ctx, kite_num *x, int64_t m) {
    int64_t *d = ctx->tmpdig.data();
    engine_to_digits(ctx, x->dev, d);
    for (uint32_t j = 0; j < ctx->L; j++) d[j] *= m;
    kite_carry_digits(ctx, d);
    engine_from_digits(ctx, x->dev, d);
}

extern "C" double kite_max_roundoff(const kite_ctx *ctx) { return engine_maxerr(const_cast<kite_ctx *>(ctx)); }
extern "C" void kite_reset_roundoff(kite_ctx *ctx) { engine_reset_maxerr(ctx); }
extern "C" void kite_set_roundoff_check(kite_ctx *ctx, uint32_t every) { ctx->err_every = every ? every : 1; }
extern "C" size_t kite_limbs(const kite_ctx *ctx) { return ctx->nlimbs; }

extern "C" void kite_get_limbs(kite_ctx *ctx, const kite_num *x, uint64_t *limbs) {
    int64_t *d = ctx->tmpdig.data();
    engine_to_digits(ctx, x->dev, d);
    digits_to_limbs(ctx, d, limbs);
}

extern "C" void kite_set_limbs(kite_ctx *ctx, kite_num *x, const uint64_t *limbs) {
    int64_t *d = ctx->tmpdig.data();
    limbs_to_digits(ctx, limbs, d);
    engine_from_digits(ctx, x->dev, d);
}

extern "C" void kite_set_u64(kite_ctx *ctx, kite_num *x, uint64_t v) {
    std::vector<uint64_t> t(ctx->nlimbs, 0);
    t[0] = v;
    kite_set_limbs(ctx, x, t.data());
}

extern "C" uint64_t kite_res64(kite_ctx *ctx, const kite_num *x) {
    std::vector<uint64_t> t(ctx->nlimbs);
    kite_get_limbs(ctx, x, t.data());
    return t[0];
}

// Add 1 to digit L/3: a valid representation of a wrong value, which only the Gerbicz check can
// catch (src/prp.c:26-34).
extern "C" void kite_debug_corrupt(kite_ctx *ctx, kite_num *x) {
    std::vector<int64_t> d(ctx->L);
    engine_to_digits(ctx, x->dev, d.data());
    d[ctx->L / 3] += 1;
    engine_from_digits(ctx, x->dev, d.data());
}
