// gfx950 squaring engine: one squaring mod N = k*2^n - 1 as three kernels over the four-step
// layout of the packed complex transform (H = L/2 = R x C complex doubles, row-major).
//
//   k_rows      per row pair (k_r, R-k_r): row FFT (length C) -> real-split square or multiply
//               -> inverse row FFT                                   (replaces the reference's
//               pass 2: src/neon.c:987-1054, src/fast2p.c:545-617)
//   k_cols_inv  per column group: conj four-step twiddle -> inverse column FFT (length R) ->
//               scale k/(4H w_j), round, int64 carry along each row chunk of 2G digits, the
//               128-bit top fold on the last chunk; writes int32 digits and one carry per chunk
//   k_cols_fwd  per column group: absorb the previous chunk's carry (the last digit of a chunk
//               keeps any remainder), weight, forward column FFT, four-step twiddle
//               (together the reference's pass 1: src/neon.c:811-877, src/fast2p.c:453-491)
//
// The reference carries sequentially across column groups with a deferred first group; here
// every chunk's carry-out is independent of its carry-in (one-step absorption), so all column
// groups of all tests run in parallel and the carries cross work-groups through a kernel
// boundary. FFTs are Stockham autosort in LDS (natural order in and out) with ping-pong buffers,
// radix 2/3/4/5/7/8 butterflies in registers, FP64 throughout; every twiddle and weight comes
// from host tables computed with exact argument reduction.
//
// Work is described by per-work-group job records (slot, unit) over a table of slots, so the
// same kernels run one test (kite.h API) or many tests of different lengths in one launch.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "engine_dev.h"
#include "kernels.h"
#include "kite_batch.h"
#include "kite_internal.h"


namespace kite {

typedef unsigned __int128 u128;

static thread_local std::string g_err;
const char *engine_last_error() { return g_err.c_str(); }

#define HIP_OK(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            g_err = std::string(#expr) + ": " + hipGetErrorString(e_);                    \
            fprintf(stderr, "kite: %s\n", g_err.c_str());                                 \
            return -1;                                                                    \
        }                                                                                 \
    } while (0)

// ================================================================================================
// Utility kernels (the pass kernels live in fft_dev.h, kernels_generic.hip, kernels_fast.hip)
// ================================================================================================

extern "C" __global__ void k_addsub(double2 *x, const double2 *y, int n, double sign) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        double2 a = x[i], b = y[i];
        x[i] = make_double2(a.x + sign * b.x, a.y + sign * b.y);
    }
}

extern "C" __global__ void k_set_u64(unsigned long long *p, unsigned long long v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Digit equality of two residues of each slot (kite_batch_equal): the first residue's digits were
// written to the slot's operand-spectrum buffer, the second's to its digit buffer. One work-group
// per slot; eq[s] = 1 if every digit agrees, 0 if any differs, -1 for a slot not compared.
__global__ void k_digits_equal(const kite::SlotDev *__restrict__ slots, int nslots, int *__restrict__ eq) {
    const int s = blockIdx.x;
    if (s >= nslots) return;
    const kite::SlotDev &d = slots[s];
    __shared__ int differ;
    if (threadIdx.x == 0) differ = 0;
    __syncthreads();
    if (d.count > 0) {
        const int32_t *a = reinterpret_cast<const int32_t *>(d.yfreq), *b = d.dig;
        const int L = 2 * d.R * d.C;
        int mine = 0;
        for (int i = threadIdx.x; i < L; i += blockDim.x) mine |= a[i] != b[i];
        if (mine) differ = 1;  // every writer stores the same value
    }
    __syncthreads();
    if (threadIdx.x == 0) eq[s] = d.count > 0 ? !differ : -1;
}

// ================================================================================================
// Host side
// ================================================================================================

// cos/sin(2 pi j / n): reduce j/n to [0, 1/8] with exact integer steps, evaluate in long double
static void root_of_unity(uint64_t j, uint64_t n, double *c, double *s) {
    j %= n;
    bool ns = false, nc = false, sw = false;
    uint64_t num = j, den = n;
    if (2 * num > den) { num = den - num; ns = true; }          // 2pi - x
    if (4 * num > den) { num = den - 2 * num; den *= 2; nc = true; }  // pi - x
    if (8 * num > den) { num = den - 4 * num; den *= 4; sw = true; }  // pi/2 - x
    const long double x = 6.283185307179586476925286766559005768L * (long double)num / (long double)den;
    long double cc = cosl(x), ss = sinl(x);
    if (sw) std::swap(cc, ss);
    if (nc) cc = -cc;
    if (ns) ss = -ss;
    *c = (double)cc;
    *s = (double)ss;
}

// Testing hook: the root of unity every device twiddle table is built from.
extern "C" void kite_debug_root_of_unity(uint64_t j, uint64_t n, double *c, double *s) { root_of_unity(j, n, c, s); }

// forward roots w_n^t = e^{-2 pi i t/n}
static std::vector<double2> roots_table(uint64_t n, uint64_t count, uint64_t step) {
    std::vector<double2> t(count);
    for (uint64_t i = 0; i < count; i++) {
        double c, s;
        root_of_unity(i * step % n, n, &c, &s);
        t[i] = make_double2(c, -s);
    }
    return t;
}

static bool make_plan(uint32_t n, Plan *pl) {
    pl->n = (int)n;
    pl->nst = 0;
    uint32_t m = n;
    while (m % 8 == 0 && pl->nst < kMaxStages) { pl->rad[pl->nst++] = 8; m /= 8; }
    for (uint32_t p : {4u, 2u, 3u, 5u, 7u})
        while (m % p == 0 && pl->nst < kMaxStages) { pl->rad[pl->nst++] = (int)p; m /= p; }
    return m == 1;
}

static bool smooth7(uint32_t n) {
    for (uint32_t p : {2u, 3u, 5u, 7u})
        while (n % p == 0) n /= p;
    return n == 1;
}

// Choose H = R x C for length L (top_digits = L - top_t, 0 if unknown). The column group width
// is g_for_R(R); the top region must fit the last 16-digit carry sub-chunk. Preference: C about
// 4R (balanced passes), and sizes with compile-time kernels (kernels_fast.hip).
struct ShapeChoice {
    uint32_t R = 0, C = 0, G = 0;
    bool fast = false;
};

static bool choose_shape(uint32_t L, uint32_t top_digits, uint32_t variant, uint32_t min_sub, ShapeChoice *out) {
    if (L % 2 || L < 64) return false;
    const uint32_t H = L / 2;
    std::vector<ShapeChoice> cands;
    std::vector<double> costs;
    for (uint32_t C = 16; C <= kMaxC && C <= H; C++) {
        if (H % C || C % 4 || !smooth7(C)) continue;
        const uint32_t R = H / C;
        if (R > kMaxR || !smooth7(R)) continue;
        uint32_t G = (uint32_t)g_for_R((int)R);
        const bool gfull = C % G == 0;  // the compile-time kernels use exactly g_for_R(R)
        while (C % G) G /= 2;           // the generic kernels take any power of two dividing C
        const uint32_t sub = G < (uint32_t)kSub ? G : (uint32_t)kSub;
        if ((top_digits && 2 * sub < top_digits) || sub < min_sub) continue;
        if ((2ull * R * G + R) * 16 + (uint64_t)R * (G / sub) * 8 > (uint64_t)kMaxLds) continue;
        const bool fast = gfull && fast_rows((int)C) && fast_cols_inv((int)R);
        double cost = fabs(log2((double)C) - log2((double)R) - 2.0) + (fast ? 0 : 4);
        cands.push_back({R, C, G, fast});
        costs.push_back(cost);
    }
    if (cands.empty()) return false;
    // rank by cost; `variant` picks the (variant)-th best (the autotuner's shape order), 0 = best
    std::vector<size_t> idx(cands.size());
    for (size_t i = 0; i < idx.size(); i++) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return costs[a] < costs[b]; });
    size_t pick = variant < idx.size() ? variant : 0;
    *out = cands[idx[pick]];
    return true;
}

bool engine_shape_ok(uint32_t L, uint32_t top_digits, uint32_t min_sub) {
    ShapeChoice sc;
    return choose_shape(L, top_digits, 0, min_sub, &sc);
}

bool engine_shape_fast(uint32_t L) {
    ShapeChoice sc;
    return choose_shape(L, 0, 0, 1, &sc) && sc.fast;
}

// Shapes (twiddle tables) are shared by every context of the same (R, C, G) on a device.
struct ShapeEntry {
    ShapeChoice ch;
    int device = 0;
    ShapeDev host{};
    ShapeDev *dev = nullptr;
    std::vector<void *> bufs;
};

static std::mutex g_shape_mu;
static std::vector<ShapeEntry *> g_shapes;

template <class T> static int upload(const std::vector<T> &v, T **out, std::vector<void *> &bufs) {
    HIP_OK(hipMalloc((void **)out, v.size() * sizeof(T) + 16));
    HIP_OK(hipMemcpy(*out, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    bufs.push_back(*out);
    return 0;
}

static ShapeEntry *get_shape(const ShapeChoice &ch) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(g_shape_mu);
    for (ShapeEntry *e : g_shapes)
        if (e->device == dev && e->ch.R == ch.R && e->ch.C == ch.C && e->ch.G == ch.G) return e;
    auto *e = new ShapeEntry();
    e->ch = ch;
    e->device = dev;
    ShapeDev &h = e->host;
    h.R = (int)ch.R;
    h.C = (int)ch.C;
    h.G = (int)ch.G;
    h.NG = (int)(ch.C / ch.G);
    const uint64_t H = (uint64_t)ch.R * ch.C;
    if (!make_plan(ch.R, &h.pr) || !make_plan(ch.C, &h.pc)) {
        delete e;
        return nullptr;
    }
    int sh = 0;
    while ((1ull << (2 * sh)) < H) sh++;
    h.sshift = sh;
    const uint64_t S = 1ull << sh, nhi = (H + S - 1) / S;
    // exact tables in the order the compile-time kernels read them (coalesced, shared by every
    // test of this shape)
    const std::vector<double2> wH = roots_table(H, H, 1);
    std::vector<double2> trows(H), tcols(H);
    for (uint64_t kr = 0; kr < ch.R; kr++)
        for (uint64_t kc = 0; kc < ch.C; kc++) trows[kr * ch.C + kc] = wH[(kr + (uint64_t)ch.R * kc) % H];
    for (uint64_t g = 0; g < ch.C / ch.G; g++)
        for (uint64_t kr = 0; kr < ch.R; kr++)
            for (uint64_t l = 0; l < ch.G; l++)
                tcols[(g * ch.R + kr) * ch.G + l] = wH[((g * ch.G + l) * kr) % H];
    // 16 x 16 row FFTs: [j][t] = w_C^{t j}
    std::vector<double2> t16(256, make_double2(1, 0));
    if (ch.C == 256) {
        const std::vector<double2> wC = roots_table(ch.C, ch.C, 1);
        for (uint64_t j = 0; j < 16; j++)
            for (uint64_t t = 0; t < 16; t++) t16[j * 16 + t] = wC[(t * j) % ch.C];
    }
    double2 *twR, *twC, *tlo, *thi, *twRows, *twCols, *tw16;
    if (upload(roots_table(ch.R, ch.R, 1), &twR, e->bufs) || upload(roots_table(ch.C, ch.C, 1), &twC, e->bufs) ||
        upload(roots_table(H, S, 1), &tlo, e->bufs) || upload(roots_table(H, nhi, S), &thi, e->bufs) ||
        upload(trows, &twRows, e->bufs) || upload(tcols, &twCols, e->bufs) || upload(t16, &tw16, e->bufs)) {
        delete e;
        return nullptr;
    }
    h.twR = twR;
    h.twC = twC;
    h.tlo = tlo;
    h.thi = thi;
    h.twRows = twRows;
    h.twCols = twCols;
    h.tw16 = tw16;
    std::vector<ShapeDev> one(1, h);
    if (upload(one, &e->dev, e->bufs)) {
        delete e;
        return nullptr;
    }
    g_shapes.push_back(e);
    return e;
}

// Testing hook (kite_debug_fail_hip): the n-th HIP call checked after it is set reports a
// failure, so that error paths can be exercised without a device fault.
static std::atomic<long> g_fail_countdown{0};

// A HIP call's result, with the testing hook applied.
static hipError_t hip_result(hipError_t e) {
    if (e == hipSuccess && g_fail_countdown.load(std::memory_order_relaxed) > 0 &&
        g_fail_countdown.fetch_sub(1) == 1)
        return hipErrorUnknown;
    return e;
}

struct Engine {
    hipStream_t stream = nullptr;
    ShapeEntry *shape = nullptr;
    SlotDev slot{};
    double *ta = nullptr, *tb = nullptr;
    uint8_t *dcode = nullptr;
    int32_t *dig = nullptr;
    int64_t *cout = nullptr;
    double2 *yfreq = nullptr;
    unsigned long long *maxerr = nullptr;
    std::vector<int32_t> hdig;
    int nrowjobs = 0, ncoljobs = 0;
    size_t lds_rows = 0, lds_cols = 0, lds_fused = 0;
    PassKernel krows = nullptr, kcinv = nullptr, kcfwd = nullptr;
    PassKernel krows_sq = nullptr;  // square-mode rows with registers, if any
    PassKernel krows1 = nullptr;    // one row pair per work-group, for this context's own squarings
    int nrowjobs1 = 0;
    size_t lds_rows1 = 0;
    PassKernel kcfused = nullptr;   // square-mode fused column pass, if any
    int nt_fused = kNT;             // its threads per work-group
    int nt_rows = kNT;              // threads per work-group of krows and krows1
    unsigned long long *flags = nullptr;  // its per-chunk carry flags
    uint32_t seq = 0;                     // its launches so far
    SlotDev *d_slot = nullptr;            // the slot table of this context's own launches (one entry)
    hipGraphExec_t sq_graph = nullptr;    // kSqGraphSteps squarings of the number in d_slot
    bool sq_graph_failed = false;
    std::string name[3];  // kernel names of the three passes, for timing reports
    uint32_t H = 0;
    // first HIP failure of this context (sticky): every later operation is skipped, the round-off
    // reads +inf and the drivers return a negative code
    hipError_t err = hipSuccess;
    std::string err_what;
};

// Records the first failure of E; true if `e` (and every earlier call) succeeded.
static bool chk(Engine *E, hipError_t e, const char *what) {
    e = hip_result(e);
    if (e != hipSuccess && E->err == hipSuccess) {
        E->err = e;
        E->err_what = std::string(what) + ": " + hipGetErrorString(e);
        g_err = E->err_what;
        fprintf(stderr, "kite: %s\n", g_err.c_str());
    }
    return E->err == hipSuccess;
}

static std::once_flag g_attr_once;

static void set_lds_limits() {
    std::call_once(g_attr_once, [] {
        PassKernel ks[512];
        int n = 0;
        all_kernels(ks, &n, 512);
        for (int i = 0; i < n; i++)
            if (ks[i] && hipFuncSetAttribute(ks[i], hipFuncAttributeMaxDynamicSharedMemorySize, kMaxLds) != hipSuccess)
                (void)hipGetLastError();  // a refused limit must not surface at a later launch
    });
}

// Per-test tables: weights factor over digit j = 2rC + b (b = 2c + h < 2C) as
// w_j = 2^{ceil(fa+fb)} 2^{-fa} 2^{-fb}, fa = frac(2rC T'/L), fb = frac(b T'/L), with the integer
// ceil(fa + fb) in {0,1,2} decided exactly from the 128-bit numerators.
int build_slot_tables(kite_ctx *ctx, const ShapeDev &sh, std::vector<double> &ta, std::vector<double> &tb,
                      std::vector<uint8_t> &dcode) {
    const uint32_t R = sh.R, C = sh.C, G = sh.G, NG = sh.NG;
    const u128 D = (u128)ctx->L << 64;
    std::vector<u128> na(R), nb(2 * C);
    ta.assign(2 * R, 0);
    tb.assign(4 * C, 0);
    for (uint32_t r = 0; r < R; r++) {
        na[r] = phase_num(ctx, 2ull * r * C);
        ta[r] = exp2_neg_frac(ctx, na[r]);
        long double f = (long double)na[r] / ((long double)ctx->L * 18446744073709551616.0L);
        ta[R + r] = (double)exp2l(f);
    }
    for (uint32_t b = 0; b < 2 * C; b++) {
        nb[b] = phase_num(ctx, b);
        tb[b] = exp2_neg_frac(ctx, nb[b]);
        long double f = (long double)nb[b] / ((long double)ctx->L * 18446744073709551616.0L);
        tb[2 * C + b] = (double)exp2l(f);
    }
    dcode.assign((size_t)R * C, 0);
    for (uint32_t g = 0; g < NG; g++)
        for (uint32_t r = 0; r < R; r++)
            for (uint32_t l = 0; l < G; l++) {
                uint8_t code = 0;
                for (int h = 0; h < 2; h++) {
                    const uint32_t b = 2 * (g * G + l) + h;
                    const uint64_t j = 2ull * r * C + b;
                    const u128 sum = na[r] + nb[b];
                    const int ce = sum == 0 ? 0 : (sum <= D ? 1 : 2);
                    const int big = (int)ctx->bits[j] - (int)ctx->b0;
                    if (big < 0 || big > 1) return -1;
                    code |= (uint8_t)((ce | (big << 2)) << (3 * h));
                }
                dcode[((size_t)g * R + r) * G + l] = code;
            }
    return 0;
}

int fill_slot_fold(kite_ctx *ctx, const ShapeDev &sh, SlotDev *s) {
    const uint32_t sub = sh.G < kSub ? (uint32_t)sh.G : (uint32_t)kSub;
    const uint32_t L = ctx->L, chunk0 = L - 2 * sub;  // the fold works on the last sub-chunk
    s->k = ctx->k;
    s->kinv = ~0ull / ctx->k;
    s->b0 = (int)ctx->b0;
    s->top_n = (int)(L - ctx->top_t);
    if (s->top_n > (int)(2 * sub) || s->top_n > kMaxTop) return -1;
    s->top_i0 = (int)(ctx->top_t - chunk0);
    if (ctx->top_t == L) s->top_i0 = (int)(2 * sub);
    s->top_m = (int)ctx->top_m;
    s->top_shift = (int)(ctx->sL - ctx->top_st);
    s->top_big = 0;
    for (int i = 0; i < s->top_n; i++) {
        const int big = (int)ctx->bits[ctx->top_t + i] - (int)ctx->b0;
        if (big < 0 || big > 1) return -1;
        s->top_big |= (uint64_t)big << i;
    }
    s->kscale = (double)ctx->k / (2.0 * L);  // k / (4H)
    s->rscale = 1.0 / sh.R;
    return 0;
}

int engine_init(kite_ctx *ctx) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        g_err = "no HIP device";
        return -1;
    }
    set_lds_limits();
    ShapeChoice ch;
    if (!choose_shape(ctx->L, ctx->L - ctx->top_t, ctx->plan_order, ctx->min_sub, &ch)) {
        g_err = "no shape for this length";
        return -1;
    }
    Engine *E = new Engine();
    ctx->eng = E;
    E->shape = get_shape(ch);
    if (!E->shape) return -1;
    const ShapeDev &sh = E->shape->host;
    E->H = ctx->L / 2;
    std::vector<double> ta, tb;
    std::vector<uint8_t> dcode;
    if (build_slot_tables(ctx, sh, ta, tb, dcode)) {
        g_err = "digit widths outside {b0, b0+1}";
        return -1;
    }
    HIP_OK(hipStreamCreateWithFlags(&E->stream, hipStreamNonBlocking));
    HIP_OK(hipMalloc(&E->ta, ta.size() * 8));
    HIP_OK(hipMemcpy(E->ta, ta.data(), ta.size() * 8, hipMemcpyHostToDevice));
    HIP_OK(hipMalloc(&E->tb, tb.size() * 8));
    HIP_OK(hipMemcpy(E->tb, tb.data(), tb.size() * 8, hipMemcpyHostToDevice));
    HIP_OK(hipMalloc(&E->dcode, dcode.size()));
    HIP_OK(hipMemcpy(E->dcode, dcode.data(), dcode.size(), hipMemcpyHostToDevice));
    HIP_OK(hipMalloc(&E->dig, (size_t)ctx->L * 4));
    HIP_OK(hipMalloc(&E->cout, (size_t)sh.NG * sh.R * 8));
    HIP_OK(hipMalloc(&E->flags, (size_t)sh.NG * sh.R * 8));
    HIP_OK(hipMemset(E->flags, 0, (size_t)sh.NG * sh.R * 8));
    HIP_OK(hipMalloc(&E->yfreq, (size_t)E->H * sizeof(double2)));
    HIP_OK(hipMalloc(&E->maxerr, 16));
    HIP_OK(hipMemset(E->maxerr, 0, 16));
    HIP_OK(hipMalloc(&E->d_slot, sizeof(SlotDev)));
    SlotDev &s = E->slot;
    s.R = sh.R;
    s.C = sh.C;
    s.G = sh.G;
    s.NG = sh.NG;
    s.twR = sh.twR;
    s.twC = sh.twC;
    s.twRows = sh.twRows;
    s.twCols = sh.twCols;
    s.tw16 = sh.tw16;
    s.tlo = sh.tlo;
    s.thi = sh.thi;
    s.sshift = sh.sshift;
    s.shp = E->shape->dev;
    s.ta = E->ta;
    s.tb = E->tb;
    s.dcode = E->dcode;
    s.dig = E->dig;
    s.cout = E->cout;
    s.flags = E->flags;
    s.seq = 0;
    s.yfreq = E->yfreq;
    s.maxerr = E->maxerr;
    s.count = 1;
    if (fill_slot_fold(ctx, sh, &s)) {
        g_err = "top region does not fit the last chunk";
        return -1;
    }
    {
        const int npairs = sh.R % 2 == 0 ? sh.R / 2 : (sh.R + 1) / 2;
        const bool gct = sh.G == g_for_R(sh.R);  // compile-time column kernels assume this width
        PassKernel fr = fast_rows(sh.C), fi = gct ? fast_cols_inv(sh.R) : nullptr,
                   ff = gct ? fast_cols_fwd(sh.R) : nullptr;
        const int ppw = fr ? ppw_for_C(sh.C) : 1;
        E->krows = fr ? fr : generic_rows();
        E->krows_sq = (fr && ppw_for_C(sh.C) == 4) ? fast_rows_sq(sh.C) : nullptr;
        E->kcinv = fi ? fi : generic_cols_inv();
        E->kcfwd = ff ? ff : generic_cols_fwd();
        // the fused pass passes carries between work-groups in 48-bit flag fields, which hold every
        // carry when digits have at least kFusedMinBits bits; narrower digits take the split passes,
        // whose carries are int64
        const bool fused_ok = ff && ctx->b0 >= (uint32_t)kFusedMinBits;
        E->kcfused = fused_ok ? fast_cols_fused(sh.R) : nullptr;
        E->lds_fused = fused_ok ? fast_cols_fused_lds(sh.R) : 0;
        E->nt_fused = fused_ok ? fast_cols_fused_threads(sh.R) : kNT;
        E->nt_rows = fr ? rows_nt(sh.C) : kNT;
        E->name[0] = E->krows_sq ? "k_rows_sq" + std::to_string(sh.C)
                     : fr       ? "k_rows_c<" + std::to_string(sh.C) + ">"
                                : std::string("k_rows_g");
        E->name[1] = fi ? "k_cols_inv_r<" + std::to_string(sh.R) + ">" : std::string("k_cols_inv_g");
        E->name[2] = ff ? "k_cols_fwd_r<" + std::to_string(sh.R) + ">" : std::string("k_cols_fwd_g");
        E->nrowjobs = (npairs + ppw - 1) / ppw;
        E->ncoljobs = sh.NG;
        if (fr && ppw > 1 && (E->krows1 = fast_rows1(sh.C))) {
            E->nrowjobs1 = npairs;
            E->lds_rows1 = rows_lds_bytes(sh.C, 1);
        }
        const int sub = sh.G < kSub ? sh.G : kSub;
        if (fr)  // single buffer of 2*ppw rows, then the row twiddles
            E->lds_rows = rows_lds_bytes(sh.C, ppw);
        else
            E->lds_rows = (size_t)16 * (2 * 2 * sh.C);
        if (fi)
            E->lds_cols = fast_cols_lds(sh.R);
        else
            E->lds_cols = (size_t)16 * (2 * sh.R * sh.G + sh.R) + (size_t)8 * sh.R * (sh.G / sub);
    }
    E->hdig.resize(ctx->L);
    return 0;
}

void engine_free(kite_ctx *ctx) {
    Engine *E = ctx->eng;
    if (!E) return;
    if (E->stream) hipStreamSynchronize(E->stream);
    if (E->sq_graph) hipGraphExecDestroy(E->sq_graph);
    void *ps[] = {E->ta, E->tb, E->dcode, E->dig, E->cout, E->flags, E->yfreq, E->maxerr, E->d_slot};
    for (void *p : ps)
        if (p) hipFree(p);
    if (E->stream) hipStreamDestroy(E->stream);
    delete E;
    ctx->eng = nullptr;
}

int engine_error(const kite_ctx *ctx) { return ctx->eng && ctx->eng->err != hipSuccess ? -1 : 0; }

void *engine_num_alloc(kite_ctx *ctx) {
    Engine *E = ctx->eng;
    void *p = nullptr;
    if (E->err != hipSuccess) return nullptr;
    if (!chk(E, hipMalloc(&p, (size_t)E->H * sizeof(double2)), "hipMalloc(number)")) return nullptr;
    if (!chk(E, hipMemsetAsync(p, 0, (size_t)E->H * sizeof(double2), E->stream), "hipMemsetAsync(number)")) {
        hipFree(p);
        return nullptr;
    }
    return p;
}

void engine_num_free(kite_ctx *ctx, void *p) {
    if (!p) return;
    hipStreamSynchronize(ctx->eng->stream);
    hipFree(p);
}

void engine_copy(kite_ctx *ctx, void *dst, const void *src) {
    Engine *E = ctx->eng;
    if (E->err != hipSuccess) return;
    chk(E, hipMemcpyAsync(dst, src, (size_t)E->H * sizeof(double2), hipMemcpyDeviceToDevice, E->stream),
        "hipMemcpyAsync(copy)");
}

// With a start/stop event pair the launch goes through hipExtLaunchKernel, which stamps the events
// from the dispatch packet itself (the kernel's own start and end, as a profiler's trace reads them),
// not from marker packets of their own around it (those add the dispatch and completion latencies,
// 5-7 us per launch).
static hipError_t launch_k(const void *fn, int grid, int threads, void **args, size_t lds, hipStream_t st,
                           hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
    if (e0) return hipExtLaunchKernel(fn, dim3(grid), dim3(threads), args, lds, st, e0, e1, 0);
    return hipLaunchKernel(fn, dim3(grid), dim3(threads), args, lds, st);
}

static hipError_t launch_pass(PassKernel fn, int grid, size_t lds, hipStream_t st, const SlotDev *slots,
                              const int2 *jobs, int mode, int step, int threads = kNT, hipEvent_t e0 = nullptr,
                              hipEvent_t e1 = nullptr) {
    void *args[] = {(void *)&slots, (void *)&jobs, (void *)&mode, (void *)&step};
    return launch_k(fn, grid, threads, args, lds, st, e0, e1);
}

__global__ void k_set_slot(SlotDev *dst, SlotDev v) {
    static_assert(sizeof(SlotDev) % 8 == 0 && sizeof(SlotDev) / 8 <= 64, "slot copied as 64-bit words");
    const uint32_t i = threadIdx.x;
    if (i < sizeof(SlotDev) / 8) reinterpret_cast<uint64_t *>(dst)[i] = reinterpret_cast<const uint64_t *>(&v)[i];
}

// This context's launches read their slot from E->d_slot, written here (in stream order) before
// each operation: the kernels then load the slot's fields with scalar loads from one table.
static bool set_slot(Engine *E, const SlotDev &s) {
    if (E->err != hipSuccess) return false;
    hipLaunchKernelGGL(k_set_slot, dim3(1), dim3(64), 0, E->stream, E->d_slot, s);
    return chk(E, hipGetLastError(), "k_set_slot");
}

// The passes of this context read their slot from `tbl` (E->d_slot, or the static scratch of
// kite_square_n_static).
static bool launch_rows(Engine *E, int mode, int step = 0, const SlotDev *tbl = nullptr) {
    if (!tbl) tbl = E->d_slot;
    if (mode == kModeSquare && E->krows_sq)
        return chk(E, launch_pass(E->krows_sq, E->nrowjobs, 0, E->stream, tbl, nullptr, mode, step, kRowsSqThreads),
                   "rows pass");
    return chk(E, launch_pass(E->krows, E->nrowjobs, E->lds_rows, E->stream, tbl, nullptr, mode, step, E->nt_rows),
               "rows pass");
}

static bool launch_cols(Engine *E, int mode, int step = 0, const SlotDev *tbl = nullptr) {
    if (!tbl) tbl = E->d_slot;
    return chk(E, launch_pass(E->kcinv, E->ncoljobs, E->lds_cols, E->stream, tbl, nullptr, mode, step),
               "column pass (inverse)") &&
           chk(E, launch_pass(E->kcfwd, E->ncoljobs, E->lds_cols, E->stream, tbl, nullptr, mode, step),
               "column pass (forward)");
}

// One squaring (step `step` of the slot in E->d_slot) of this context's number alone: rows (one
// pair per work-group where that kernel exists, since a single test leaves most CUs idle and
// latency decides), then the fused column pass (step s uses seq + s + 1) or the two column passes.
static bool launch_square(Engine *E, int step, const SlotDev *tbl = nullptr) {
    if (!tbl) tbl = E->d_slot;
    bool ok;
    if (E->krows1)
        ok = chk(E, launch_pass(E->krows1, E->nrowjobs1, E->lds_rows1, E->stream, tbl, nullptr, kModeSquare, step,
                                E->nt_rows),
                 "rows pass");
    else
        ok = launch_rows(E, kModeSquare, step, tbl);
    if (!ok) return false;
    if (E->kcfused)
        return chk(E, launch_pass(E->kcfused, E->ncoljobs, E->lds_fused, E->stream, tbl, nullptr, kModeSquare, step,
                                  E->nt_fused),
                   "fused column pass");
    return launch_cols(E, kModeSquare, step, tbl);
}

// Squarings of one context in runs of kSqGraphSteps are replayed from a captured HIP graph whose
// kernels read the number, the launch counter and the step count from the slot in E->d_slot,
// written before each replay. Replays remove the host's per-launch cost, which bounds a single
// test's squaring otherwise.
constexpr uint32_t kSqGraphSteps = 32;

static bool build_square_graph(Engine *E) {
    if (E->sq_graph) return true;
    if (E->sq_graph_failed || E->err != hipSuccess) return false;
    E->sq_graph_failed = true;  // until the capture below succeeds
    hipGraph_t g = nullptr;
    bool ok = hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
        bool launched = true;
        for (uint32_t st = 0; st < kSqGraphSteps && launched; st++) {
            const hipError_t saved = E->err;
            launched = launch_square(E, (int)st);
            E->err = saved;  // a launch refused during capture is not an engine failure
        }
        ok = hipStreamEndCapture(E->stream, &g) == hipSuccess && g && launched &&
             hipGraphInstantiate(&E->sq_graph, g, nullptr, nullptr, 0) == hipSuccess;
        if (g) hipGraphDestroy(g);
    }
    // a capture that failed (this clears only the capture's own error: the engine has none
    // latched, checked above) falls back to direct launches
    (void)hipGetLastError();
    if (!ok) E->sq_graph = nullptr;
    E->sq_graph_failed = !ok;
    return ok;
}

void engine_square(kite_ctx *ctx, void *x, uint64_t count, int32_t xadd) {
    Engine *E = ctx->eng;
    if (E->err != hipSuccess) return;
    SlotDev s = E->slot;
    s.x = (double2 *)x;
    s.xadd = xadd;
    if (count >= kSqGraphSteps && build_square_graph(E)) {
        s.count = kSqGraphSteps;
        for (; count >= kSqGraphSteps; count -= kSqGraphSteps) {
            s.seq = E->seq;
            if (!set_slot(E, s)) return;
            if (!chk(E, hipGraphLaunch(E->sq_graph, E->stream), "hipGraphLaunch(square)")) return;
            E->seq += kSqGraphSteps;
        }
    }
    // the rest by direct launches, in runs of up to 2^20 steps per slot write
    while (count > 0) {
        const uint32_t run = (uint32_t)std::min<uint64_t>(count, 1u << 20);
        s.seq = E->seq;
        s.count = run;
        if (!set_slot(E, s)) return;
        for (uint32_t i = 0; i < run; i++)
            if (!launch_square(E, (int)i)) return;
        E->seq += run;
        count -= run;
    }
}

// ---- squarings without allocation (kite_square_n_static) -------------------------------------
//
// Callers that may not allocate device memory (code under stream capture, tight serving loops,
// harnesses that forbid it) square through this entry point: no graph, no allocation, the device
// scratch of the squaring (the slot record its kernels read, and the carry flags of the fused
// column pass) in this file-scope array. Every call on a GPU shares that GPU's copy, and each call
// clears the flag words it uses, so calls that overlap on one GPU race; calls in sequence do not.
constexpr size_t kScratchSlotBytes = 256;
constexpr size_t kScratchBytes = (size_t)2 << 20;  // 2 MiB: the slot record and 262112 flag words
constexpr size_t kScratchFlags = (kScratchBytes - kScratchSlotBytes) / 8;
static_assert(sizeof(SlotDev) <= kScratchSlotBytes, "slot record fits its part of the scratch");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];  // the device address of `scratch`, per device

// The address of this device's `scratch`, resolved once per device (the device must be current).
static unsigned char *scratch_of_device(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            g_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            fprintf(stderr, "kite: %s\n", g_err.c_str());
            return nullptr;
        }
        g_scratch_ptr[dev] = p;
    }
    return static_cast<unsigned char *>(g_scratch_ptr[dev]);
}

int engine_square_static(kite_ctx *ctx, void *x, uint64_t count) {
    Engine *E = ctx->eng;
    if (E->err != hipSuccess) return -1;
    int dev = -1;
    if (!chk(E, hipGetDevice(&dev), "hipGetDevice")) return -1;
    if (dev != E->shape->device) {
        g_err = "kite_square_n_static: the context's device is not the current device";
        fprintf(stderr, "kite: %s\n", g_err.c_str());
        return -1;
    }
    unsigned char *ptr = scratch_of_device(dev);
    if (!ptr) return -1;
    SlotDev *tbl = reinterpret_cast<SlotDev *>(ptr);
    unsigned long long *flags = reinterpret_cast<unsigned long long *>(ptr + kScratchSlotBytes);
    const size_t nflags = (size_t)E->shape->host.NG * E->shape->host.R;
    if (E->kcfused && nflags > kScratchFlags) {
        g_err = "kite_square_n_static: this FFT length needs " + std::to_string(nflags) + " carry flags, the " +
                "static scratch holds " + std::to_string(kScratchFlags);
        fprintf(stderr, "kite: %s\n", g_err.c_str());
        return -2;
    }
    if (E->kcfused) {  // the words this call waits on start below its first launch tag (1)
        hipLaunchKernelGGL(k_set_u64, dim3((unsigned)((nflags + 255) / 256)), dim3(256), 0, E->stream, flags, 0ull,
                           (int)nflags);
        if (!chk(E, hipGetLastError(), "k_set_u64(scratch flags)")) return -1;
    }
    SlotDev s = E->slot;
    s.x = (double2 *)x;
    s.xadd = 0;
    s.flags = flags;
    for (uint64_t done = 0; done < count;) {
        const uint32_t run = (uint32_t)std::min<uint64_t>(count - done, 1u << 20);
        s.seq = (uint32_t)done;  // step i of this run waits for tag done + i + 1, as it was written
        s.count = run;
        hipLaunchKernelGGL(k_set_slot, dim3(1), dim3(64), 0, E->stream, tbl, s);
        if (!chk(E, hipGetLastError(), "k_set_slot(scratch)")) return -1;
        for (uint32_t i = 0; i < run; i++)
            if (!launch_square(E, (int)i, tbl)) return -1;
        done += run;
    }
    return chk(E, hipStreamSynchronize(E->stream), "hipStreamSynchronize(square_n_static)") ? 0 : -1;
}

void engine_mul(kite_ctx *ctx, void *x, const void *y) {
    Engine *E = ctx->eng;
    SlotDev s = E->slot;
    s.x = (double2 *)x;
    s.y_src = (const double2 *)y;
    if (!set_slot(E, s)) return;
    if (launch_rows(E, kModeRowFft) && launch_rows(E, kModeMul)) launch_cols(E, kModeMul);
}

void engine_addsub(kite_ctx *ctx, void *x, const void *y, int sign) {
    Engine *E = ctx->eng;
    if (E->err != hipSuccess) return;
    const int n = (int)E->H;
    hipLaunchKernelGGL(k_addsub, dim3((n + 255) / 256), dim3(256), 0, E->stream, (double2 *)x, (const double2 *)y, n,
                       (double)sign);
    chk(E, hipGetLastError(), "k_addsub");
}

// digit j = 2 (r C + g G) + i  <->  blocked index (g R + r) 2G + i
void engine_from_digits(kite_ctx *ctx, void *x, const int64_t *dig) {
    Engine *E = ctx->eng;
    if (E->err != hipSuccess) return;
    const ShapeDev &sh = E->shape->host;
    const uint32_t R = sh.R, C = sh.C, G = sh.G, NG = sh.NG;
    for (uint32_t g = 0; g < NG; g++)
        for (uint32_t r = 0; r < R; r++)
            for (uint32_t i = 0; i < 2 * G; i++)
                E->hdig[((size_t)g * R + r) * 2 * G + i] = (int32_t)dig[2 * ((size_t)r * C + g * G) + i];
    SlotDev s = E->slot;
    s.x = (double2 *)x;
    if (chk(E, hipMemcpyAsync(E->dig, E->hdig.data(), (size_t)ctx->L * 4, hipMemcpyHostToDevice, E->stream),
            "hipMemcpyAsync(digits to device)") &&
        set_slot(E, s))
        chk(E, launch_pass(E->kcfwd, E->ncoljobs, E->lds_cols, E->stream, E->d_slot, nullptr, kModeDigits, 0),
            "column pass (from digits)");
    chk(E, hipStreamSynchronize(E->stream), "hipStreamSynchronize(from digits)");  // hdig is reused
}

void engine_to_digits(kite_ctx *ctx, const void *x, int64_t *dig) {
    Engine *E = ctx->eng;
    const ShapeDev &sh = E->shape->host;
    SlotDev s = E->slot;
    s.x = (double2 *)const_cast<void *>(x);
    // the conversion's own rounding error is not an operation's round-off: count it aside
    s.maxerr = E->maxerr + 1;
    bool ok = set_slot(E, s) &&
              chk(E, launch_pass(E->kcinv, E->ncoljobs, E->lds_cols, E->stream, E->d_slot, nullptr, kModeDigits, 0),
                  "column pass (to digits)") &&
              chk(E, hipMemcpyAsync(E->hdig.data(), E->dig, (size_t)ctx->L * 4, hipMemcpyDeviceToHost, E->stream),
                  "hipMemcpyAsync(digits to host)");
    ok = chk(E, hipStreamSynchronize(E->stream), "hipStreamSynchronize(to digits)") && ok;
    if (!ok) {  // nothing valid came back: zeros, with the failure latched for the caller
        memset(dig, 0, (size_t)ctx->L * sizeof(int64_t));
        return;
    }
    const uint32_t R = sh.R, C = sh.C, G = sh.G, NG = sh.NG;
    for (uint32_t g = 0; g < NG; g++)
        for (uint32_t r = 0; r < R; r++)
            for (uint32_t i = 0; i < 2 * G; i++)
                dig[2 * ((size_t)r * C + g * G) + i] = E->hdig[((size_t)g * R + r) * 2 * G + i];
}

// Max round-off since the last reset; +inf once the context has failed, so that every caller's
// round-off test fails as well.
double engine_maxerr(kite_ctx *ctx) {
    Engine *E = ctx->eng;
    unsigned long long v = 0;
    if (E->err == hipSuccess) {
        chk(E, hipMemcpyAsync(&v, E->maxerr, 8, hipMemcpyDeviceToHost, E->stream), "hipMemcpyAsync(round-off)");
        chk(E, hipStreamSynchronize(E->stream), "hipStreamSynchronize(round-off)");
    }
    if (E->err != hipSuccess) return INFINITY;
    double d;
    memcpy(&d, &v, 8);
    return d;
}

void engine_reset_maxerr(kite_ctx *ctx) {
    Engine *E = ctx->eng;
    if (E->err != hipSuccess) return;
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, E->stream, E->maxerr, 0ull, 1);
    chk(E, hipGetLastError(), "k_set_u64");
}

void engine_sync(kite_ctx *ctx) { chk(ctx->eng, hipStreamSynchronize(ctx->eng->stream), "hipStreamSynchronize"); }

}  // namespace kite

// ================================================================================================
// Batch: many contexts advanced by shared launches (include/kite_batch.h)
// ================================================================================================

namespace kite {

struct BatchSlot {
    kite_ctx *ctx = nullptr;
    std::vector<kite_num *> nums;
    int32_t xadd = 0;  // added to x after each squaring of this slot (LLR: -2)
};

}  // namespace kite

struct kite_batch {
    std::vector<kite::BatchSlot> slots;
    hipStream_t stream = nullptr;
    kite::SlotDev *d_slots = nullptr, *h_slots = nullptr;  // h_slots is pinned
    size_t slot_cap = 0;
    int2 *d_rows = nullptr, *d_cols = nullptr;
    int nrows = 0, ncols = 0;
    size_t jobs_cap = 0;
    struct Group {
        kite::PassKernel k1, k2;  // rows: k1 (k2: square-mode kernel, if any); columns: k1 = inverse, k2 = forward
        int off, count;
        size_t lds;
        double sumL;        // sum of the FFT lengths of the group's tests
        int stat[2];        // index into kstats of k1, k2
        kite::PassKernel kf = nullptr;  // columns: square-mode fused pass, if any
        int stat_f = -1;
        size_t lds_f = 0;
        int nt_f = kite::kNT;  // its threads per work-group
        int nt_r = kite::kNT;  // threads per work-group of the row kernels k1 (and k2 unless rsq)
        int part = 0;  // the stream the group's launches go to
        // the persistent variant of the square-mode pass (rows: k2, columns: kf), if any; its job
        // lists (index into d_lists, the list count, the longest list) and the work-groups each list
        // needs at least (kernels_fast.h: persist_loop)
        kite::PersistKernel kp = nullptr;
        int ls = -1, nl = 1, maxcnt = 0, pmin = 1, stat_p = -1;
    };
    std::vector<Group> grows, gcols;
    // The slots are split into `nparts` parts with a stream each: the parts' launches run
    // concurrently, so their kernels overlap in different phases (loads, compute, stores).
    static constexpr int kMaxParts = 4;
    int nparts = 2;
    // kite_batch_set_shape_parts: the most common shape's slots go to the nparts parts, each other
    // shape's to a stream of its own while streams are left; nused = streams in use (refresh)
    int shape_parts = 0;
    int nused = 2;
    hipStream_t ps[kMaxParts] = {};  // the set in use; ps[0] == stream
    // Two stream sets: `pooled` (hipStreamCreateWithFlags: the runtime shares a few hardware queues
    // among all of the process's streams) for graph replay, and `own` (a full CU mask: each stream
    // on a hardware queue of its own, created when direct launches are first chosen). Direct
    // launches on pooled streams can find two parts on one queue, where they run one after the
    // other (320 tests at n ~ 660K: 110.7 against 92-94 us per step); graph replay on the own set
    // measured slower than on the pooled one (125-130 against 96 us).
    hipStream_t pooled[kMaxParts] = {}, own[kMaxParts] = {};
    hipEvent_t evfork = nullptr, evtable = nullptr, evjoin[kMaxParts] = {};
    std::vector<int> slot_part;
    unsigned long long *d_maxerr = nullptr;
    unsigned long long *d_cvterr = nullptr;  // round-off of digit conversions (kept apart)
    int *d_eq = nullptr, *h_eq = nullptr;     // kite_batch_equal's answers (h_eq pinned)
    bool dirty = true;
    uint64_t gen = 0;  // bumped whenever the job arrays or groups are rebuilt
    // square calls replayed from a captured graph: valid for gsteps steps of generation ggen
    bool graphs = true;
    hipGraphExec_t gexec = nullptr;
    uint32_t gsteps = 0;
    uint64_t ggen = 0;
    // column jobs are laid out for this many XCDs (work-group i runs on XCD i mod xcds): each
    // test's column groups go to one XCD, in order (1: plain order)
    int xcds = kite::kXcds;
    // persistent launches (kite_batch_set_persistent): square-mode passes with a persistent kernel
    // run on at most the device's resident capacity (divided among the parts), each work-group
    // running a fixed share of the jobs
    int persist = 0;
    unsigned int *d_lists = nullptr;  // kPersistListWords per group with a persistent kernel
    size_t lists_cap = 0;
    int2 *d_cols_p = nullptr;         // column jobs list by list (the persistent launches' order)
    int timing = 0;  // 0: off; N: every N-th step of a square call is timed
    // timing: one event pair per launch of the last timed call, and what each pair measured
    std::vector<hipEvent_t> ev;
    std::vector<int> ev_stat;
    std::vector<double> ev_bytes;  // algorithmic bytes of each timed launch
    double ms[3] = {0, 0, 0};
    uint64_t launches[3] = {0, 0, 0};
    struct KStat {
        std::string name;
        int pass;
        double ms = 0, bytes = 0;  // summed over launches: kernel time, algorithmic bytes
        uint64_t launches = 0;
    };
    std::vector<KStat> kstats;
    // first HIP failure of the batch (sticky): every later call returns -1
    hipError_t err = hipSuccess;
};

using namespace kite;

// Records the first failure of b; true if `e` (and every earlier call) succeeded.
static bool bchk(kite_batch *b, hipError_t e, const char *what) {
    e = hip_result(e);
    if (e != hipSuccess && b->err == hipSuccess) {
        b->err = e;
        g_err = std::string(what) + ": " + hipGetErrorString(e);
        fprintf(stderr, "kite: batch: %s\n", g_err.c_str());
    }
    return b->err == hipSuccess;
}

#define BATCH_OK(b, expr)                                 \
    do {                                                  \
        if (!bchk((b), (expr), #expr)) return -1;         \
    } while (0)

static bool batch_sync_all(kite_batch *b) {
    bool ok = true;
    for (int p = 0; p < kite_batch::kMaxParts; p++)
        if (b->ps[p]) ok = bchk(b, hipStreamSynchronize(b->ps[p]), "hipStreamSynchronize(batch)") && ok;
    return ok;
}

// The row kernel and job count a slot's row pass uses: a batch of one test spreads its rows over
// one pair per work-group where that kernel exists (latency decides when the device is mostly idle).
static void rows_kernel_of(const Engine *E, bool alone, PassKernel *k, PassKernel *ksq, int *njobs, size_t *lds) {
    if (alone && E->krows1) {
        *k = E->krows1;
        *ksq = nullptr;
        *njobs = E->nrowjobs1;
        *lds = E->lds_rows1;
    } else {
        *k = E->krows;
        *ksq = E->krows_sq;
        *njobs = E->nrowjobs;
        *lds = E->lds_rows;
    }
}

static int batch_refresh(kite_batch *b) {
    if (b->err != hipSuccess) return -1;
    if (!b->dirty) return 0;
    if (!batch_sync_all(b)) return -1;  // the tables below may be reallocated
    const size_t cap = b->slots.size();
    if (cap > b->slot_cap) {
        const size_t nc = std::max<size_t>(cap, 2 * b->slot_cap);
        if (b->d_slots) hipFree(b->d_slots);
        if (b->h_slots) hipHostFree(b->h_slots);
        b->d_slots = b->h_slots = nullptr;
        unsigned long long *old = b->d_maxerr;
        b->d_maxerr = nullptr;
        BATCH_OK(b, hipMalloc(&b->d_slots, nc * sizeof(SlotDev)));
        BATCH_OK(b, hipHostMalloc((void **)&b->h_slots, nc * sizeof(SlotDev), hipHostMallocDefault));
        BATCH_OK(b, hipMalloc(&b->d_maxerr, nc * 8));
        BATCH_OK(b, hipMemsetAsync(b->d_maxerr, 0, nc * 8, b->stream));
        if (b->d_cvterr) hipFree(b->d_cvterr);
        if (b->d_eq) hipFree(b->d_eq);
        if (b->h_eq) hipHostFree(b->h_eq);
        b->d_cvterr = nullptr;
        b->d_eq = b->h_eq = nullptr;
        BATCH_OK(b, hipMalloc(&b->d_cvterr, nc * 8));
        BATCH_OK(b, hipMalloc(&b->d_eq, nc * sizeof(int)));
        BATCH_OK(b, hipHostMalloc((void **)&b->h_eq, nc * sizeof(int), hipHostMallocDefault));
        if (old) {
            BATCH_OK(b, hipMemcpyAsync(b->d_maxerr, old, b->slot_cap * 8, hipMemcpyDeviceToDevice, b->stream));
            BATCH_OK(b, hipStreamSynchronize(b->stream));
            hipFree(old);
        }
        b->slot_cap = nc;
    }
    // jobs grouped by kernel (one launch per group), groups contiguous in the job arrays; cols_orig:
    // the column jobs before the XCD layout (each test's jobs consecutive)
    std::vector<int2> rows, cols, cols_orig;
    b->grows.clear();
    b->gcols.clear();
    std::vector<int> order;
    for (size_t s = 0; s < cap; s++)
        if (b->slots[s].ctx) order.push_back((int)s);
    const bool alone = order.size() == 1;
    auto stat_of = [&](const std::string &nm, int pass) {
        for (size_t q = 0; q < b->kstats.size(); q++)
            if (b->kstats[q].name == nm && b->kstats[q].pass == pass) return (int)q;
        kite_batch::KStat ks;
        ks.name = nm;
        ks.pass = pass;
        b->kstats.push_back(ks);
        return (int)b->kstats.size() - 1;
    };
    // parts: active slots dealt round-robin (by slot order) to the parts; with shape_parts, only the
    // most common shape's, each other shape taking a part of its own while streams are left (its
    // short launches then run beside the main parts' rather than after them in one stream, and
    // lengths that would share a pair kernel each keep their own kernel)
    b->slot_part.assign(cap, 0);
    b->nused = std::min<int>(b->nparts, std::max<int>(1, (int)order.size()));
    {
        std::vector<const void *> own(cap, nullptr);  // a shape with a part of its own -> that part
        std::vector<int> own_part;
        std::vector<const void *> own_key;
        int avail = 0;  // streams that exist (ps[0] always)
        while (avail < kite_batch::kMaxParts && b->ps[avail]) avail++;
        if (b->shape_parts && b->nparts < avail) {
            std::vector<std::pair<int, const void *>> cnt;  // (slots, shape)
            for (int s : order) {
                const void *key = b->slots[s].ctx->eng->shape;
                auto it = std::find_if(cnt.begin(), cnt.end(), [&](const std::pair<int, const void *> &c) { return c.second == key; });
                if (it == cnt.end()) cnt.push_back({1, key});
                else it->first++;
            }
            std::stable_sort(cnt.begin(), cnt.end(), [](const std::pair<int, const void *> &x, const std::pair<int, const void *> &y) {
                return x.first > y.first;
            });
            for (size_t i = 1; i < cnt.size() && b->nparts + (int)own_key.size() < avail; i++) {
                own_key.push_back(cnt[i].second);
                own_part.push_back(b->nparts + (int)own_key.size() - 1);
            }
            if (!own_key.empty()) b->nused = b->nparts + (int)own_key.size();
        }
        int i = 0;
        for (int s : order) {
            const void *key = b->slots[s].ctx->eng->shape;
            const auto it = std::find(own_key.begin(), own_key.end(), key);
            b->slot_part[s] = it != own_key.end() ? own_part[it - own_key.begin()] : (i++) % b->nparts;
        }
    }
    auto build = [&](bool isrows, int part) {
        std::vector<int2> &jobs = isrows ? rows : cols;
        std::vector<kite_batch::Group> grp;
        std::vector<int> gR;  // column length of each group
        std::vector<bool> done(cap, false);
        std::vector<int> ord;
        for (int s : order)
            if (b->slot_part[s] == part) ord.push_back(s);
        if (!isrows)  // column groups in ascending R, so that adjacent lengths sit side by side
            std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
                return b->slots[x].ctx->eng->shape->ch.R < b->slots[y].ctx->eng->shape->ch.R;
            });
        // per-slot job lists of this part, so that a group's column jobs can be interleaved
        for (int s0 : ord) {
            if (done[s0]) continue;
            Engine *E0 = b->slots[s0].ctx->eng;
            PassKernel rk0, rsq0;
            int rj0;
            size_t rl0;
            rows_kernel_of(E0, alone, &rk0, &rsq0, &rj0, &rl0);
            const PassKernel key = isrows ? rk0 : E0->kcinv;
            kite_batch::Group g{key, isrows ? rsq0 : E0->kcfwd, (int)jobs.size(), 0, 0, 0.0, {-1, -1}};
            for (int s : ord) {
                Engine *E = b->slots[s].ctx->eng;
                PassKernel rk, rsq;
                int rj;
                size_t rl;
                rows_kernel_of(E, alone, &rk, &rsq, &rj, &rl);
                if (done[s] || (isrows ? rk : E->kcinv) != key || (!isrows && E->kcfused != E0->kcfused)) continue;
                done[s] = true;
                const int nj = isrows ? rj : E->ncoljobs;
                for (int u = 0; u < nj; u++) jobs.push_back(make_int2(s, u));
                g.lds = std::max(g.lds, isrows ? rl : E->lds_cols);
                if (!isrows) g.lds_f = std::max(g.lds_f, E->lds_fused);
                g.sumL += 2.0 * E->H;
            }
            g.count = (int)jobs.size() - g.off;
            if (isrows) {
                g.stat[0] = stat_of(rk0 == E0->krows1 && alone ? "k_rows_c1<" + std::to_string(E0->shape->ch.C) + ">"
                                                               : E0->name[0], 0);
                g.kp = rsq0 ? fast_rows_sq_p(E0->shape->ch.C) : nullptr;
                g.nt_r = E0->nt_rows;
                if (g.kp) g.stat_p = stat_of(E0->name[0] + "_p", 0);
            } else {
                g.stat[0] = stat_of(E0->name[1], 1);
                g.stat[1] = stat_of(E0->name[2], 2);
                g.kf = E0->kcfused;
                g.nt_f = E0->nt_fused;
                if (g.kf) g.stat_f = stat_of("k_cols_fused_r<" + std::to_string(E0->shape->ch.R) + ">", 1);
                g.kp = g.kf ? fast_cols_fused_p(E0->shape->ch.R) : nullptr;
                if (g.kp) g.stat_p = stat_of("k_cols_fused_r_p<" + std::to_string(E0->shape->ch.R) + ">", 1);
            }
            g.part = part;
            grp.push_back(g);
            gR.push_back((int)E0->shape->ch.R);
        }
        std::vector<kite_batch::Group> &out = isrows ? b->grows : b->gcols;
        if (isrows) {
            out.insert(out.end(), grp.begin(), grp.end());
            return;
        }
        // Work-group i of a launch runs on XCD i mod xcds. The jobs of each column group are
        // reordered so that every test's column groups (whose carries pass between neighbours
        // through flags) land on one XCD, consecutively: neighbours then start together and
        // exchange their flags through that XCD's L2. Each test's jobs keep their relative order,
        // so every flag a work-group waits for is still published by one dispatched before it or
        // by its own test's last group, as before.
        cols_orig.resize(jobs.size());
        for (const kite_batch::Group &g : grp)
            std::copy(jobs.begin() + g.off, jobs.begin() + g.off + g.count, cols_orig.begin() + g.off);
        if (b->xcds > 1)
            for (const kite_batch::Group &g : grp) {
                // each test's jobs (first, count), in order
                std::vector<std::pair<int, int>> runs;
                for (int i = g.off; i < g.off + g.count;) {
                    int j = i;
                    while (j < g.off + g.count && jobs[j].x == jobs[i].x) j++;
                    runs.push_back({i, j - i});
                    i = j;
                }
                // Fewer tests than XCDs: each test's jobs are cut into xcds / tests contiguous
                // chunks, one XCD each, so that neighbouring column groups still share an L2 (with
                // G = 4 a group's row segments are half cache lines: on different XCDs both halves
                // would be fetched twice and written back partially). A chunk's first job waits for
                // the previous chunk's last one, dispatched later; those are at most xcds jobs per
                // test, far fewer than the device holds at once, so the rest keep running.
                const int ntests = (int)runs.size();
                const int per = ntests < b->xcds ? b->xcds / ntests : 1;
                std::vector<std::vector<int2>> lists(b->xcds);
                for (int t = 0; t < ntests; t++) {
                    const int first = runs[t].first, cnt = runs[t].second;
                    for (int u = 0; u < cnt; u++) {
                        const int x = per > 1 ? t * per + (int)((int64_t)u * per / cnt) : t % b->xcds;
                        lists[x].push_back(jobs[first + u]);
                    }
                }
                std::vector<size_t> pos(b->xcds, 0);
                int i = g.off;
                for (size_t q = 0; i < g.off + g.count; q++)
                    for (int x = 0; x < b->xcds && i < g.off + g.count; x++) {
                        int xx = x;  // a drained list lends its position to the next one with jobs
                        while (pos[xx] >= lists[xx].size()) xx = (xx + 1) % b->xcds;
                        jobs[i++] = lists[xx][pos[xx]++];
                    }
            }
        // adjacent compile-time column lengths share one launch per pass (the fused pass only if
        // both lengths have it)
        std::vector<kite_batch::Group> merged;
        for (size_t i = 0; i < grp.size(); i++) {
            const int r1 = gR[i];
            if (i + 1 < grp.size() && grp[i].k1 == fast_cols_inv(r1) && fast_cols_next(r1) == gR[i + 1] &&
                grp[i + 1].k1 == fast_cols_inv(gR[i + 1]) && (grp[i].kf != nullptr) == (grp[i + 1].kf != nullptr)) {
                kite_batch::Group g = grp[i];
                g.k1 = fast_cols_pair_inv(r1);
                g.k2 = fast_cols_pair_fwd(r1);
                g.count += grp[i + 1].count;
                g.lds = std::max(g.lds, grp[i + 1].lds);
                g.lds_f = std::max(g.lds_f, grp[i + 1].lds_f);
                g.sumL += grp[i + 1].sumL;
                const std::string rr = "<" + std::to_string(r1) + "," + std::to_string(gR[i + 1]) + ">";
                g.stat[0] = stat_of("k_cols_inv_rr" + rr, 1);
                g.stat[1] = stat_of("k_cols_fwd_rr" + rr, 2);
                g.kf = (grp[i].kf && grp[i + 1].kf) ? fast_cols_pair_fused(r1) : nullptr;
                g.nt_f = fast_cols_pair_fused_threads(r1);
                g.stat_f = g.kf ? stat_of("k_cols_fused_rr" + rr, 1) : -1;
                g.kp = g.kf ? fast_cols_pair_fused_p(r1) : nullptr;
                g.stat_p = g.kp ? stat_of("k_cols_fused_rr_p" + rr, 1) : -1;
                merged.push_back(g);
                i++;
            } else {
                merged.push_back(grp[i]);
            }
        }
        out.insert(out.end(), merged.begin(), merged.end());
    };
    for (int part = 0; part < b->nused; part++) {
        build(true, part);
        build(false, part);
    }
    // Persistent launches: the job lists of each group that has a persistent kernel (each list's
    // first job and count), and the column jobs list by list, whole tests dealt round-robin to
    // min(xcds, tests, 8) lists.
    std::vector<int2> cols_p(cols);
    std::vector<uint32_t> lsh;
    auto add_lists = [&](kite_batch::Group &g, int nl, const uint32_t *off, const uint32_t *cnt, int pmin) {
        g.ls = (int)(lsh.size() / kPersistListWords);
        g.nl = nl;
        g.maxcnt = 0;
        lsh.resize(lsh.size() + kPersistListWords, 0);
        uint32_t *t = &lsh[(size_t)g.ls * kPersistListWords];
        for (int x = 0; x < nl; x++) {
            t[kLsOff + x] = off[x];
            t[kLsCnt + x] = cnt[x];
            g.maxcnt = std::max(g.maxcnt, (int)cnt[x]);
        }
        g.pmin = std::min(pmin, g.maxcnt);
    };
    for (kite_batch::Group &g : b->grows)
        if (g.kp) {
            const uint32_t off = 0, cnt = (uint32_t)g.count;
            add_lists(g, 1, &off, &cnt, 1);
        }
    for (kite_batch::Group &g : b->gcols) {
        if (!g.kp) continue;
        std::vector<std::pair<int, int>> runs;  // each test's jobs in cols_orig: (first, count)
        int ng = 1;
        for (int i = g.off; i < g.off + g.count;) {
            int j = i;
            while (j < g.off + g.count && cols_orig[j].x == cols_orig[i].x) j++;
            runs.push_back({i, j - i});
            ng = std::max(ng, j - i);
            i = j;
        }
        const int nl = std::max(1, std::min({b->xcds, kXcds, (int)runs.size()}));
        uint32_t off[kXcds], cnt[kXcds];
        int pos = g.off;
        for (int x = 0; x < nl; x++) {
            off[x] = (uint32_t)(pos - g.off);
            for (size_t t = (size_t)x; t < runs.size(); t += (size_t)nl)
                for (int u = 0; u < runs[t].second; u++) cols_p[pos++] = cols_orig[runs[t].first + u];
            cnt[x] = (uint32_t)(pos - g.off) - off[x];
        }
        add_lists(g, nl, off, cnt, ng);  // a column job waits for jobs up to NG - 1 positions later
    }
    const size_t need = std::max(rows.size(), cols.size());
    if (need > b->jobs_cap) {
        if (b->d_rows) hipFree(b->d_rows);
        if (b->d_cols) hipFree(b->d_cols);
        if (b->d_cols_p) hipFree(b->d_cols_p);
        b->d_rows = b->d_cols = b->d_cols_p = nullptr;
        const size_t nc = std::max(need, 2 * b->jobs_cap);
        BATCH_OK(b, hipMalloc(&b->d_rows, nc * sizeof(int2)));
        BATCH_OK(b, hipMalloc(&b->d_cols, nc * sizeof(int2)));
        BATCH_OK(b, hipMalloc(&b->d_cols_p, nc * sizeof(int2)));
        b->jobs_cap = nc;
    }
    if (lsh.size() > b->lists_cap) {
        if (b->d_lists) hipFree(b->d_lists);
        b->d_lists = nullptr;
        const size_t nc = std::max(lsh.size(), 2 * b->lists_cap);
        BATCH_OK(b, hipMalloc(&b->d_lists, nc * sizeof(uint32_t)));
        b->lists_cap = nc;
    }
    if (!lsh.empty())
        BATCH_OK(b, hipMemcpy(b->d_lists, lsh.data(), lsh.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    if (!cols_p.empty())
        BATCH_OK(b, hipMemcpy(b->d_cols_p, cols_p.data(), cols_p.size() * sizeof(int2), hipMemcpyHostToDevice));
    if (!rows.empty()) BATCH_OK(b, hipMemcpy(b->d_rows, rows.data(), rows.size() * sizeof(int2), hipMemcpyHostToDevice));
    if (!cols.empty()) BATCH_OK(b, hipMemcpy(b->d_cols, cols.data(), cols.size() * sizeof(int2), hipMemcpyHostToDevice));
    b->nrows = (int)rows.size();
    b->ncols = (int)cols.size();
    b->dirty = false;
    b->gen++;
    return 0;
}

// Fills the slot table: residue `x` (and operand `y`) per slot, step count per slot (0 = idle).
// digits_aside: a digits-mode pass writes into the slot's operand-spectrum buffer instead of its
// digit buffer, and its rounding error into the conversion accumulator.
static int batch_upload(kite_batch *b, const int32_t *xi, const int32_t *yi, const uint32_t *count,
                        bool digits_aside = false) {
    if (batch_refresh(b)) return -1;
    if (!batch_sync_all(b)) return -1;  // the previous table may still be read
    for (size_t s = 0; s < b->slots.size(); s++) {
        SlotDev &d = b->h_slots[s];
        kite_ctx *ctx = b->slots[s].ctx;
        if (!ctx) {
            memset(&d, 0, sizeof d);
            continue;
        }
        d = ctx->eng->slot;
        d.maxerr = b->d_maxerr + s;
        const bool on = xi[s] >= 0 && (size_t)xi[s] < b->slots[s].nums.size() && count[s] > 0;
        d.count = on ? count[s] : 0;
        d.seq = ctx->eng->seq;  // one counter per engine: its own squarings use the same flags
        d.xadd = yi ? 0 : b->slots[s].xadd;  // squarings only
        d.x = on ? (double2 *)b->slots[s].nums[xi[s]]->dev : nullptr;
        d.y_src = (on && yi && yi[s] >= 0) ? (const double2 *)b->slots[s].nums[yi[s]]->dev : nullptr;
        if (digits_aside) {
            d.dig = reinterpret_cast<int32_t *>(d.yfreq);  // 8 L bytes: room for L int32 digits
            d.maxerr = b->d_cvterr + s;
        }
    }
    BATCH_OK(b, hipMemcpyAsync(b->d_slots, b->h_slots, b->slots.size() * sizeof(SlotDev), hipMemcpyHostToDevice,
                               b->stream));
    // the other parts' launches read the table too: they wait for the copy
    if (b->nused > 1) {
        BATCH_OK(b, hipEventRecord(b->evtable, b->stream));
        for (int p = 1; p < b->nused; p++) BATCH_OK(b, hipStreamWaitEvent(b->ps[p], b->evtable, 0));
    }
    return 0;
}

// Work-groups of a persistent launch (0: launch one work-group per job instead): at most the
// kernel's resident capacity on the device (the occupancy its registers and LDS allow, times the
// CUs) divided among the parts whose launches run at the same time, nl lists of P work-groups each,
// P no more than the longest list and at least what the lists' waits need (persist_loop).
static int persist_grid(kite_batch *b, const kite_batch::Group &g, int threads, size_t lds) {
    static std::mutex mu;
    static std::vector<std::pair<std::pair<const void *, size_t>, int>> cache;  // (kernel, lds) -> WGs resident
    int cap = 0;
    {
        std::lock_guard<std::mutex> lk(mu);
        for (const auto &c : cache)
            if (c.first.first == g.kp && c.first.second == lds) cap = c.second;
        if (!cap) {
            int per_cu = 0, dev = 0, cus = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, g.kp, threads, lds) != hipSuccess ||
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
                return 0;
            cap = std::max(per_cu, 0) * std::max(cus, 0);
            cache.push_back({{g.kp, lds}, cap});
        }
    }
    const int P = std::min(cap / std::max(1, b->nused) / g.nl, g.maxcnt);
    return P >= g.pmin && P > 0 ? P * g.nl : 0;
}

// One launch per group of the pass; with `timed`, each launch carries a start/stop event pair.
// In square mode a column group with a fused kernel runs it as pass 1 and skips pass 2.
static hipError_t batch_launch(kite_batch *b, int pass, int mode, int step, bool timed = false) {
    const auto &groups = pass == 0 ? b->grows : b->gcols;
    const int2 *jobs = pass == 0 ? b->d_rows : b->d_cols;
    for (const auto &g : groups) {
        if (!g.count) continue;
        const bool rsq = pass == 0 && mode == kModeSquare && g.k2;
        const bool fused = pass > 0 && mode == kModeSquare && g.kf;
        if (fused && pass == 2) continue;
        const PassKernel fn = fused ? g.kf : (pass == 2 || rsq) ? g.k2 : g.k1;
        const size_t lds = rsq ? 0 : fused ? g.lds_f : g.lds;
        const int threads = rsq ? kRowsSqThreads : fused ? g.nt_f : pass == 0 ? g.nt_r : kNT;
        const int pgrid = b->persist && g.kp && (rsq || fused) ? persist_grid(b, g, threads, lds) : 0;
        const bool persist = pgrid > 0;
        // a call with timing on runs every launch on the first stream, so that each timed launch
        // has the device to itself (every call starts after all streams' earlier work)
        const hipStream_t st = b->timing ? b->ps[0] : b->ps[g.part];
        auto launch = [&](hipEvent_t e0, hipEvent_t e1) -> hipError_t {
            if (!persist)
                return launch_pass(fn, g.count, lds, st, b->d_slots, jobs + g.off, mode, step, threads, e0, e1);
            const int2 *pj = (pass == 0 ? b->d_rows : b->d_cols_p) + g.off;
            const unsigned int *ls = b->d_lists + (size_t)g.ls * kPersistListWords;
            int nl = g.nl;
            void *args[] = {(void *)&b->d_slots, (void *)&pj, (void *)&step, (void *)&ls, (void *)&nl};
            return launch_k(g.kp, pgrid, threads, args, lds, st, e0, e1);
        };
        hipError_t e;
        if (timed) {
            const size_t ei = 2 * b->ev_stat.size();
            if (b->ev.size() < ei + 2) {
                b->ev.resize(ei + 2);
                hipEventCreate(&b->ev[ei]);
                hipEventCreate(&b->ev[ei + 1]);
            }
            b->ev_stat.push_back(persist ? g.stat_p : fused ? g.stat_f : g.stat[pass == 2 ? 1 : 0]);
            // algorithmic bytes: each test's residue (8 L) read and written by the row pass and the
            // fused column pass, read or written by each split column pass
            b->ev_bytes.push_back((pass == 0 || fused ? 16.0 : 8.0) * g.sumL);
            e = launch(b->ev[ei], b->ev[ei + 1]);
        } else {
            e = launch(nullptr, nullptr);
        }
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

extern "C" kite_batch *kite_batch_new(void) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return nullptr;
    set_lds_limits();
    kite_batch *b = new kite_batch();
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        return nullptr;
    }
    b->ps[0] = b->stream;
    for (int p = 1; p < kite_batch::kMaxParts; p++)
        if (hipStreamCreateWithFlags(&b->ps[p], hipStreamNonBlocking) != hipSuccess) b->ps[p] = nullptr;
    for (int p = 0; p < kite_batch::kMaxParts; p++) b->pooled[p] = b->ps[p];
    hipEventCreateWithFlags(&b->evfork, hipEventDisableTiming);
    hipEventCreateWithFlags(&b->evtable, hipEventDisableTiming);
    for (int p = 0; p < kite_batch::kMaxParts; p++) hipEventCreateWithFlags(&b->evjoin[p], hipEventDisableTiming);
    return b;
}

extern "C" void kite_batch_free(kite_batch *b) {
    if (!b) return;
    batch_sync_all(b);
    for (size_t s = 0; s < b->slots.size(); s++) kite_batch_remove(b, (int)s);
    for (hipEvent_t e : b->ev) hipEventDestroy(e);
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    if (b->d_slots) hipFree(b->d_slots);
    if (b->h_slots) hipHostFree(b->h_slots);
    if (b->d_rows) hipFree(b->d_rows);
    if (b->d_cols) hipFree(b->d_cols);
    if (b->d_cols_p) hipFree(b->d_cols_p);
    if (b->d_lists) hipFree(b->d_lists);
    if (b->d_maxerr) hipFree(b->d_maxerr);
    if (b->d_cvterr) hipFree(b->d_cvterr);
    if (b->d_eq) hipFree(b->d_eq);
    if (b->h_eq) hipHostFree(b->h_eq);
    for (int p = 0; p < kite_batch::kMaxParts; p++) {
        if (b->pooled[p]) hipStreamDestroy(b->pooled[p]);
        if (b->own[p]) hipStreamDestroy(b->own[p]);
        if (b->evjoin[p]) hipEventDestroy(b->evjoin[p]);
    }
    if (b->evfork) hipEventDestroy(b->evfork);
    if (b->evtable) hipEventDestroy(b->evtable);
    delete b;
}

int kite::batch_add_opts(kite_batch *b, uint64_t k, uint64_t n, const kite_opts *o, uint32_t nnums) {
    if (b->err != hipSuccess) return -1;
    kite_ctx *ctx = kite_ctx_new(k, n, -1, o);
    if (!ctx) return -1;
    BatchSlot bs;
    bs.ctx = ctx;
    for (uint32_t i = 0; i < nnums; i++) {
        kite_num *x = kite_num_new(ctx);
        if (!x) {
            for (kite_num *y : bs.nums) kite_num_free(ctx, y);
            kite_ctx_free(ctx);
            return -1;
        }
        bs.nums.push_back(x);
    }
    engine_sync(ctx);
    if (engine_error(ctx)) {
        for (kite_num *y : bs.nums) kite_num_free(ctx, y);
        kite_ctx_free(ctx);
        return -1;
    }
    size_t s = 0;
    while (s < b->slots.size() && b->slots[s].ctx) s++;
    if (s == b->slots.size()) b->slots.emplace_back();
    b->slots[s] = std::move(bs);
    b->dirty = true;
    if (batch_refresh(b)) return -1;
    // clear this slot's round-off accumulator
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, b->stream, b->d_maxerr + s, 0ull, 1);
    BATCH_OK(b, hipGetLastError());
    BATCH_OK(b, hipStreamSynchronize(b->stream));
    return (int)s;
}

extern "C" int kite_batch_add(kite_batch *b, uint64_t k, uint64_t n, uint32_t fft_len, uint32_t nnums) {
    kite_opts o{};
    o.backend = KITE_BACKEND_HIP;
    o.fft_len = fft_len;
    return batch_add_opts(b, k, n, &o, nnums);
}

extern "C" void kite_batch_remove(kite_batch *b, int slot) {
    if (slot < 0 || (size_t)slot >= b->slots.size() || !b->slots[slot].ctx) return;
    batch_sync_all(b);
    BatchSlot &bs = b->slots[slot];
    for (kite_num *x : bs.nums) kite_num_free(bs.ctx, x);
    kite_ctx_free(bs.ctx);
    bs = BatchSlot();
    b->dirty = true;
}

extern "C" int kite_batch_capacity(const kite_batch *b) { return (int)b->slots.size(); }

extern "C" kite_ctx *kite_batch_ctx(kite_batch *b, int slot) {
    return slot >= 0 && (size_t)slot < b->slots.size() ? b->slots[slot].ctx : nullptr;
}

extern "C" kite_num *kite_batch_num(kite_batch *b, int slot, uint32_t i) {
    if (slot < 0 || (size_t)slot >= b->slots.size() || i >= b->slots[slot].nums.size()) return nullptr;
    return b->slots[slot].nums[i];
}

// 0 if the batch and every slot's context (whose streams carry the limb conversions) are intact.
extern "C" int kite_batch_sync(kite_batch *b) {
    if (!batch_sync_all(b)) return -1;
    for (const BatchSlot &bs : b->slots)
        if (bs.ctx && engine_error(bs.ctx)) return -1;
    return 0;
}

extern "C" int kite_batch_set_parts(kite_batch *b, int parts) {
    int avail = 0;
    while (avail < kite_batch::kMaxParts && b->ps[avail]) avail++;
    parts = std::max(1, std::min(parts, avail));
    if (parts != b->nparts) {
        b->nparts = parts;
        b->dirty = true;
    }
    return parts;
}

extern "C" int kite_batch_square(kite_batch *b, const int32_t *which, const uint32_t *count) {
    if (b->err != hipSuccess) return -1;
    uint32_t maxc = 0;
    for (size_t s = 0; s < b->slots.size(); s++)
        if (b->slots[s].ctx && which[s] >= 0) maxc = std::max(maxc, count[s]);
    if (!maxc) return 0;
    if (batch_upload(b, which, nullptr, count)) return -1;
    b->ev_stat.clear();
    b->ev_bytes.clear();
    for (size_t s = 0; s < b->slots.size(); s++)
        if (b->slots[s].ctx) b->slots[s].ctx->eng->seq += b->h_slots[s].count;  // the table holds the old value
    auto launch_all = [&]() -> hipError_t {
        for (uint32_t st = 0; st < maxc; st++) {
            const bool timed = b->timing && st % (uint32_t)b->timing == 0;
            for (int pass = 0; pass < 3; pass++) {
                const hipError_t e = batch_launch(b, pass, kModeSquare, (int)st, timed);
                if (e != hipSuccess) return e;
            }
        }
        return hipSuccess;
    };
    // Kernels read the slot table at run time, so a captured sequence of maxc steps stays valid
    // until the job arrays change (a new generation): the usual Gerbicz-block loop replays it.
    bool replay = b->graphs && !b->timing && maxc > 1;
    if (replay && (!b->gexec || b->gsteps != maxc || b->ggen != b->gen)) {
        if (b->gexec) hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
        hipGraph_t g = nullptr;
        replay = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (replay) {
            // the other parts' streams join the capture (fork) and rejoin the first (join)
            hipEventRecord(b->evfork, b->stream);
            for (int p = 1; p < b->nused; p++) hipStreamWaitEvent(b->ps[p], b->evfork, 0);
            const bool launched = launch_all() == hipSuccess;
            for (int p = 1; p < b->nused; p++) {
                hipEventRecord(b->evjoin[p], b->ps[p]);
                hipStreamWaitEvent(b->stream, b->evjoin[p], 0);
            }
            replay = hipStreamEndCapture(b->stream, &g) == hipSuccess && g && launched &&
                     hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0) == hipSuccess;
            if (g) hipGraphDestroy(g);
            if (!replay) b->gexec = nullptr;
        }
        // a capture that failed falls back to direct launches (this clears only the capture's own
        // error: every earlier call of the batch was checked)
        (void)hipGetLastError();
        b->gsteps = maxc;
        b->ggen = b->gen;
    }
    if (replay && b->gexec) {
        // the graph runs every part on the first stream: order it after the other streams' work,
        // and their later work (copies) after it
        for (int p = 1; p < b->nused; p++) {
            BATCH_OK(b, hipEventRecord(b->evjoin[p], b->ps[p]));
            BATCH_OK(b, hipStreamWaitEvent(b->stream, b->evjoin[p], 0));
        }
        BATCH_OK(b, hipGraphLaunch(b->gexec, b->stream));
        BATCH_OK(b, hipEventRecord(b->evfork, b->stream));
        for (int p = 1; p < b->nused; p++) BATCH_OK(b, hipStreamWaitEvent(b->ps[p], b->evfork, 0));
    } else {
        BATCH_OK(b, launch_all());
    }
    if (b->timing) {
        // per-pass and per-kernel totals of this call; algorithmic bytes per test and launch are
        // 16 L for the row pass (residue read + written) and 8 L for each column pass
        if (!batch_sync_all(b)) return -1;
        for (int p = 0; p < 3; p++) {
            b->ms[p] = 0;
            b->launches[p] = (maxc + b->timing - 1) / b->timing;
        }
        for (auto &ks : b->kstats) ks = kite_batch::KStat{ks.name, ks.pass};
        for (size_t i = 0; i < b->ev_stat.size(); i++) {
            float ms = 0;
            hipEventElapsedTime(&ms, b->ev[2 * i], b->ev[2 * i + 1]);
            kite_batch::KStat &ks = b->kstats[b->ev_stat[i]];
            ks.ms += ms;
            ks.launches++;
            ks.bytes += b->ev_bytes[i];
            b->ms[ks.pass] += ms;
        }
    }
    return 0;
}

extern "C" int kite_batch_mul(kite_batch *b, const int32_t *dst, const int32_t *src) {
    if (b->err != hipSuccess) return -1;
    std::vector<uint32_t> one(b->slots.size(), 1);
    bool any = false;
    for (size_t s = 0; s < b->slots.size(); s++) any |= b->slots[s].ctx && dst[s] >= 0;
    if (!any) return 0;
    if (batch_upload(b, dst, src, one.data())) return -1;
    BATCH_OK(b, batch_launch(b, 0, kModeRowFft, 0));
    BATCH_OK(b, batch_launch(b, 0, kModeMul, 0));
    BATCH_OK(b, batch_launch(b, 1, kModeMul, 0));
    BATCH_OK(b, batch_launch(b, 2, kModeMul, 0));
    return 0;
}

extern "C" int kite_batch_copy(kite_batch *b, const int32_t *dst, const int32_t *src) {
    if (batch_refresh(b)) return -1;
    for (size_t s = 0; s < b->slots.size(); s++) {
        BatchSlot &bs = b->slots[s];
        if (!bs.ctx || dst[s] < 0 || dst[s] == src[s]) continue;
        // ordered with the slot's launches
        BATCH_OK(b, hipMemcpyAsync(bs.nums[dst[s]]->dev, bs.nums[src[s]]->dev, (size_t)bs.ctx->eng->H * sizeof(double2),
                                   hipMemcpyDeviceToDevice, b->ps[b->slot_part[s]]));
    }
    return 0;
}

extern "C" int kite_batch_equal(kite_batch *b, const int32_t *a, const int32_t *c, int32_t *eq) {
    if (b->err != hipSuccess) return -1;
    const size_t cap = b->slots.size();
    std::vector<uint32_t> one(cap, 0);
    bool any = false;
    for (size_t s = 0; s < cap; s++) {
        one[s] = b->slots[s].ctx && a[s] >= 0 && c[s] >= 0;
        any |= one[s] != 0;
        eq[s] = -1;
    }
    if (!any) return 0;
    // the digits of a[s] into the operand-spectrum buffer, then those of c[s] into the digit buffer
    // (digit mode of the inverse column pass: the stored form's digits, each rounded to an integer)
    if (batch_upload(b, a, nullptr, one.data(), true)) return -1;
    BATCH_OK(b, batch_launch(b, 1, kModeDigits, 0));
    if (batch_upload(b, c, nullptr, one.data())) return -1;
    BATCH_OK(b, batch_launch(b, 1, kModeDigits, 0));
    if (!batch_sync_all(b)) return -1;
    hipLaunchKernelGGL(k_digits_equal, dim3((unsigned)cap), dim3(256), 0, b->stream, b->d_slots, (int)cap, b->d_eq);
    BATCH_OK(b, hipGetLastError());
    BATCH_OK(b, hipMemcpyAsync(b->h_eq, b->d_eq, cap * sizeof(int), hipMemcpyDeviceToHost, b->stream));
    BATCH_OK(b, hipStreamSynchronize(b->stream));
    for (size_t s = 0; s < cap; s++) eq[s] = b->h_eq[s];
    return 0;
}

extern "C" int kite_batch_maxerr(kite_batch *b, double *out, int reset) {
    if (batch_refresh(b)) return -1;
    const size_t cap = b->slots.size();
    if (!cap) return 0;
    std::vector<unsigned long long> v(cap);
    if (!batch_sync_all(b)) return -1;
    BATCH_OK(b, hipMemcpyAsync(v.data(), b->d_maxerr, cap * 8, hipMemcpyDeviceToHost, b->stream));
    BATCH_OK(b, hipStreamSynchronize(b->stream));
    for (size_t s = 0; s < cap; s++) memcpy(&out[s], &v[s], 8);
    if (reset) {
        hipLaunchKernelGGL(k_set_u64, dim3((unsigned)((cap + 63) / 64)), dim3(64), 0, b->stream, b->d_maxerr, 0ull,
                           (int)cap);
        BATCH_OK(b, hipGetLastError());
        BATCH_OK(b, hipStreamSynchronize(b->stream));
    }
    return 0;
}

extern "C" int kite_batch_set_xcds(kite_batch *b, int xcds) {
    xcds = std::max(1, std::min(xcds, 64));
    if (xcds != b->xcds) {
        b->xcds = xcds;
        b->dirty = true;
    }
    return xcds;
}

extern "C" int kite_batch_set_shape_parts(kite_batch *b, int on) {
    on = on != 0;
    if (on != b->shape_parts) {
        b->shape_parts = on;
        b->dirty = true;
    }
    return on;
}

extern "C" int kite_batch_set_persistent(kite_batch *b, int on) {
    on = on != 0;
    if (on != b->persist) {
        b->persist = on;
        b->dirty = true;  // the captured graph holds the other launches
    }
    return on;
}

// Graph replay runs on the pooled streams, direct launches on streams of their own (the struct's
// stream sets): the streams in use drain before the other set takes over.
extern "C" int kite_batch_set_graphs(kite_batch *b, int on) {
    const bool g = on != 0;
    if (g == b->graphs) return 0;
    if (!batch_sync_all(b)) return -1;
    if (!g && !b->own[0]) {
        int dev = 0, ncu = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0) {
            const std::vector<uint32_t> mask((size_t)(ncu + 31) / 32, 0xFFFFFFFFu);
            bool made = true;
            for (int p = 0; p < kite_batch::kMaxParts && made; p++)
                made = hipExtStreamCreateWithCUMask(&b->own[p], (uint32_t)mask.size(), mask.data()) == hipSuccess;
            if (!made) {  // all or none: the pooled set stays in use (every earlier call was checked)
                for (int p = 0; p < kite_batch::kMaxParts; p++) {
                    if (b->own[p]) hipStreamDestroy(b->own[p]);
                    b->own[p] = nullptr;
                }
                (void)hipGetLastError();
            }
        }
    }
    const hipStream_t *set = (g || !b->own[0]) ? b->pooled : b->own;
    for (int p = 0; p < kite_batch::kMaxParts; p++) b->ps[p] = set[p];
    b->stream = b->ps[0];
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
    b->graphs = g;
    return 0;
}

extern "C" int kite_batch_set_xadd(kite_batch *b, int slot, int32_t xadd) {
    if (slot < 0 || (size_t)slot >= b->slots.size() || !b->slots[slot].ctx) return -1;
    b->slots[slot].xadd = xadd;
    return 0;
}

extern "C" int kite_batch_set_timing(kite_batch *b, int every) {
    b->timing = every > 0 ? every : 0;
    return 0;
}

extern "C" int kite_batch_pass_times(kite_batch *b, double *out3, uint64_t *launches3) {
    for (int p = 0; p < 3; p++) {
        out3[p] = b->ms[p];
        if (launches3) launches3[p] = b->launches[p];
    }
    return 0;
}

extern "C" int kite_batch_kernel_times(kite_batch *b, int i, char *name, size_t name_len, double *ms,
                                       uint64_t *launches, double *bytes) {
    if (i < 0) return (int)b->kstats.size();
    if ((size_t)i >= b->kstats.size()) return -1;
    const kite_batch::KStat &ks = b->kstats[i];
    if (name && name_len) {
        strncpy(name, ks.name.c_str(), name_len - 1);
        name[name_len - 1] = 0;
    }
    if (ms) *ms = ks.ms;
    if (launches) *launches = ks.launches;
    if (bytes) *bytes = ks.bytes;
    return (int)b->kstats.size();
}

// Testing hook: the n-th HIP call checked from now on (by any context or batch) reports a failure
// (n <= 0 turns it off). No device operation fails; the library only takes the failure path.
extern "C" void kite_debug_fail_hip(long n) { g_fail_countdown.store(n > 0 ? n : 0); }
