// Compile-time-size row kernels (row lengths C = 64..2048) and the registry of every compile-time
// pass kernel; the column kernels live in kernels_cols_{1,2,3,4}.hip (kernels_fast.h lists them).
#include "kernels_fast.h"

namespace kite {

__global__ void __launch_bounds__(128, 3) k_rows_sq256(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs,
                                                       int mode, int step) {
    (void)mode;
    rows256_sq(slots, jobs ? jobs[blockIdx.x] : make_int2(0, (int)blockIdx.x), step);
}

// Persistent launch of the same (kernels_fast.h: persist_loop).
__global__ void __launch_bounds__(128, 3) k_rows_sq256_p(const SlotDev *__restrict__ slots,
                                                         const int2 *__restrict__ jobs, int step,
                                                         const unsigned int *lists, int nl) {
    persist_loop(jobs, lists, nl, [&](int2 jb) { rows256_sq<true>(slots, jb, step); });
}

namespace {
PassKernel col(int kind, int R) {
    if (PassKernel k = cols_slice_1(kind, R)) return k;
    if (PassKernel k = cols_slice_2(kind, R)) return k;
    if (PassKernel k = cols_slice_3(kind, R)) return k;
    return cols_slice_4(kind, R);
}
}  // namespace

size_t fast_cols_fused_lds(int R) {
    switch (R) {
#define KITE_FUSED_LDS(r, next) case r: return FusedGeom<r>::lds_bytes();
        KITE_COLS_ALL(KITE_FUSED_LDS)
#undef KITE_FUSED_LDS
    default: return 0;
    }
}

size_t fast_cols_lds(int R) {
    switch (R) {
#define KITE_COLS_LDS(r, next) case r: return ColGeom<r>::lds_bytes();
        KITE_COLS_ALL(KITE_COLS_LDS)
#undef KITE_COLS_LDS
    default: return 0;
    }
}

int fast_cols_fused_threads(int R) {
    switch (R) {
#define KITE_FUSED_NT(r, next) case r: return fused_nt<r>();
        KITE_COLS_ALL(KITE_FUSED_NT)
#undef KITE_FUSED_NT
    default: return kNT;
    }
}

int fast_cols_pair_fused_threads(int R1) {
    switch (R1) {
#define KITE_PAIR_NT(r, next) case r: return next ? fused_nt2<r, next ? next : r>() : fused_nt<r>();
        KITE_COLS_ALL(KITE_PAIR_NT)
#undef KITE_PAIR_NT
    default: return kNT;
    }
}

int fast_cols_next(int R) {
    switch (R) {
#define KITE_COLS_NEXT(r, next) case r: return next;
        KITE_COLS_ALL(KITE_COLS_NEXT)
#undef KITE_COLS_NEXT
    default: return 0;
    }
}

PassKernel fast_rows(int C) {
    switch (C) {
    case 64: return (PassKernel)k_rows_c<64>;
    case 128: return (PassKernel)k_rows_c<128>;
    case 256: return (PassKernel)k_rows_c<256>;
    case 512: return (PassKernel)k_rows_c<512>;
    case 1024: return (PassKernel)k_rows_c<1024>;
    case 2048: return (PassKernel)k_rows_c<2048>;
    default: return nullptr;
    }
}

PassKernel fast_rows1(int C) {
    switch (C) {
    case 64: return (PassKernel)k_rows_c1<64>;
    case 128: return (PassKernel)k_rows_c1<128>;
    case 256: return (PassKernel)k_rows_c1<256>;
    case 512: return (PassKernel)k_rows_c1<512>;
    default: return nullptr;  // longer rows already take one pair per work-group
    }
}

PassKernel fast_rows_sq(int C) { return C == kRowsRegC ? (PassKernel)k_rows_sq256 : nullptr; }
PersistKernel fast_rows_sq_p(int C) { return C == kRowsRegC ? (PersistKernel)k_rows_sq256_p : nullptr; }

PassKernel fast_cols_inv(int R) { return col(kColInv, R); }
PassKernel fast_cols_fwd(int R) { return col(kColFwd, R); }
PassKernel fast_cols_fused(int R) { return col(kColFused, R); }
PassKernel fast_cols_pair_inv(int R1) { return fast_cols_next(R1) ? col(kColPairInv, R1) : nullptr; }
PassKernel fast_cols_pair_fwd(int R1) { return fast_cols_next(R1) ? col(kColPairFwd, R1) : nullptr; }
PassKernel fast_cols_pair_fused(int R1) { return fast_cols_next(R1) ? col(kColPairFused, R1) : nullptr; }
PersistKernel fast_cols_fused_p(int R) { return col(kColFusedP, R); }
PersistKernel fast_cols_pair_fused_p(int R1) { return fast_cols_next(R1) ? col(kColPairFusedP, R1) : nullptr; }

void all_kernels(PassKernel *out, int *count, int cap) {
    int n = 0;
    PassKernel g[3] = {generic_rows(), generic_cols_inv(), generic_cols_fwd()};
    for (int i = 0; i < 3 && n < cap; i++) out[n++] = g[i];
    const int Cs[] = {64, 128, 256, 512, 1024, 2048};
    for (int c : Cs) {
        if (n < cap) out[n++] = fast_rows(c);
        if (fast_rows1(c) && n < cap) out[n++] = fast_rows1(c);
    }
#define KITE_COLS_ALL_KERNELS(r, next)                                         \
    for (int kind = kColInv; kind <= kColPairFusedP; kind++)                   \
        if (PassKernel k = ((kind == kColPairInv || kind == kColPairFwd || kind == kColPairFused || \
                             kind == kColPairFusedP) && !next) ? nullptr : col(kind, r)) \
            if (n < cap) out[n++] = k;
    KITE_COLS_ALL(KITE_COLS_ALL_KERNELS)
#undef KITE_COLS_ALL_KERNELS
    *count = n;
}

}  // namespace kite
