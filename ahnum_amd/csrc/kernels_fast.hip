// Compile-time-size pass kernels: row kernels for row lengths C = 64..2048 and column kernels for
// column lengths R = 2^b m (m in 1, 3, 5, 7; 8 <= R <= 768). Generated list; the bodies are the
// templates of fft_dev.h.
#include "fft_dev.h"
#include "kernels.h"

namespace kite {

template <int C>
__global__ void __launch_bounds__(kNT) k_rows_c(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs, int mode, int step) {
    rows_fast<C>(slots, jobs, mode, step);
}
// One row pair per work-group: a single test's row pass on four times as many work-groups.
template <int C>
__global__ void __launch_bounds__(kNT) k_rows_c1(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs, int mode, int step) {
    rows_fast<C, 1>(slots, jobs, mode, step);
}
__global__ void __launch_bounds__(128, 3) k_rows_sq256(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs, int mode, int step) {
    (void)mode;
    rows256_sq(slots, jobs, step);
}
template <int R>
__global__ void __launch_bounds__(kNT) k_cols_inv_r(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs, int mode, int step) {
    cols_inv_fast<R>(slots, jobs, mode, step);
}
template <int R>
__global__ void __launch_bounds__(kNT) k_cols_fwd_r(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs, int mode, int step) {
    cols_fwd_fast<R>(slots, jobs, mode, step);
}

template <int R>
__global__ void __launch_bounds__(kNT, 4) k_cols_fused_r(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs, int mode, int step) {
    (void)mode;
    const int2 jb = jobs ? jobs[blockIdx.x] : make_int2(0, (int)blockIdx.x);
    cols_fused_fast<R>(slots[jb.x], jb.y, step);
}

// Two adjacent column lengths in one launch: tests of both shapes advance together (each work-
// group runs the body of its test's R), so a batch straddling an FFT-length boundary fills the
// device with one launch per column pass instead of two.
template <int R1, int R2>
__global__ void __launch_bounds__(kNT) k_cols_inv_rr(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs, int mode, int step) {
    const int2 jb = jobs ? jobs[blockIdx.x] : make_int2(0, (int)blockIdx.x);
    if (slots[jb.x].R == R1) cols_inv_fast<R1>(slots, jobs, mode, step);
    else cols_inv_fast<R2>(slots, jobs, mode, step);
}
template <int R1, int R2>
__global__ void __launch_bounds__(kNT) k_cols_fwd_rr(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs, int mode, int step) {
    const int2 jb = jobs ? jobs[blockIdx.x] : make_int2(0, (int)blockIdx.x);
    if (slots[jb.x].R == R1) cols_fwd_fast<R1>(slots, jobs, mode, step);
    else cols_fwd_fast<R2>(slots, jobs, mode, step);
}

template <int R1, int R2>
__global__ void __launch_bounds__(kNT, 4) k_cols_fused_rr(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs, int mode, int step) {
    (void)mode;
    const int2 jb = jobs ? jobs[blockIdx.x] : make_int2(0, (int)blockIdx.x);
    const SlotDev &s = slots[jb.x];
    if (s.R == R1) cols_fused_fast<R1>(s, jb.y, step);
    else cols_fused_fast<R2>(s, jb.y, step);
}

size_t fast_cols_fused_lds(int R) {
    switch (R) {
#define KITE_FUSED_LDS(r) case r: return FusedGeom<r>::lds_bytes();
    KITE_FUSED_LDS(8) KITE_FUSED_LDS(12) KITE_FUSED_LDS(16) KITE_FUSED_LDS(20) KITE_FUSED_LDS(24)
    KITE_FUSED_LDS(28) KITE_FUSED_LDS(32) KITE_FUSED_LDS(40) KITE_FUSED_LDS(48) KITE_FUSED_LDS(56)
    KITE_FUSED_LDS(64) KITE_FUSED_LDS(80) KITE_FUSED_LDS(96) KITE_FUSED_LDS(112) KITE_FUSED_LDS(128)
    KITE_FUSED_LDS(160) KITE_FUSED_LDS(192) KITE_FUSED_LDS(224) KITE_FUSED_LDS(256) KITE_FUSED_LDS(320)
    KITE_FUSED_LDS(384) KITE_FUSED_LDS(448) KITE_FUSED_LDS(512) KITE_FUSED_LDS(640) KITE_FUSED_LDS(768)
#undef KITE_FUSED_LDS
    default: return 0;
    }
}

size_t fast_cols_lds(int R) {
    switch (R) {
#define KITE_COLS_LDS(r) case r: return ColGeom<r>::lds_bytes();
    KITE_COLS_LDS(8)
    KITE_COLS_LDS(12)
    KITE_COLS_LDS(16)
    KITE_COLS_LDS(20)
    KITE_COLS_LDS(24)
    KITE_COLS_LDS(28)
    KITE_COLS_LDS(32)
    KITE_COLS_LDS(40)
    KITE_COLS_LDS(48)
    KITE_COLS_LDS(56)
    KITE_COLS_LDS(64)
    KITE_COLS_LDS(80)
    KITE_COLS_LDS(96)
    KITE_COLS_LDS(112)
    KITE_COLS_LDS(128)
    KITE_COLS_LDS(160)
    KITE_COLS_LDS(192)
    KITE_COLS_LDS(224)
    KITE_COLS_LDS(256)
    KITE_COLS_LDS(320)
    KITE_COLS_LDS(384)
    KITE_COLS_LDS(448)
    KITE_COLS_LDS(512)
    KITE_COLS_LDS(640)
    KITE_COLS_LDS(768)
#undef KITE_COLS_LDS
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

PassKernel fast_cols_inv(int R) {
    switch (R) {
    case 8: return (PassKernel)k_cols_inv_r<8>;
    case 12: return (PassKernel)k_cols_inv_r<12>;
    case 16: return (PassKernel)k_cols_inv_r<16>;
    case 20: return (PassKernel)k_cols_inv_r<20>;
    case 24: return (PassKernel)k_cols_inv_r<24>;
    case 28: return (PassKernel)k_cols_inv_r<28>;
    case 32: return (PassKernel)k_cols_inv_r<32>;
    case 40: return (PassKernel)k_cols_inv_r<40>;
    case 48: return (PassKernel)k_cols_inv_r<48>;
    case 56: return (PassKernel)k_cols_inv_r<56>;
    case 64: return (PassKernel)k_cols_inv_r<64>;
    case 80: return (PassKernel)k_cols_inv_r<80>;
    case 96: return (PassKernel)k_cols_inv_r<96>;
    case 112: return (PassKernel)k_cols_inv_r<112>;
    case 128: return (PassKernel)k_cols_inv_r<128>;
    case 160: return (PassKernel)k_cols_inv_r<160>;
    case 192: return (PassKernel)k_cols_inv_r<192>;
    case 224: return (PassKernel)k_cols_inv_r<224>;
    case 256: return (PassKernel)k_cols_inv_r<256>;
    case 320: return (PassKernel)k_cols_inv_r<320>;
    case 384: return (PassKernel)k_cols_inv_r<384>;
    case 448: return (PassKernel)k_cols_inv_r<448>;
    case 512: return (PassKernel)k_cols_inv_r<512>;
    case 640: return (PassKernel)k_cols_inv_r<640>;
    case 768: return (PassKernel)k_cols_inv_r<768>;
    default: return nullptr;
    }
}

PassKernel fast_cols_fwd(int R) {
    switch (R) {
    case 8: return (PassKernel)k_cols_fwd_r<8>;
    case 12: return (PassKernel)k_cols_fwd_r<12>;
    case 16: return (PassKernel)k_cols_fwd_r<16>;
    case 20: return (PassKernel)k_cols_fwd_r<20>;
    case 24: return (PassKernel)k_cols_fwd_r<24>;
    case 28: return (PassKernel)k_cols_fwd_r<28>;
    case 32: return (PassKernel)k_cols_fwd_r<32>;
    case 40: return (PassKernel)k_cols_fwd_r<40>;
    case 48: return (PassKernel)k_cols_fwd_r<48>;
    case 56: return (PassKernel)k_cols_fwd_r<56>;
    case 64: return (PassKernel)k_cols_fwd_r<64>;
    case 80: return (PassKernel)k_cols_fwd_r<80>;
    case 96: return (PassKernel)k_cols_fwd_r<96>;
    case 112: return (PassKernel)k_cols_fwd_r<112>;
    case 128: return (PassKernel)k_cols_fwd_r<128>;
    case 160: return (PassKernel)k_cols_fwd_r<160>;
    case 192: return (PassKernel)k_cols_fwd_r<192>;
    case 224: return (PassKernel)k_cols_fwd_r<224>;
    case 256: return (PassKernel)k_cols_fwd_r<256>;
    case 320: return (PassKernel)k_cols_fwd_r<320>;
    case 384: return (PassKernel)k_cols_fwd_r<384>;
    case 448: return (PassKernel)k_cols_fwd_r<448>;
    case 512: return (PassKernel)k_cols_fwd_r<512>;
    case 640: return (PassKernel)k_cols_fwd_r<640>;
    case 768: return (PassKernel)k_cols_fwd_r<768>;
    default: return nullptr;
    }
}

int fast_cols_next(int R) {
    switch (R) {
    case 8: return 12;
    case 12: return 16;
    case 16: return 20;
    case 20: return 24;
    case 24: return 28;
    case 28: return 32;
    case 32: return 40;
    case 40: return 48;
    case 48: return 56;
    case 56: return 64;
    case 64: return 80;
    case 80: return 96;
    case 96: return 112;
    case 112: return 128;
    case 128: return 160;
    case 160: return 192;
    case 192: return 224;
    case 224: return 256;
    case 256: return 320;
    case 320: return 384;
    case 384: return 448;
    case 448: return 512;
    case 512: return 640;
    case 640: return 768;
    default: return 0;
    }
}

PassKernel fast_cols_pair_inv(int R1) {
    switch (R1) {
    case 8: return (PassKernel)k_cols_inv_rr<8, 12>;
    case 12: return (PassKernel)k_cols_inv_rr<12, 16>;
    case 16: return (PassKernel)k_cols_inv_rr<16, 20>;
    case 20: return (PassKernel)k_cols_inv_rr<20, 24>;
    case 24: return (PassKernel)k_cols_inv_rr<24, 28>;
    case 28: return (PassKernel)k_cols_inv_rr<28, 32>;
    case 32: return (PassKernel)k_cols_inv_rr<32, 40>;
    case 40: return (PassKernel)k_cols_inv_rr<40, 48>;
    case 48: return (PassKernel)k_cols_inv_rr<48, 56>;
    case 56: return (PassKernel)k_cols_inv_rr<56, 64>;
    case 64: return (PassKernel)k_cols_inv_rr<64, 80>;
    case 80: return (PassKernel)k_cols_inv_rr<80, 96>;
    case 96: return (PassKernel)k_cols_inv_rr<96, 112>;
    case 112: return (PassKernel)k_cols_inv_rr<112, 128>;
    case 128: return (PassKernel)k_cols_inv_rr<128, 160>;
    case 160: return (PassKernel)k_cols_inv_rr<160, 192>;
    case 192: return (PassKernel)k_cols_inv_rr<192, 224>;
    case 224: return (PassKernel)k_cols_inv_rr<224, 256>;
    case 256: return (PassKernel)k_cols_inv_rr<256, 320>;
    case 320: return (PassKernel)k_cols_inv_rr<320, 384>;
    case 384: return (PassKernel)k_cols_inv_rr<384, 448>;
    case 448: return (PassKernel)k_cols_inv_rr<448, 512>;
    case 512: return (PassKernel)k_cols_inv_rr<512, 640>;
    case 640: return (PassKernel)k_cols_inv_rr<640, 768>;
    default: return nullptr;
    }
}

PassKernel fast_cols_pair_fwd(int R1) {
    switch (R1) {
    case 8: return (PassKernel)k_cols_fwd_rr<8, 12>;
    case 12: return (PassKernel)k_cols_fwd_rr<12, 16>;
    case 16: return (PassKernel)k_cols_fwd_rr<16, 20>;
    case 20: return (PassKernel)k_cols_fwd_rr<20, 24>;
    case 24: return (PassKernel)k_cols_fwd_rr<24, 28>;
    case 28: return (PassKernel)k_cols_fwd_rr<28, 32>;
    case 32: return (PassKernel)k_cols_fwd_rr<32, 40>;
    case 40: return (PassKernel)k_cols_fwd_rr<40, 48>;
    case 48: return (PassKernel)k_cols_fwd_rr<48, 56>;
    case 56: return (PassKernel)k_cols_fwd_rr<56, 64>;
    case 64: return (PassKernel)k_cols_fwd_rr<64, 80>;
    case 80: return (PassKernel)k_cols_fwd_rr<80, 96>;
    case 96: return (PassKernel)k_cols_fwd_rr<96, 112>;
    case 112: return (PassKernel)k_cols_fwd_rr<112, 128>;
    case 128: return (PassKernel)k_cols_fwd_rr<128, 160>;
    case 160: return (PassKernel)k_cols_fwd_rr<160, 192>;
    case 192: return (PassKernel)k_cols_fwd_rr<192, 224>;
    case 224: return (PassKernel)k_cols_fwd_rr<224, 256>;
    case 256: return (PassKernel)k_cols_fwd_rr<256, 320>;
    case 320: return (PassKernel)k_cols_fwd_rr<320, 384>;
    case 384: return (PassKernel)k_cols_fwd_rr<384, 448>;
    case 448: return (PassKernel)k_cols_fwd_rr<448, 512>;
    case 512: return (PassKernel)k_cols_fwd_rr<512, 640>;
    case 640: return (PassKernel)k_cols_fwd_rr<640, 768>;
    default: return nullptr;
    }
}

PassKernel fast_cols_fused(int R) {
    switch (R) {
    case 8: return (PassKernel)k_cols_fused_r<8>;
    case 12: return (PassKernel)k_cols_fused_r<12>;
    case 16: return (PassKernel)k_cols_fused_r<16>;
    case 20: return (PassKernel)k_cols_fused_r<20>;
    case 24: return (PassKernel)k_cols_fused_r<24>;
    case 28: return (PassKernel)k_cols_fused_r<28>;
    case 32: return (PassKernel)k_cols_fused_r<32>;
    case 40: return (PassKernel)k_cols_fused_r<40>;
    case 48: return (PassKernel)k_cols_fused_r<48>;
    case 56: return (PassKernel)k_cols_fused_r<56>;
    case 64: return (PassKernel)k_cols_fused_r<64>;
    case 80: return (PassKernel)k_cols_fused_r<80>;
    case 96: return (PassKernel)k_cols_fused_r<96>;
    case 112: return (PassKernel)k_cols_fused_r<112>;
    case 128: return (PassKernel)k_cols_fused_r<128>;
    case 160: return (PassKernel)k_cols_fused_r<160>;
    case 192: return (PassKernel)k_cols_fused_r<192>;
    case 224: return (PassKernel)k_cols_fused_r<224>;
    case 256: return (PassKernel)k_cols_fused_r<256>;
    case 320: return (PassKernel)k_cols_fused_r<320>;
    case 384: return (PassKernel)k_cols_fused_r<384>;
    case 448: return (PassKernel)k_cols_fused_r<448>;
    case 512: return (PassKernel)k_cols_fused_r<512>;
    case 640: return (PassKernel)k_cols_fused_r<640>;
    case 768: return (PassKernel)k_cols_fused_r<768>;
    default: return nullptr;
    }
}

PassKernel fast_cols_pair_fused(int R1) {
    switch (R1) {
    case 8: return (PassKernel)k_cols_fused_rr<8, 12>;
    case 12: return (PassKernel)k_cols_fused_rr<12, 16>;
    case 16: return (PassKernel)k_cols_fused_rr<16, 20>;
    case 20: return (PassKernel)k_cols_fused_rr<20, 24>;
    case 24: return (PassKernel)k_cols_fused_rr<24, 28>;
    case 28: return (PassKernel)k_cols_fused_rr<28, 32>;
    case 32: return (PassKernel)k_cols_fused_rr<32, 40>;
    case 40: return (PassKernel)k_cols_fused_rr<40, 48>;
    case 48: return (PassKernel)k_cols_fused_rr<48, 56>;
    case 56: return (PassKernel)k_cols_fused_rr<56, 64>;
    case 64: return (PassKernel)k_cols_fused_rr<64, 80>;
    case 80: return (PassKernel)k_cols_fused_rr<80, 96>;
    case 96: return (PassKernel)k_cols_fused_rr<96, 112>;
    case 112: return (PassKernel)k_cols_fused_rr<112, 128>;
    case 128: return (PassKernel)k_cols_fused_rr<128, 160>;
    case 160: return (PassKernel)k_cols_fused_rr<160, 192>;
    case 192: return (PassKernel)k_cols_fused_rr<192, 224>;
    case 224: return (PassKernel)k_cols_fused_rr<224, 256>;
    case 256: return (PassKernel)k_cols_fused_rr<256, 320>;
    case 320: return (PassKernel)k_cols_fused_rr<320, 384>;
    case 384: return (PassKernel)k_cols_fused_rr<384, 448>;
    case 448: return (PassKernel)k_cols_fused_rr<448, 512>;
    case 512: return (PassKernel)k_cols_fused_rr<512, 640>;
    case 640: return (PassKernel)k_cols_fused_rr<640, 768>;
    default: return nullptr;
    }
}

void all_kernels(PassKernel *out, int *count, int cap) {
    int n = 0;
    PassKernel g[3] = {generic_rows(), generic_cols_inv(), generic_cols_fwd()};
    for (int i = 0; i < 3 && n < cap; i++) out[n++] = g[i];
    const int Cs[] = {64, 128, 256, 512, 1024, 2048};
    const int Rs[] = {8, 12, 16, 20, 24, 28, 32, 40, 48, 56, 64, 80, 96, 112, 128, 160, 192, 224, 256, 320, 384, 448, 512, 640, 768};
    for (int c : Cs) {
        if (n < cap) out[n++] = fast_rows(c);
        if (fast_rows1(c) && n < cap) out[n++] = fast_rows1(c);
    }
    for (int r : Rs) {
        if (n < cap) out[n++] = fast_cols_inv(r);
        if (n < cap) out[n++] = fast_cols_fwd(r);
        if (n < cap) out[n++] = fast_cols_fused(r);
        if (fast_cols_next(r) && n + 2 < cap) {
            out[n++] = fast_cols_pair_inv(r);
            out[n++] = fast_cols_pair_fwd(r);
            out[n++] = fast_cols_pair_fused(r);
        }
    }
    *count = n;
}

}  // namespace kite
