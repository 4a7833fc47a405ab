// Compile-time-size pass kernels (the bodies are the templates of fft_dev.h) and the list of column
// lengths that have them. The column kernels are instantiated by kernels_cols_{1,2,3,4}.hip, one
// slice of the list each, so that the slices compile in parallel; kernels_fast.hip holds the row
// kernels and the registry of kernels.h.
#ifndef KITE_KERNELS_FAST_H
#define KITE_KERNELS_FAST_H

#include "fft_dev.h"
#include "kernels.h"

// Column lengths R = 2^a m (m in 1, 3, 5, 7, 9, 15, 25, 27; 8 <= R <= 768, R divisible by 4) in
// ascending order, as X(R, next R): a batch whose tests straddle two adjacent lengths runs both in
// one launch of the pair kernel <R, next>.
#define KITE_COLS_SLICE_1(X) \
    X(8, 12) X(12, 16) X(16, 20) X(20, 24) X(24, 28) X(28, 32) X(32, 36) X(36, 40) X(40, 48) X(48, 56) X(56, 60) \
    X(60, 64) X(64, 72) X(72, 80)
#define KITE_COLS_SLICE_2(X) \
    X(80, 96) X(96, 100) X(100, 108) X(108, 112) X(112, 120) X(120, 128) X(128, 144) X(144, 160) X(160, 192)
#define KITE_COLS_SLICE_3(X) \
    X(192, 200) X(200, 216) X(216, 224) X(224, 240) X(240, 256) X(256, 288) X(288, 320) X(320, 384)
#define KITE_COLS_SLICE_4(X) \
    X(384, 400) X(400, 432) X(432, 448) X(448, 480) X(480, 512) X(512, 576) X(576, 640) X(640, 768) X(768, 0)
#define KITE_COLS_ALL(X) KITE_COLS_SLICE_1(X) KITE_COLS_SLICE_2(X) KITE_COLS_SLICE_3(X) KITE_COLS_SLICE_4(X)

namespace kite {

template <int C>
__global__ void __launch_bounds__(rows_nt(C)) k_rows_c(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs,
                                                     int mode, int step) {
    rows_fast<C>(slots, jobs, mode, step);
}
// One row pair per work-group: a single test's row pass on four times as many work-groups.
template <int C>
__global__ void __launch_bounds__(rows_nt(C)) k_rows_c1(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs,
                                               int mode, int step) {
    rows_fast<C, 1>(slots, jobs, mode, step);
}
template <int R>
__global__ void __launch_bounds__(kNT) k_cols_inv_r(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs,
                                                  int mode, int step) {
    cols_inv_fast<R>(slots, jobs, mode, step);
}
template <int R>
__global__ void __launch_bounds__(kNT) k_cols_fwd_r(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs,
                                                  int mode, int step) {
    cols_fwd_fast<R>(slots, jobs, mode, step);
}
// The fused kernels run fused_nt<R>() threads per work-group, the pair kernels the larger of their
// two lengths' counts (fewer would make the longer length iterate, which spilled registers);
// fast_cols_fused_threads and fast_cols_pair_fused_threads tell the host.
// Waves per SIMD the fused kernels are compiled for (__launch_bounds__): 4 (128 VGPRs); 2 where
// the groups are 8 columns of R >= 384 (a single record-size test's pass: one work-group per CU).
constexpr int fused_waves_eu(int R) { return g_for_R(R) * R > 2048 ? 2 : 4; }
template <int R1, int R2> constexpr int fused_nt2() { return fused_nt<R1>() > fused_nt<R2>() ? fused_nt<R1>() : fused_nt<R2>(); }
template <int R>
__global__ void __launch_bounds__(fused_nt<R>(), fused_waves_eu(R)) k_cols_fused_r(const SlotDev *__restrict__ slots,
                                                                 const int2 *__restrict__ jobs, int mode, int step) {
    (void)mode;
    const int2 jb = jobs ? jobs[blockIdx.x] : make_int2(0, (int)blockIdx.x);
    cols_fused_fast<R, false, fused_nt<R>()>(slots[jb.x], jb.y, step);
}

// Two adjacent column lengths in one launch: tests of both shapes advance together (each work-
// group runs the body of its test's R), so a batch straddling an FFT-length boundary fills the
// device with one launch per column pass instead of two.
template <int R1, int R2>
__global__ void __launch_bounds__(kNT) k_cols_inv_rr(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs,
                                                   int mode, int step) {
    const int2 jb = jobs ? jobs[blockIdx.x] : make_int2(0, (int)blockIdx.x);
    if (slots[jb.x].R == R1) cols_inv_fast<R1>(slots, jobs, mode, step);
    else cols_inv_fast<R2>(slots, jobs, mode, step);
}
template <int R1, int R2>
__global__ void __launch_bounds__(kNT) k_cols_fwd_rr(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs,
                                                   int mode, int step) {
    const int2 jb = jobs ? jobs[blockIdx.x] : make_int2(0, (int)blockIdx.x);
    if (slots[jb.x].R == R1) cols_fwd_fast<R1>(slots, jobs, mode, step);
    else cols_fwd_fast<R2>(slots, jobs, mode, step);
}
template <int R1, int R2>
__global__ void __launch_bounds__((fused_nt2<R1, R2>()), fused_waves_eu(R2)) k_cols_fused_rr(const SlotDev *__restrict__ slots,
                                                                   const int2 *__restrict__ jobs, int mode, int step) {
    (void)mode;
    const int2 jb = jobs ? jobs[blockIdx.x] : make_int2(0, (int)blockIdx.x);
    const SlotDev &s = slots[jb.x];
    if (s.R == R1) cols_fused_fast<R1, false, fused_nt2<R1, R2>()>(s, jb.y, step);
    else cols_fused_fast<R2, false, fused_nt2<R1, R2>()>(s, jb.y, step);
}

// ---- persistent launches ----------------------------------------------------------------------
// A persistent launch runs no more work-groups than the device holds at once, each running a fixed
// share of the jobs in turn, instead of one work-group per job: the dispatcher no longer has to
// refill the CUs as work-groups finish (with one work-group per job about a fifth of the resident
// slots stand empty between a work-group's end and the next one's start). Jobs are not handed out
// by atomic tickets: same-address device-scope atomics serialize (tens of ns each), which cost
// more than the empty slots.
//
// Job lists (kPersistListWords words per launch, written by the host): words 0..7 each list's
// first job, 8..15 its job count. The jobs are laid out list by list, each test's jobs consecutive
// within one list (the column jobs' XCD lists); the launch has nl * P work-groups and work-group
// w = x + nl k runs positions k, k + P, k + 2P, ... of list x.
//
// Column jobs wait for each other (the carry flags): the job at position q waits for position
// q - 1 or, for a test's column group 0, for its own test's last group at q + NG - 1; each publishes
// its carry before it waits. The job at q starts when its work-group's job at q - P has finished
// waiting, i.e. once q - P - 1 (or q - P + NG - 1) has started: with P >= NG (or P at least the
// list's length) that is an earlier position, so every job starts, provided every work-group of
// the launch becomes resident. The host keeps each launch within the resident capacity divided
// among the concurrently running parts; row jobs never wait.
template <class BODY>
__device__ __forceinline__ void persist_loop(const int2 *__restrict__ jobs, const unsigned int *lists, int nl,
                                             BODY body) {
    const int list = (int)blockIdx.x % nl, k = (int)blockIdx.x / nl, P = (int)gridDim.x / nl;
    const int off = (int)ldg(lists + kLsOff + list), cnt = (int)ldg(lists + kLsCnt + list);
    for (int t = k; t < cnt; t += P) {
        body(jobs[off + t]);
        __syncthreads();  // this job's last LDS reads before the next job's writes
    }
}

template <int R>
__global__ void __launch_bounds__(fused_nt<R>(), fused_waves_eu(R)) k_cols_fused_r_p(const SlotDev *__restrict__ slots,
                                                                   const int2 *__restrict__ jobs, int step,
                                                                   const unsigned int *lists, int nl) {
    persist_loop(jobs, lists, nl,
                 [&](int2 jb) { cols_fused_fast<R, true, fused_nt<R>()>(slots[jb.x], jb.y, step); });
}
template <int R1, int R2>
__global__ void __launch_bounds__((fused_nt2<R1, R2>()), fused_waves_eu(R2)) k_cols_fused_rr_p(const SlotDev *__restrict__ slots,
                                                                     const int2 *__restrict__ jobs, int step,
                                                                     const unsigned int *lists, int nl) {
    persist_loop(jobs, lists, nl, [&](int2 jb) {
        const SlotDev &s = slots[jb.x];
        if (s.R == R1) cols_fused_fast<R1, true, fused_nt2<R1, R2>()>(s, jb.y, step);
        else cols_fused_fast<R2, true, fused_nt2<R1, R2>()>(s, jb.y, step);
    });
}

// The column kernels of one slice: kind is one of ColKind, R the (first) column length; nullptr if
// the slice has no such kernel.
enum ColKind : int { kColInv, kColFwd, kColFused, kColPairInv, kColPairFwd, kColPairFused, kColFusedP, kColPairFusedP };
PassKernel cols_slice_1(int kind, int R);
PassKernel cols_slice_2(int kind, int R);
PassKernel cols_slice_3(int kind, int R);
PassKernel cols_slice_4(int kind, int R);

// A slice's lookup for one column length: the single-length kernels, and the pair kernels
// <R, NEXT> when NEXT is nonzero.
template <int R, int NEXT> PassKernel col_kernel(int kind) {
    switch (kind) {
    case kColInv: return (PassKernel)k_cols_inv_r<R>;
    case kColFwd: return (PassKernel)k_cols_fwd_r<R>;
    case kColFused: return (PassKernel)k_cols_fused_r<R>;
    case kColFusedP: return (PassKernel)k_cols_fused_r_p<R>;
    default: break;
    }
    if constexpr (NEXT != 0) {
        switch (kind) {
        case kColPairInv: return (PassKernel)k_cols_inv_rr<R, NEXT>;
        case kColPairFwd: return (PassKernel)k_cols_fwd_rr<R, NEXT>;
        case kColPairFused: return (PassKernel)k_cols_fused_rr<R, NEXT>;
        case kColPairFusedP: return (PassKernel)k_cols_fused_rr_p<R, NEXT>;
        default: break;
        }
    }
    return nullptr;
}
#define KITE_COLS_CASE(R, NEXT) \
    case R: return col_kernel<R, NEXT>(kind);

}  // namespace kite

#endif
