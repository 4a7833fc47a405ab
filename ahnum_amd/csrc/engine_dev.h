// Structures shared by the gfx950 kernels and the host code that launches them.
#ifndef KITE_ENGINE_DEV_H
#define KITE_ENGINE_DEV_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kite {

constexpr int kNT = 256;            // threads per work-group
constexpr int kMaxStages = 16;
constexpr uint32_t kMaxC = 2048;    // row length (row pass keeps 2 rows x 2 buffers in LDS)
constexpr uint32_t kMaxR = 1024;    // column length
constexpr uint32_t kMaxColLds = 96 * 1024;
constexpr int kMaxLds = 160 * 1024; // gfx950 LDS per work-group
constexpr int kMaxTop = 64;         // top-region digits handled by the device fold
constexpr int kSub = 8;             // elements (16 digits) per carry sub-chunk
constexpr int kTwLdsMaxC = 1024;    // row twiddles are staged in LDS up to this row length
constexpr int kRowTwPad = 3;        // row-kernel LDS twiddle table: one padding element per 8
constexpr int kXcds = 8;            // MI355X: 8 XCDs; work-group i of a launch runs on XCD i mod 8
constexpr int kFusedMinBits = 6;    // fused column pass: digits of at least this many bits (carry < 2^47)
constexpr int kRowsSqThreads = 128; // threads per work-group of the register row kernel (C = 256)

// Radix plan of a length-n FFT: 8s first, then 4, 2, 3, 5, 7 (host make_plan uses the same order).
__host__ __device__ constexpr int ct_nstages(int n) {
    int c = 0;
    while (n % 8 == 0) { n /= 8; c++; }
    const int pr[5] = {4, 2, 3, 5, 7};
    for (int i = 0; i < 5; i++)
        while (n % pr[i] == 0) { n /= pr[i]; c++; }
    return c;
}
__host__ __device__ constexpr int ct_radix(int n, int idx) {
    int c = 0;
    while (n % 8 == 0) {
        if (c == idx) return 8;
        n /= 8;
        c++;
    }
    const int pr[5] = {4, 2, 3, 5, 7};
    for (int i = 0; i < 5; i++)
        while (n % pr[i] == 0) {
            if (c == idx) return pr[i];
            n /= pr[i];
            c++;
        }
    return 0;
}
// Column group width for column length R: about 2048 elements per work-group, 4..64 columns; for
// R 288..512, 8 columns (up to 4096 elements): 128-byte row segments, whole cache lines, where
// 4 columns wrote half lines (the record's fused pass at R = 400: 40.6 -> 32.4 us).
__host__ __device__ constexpr int g_for_R(int R) {
    int G = 64;
    while (G > 4 && G * R > (R > 256 ? 4096 : 2048)) G /= 2;
    return G;
}
// Row pairs per work-group for row length C in the compile-time row kernels (~2048 elements, so
// that every radix-8 stage keeps all threads busy).
__host__ __device__ constexpr int ppw_for_C(int C) { return C >= 1024 ? 1 : 1024 / C; }
// Threads per work-group of the LDS row kernels: 512 at C = 2048 (one row pair of 4096 elements:
// each radix-16 or radix-8 stage in one round of butterflies), else kNT.
__host__ __device__ constexpr int rows_nt(int C) { return C >= 2048 ? 512 : kNT; }
// LDS of the compile-time row kernels: 2 ppw rows of C with one padding element per 2^rows_pad(C)
// (per 16 where the FFT starts with radix 16, else per 8: the first stage's strided writes then
// spread over all banks), then, for C <= kTwLdsMaxC, the w_C table, for C >= 512 with one padding
// element per 2^kRowTwPad (the stages' twiddle reads jm r TS, with strides of whole bank rows,
// spread the same way: k_rows_c<1024> 61.7 -> 56.8 us; at C = 128 it measured 1.5% slower).
__host__ __device__ constexpr int rows_pad(int C) { return C >= 1024 ? 4 : 3; }
__host__ __device__ constexpr int rows_twpad(int C) { return C >= 512 && C <= (int)kTwLdsMaxC ? kRowTwPad : 0; }
__host__ __device__ constexpr size_t rows_lds_bytes(int C, int ppw) {
    return (size_t)16 * (2 * ppw * C + ((2 * ppw * C) >> rows_pad(C)) +
                         (C <= (int)kTwLdsMaxC ? C + (rows_twpad(C) ? C >> rows_twpad(C) : 0) : 0));
}

enum : int { kModeSquare = 0, kModeMul = 1, kModeDigits = 2, kModeRowFft = 3 };

struct Plan {  // radix sequence of one FFT length
    int n, nst;
    int rad[kMaxStages];
};

struct ShapeDev {  // one (R, C, G) shape: H = R*C complex, NG = C/G column groups
    int R, C, G, NG;
    int sshift;            // two-level w_H table split: t = hi*2^sshift + lo
    Plan pr, pc;           // column (length R) and row (length C) plans
    const double2 *twR;    // w_R^t, t < R
    const double2 *twC;    // w_C^t, t < C
    const double2 *tlo;    // w_H^t, t < 2^sshift
    const double2 *thi;    // w_H^(t 2^sshift)
    const double2 *twRows; // [k_r][k_c] = w_H^(k_r + R k_c): split twiddles in row order
    const double2 *twCols; // [g][k_r][l] = w_H^((g G + l) k_r): four-step twiddles in group order
    const double2 *tw16;   // C = 256: [j][t] = w_C^(t j), j, t < 16 (the 16 x 16 row FFT's twiddles)
};

struct SlotDev {  // one number being operated on (one test in a batch)
    // the shape's sizes and tables, copied here so that a kernel reads them with the slot (one
    // dependent load fewer than through shp)
    int R, C, G, NG;
    const double2 *twR, *twC, *twRows, *twCols, *tw16;
    const double2 *tlo, *thi;   // two-level w_H: w_H^t = thi[t >> sshift] tlo[t mod 2^sshift]
    int sshift;
    const ShapeDev *shp;        // the whole shape (plans, two-level w_H table) for the generic kernels
    double2 *x;                 // the residue, H complex (after forward column pass)
    const double2 *y_src;       // multiplication operand (stored form), for kModeRowFft
    double2 *yfreq;             // operand row spectra, H complex
    int32_t *dig;               // L digits, blocked [g][r][2G]
    int64_t *cout;              // carry out of each (g, r) chunk; the last holds the top fold
    const uint8_t *dcode;       // per element [g][r][l]: ceil and big-digit bits of both digits
    const double *ta;           // [R] 2^-fa(r), [R] 2^+fa(r)
    const double *tb;           // [2C] 2^-fb(b), [2C] 2^+fb(b)
    unsigned long long *maxerr; // max round-off (bits of a non-negative double)
    unsigned long long *flags;  // fused column pass: per (g, r) chunk, seq << 48 | carry (48 bits)
    uint32_t seq;               // fused column launches of this slot so far (step s uses seq + s + 1)
    int32_t xadd;               // square mode: x -> x^2 + xadd (V += xadd 2^n, added in the top fold)
    double kscale;              // k / (4H)
    double rscale;              // 1 / R
    uint64_t k;
    uint64_t kinv;              // floor((2^64 - 1) / k), for the fold's division by k
    uint32_t count;             // the slot takes part in launches with step < count
    int b0;                     // digits have b0 or b0+1 bits
    int top_i0, top_n, top_m, top_shift;  // fold: first top digit within the last chunk, count,
                                          // m = n - s_t, s_L - s_t
    uint64_t top_big;           // bit i: top digit i has b0 + 1 bits (else b0)
};

}  // namespace kite

#endif
