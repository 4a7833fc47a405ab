// Registry of the pass kernels (host side): generic kernels for any shape, compile-time-size
// kernels for the preferred row lengths C and column lengths R.
#ifndef KITE_KERNELS_H
#define KITE_KERNELS_H

#include <stddef.h>

#include "engine_dev.h"

namespace kite {

typedef const void *PassKernel;  // a __global__ function (const SlotDev *slots, const int2 *jobs, int mode, int step)
// a persistent square-mode pass: (const SlotDev *slots, const int2 *jobs, int step, const unsigned
// int *lists, int nl); kernels_fast.h describes the job lists and which work-group runs which job
typedef const void *PersistKernel;
constexpr int kPersistListWords = 16;
constexpr int kLsOff = 0, kLsCnt = 8;  // words of a launch's job lists: first job, count (8 lists)

PassKernel generic_rows();
PassKernel generic_cols_inv();
PassKernel generic_cols_fwd();
// nullptr when no compile-time kernel exists for that size
PassKernel fast_rows(int C);
// the same with one row pair per work-group (C < 1024; nullptr otherwise)
PassKernel fast_rows1(int C);
// square-mode row kernel with the rows held in registers (128 threads per work-group, four row
// pairs like the LDS row kernel, static LDS only); nullptr for row lengths without one
PassKernel fast_rows_sq(int C);
PassKernel fast_cols_inv(int R);
PassKernel fast_cols_fwd(int R);
size_t fast_cols_lds(int R);  // LDS bytes of the compile-time column kernels
// the column length after R with compile-time kernels (0: none), and the kernels that run
// column lengths R and fast_cols_next(R) in one launch
int fast_cols_next(int R);
PassKernel fast_cols_pair_inv(int R);
PassKernel fast_cols_pair_fwd(int R);
// square mode: both column passes in one kernel, carries between work-groups through flags
PassKernel fast_cols_fused(int R);
size_t fast_cols_fused_lds(int R);  // LDS bytes of the fused column kernel
// threads per work-group of the fused column kernels of length R (and of the pair kernels whose
// first length is R)
int fast_cols_fused_threads(int R);
int fast_cols_pair_fused_threads(int R1);
PassKernel fast_cols_pair_fused(int R);
// persistent variants (nullptr: none): the register row kernel, the fused column kernels
PersistKernel fast_rows_sq_p(int C);
PersistKernel fast_cols_fused_p(int R);
PersistKernel fast_cols_pair_fused_p(int R);
// every kernel with dynamic LDS, for raising its LDS limit once
void all_kernels(PassKernel *out, int *count, int cap);

}  // namespace kite

#endif
