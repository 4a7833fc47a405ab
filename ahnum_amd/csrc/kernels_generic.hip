// Generic pass kernels: every size is read from the shape at run time. They run any 7-smooth
// length the host can shape; the compile-time kernels of kernels_fast.hip cover the lengths the
// length chooser prefers.
#include "fft_dev.h"
#include "kernels.h"

namespace kite {

__global__ void __launch_bounds__(kNT) k_rows_g(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs, int mode,
                                              int step) {
    rows_body<0>(slots, jobs, mode, step);
}
__global__ void __launch_bounds__(kNT) k_cols_inv_g(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs, int mode,
                                                  int step) {
    cols_inv_body<0>(slots, jobs, mode, step);
}
__global__ void __launch_bounds__(kNT) k_cols_fwd_g(const SlotDev *__restrict__ slots, const int2 *__restrict__ jobs, int mode,
                                                  int step) {
    cols_fwd_body<0>(slots, jobs, mode, step);
}

PassKernel generic_rows() { return (PassKernel)k_rows_g; }
PassKernel generic_cols_inv() { return (PassKernel)k_cols_inv_g; }
PassKernel generic_cols_fwd() { return (PassKernel)k_cols_fwd_g; }

}  // namespace kite
