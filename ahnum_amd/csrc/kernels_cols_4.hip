// Column kernels of slice 4 of the compile-time column lengths (kernels_fast.h).
#include "kernels_fast.h"

namespace kite {

PassKernel cols_slice_4(int kind, int R) {
    switch (R) {
        KITE_COLS_SLICE_4(KITE_COLS_CASE)
    default: return nullptr;
    }
}

}  // namespace kite
