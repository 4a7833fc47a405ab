// Column kernels of slice 3 of the compile-time column lengths (kernels_fast.h).
#include "kernels_fast.h"

namespace kite {

PassKernel cols_slice_3(int kind, int R) {
    switch (R) {
        KITE_COLS_SLICE_3(KITE_COLS_CASE)
    default: return nullptr;
    }
}

}  // namespace kite
