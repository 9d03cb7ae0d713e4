// Register line-FFT kernels for the lengths of group A (BDDA_REG2_LENGTHS_A).
#include "fft_impl.h"

namespace bdda {
BDDA_REG2_LENGTHS_A(BDDA_REG2_INSTANTIATE)
}  // namespace bdda
