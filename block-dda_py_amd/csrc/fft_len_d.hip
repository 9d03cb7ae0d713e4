// Register line-FFT kernels for the lengths of group D (BDDA_REG2_LENGTHS_D).
#include "fft_impl.h"

namespace bdda {
BDDA_REG2_LENGTHS_D(BDDA_REG2_INSTANTIATE)
}  // namespace bdda
