// Register line-FFT kernels for the lengths of group C (BDDA_REG2_LENGTHS_C).
#include "fft_impl.h"

namespace bdda {
BDDA_REG2_LENGTHS_C(BDDA_REG2_INSTANTIATE)
}  // namespace bdda
