// Register line-FFT kernels for the lengths of group B (BDDA_REG2_LENGTHS_B).
#include "fft_impl.h"

namespace bdda {
BDDA_REG2_LENGTHS_B(BDDA_REG2_INSTANTIATE)
}  // namespace bdda
