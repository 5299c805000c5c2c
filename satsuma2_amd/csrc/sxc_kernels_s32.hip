// sxc_kernels_s32.hip -- FFT kernels with a 32-wide stage A (half the threads, twice the
// positions per thread) and three-level LDS rounds, for the lengths where the default stage
// A is 16 wide (one of the geometries the autotuner times).
#include "sxc_fft_kernels.h"

namespace sxc {

const FftEntry* fft_table_s32(int* n) {
  static const FftEntry t[] = {
      SXC_FFT_ENTRY(geo_code(11, 32, 3)) SXC_FFT_ENTRY(geo_code(12, 32, 3)) SXC_FFT_ENTRY(geo_code(13, 32, 3))
      SXC_FFT_ENTRY(geo_code(14, 32, 3))};
  *n = (int)(sizeof(t) / sizeof(t[0]));
  return t;
}

}  // namespace sxc
