// sxc_kernels_r2.hip -- FFT kernels with a 16-wide stage A and LDS rounds of at most two
// levels (one of the geometries the autotuner times; the default rounds carry three).
#include "sxc_fft_kernels.h"

namespace sxc {

const FftEntry* fft_table_r2(int* n) {
  static const FftEntry t[] = {
      SXC_FFT_ENTRY(geo_code(9, 16, 2)) SXC_FFT_ENTRY(geo_code(10, 16, 2)) SXC_FFT_ENTRY(geo_code(11, 16, 2))
      SXC_FFT_ENTRY(geo_code(12, 16, 2)) SXC_FFT_ENTRY(geo_code(13, 16, 2)) SXC_FFT_ENTRY(geo_code(14, 16, 2))
      SXC_FFT_ENTRY(geo_code(15, 32, 2))};
  *n = (int)(sizeof(t) / sizeof(t[0]));
  return t;
}

}  // namespace sxc
