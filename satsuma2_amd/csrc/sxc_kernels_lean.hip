// sxc_kernels_lean.hip -- FFT kernels with the lean LDS plan (twiddles and chunk bytes
// read through the caches), so that more of the 32-wide stage-A workgroups share a CU (at
// N = 16384 also with a register budget for two); geometries the autotuner times.
#include "sxc_fft_kernels.h"

namespace sxc {

const FftEntry* fft_table_lean(int* n) {
  static const FftEntry t[] = {
      SXC_FFT_ENTRY(geo_code(11, 32, 3, 0, 1)) SXC_FFT_ENTRY(geo_code(11, 16, 2, 0, 1))
      SXC_FFT_ENTRY(geo_code(12, 32, 3, 0, 1)) SXC_FFT_ENTRY(geo_code(12, 16, 2, 0, 1))
      SXC_FFT_ENTRY(geo_code(13, 32, 3, 0, 1)) SXC_FFT_ENTRY(geo_code(13, 32, 2, 0, 1))
      SXC_FFT_ENTRY(geo_code(14, 32, 3, 0, 1)) SXC_FFT_ENTRY(geo_code(14, 32, 3, 1, 1))};
  *n = (int)(sizeof(t) / sizeof(t[0]));
  return t;
}

}  // namespace sxc
