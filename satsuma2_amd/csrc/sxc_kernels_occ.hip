// sxc_kernels_occ.hip -- FFT kernels compiled for three workgroups per CU (a register
// budget of 6 waves per SIMD for 512-thread workgroups) at N = 8192, the length where the
// default kernels hold two; one of the geometries the autotuner times.
#include "sxc_fft_kernels.h"

namespace sxc {

const FftEntry* fft_table_occ(int* n) {
  static const FftEntry t[] = {SXC_FFT_ENTRY(geo_code(13, 16, 3, 1)) SXC_FFT_ENTRY(geo_code(13, 16, 2, 1))};
  *n = (int)(sizeof(t) / sizeof(t[0]));
  return t;
}

}  // namespace sxc
