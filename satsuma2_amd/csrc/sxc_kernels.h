// sxc_kernels.h -- launchers of the gfx950 kernels (defined in sxc_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "sxc_pipeline.h"
#include "sxc_scan.h"

namespace sxc {

// The FFT kernels take a geometry code (geo_code in sxc_fft.h; a plain log2 length is the
// default geometry for that length); hipErrorInvalidValue if it was not compiled.
hipError_t launch_correlate(int geo, const PairArgs& a, int total, hipStream_t s);
// Compiled geometries of length 2^nb (distinct stage-A width / round depth), default first.
int fft_geometries(int nb, int* out, int cap);
hipError_t launch_scan(int nb, const ScanArgs& a, int total, hipStream_t s);
hipError_t launch_peaks(int nb, const float* xc, double cutoff, uint16_t* cands, int32_t* ncand, hipStream_t s);
hipError_t launch_spectra(int geo, const SpecArgs& a, hipStream_t s);
hipError_t launch_encode(int nb, const EncodeArgs& a, int n, hipStream_t s);
hipError_t launch_xcorr(int nb, const XcorrArgs& a, int n, hipStream_t s);
hipError_t launch_prob(const ProbArgs& a, hipStream_t s);
hipError_t launch_order_runs(const int32_t* runs, const uint64_t* keys, int n, int32_t* out, hipStream_t s);
// Exclusive scan of the per-(pair, strand) record counts (clipped at cap) into offsets, their
// total and the largest unclipped count.  With nflag (a counter the scan raises for each record
// it flags near min_prob), total[1] receives the counter and the counter is cleared.
hipError_t launch_count_offsets(const int32_t* counts, int n, int cap, int64_t* offsets, int64_t* total,
                                int32_t* max_count, hipStream_t s, int32_t* nflag = nullptr);
// Orders each (pair, strand)'s slot records and packs them at its offset; records at or past
// out_cap are dropped.
hipError_t launch_order_records(const int32_t* counts, int n, int cap, const int64_t* offsets,
                                const sxc_result_dev* slots, const uint64_t* keys, sxc_result_dev* out,
                                int64_t out_cap, hipStream_t s);

}  // namespace sxc
