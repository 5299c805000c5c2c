// sxc_kernels.h -- launchers of the gfx950 kernels (defined in sxc_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "sxc_pipeline.h"
#include "sxc_scan.h"

namespace sxc {

hipError_t launch_correlate(int nb, const PairArgs& a, int total, hipStream_t s);
hipError_t launch_scan(int nb, const ScanArgs& a, int total, hipStream_t s);
hipError_t launch_peaks(int nb, const float* xc, double cutoff, uint16_t* cands, int32_t* ncand, hipStream_t s);
hipError_t launch_spectra(int nb, const SpecArgs& a, hipStream_t s);
hipError_t launch_encode(int nb, const EncodeArgs& a, int n, hipStream_t s);
hipError_t launch_xcorr(int nb, const XcorrArgs& a, int n, hipStream_t s);
hipError_t launch_prob(const ProbArgs& a, hipStream_t s);
hipError_t launch_order_runs(const int32_t* runs, const uint64_t* keys, int n, int32_t* out, hipStream_t s);
// scan_only: exclusive scan of counts into offsets/total/max_count; else pack slots into out.
hipError_t launch_finalize(const int32_t* counts, int n, int cap, int64_t* offsets, int64_t* total,
                           int32_t* max_count, const sxc_result_dev* slots, const uint64_t* keys,
                           sxc_result_dev* out, hipStream_t s, bool scan_only);

}  // namespace sxc
