// sxc_fft_kernels.h -- the FFT kernels (k_correlate, k_spectra) as templates over a
// geometry code (see geo_code in sxc_fft.h), the device executor they run on, and the
// launch table each translation unit exports for the geometries it instantiates.
// The default geometries live in sxc_kernels.hip; the alternatives the autotuner times are
// spread over further translation units so that they compile in parallel.
#pragma once
#include <hip/hip_runtime.h>

#include "sxc_pipeline.h"

namespace sxc {

// Device executor: one HIP thread per emulated thread; per-thread state in registers.
// Each phase receives an opaque copy of the thread index, so the compiler recomputes the
// phase's (cheap) LDS addresses instead of hoisting every round's addresses out of the
// channel loop and keeping them all live.
template <class R>
struct DevExec {
  R& r;
  template <class F>
  __device__ __forceinline__ void each(F&& f) {
    int t = (int)threadIdx.x;
    asm volatile("" : "+v"(t));
    f(t);
  }
  __device__ __forceinline__ void sync() { __syncthreads(); }
  template <class T>
  __device__ __forceinline__ T& reg(int) {
    return r;
  }
  __device__ __forceinline__ int atomic_add_shared(int* p, int v) { return atomicAdd(p, v); }
  __device__ __forceinline__ int atomic_add_global(int* p, int v) { return atomicAdd(p, v); }
};

// Workgroup -> (pair, strand): the two strands of a pair read the same resident target
// spectra, so they are placed 8 workgroups apart, which the dispatcher sends to the same
// XCD (and its L2) under round-robin XCD assignment.
__device__ __forceinline__ int xcd_pair_strand(int b, int strands, int total) {
  if (strands != 2) return b;
  const int grp = b / 16, r = b % 16;
  const int ps = (grp * 8 + (r % 8)) * 2 + (r / 8);
  return ps < total ? ps : -1;
}

template <int G>
__global__ __launch_bounds__(Geo<G>::NT) void k_correlate(PairArgs a, int total) {
  __shared__ __align__(16) char sm[Smem<G>::BYTES];
  Regs<G> r;
  DevExec<Regs<G>> ex{r};
  // the grid is rounded up to a multiple of 16 for the strand interleave
  const int ps = xcd_pair_strand((int)blockIdx.x, a.strands, total);
  if (ps < 0) return;
  correlate_body<G>(ex, sm, a, ps);
}

template <int G>
__global__ __launch_bounds__(Geo<G>::NT) void k_spectra(SpecArgs a) {
  __shared__ __align__(16) char sm[Smem<G>::BYTES];
  Regs<G> r;
  DevExec<Regs<G>> ex{r};
  spectra_body<G>(ex, sm, a, (int)blockIdx.x);
}

// The same kernels with a register budget that lets as many workgroups share a CU as its
// LDS plan allows, up to three (geometry codes with bit 8; whether the extra occupancy
// beats the spills it may cost is for the autotuner to measure).
template <int G>
constexpr int occ_waves() {
  // waves per SIMD for as many workgroups as the LDS plan lets share a CU (3 at most)
  constexpr int by_lds = 163840 / Smem<G>::BYTES;
  constexpr int wg = by_lds < 1 ? 1 : (by_lds > 3 ? 3 : by_lds);
  constexpr int w = (wg * (Geo<G>::NT / 64) + 3) / 4;
  return w < 1 ? 1 : (w > 8 ? 8 : w);
}

template <int G>
__global__ __launch_bounds__(Geo<G>::NT) __attribute__((amdgpu_waves_per_eu(occ_waves<G>())))
void k_correlate_occ(PairArgs a, int total) {
  __shared__ __align__(16) char sm[Smem<G>::BYTES];
  Regs<G> r;
  DevExec<Regs<G>> ex{r};
  const int ps = xcd_pair_strand((int)blockIdx.x, a.strands, total);
  if (ps < 0) return;
  correlate_body<G>(ex, sm, a, ps);
}

template <int G>
__global__ __launch_bounds__(Geo<G>::NT) __attribute__((amdgpu_waves_per_eu(occ_waves<G>())))
void k_spectra_occ(SpecArgs a) {
  __shared__ __align__(16) char sm[Smem<G>::BYTES];
  Regs<G> r;
  DevExec<Regs<G>> ex{r};
  spectra_body<G>(ex, sm, a, (int)blockIdx.x);
}

template <int G>
hipError_t launch_correlate_g(const PairArgs& a, int total, hipStream_t s) {
  const int grid = (a.strands == 2) ? ((total + 15) / 16) * 16 : total;
  if constexpr (geo_occ(G)) {
    hipLaunchKernelGGL(k_correlate_occ<G>, dim3(grid), dim3(Geo<G>::NT), 0, s, a, total);
  } else {
    hipLaunchKernelGGL(k_correlate<G>, dim3(grid), dim3(Geo<G>::NT), 0, s, a, total);
  }
  return hipGetLastError();
}

template <int G>
hipError_t launch_spectra_g(const SpecArgs& a, hipStream_t s) {
  if constexpr (geo_occ(G)) {
    hipLaunchKernelGGL(k_spectra_occ<G>, dim3(a.n), dim3(Geo<G>::NT), 0, s, a);
  } else {
    hipLaunchKernelGGL(k_spectra<G>, dim3(a.n), dim3(Geo<G>::NT), 0, s, a);
  }
  return hipGetLastError();
}

// One compiled geometry: its code and launchers.
struct FftEntry {
  int geo;
  hipError_t (*correlate)(const PairArgs&, int, hipStream_t);
  hipError_t (*spectra)(const SpecArgs&, hipStream_t);
};
#define SXC_FFT_ENTRY(G) ::sxc::FftEntry{(G), &::sxc::launch_correlate_g<(G)>, &::sxc::launch_spectra_g<(G)>},

// Launch tables of the translation units (sxc_kernels.hip: defaults; sxc_kernels_r2.hip:
// 2-level LDS rounds; sxc_kernels_s32.hip: 32-wide stage A; sxc_kernels_occ.hip: three
// workgroups per CU; sxc_kernels_lean.hip: lean LDS plan).  Host functions returning a
// function-local table: a namespace-scope constant table would also be emitted for the
// device, where the host launchers it points to do not exist.
const FftEntry* fft_table_default(int* n);
const FftEntry* fft_table_r2(int* n);
const FftEntry* fft_table_s32(int* n);
const FftEntry* fft_table_occ(int* n);
const FftEntry* fft_table_lean(int* n);

}  // namespace sxc
