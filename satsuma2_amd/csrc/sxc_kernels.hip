// sxc_kernels.hip -- gfx950 kernels of the chunk-pair path and their launchers.
//
// Build: hipcc --offload-arch=gfx950 -O3 -ffp-contract=off (FFTReal and the reference's
// double arithmetic are never contracted into FMA; see sxc_fft.h).
#include <hip/hip_runtime.h>

#include "sxc_fft_kernels.h"
#include "sxc_kernels.h"
#include "sxc_pipeline.h"
#include "sxc_scan.h"

namespace sxc {

struct NoRegs {};

// The pure-A/C/G/T scan kernel gets a register budget for as many of its 4-wave
// workgroups per CU as its LDS plan admits, up to 8 waves per SIMD (measured on the 4096-pair
// workload at N = 8192 with byte-parallel flags: 5 waves 1.05 ms, 6: 1.00 ms, 7: 0.94 ms,
// 8: 0.93 ms; the spills the budget costs are cheaper than the latency the extra waves hide).
template <int NB>
constexpr int scan_waves() {
  if constexpr (!ScanSmem<NB>::PURE) {
    return 1;
  } else {
    constexpr int wg = 163840 / ScanSmem<NB>::BYTES;
    return wg < 1 ? 1 : (wg > 8 ? 8 : wg);
  }
}
static_assert(scan_waves<13 | 32>() == 8, "the pure scan's LDS plan at N = 8192 fits 8 workgroups per CU");

template <int NB>
__global__ __launch_bounds__(SCAN_NT) __attribute__((amdgpu_waves_per_eu(scan_waves<NB>()))) void k_scan(ScanArgs a) {
  __shared__ __align__(16) char sm[ScanSmem<NB>::BYTES];
  NoRegs r;
  DevExec<NoRegs> ex{r};
  scan_body<NB>(ex, sm, a, (int)blockIdx.x);
}

template <int NB>
__global__ __launch_bounds__(Geo<NB>::NT) void k_peaks(const float* xc, double cutoff, uint16_t* cands,
                                                     int32_t* ncand) {
  __shared__ __align__(16) char sm[Smem<NB>::BYTES];
  Regs<NB> r;
  DevExec<Regs<NB>> ex{r};
  peaks_body<NB>(ex, sm, xc, cutoff, cands, ncand);
}

// Plain-API kernels ----------------------------------------------------------------

// CCSignal::SetSequence: the 4*N encoded signal of one chunk.
template <int NB>
__global__ __launch_bounds__(Geo<NB>::NT) void k_encode(EncodeArgs a) {
  __shared__ __align__(16) char sm[Smem<NB>::BYTES];
  Regs<NB> r;
  DevExec<Regs<NB>> ex{r};
  encode_body<NB>(ex, sm, a, (int)blockIdx.x);
}

// CrossCorrelation::CrossCorrelate of given signals against resident target spectra.
template <int NB>
__global__ __launch_bounds__(Geo<NB>::NT) void k_xcorr(XcorrArgs a) {
  __shared__ __align__(16) char sm[Smem<NB>::BYTES];
  Regs<NB> r;
  DevExec<Regs<NB>> ex{r};
  using SM = Smem<NB>;
  float* lut = (float*)(sm + SM::LUT);
  for (int i = threadIdx.x; i < SXC_LUT_LEN; i += Geo<NB>::NT) lut[i] = a.tab->lut[i];
  __syncthreads();
  const Twiddle tw{lut, a.osc};
  const int k = blockIdx.x;
  correlate<NB>(ex, sm, a.tab, tw, SeqView{nullptr, nullptr, 0, 0}, a.spectra + (size_t)k * 4 * Geo<NB>::N,
                a.qsig + (size_t)k * 4 * Geo<NB>::N, TailRef{nullptr, 0, nullptr});
  const float* Y = (const float*)(sm + SM::Y);
  for (int p = threadIdx.x; p < Geo<NB>::N; p += Geo<NB>::NT) a.xc[(size_t)k * Geo<NB>::N + p] = Y[pad(p)];
}

// GetMatchProbabilityEx + PrintMatch identity for given runs (one thread per run).
__global__ void k_prob(ProbArgs a) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.n) return;
  const int32_t* r = a.runs + 3 * k;
  const int st = r[0], sq = r[1], len = r[2];
  const sxc_tables* tab = a.tab;
  double matches = 0, gct = 0, gcq = 0;
  for (int i = 0; i < len; ++i) {
    const int ca = tab->cls[a.t[st + i] & 127], cb = tab->cls[a.q[sq + i] & 127];
    matches += tab->equal[ca][cb];
    gct += tab->gc[ca];
    gcq += tab->gc[cb];
  }
  const double ident = matches / (double)len;
  const double gc_target = gct / (double)len;
  const double at_target = 1. - gc_target;
  double rr = gcq * gc_target;
  rr += ((double)len - gcq) * at_target;
  const double p = rr / (double)len / 2.;
  const double s = sqrt(p * (1. - p) * (double)len);
  const double m = p * (double)len;
  const double x = (double)len * ident;
  const double cdf = 0.5 * (1. + erf((m - x) / s / 1.414213562));
  a.prob[k] = exp(-(cdf * a.target_size));
  a.ident[k] = ident;
}

// Exclusive scan of per-(pair, strand) kept counts (clipped to cap) -> offsets, total.
// With nflag: total[1] receives the scan's count of records flagged near min_prob, and the
// counter is cleared for the next scan.
__global__ __launch_bounds__(1024) void k_scan_counts(const int32_t* counts, int n, int cap, int64_t* offsets,
                                                      int64_t* total, int32_t* max_count, int32_t* nflag) {
  __shared__ int64_t part[1024];
  __shared__ int mx[1024];
  const int t = threadIdx.x;
  const int per = (n + 1023) / 1024;
  const int b = t * per, e = (b + per < n) ? b + per : n;
  int64_t s = 0;
  int m = 0;
  for (int i = b; i < e; ++i) {
    const int c = counts[i];
    s += c < cap ? c : cap;
    m = c > m ? c : m;
  }
  part[t] = s;
  mx[t] = m;
  __syncthreads();
  if (t == 0) {
    int64_t acc = 0;
    int mm = 0;
    for (int i = 0; i < 1024; ++i) {
      const int64_t v = part[i];
      part[i] = acc;
      acc += v;
      mm = mx[i] > mm ? mx[i] : mm;
    }
    *total = acc;
    *max_count = mm;
    if (nflag) {
      total[1] = *nflag;
      *nflag = 0;
    }
  }
  __syncthreads();
  int64_t acc = part[t];
  for (int i = b; i < e; ++i) {
    offsets[i] = acc;
    const int c = counts[i];
    acc += c < cap ? c : cap;
  }
}

// Orders each (pair, strand)'s kept records by (lag rank, run start) -- the order
// FilterMatches pushes them -- and packs them at their offsets.
// Records that would land at or beyond out_cap are not written (the host compares the
// total with out_cap and packs again into a larger buffer).
__global__ __launch_bounds__(256) void k_finalize(const sxc_result_dev* slots, const uint64_t* keys,
                                                  const int32_t* counts, const int64_t* offsets, int cap,
                                                  sxc_result_dev* out, int64_t out_cap) {
  const int ps = blockIdx.x;
  const int n = counts[ps] < cap ? counts[ps] : cap;
  const size_t base = (size_t)ps * cap;
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    const uint64_t ke = keys[base + e];
    int rank = 0;
    for (int k = 0; k < n; ++k) rank += keys[base + k] < ke;
    const int64_t at = offsets[ps] + rank;
    if (at < out_cap) out[at] = slots[base + e];
  }
}

// Orders MatchUp runs by (lag rank, run start): rank by counting smaller keys.
__global__ __launch_bounds__(256) void k_order_runs(const int32_t* runs, const uint64_t* keys, int n, int32_t* out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const uint64_t ke = keys[e];
  int rank = 0;
  for (int k = 0; k < n; ++k) rank += keys[k] < ke;
  out[3 * rank] = runs[3 * e];
  out[3 * rank + 1] = runs[3 * e + 1];
  out[3 * rank + 2] = runs[3 * e + 2];
}

// ---------------------------------------------------------------- launchers

hipError_t launch_order_runs(const int32_t* runs, const uint64_t* keys, int n, int32_t* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_order_runs, dim3((n + 255) / 256), dim3(256), 0, s, runs, keys, n, out);
  return hipGetLastError();
}

#define SXC_NB_SWITCH(NB, CALL)      \
  switch (NB) {                      \
    case 9: CALL(9); break;          \
    case 10: CALL(10); break;        \
    case 11: CALL(11); break;        \
    case 12: CALL(12); break;        \
    case 13: CALL(13); break;        \
    case 14: CALL(14); break;        \
    case 15: CALL(15); break;        \
    default: return hipErrorInvalidValue; \
  }

const FftEntry* fft_table_default(int* n) {
  static const FftEntry t[] = {SXC_FFT_ENTRY(9) SXC_FFT_ENTRY(10) SXC_FFT_ENTRY(11) SXC_FFT_ENTRY(12)
                                   SXC_FFT_ENTRY(13) SXC_FFT_ENTRY(14) SXC_FFT_ENTRY(15)};
  *n = (int)(sizeof(t) / sizeof(t[0]));
  return t;
}

const FftEntry* fft_entry(int geo) {
  for (auto tf : {fft_table_default, fft_table_r2, fft_table_s32, fft_table_occ, fft_table_lean}) {
    int n = 0;
    const FftEntry* t = tf(&n);
    for (int i = 0; i < n; ++i)
      if (t[i].geo == geo) return &t[i];
  }
  return nullptr;
}

int fft_geometries(int nb, int* out, int cap) {
  int n = 0;
  for (auto tf : {fft_table_default, fft_table_r2, fft_table_s32, fft_table_occ, fft_table_lean}) {
    int nt = 0;
    const FftEntry* t = tf(&nt);
    for (int i = 0; i < nt; ++i) {
      const int g = t[i].geo;
      if (geo_nb(g) != nb) continue;
      bool dup = false;  // the same geometry under another code
      for (int j = 0; j < n; ++j)
        dup |= geo_sa(out[j]) == geo_sa(g) && geo_km(out[j]) == geo_km(g) && geo_occ(out[j]) == geo_occ(g) &&
               geo_lean(out[j]) == geo_lean(g);
      if (!dup && n < cap) out[n++] = g;
    }
  }
  return n;
}

hipError_t launch_correlate(int geo, const PairArgs& a, int total, hipStream_t s) {
  const FftEntry* e = fft_entry(geo);
  return e ? e->correlate(a, total, s) : hipErrorInvalidValue;
}

hipError_t launch_scan(int nb, const ScanArgs& a, int total, hipStream_t s) {
  // the pure-A/C/G/T kernel, its A/C/G/T/N variant and the general one over the same grid;
  // each workgroup exits at once on the other kernels' pairs
  const int grid = total * (a.split > 1 ? a.split : 1);
#define CALL(X)                                                            \
  hipLaunchKernelGGL(k_scan<(X) | 32>, dim3(grid), dim3(SCAN_NT), 0, s, a); \
  hipLaunchKernelGGL(k_scan<(X) | 64>, dim3(grid), dim3(SCAN_NT), 0, s, a); \
  hipLaunchKernelGGL(k_scan<X>, dim3(grid), dim3(SCAN_NT), 0, s, a)
  SXC_NB_SWITCH(nb, CALL)
#undef CALL
  return hipGetLastError();
}

hipError_t launch_peaks(int nb, const float* xc, double cutoff, uint16_t* cands, int32_t* ncand, hipStream_t s) {
#define CALL(X) hipLaunchKernelGGL(k_peaks<X>, dim3(1), dim3(Geo<X>::NT), 0, s, xc, cutoff, cands, ncand)
  SXC_NB_SWITCH(nb, CALL)
#undef CALL
  return hipGetLastError();
}

hipError_t launch_spectra(int geo, const SpecArgs& a, hipStream_t s) {
  const FftEntry* e = fft_entry(geo);
  return e ? e->spectra(a, s) : hipErrorInvalidValue;
}

hipError_t launch_encode(int nb, const EncodeArgs& a, int n, hipStream_t s) {
#define CALL(X) hipLaunchKernelGGL(k_encode<X>, dim3(n), dim3(Geo<X>::NT), 0, s, a)
  SXC_NB_SWITCH(nb, CALL)
#undef CALL
  return hipGetLastError();
}

hipError_t launch_xcorr(int nb, const XcorrArgs& a, int n, hipStream_t s) {
#define CALL(X) hipLaunchKernelGGL(k_xcorr<X>, dim3(n), dim3(Geo<X>::NT), 0, s, a)
  SXC_NB_SWITCH(nb, CALL)
#undef CALL
  return hipGetLastError();
}

hipError_t launch_prob(const ProbArgs& a, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_prob, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_count_offsets(const int32_t* counts, int n, int cap, int64_t* offsets, int64_t* total,
                                int32_t* max_count, hipStream_t s, int32_t* nflag) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, s, counts, n, cap, offsets, total, max_count, nflag);
  return hipGetLastError();
}

hipError_t launch_order_records(const int32_t* counts, int n, int cap, const int64_t* offsets,
                                const sxc_result_dev* slots, const uint64_t* keys, sxc_result_dev* out,
                                int64_t out_cap, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_finalize, dim3(n), dim3(256), 0, s, slots, keys, counts, offsets, cap, out, out_cap);
  return hipGetLastError();
}

}  // namespace sxc
