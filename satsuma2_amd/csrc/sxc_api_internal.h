// sxc_api_internal.h -- state shared by the C ABI translation units (sxc_api.hip,
// sxc_server.hip): error reporting, device buffers, the context.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>

#include "../../include/sxc.h"
#include "sxc_batch_host.h"
#include "sxc_types.h"

namespace sxc {

extern thread_local std::string g_err;
int fail(int code, const std::string& msg);

#define HIPCHK(expr)                                                                                      \
  do {                                                                                                    \
    hipError_t e_ = (expr);                                                                               \
    if (e_ != hipSuccess) return ::sxc::fail(SXC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// A device allocation that grows on demand (contents are not kept when it grows).
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap && p) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      if (e != hipSuccess) return e;
      p = nullptr;
      cap = 0;
    }
    size_t want = std::max<size_t>(bytes, 256);
    want = (want + 4095) & ~(size_t)4095;
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) cap = want;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  template <class T>
  T* as() const {
    return (T*)p;
  }
};

}  // namespace sxc

using sxc::DevBuf;

struct sxc_batch_dev {
  DevBuf seq, tch, qch, pairs, spectra, which, slots, keys, counts, offsets, total, maxc, final_, stats, cands, ncand;
  DevBuf sch, sstr, tails, segs, ttail, src_sig;  // signal tails
  DevBuf ent;    // host-computed window entropies of chunks with B/D/H/V letters
  DevBuf nflag;  // records the last scan flagged near min_prob (the counts kernel reads and clears it)
  sxc::HostBatch hb;
  sxc_params prm{};
  int nb = 0;
  int slot_cap = 64;
  int64_t n_results = 0;
  double spectra_ms = 0, correlate_ms = 0, scan_ms = 0, finalize_ms = 0;  // last run, by phase
  double device_ms = 0;  // last run, first kernel to last copy
  // Own stream, phase events and pinned landing place for the record count, so that the
  // sub-batches of different length classes run side by side.
  hipStream_t s = nullptr;
  hipEvent_t ev[5] = {};
  int64_t* h_ctl = nullptr;  // [0] largest per-(pair, strand) count, [1] total records, [2] records
                             // flagged near min_prob
  int64_t out_cap = 0;       // records final_ holds
  int64_t resolved = 0;      // records of the last run decided on the host (near min_prob)
  const double* ptable = nullptr;
  // The run's launch sequence captured as a graph, replayed while the buffers it was
  // captured with (signature) stay in place.
  bool use_graph = true;
  bool last_graph = false;  // the last run replayed the graph
  int runs = 0;
  hipGraphExec_t gexec = nullptr;
  uint64_t gsig = 0;
  void release() {
    if (s) (void)hipStreamSynchronize(s);
    if (gexec) (void)hipGraphExecDestroy(gexec), gexec = nullptr;
    for (DevBuf* b : {&seq, &tch, &qch, &pairs, &spectra, &which, &slots, &keys, &counts, &offsets, &total, &maxc,
                      &final_, &stats, &cands, &ncand, &sch, &sstr, &tails, &segs, &ttail, &src_sig, &ent, &nflag})
      b->release();
    for (hipEvent_t& e : ev)
      if (e) (void)hipEventDestroy(e), e = nullptr;
    if (h_ctl) (void)hipHostFree(h_ctl), h_ctl = nullptr;
    if (s) (void)hipStreamDestroy(s), s = nullptr;
  }
};

struct sxc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  sxc_tables* d_tab = nullptr;
  float* d_osc = nullptr;
  double* d_ent[16] = {};
  hipEvent_t ev[8] = {};
  sxc_timing last{};
  std::mutex mu;
  sxc_batch_dev scratch;          // buffers reused by the one-shot entry points
  sxc_batch_dev scratch_nb[16];  // ... per length class when N is chosen per pair
  DevBuf p0, p1, p2, p3, p4, p5, p6;  // plain-API staging
  std::map<double, DevBuf> ptables;   // ProbTable thresholds by target size (-prob_table)
  // FFT geometry code per log2 length for the pair kernel and the target-spectra kernel
  // (a plain log2 length: the default geometry; set by sxc_autotune / sxc_set_geometry)
  int geo_corr[16] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15};
  int geo_spec[16] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15};
};

namespace sxc {
// The ProbTable thresholds for `target_size` in HBM, built on the host on first use.
// Call with ctx->mu held.
int ctx_prob_table(sxc_ctx* ctx, double target_size, const double** out);
// Looks up the device address of the static scratch array of sxc_align_pair_static on the
// current device, once per device (sxc_static.hip).
int static_init(int device);
}  // namespace sxc
