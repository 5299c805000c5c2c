// sxc_api.hip -- C ABI of libsxc_hip.so (declared in include/sxc.h).
//
// Host side of the drop-in: validates and packs caller buffers, owns device memory,
// launches the gfx950 kernels on one HIP stream per context, and returns results in
// the reference's order.  All compute runs on the GPU; there is no host fallback.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sxc.h"
#include "sxc_api_internal.h"
#include "sxc_batch_host.h"
#include "sxc_kernels.h"
#include "sxc_tables.h"

namespace sxc {
thread_local std::string g_err;
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
}  // namespace sxc

using sxc::fail;

int sxc::ctx_prob_table(sxc_ctx* ctx, double target_size, const double** out) {
  auto it = ctx->ptables.find(target_size);
  if (it == ctx->ptables.end()) {
    std::vector<double> t((size_t)SXC_PT_ROWS * SXC_PT_LEN);
    sxc::build_prob_table(target_size, t.data());
    DevBuf& d = ctx->ptables[target_size];
    HIPCHK(d.ensure(t.size() * sizeof(double)));
    HIPCHK(hipMemcpy(d.p, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
    it = ctx->ptables.find(target_size);
  }
  *out = it->second.as<double>();
  return SXC_OK;
}

namespace {

int check_N(int N, int* nb) {
  *nb = sxc::nb_of(N);
  if (*nb < SXC_MIN_NB || *nb > SXC_MAX_NB) return fail(SXC_ERR_ARG, "N must be a power of two in [512, 32768]");
  return SXC_OK;
}

int set_device(sxc_ctx* ctx) {
  HIPCHK(hipSetDevice(ctx->device));
  return SXC_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
  return ms;
}

// Upload a packed batch (inputs become resident in HBM).
// The sub-batch's own stream, events and pinned control words (created on first use, on
// the context's device).
int batch_stream(sxc_batch_dev& b) {
  if (b.s) return SXC_OK;
  HIPCHK(hipStreamCreateWithFlags(&b.s, hipStreamNonBlocking));
  for (hipEvent_t& e : b.ev) HIPCHK(hipEventCreate(&e));
  HIPCHK(hipHostMalloc((void**)&b.h_ctl, 3 * sizeof(int64_t), hipHostMallocDefault));
  return SXC_OK;
}

int batch_upload(sxc_ctx* ctx, sxc_batch_dev& b) {
  (void)ctx;
  const sxc::HostBatch& hb = b.hb;
  const int N = b.prm.N;
  int rc = batch_stream(b);
  if (rc) return rc;
  hipStream_t s = b.s;
  HIPCHK(b.seq.ensure(hb.seq.size()));
  HIPCHK(b.tch.ensure(hb.tchunks.size() * sizeof(sxc_chunk)));
  HIPCHK(b.qch.ensure(hb.qchunks.size() * sizeof(sxc_chunk)));
  HIPCHK(b.pairs.ensure(hb.pairs.size() * sizeof(sxc_pair_dev)));
  HIPCHK(b.spectra.ensure(hb.tchunks.size() * 4 * (size_t)N * sizeof(float)));
  HIPCHK(b.which.ensure(hb.tchunks.size() * sizeof(int32_t)));
  if (!hb.ent.empty()) {
    HIPCHK(b.ent.ensure(hb.ent.size() * sizeof(float)));
    HIPCHK(hipMemcpyAsync(b.ent.p, hb.ent.data(), hb.ent.size() * sizeof(float), hipMemcpyHostToDevice, s));
  }
  if (!hb.seq.empty()) HIPCHK(hipMemcpyAsync(b.seq.p, hb.seq.data(), hb.seq.size(), hipMemcpyHostToDevice, s));
  if (!hb.tchunks.empty())
    HIPCHK(hipMemcpyAsync(b.tch.p, hb.tchunks.data(), hb.tchunks.size() * sizeof(sxc_chunk), hipMemcpyHostToDevice, s));
  if (!hb.qchunks.empty())
    HIPCHK(hipMemcpyAsync(b.qch.p, hb.qchunks.data(), hb.qchunks.size() * sizeof(sxc_chunk), hipMemcpyHostToDevice, s));
  if (!hb.pairs.empty())
    HIPCHK(hipMemcpyAsync(b.pairs.p, hb.pairs.data(), hb.pairs.size() * sizeof(sxc_pair_dev), hipMemcpyHostToDevice, s));
  std::vector<int32_t> idx(hb.tchunks.size());
  for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int32_t)i;
  if (!idx.empty()) HIPCHK(hipMemcpyAsync(b.which.p, idx.data(), idx.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
  if (!hb.tails.empty()) {
    HIPCHK(b.sch.ensure(hb.schunks.size() * sizeof(sxc_chunk)));
    HIPCHK(b.sstr.ensure(hb.sstrand.size() * sizeof(int32_t)));
    HIPCHK(b.tails.ensure(hb.tails.size() * sizeof(sxc_tail_dev)));
    HIPCHK(b.segs.ensure(hb.segs.size() * sizeof(sxc_tail_seg_dev)));
    HIPCHK(b.ttail.ensure(hb.ttail.size() * sizeof(int32_t)));
    HIPCHK(b.src_sig.ensure(hb.schunks.size() * 4 * (size_t)N * sizeof(float)));
    HIPCHK(hipMemcpyAsync(b.sch.p, hb.schunks.data(), hb.schunks.size() * sizeof(sxc_chunk), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(b.sstr.p, hb.sstrand.data(), hb.sstrand.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(b.tails.p, hb.tails.data(), hb.tails.size() * sizeof(sxc_tail_dev), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(b.segs.p, hb.segs.data(), hb.segs.size() * sizeof(sxc_tail_seg_dev), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(b.ttail.p, hb.ttail.data(), hb.ttail.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
  }
  HIPCHK(hipStreamSynchronize(s));  // idx is a host temporary
  return SXC_OK;
}

// Device buffers of a run, sized before anything is queued (nothing allocates while a
// launch sequence is being captured into a graph).
int batch_alloc(sxc_ctx* ctx, sxc_batch_dev& b) {
  const int total = (int)b.hb.pairs.size() * b.prm.strands;
  HIPCHK(b.cands.ensure((size_t)total * b.prm.N * sizeof(uint16_t)));
  HIPCHK(b.ncand.ensure((size_t)total * sizeof(int32_t)));
  HIPCHK(b.slots.ensure((size_t)total * b.slot_cap * sizeof(sxc_result_dev)));
  HIPCHK(b.keys.ensure((size_t)total * b.slot_cap * sizeof(uint64_t)));
  HIPCHK(b.counts.ensure((size_t)total * sizeof(int32_t)));
  HIPCHK(b.stats.ensure((size_t)total * 2 * sizeof(int32_t)));
  HIPCHK(b.offsets.ensure((size_t)total * sizeof(int64_t)));
  HIPCHK(b.total.ensure(2 * sizeof(int64_t)));
  HIPCHK(b.maxc.ensure(sizeof(int32_t)));
  const void* had = b.nflag.p;
  HIPCHK(b.nflag.ensure(sizeof(int32_t)));
  if (b.nflag.p != had) HIPCHK(hipMemsetAsync(b.nflag.p, 0, sizeof(int32_t), b.s));  // before any capture
  if (b.out_cap == 0) b.out_cap = std::max<int64_t>(1024, 2 * (int64_t)total);
  HIPCHK(b.final_.ensure((size_t)b.out_cap * sizeof(sxc_result_dev)));
  b.ptable = nullptr;
  if (b.prm.prob_table) {
    const int rc = sxc::ctx_prob_table(ctx, b.prm.target_total, &b.ptable);
    if (rc) return rc;
  }
  return SXC_OK;
}

int launch_correlate(sxc_ctx* ctx, sxc_batch_dev& b) {
  const int total = (int)b.hb.pairs.size() * b.prm.strands;
  sxc::PairArgs a{};
  a.tab = ctx->d_tab;
  a.osc = ctx->d_osc;
  a.ent_lut = ctx->d_ent[b.nb];
  a.ent_rows = b.hb.ent.empty() ? nullptr : b.ent.as<float>();
  a.seq = b.seq.as<unsigned char>();
  a.tchunks = b.tch.as<sxc_chunk>();
  a.qchunks = b.qch.as<sxc_chunk>();
  a.pairs = b.pairs.as<sxc_pair_dev>();
  a.spectra = b.spectra.as<float>();
  a.npairs = (int)b.hb.pairs.size();
  a.strands = b.prm.strands;
  a.cutoff = b.prm.cutoff;
  a.cutoff_fast = b.prm.cutoff_fast;
  a.cands = b.cands.as<uint16_t>();
  a.ncand = b.ncand.as<int32_t>();
  a.xc_out = nullptr;
  if (!b.hb.tails.empty()) {
    a.tails = b.tails.as<sxc_tail_dev>();
    a.segs = b.segs.as<sxc_tail_seg_dev>();
    a.src_sig = b.src_sig.as<float>();
  }
  HIPCHK(sxc::launch_correlate(ctx->geo_corr[b.nb], a, total, b.s));
  return SXC_OK;
}

int launch_pairs(sxc_ctx* ctx, sxc_batch_dev& b) {
  const int total = (int)b.hb.pairs.size() * b.prm.strands;
  sxc::ScanArgs a{};
  a.tab = ctx->d_tab;
  a.seq = b.seq.as<unsigned char>();
  a.tchunks = b.tch.as<sxc_chunk>();
  a.qchunks = b.qch.as<sxc_chunk>();
  a.pairs = b.pairs.as<sxc_pair_dev>();
  a.cands = b.cands.as<uint16_t>();
  a.ncand = b.ncand.as<int32_t>();
  a.n_fft = b.prm.N;
  a.strands = b.prm.strands;
  a.min_len = b.prm.min_len;
  a.q_chunk = b.prm.q_chunk;
  sxc::set_keep(a, b.prm.min_prob, b.prm.target_total);
  a.nflag = b.nflag.as<int32_t>();
  a.mode = 0;
  a.out = b.slots.as<sxc_result_dev>();
  a.runs = nullptr;
  a.keys = b.keys.as<uint64_t>();
  a.counts = b.counts.as<int32_t>();
  a.cap = b.slot_cap;
  a.stats = b.stats.as<int32_t>();
  if (b.prm.prob_table) {
    a.ptable = b.ptable;
    a.table_prob = b.prm.table_prob;
  }
  HIPCHK(sxc::launch_scan(b.nb, a, total, b.s));
  return SXC_OK;
}

// Device-only work of a resident batch, in two halves so that several sub-batches (one per
// length class) run side by side on their own streams.  batch_begin queues the whole run --
// target spectra, the fused pair kernel, the record counts, the ordered records and the copy
// of the count and total to pinned host memory -- and returns without waiting; from the
// second run on, the sequence is captured once as a graph and replayed.  batch_end waits and,
// if a (pair, strand) kept more records than its slot holds or the records overflow the
// output buffer, grows the buffer and redoes what it needs eagerly.

// The four spectra of every target chunk of the batch, into b.spectra.
int launch_target_spectra(sxc_ctx* ctx, sxc_batch_dev& b) {
  if (b.hb.tchunks.empty()) return SXC_OK;
  sxc::SpecArgs sa{};
  sa.tab = ctx->d_tab;
  sa.osc = ctx->d_osc;
  sa.ent_lut = ctx->d_ent[b.nb];
  sa.ent_rows = b.hb.ent.empty() ? nullptr : b.ent.as<float>();
  sa.seq = b.seq.as<unsigned char>();
  sa.chunks = b.tch.as<sxc_chunk>();
  sa.which = b.which.as<int32_t>();
  sa.n = (int)b.hb.tchunks.size();
  sa.spectra = b.spectra.as<float>();
  sa.dst = b.which.as<int32_t>();
  sa.signals = nullptr;
  if (!b.hb.tails.empty()) {
    sa.chunk_tail = b.ttail.as<int32_t>();
    sa.tails = b.tails.as<sxc_tail_dev>();
    sa.segs = b.segs.as<sxc_tail_seg_dev>();
    sa.src_sig = b.src_sig.as<float>();
  }
  HIPCHK(sxc::launch_spectra(ctx->geo_spec[b.nb], sa, b.s));
  return SXC_OK;
}

// Kernels and copies of one run on b.s (no allocation, no host synchronisation).
int batch_enqueue(sxc_ctx* ctx, sxc_batch_dev& b, bool events) {
  hipStream_t s = b.s;
  const int total = (int)b.hb.pairs.size() * b.prm.strands;
  if (events) HIPCHK(hipEventRecord(b.ev[0], s));
  if (!b.hb.schunks.empty()) {
    // signals of the sequences named by signal tails
    sxc::EncodeArgs ea{ctx->d_tab, ctx->d_ent[b.nb], b.hb.ent.empty() ? nullptr : b.ent.as<float>(),
                       b.seq.as<unsigned char>(), b.sch.as<sxc_chunk>(), 0, b.sstr.as<int32_t>(),
                       b.src_sig.as<float>()};
    HIPCHK(sxc::launch_encode(b.nb, ea, (int)b.hb.schunks.size(), s));
  }
  int rc = launch_target_spectra(ctx, b);
  if (rc) return rc;
  if (events) HIPCHK(hipEventRecord(b.ev[1], s));
  rc = launch_correlate(ctx, b);
  if (rc) return rc;
  if (events) HIPCHK(hipEventRecord(b.ev[2], s));
  rc = launch_pairs(ctx, b);
  if (rc) return rc;
  if (events) HIPCHK(hipEventRecord(b.ev[3], s));
  HIPCHK(sxc::launch_count_offsets(b.counts.as<int32_t>(), total, b.slot_cap, b.offsets.as<int64_t>(),
                                   b.total.as<int64_t>(), b.maxc.as<int32_t>(), s, b.nflag.as<int32_t>()));
  HIPCHK(sxc::launch_order_records(b.counts.as<int32_t>(), total, b.slot_cap, b.offsets.as<int64_t>(),
                                   b.slots.as<sxc_result_dev>(), b.keys.as<uint64_t>(),
                                   b.final_.as<sxc_result_dev>(), b.out_cap, s));
  HIPCHK(hipMemcpyAsync(&b.h_ctl[0], b.maxc.p, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(&b.h_ctl[1], b.total.p, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, s));
  if (events) HIPCHK(hipEventRecord(b.ev[4], s));
  return SXC_OK;
}

// Everything a captured graph depends on besides the batch's (fixed) inputs.
uint64_t graph_signature(const sxc_ctx* ctx, const sxc_batch_dev& b) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t v) { h = (h ^ v) * 1099511628211ull; };
  for (const DevBuf* d : {&b.seq, &b.tch, &b.qch, &b.pairs, &b.spectra, &b.which, &b.slots, &b.keys, &b.counts,
                          &b.offsets, &b.total, &b.maxc, &b.final_, &b.stats, &b.cands, &b.ncand, &b.sch, &b.sstr,
                          &b.tails, &b.segs, &b.ttail, &b.src_sig, &b.ent, &b.nflag})
    mix((uint64_t)(uintptr_t)d->p);
  mix((uint64_t)(uintptr_t)b.ptable);
  mix((uint64_t)(uintptr_t)b.h_ctl);
  mix((uint64_t)b.slot_cap);
  mix((uint64_t)b.out_cap);
  // launch shapes and scalar arguments
  mix(b.hb.pairs.size());
  mix(b.hb.tchunks.size());
  mix(b.hb.qchunks.size());
  mix(b.hb.schunks.size());
  mix(b.hb.tails.size());
  mix((uint64_t)b.nb);
  mix((uint64_t)ctx->geo_corr[b.nb]);
  mix((uint64_t)ctx->geo_spec[b.nb]);
  const unsigned char* pb = (const unsigned char*)&b.prm;
  for (size_t i = 0; i < sizeof b.prm; ++i) mix(pb[i]);
  return h;
}

int batch_begin(sxc_ctx* ctx, sxc_batch_dev& b) {
  int rc = batch_stream(b);
  if (rc) return rc;
  hipStream_t s = b.s;
  const int total = (int)b.hb.pairs.size() * b.prm.strands;
  b.last_graph = false;
  if (total == 0) {
    // no pairs (and so no chunks): nothing to run
    HIPCHK(hipEventRecord(b.ev[0], s));
    HIPCHK(hipEventRecord(b.ev[1], s));
    return SXC_OK;
  }
  rc = batch_alloc(ctx, b);
  if (rc) return rc;
  b.h_ctl[0] = b.h_ctl[1] = b.h_ctl[2] = 0;
  if (b.use_graph && b.runs > 0) {
    const uint64_t sig = graph_signature(ctx, b);
    if (!b.gexec || b.gsig != sig) {
      if (b.gexec) HIPCHK(hipGraphExecDestroy(b.gexec));
      b.gexec = nullptr;
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      rc = batch_enqueue(ctx, b, false);
      const hipError_t e = hipStreamEndCapture(s, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      HIPCHK(e);
      const hipError_t ei = hipGraphInstantiateWithFlags(&b.gexec, g, 0);
      (void)hipGraphDestroy(g);
      HIPCHK(ei);
      b.gsig = sig;
    }
    HIPCHK(hipEventRecord(b.ev[0], s));
    HIPCHK(hipGraphLaunch(b.gexec, s));
    HIPCHK(hipEventRecord(b.ev[4], s));
    b.last_graph = true;
  } else {
    rc = batch_enqueue(ctx, b, true);
    if (rc) return rc;
  }
  ++b.runs;
  return SXC_OK;
}

// The scan flagged records whose device probability is within the keep band of min_prob:
// decide them on the host with its libm (resolve_record), drop those below min_prob, and put
// the ordered records and the per-(pair, strand) offsets back in place.  Flagged records are
// rare (none in most runs), so this costs one copy each way when it happens.
int batch_resolve(sxc_batch_dev& b) {
  hipStream_t s = b.s;
  const int total = (int)b.hb.pairs.size() * b.prm.strands;
  const int64_t n = b.n_results;
  std::vector<sxc_result_dev> rec((size_t)n);
  std::vector<int64_t> off((size_t)total + 1);
  if (n > 0) HIPCHK(hipMemcpyAsync(rec.data(), b.final_.p, (size_t)n * sizeof(sxc_result_dev), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(off.data(), b.offsets.p, (size_t)total * sizeof(int64_t), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  off[(size_t)total] = n;
  const sxc::ResolveCtx rc{b.hb.seq.data(), b.hb.tchunks.data(), b.hb.qchunks.data(), b.hb.pairs.data(),
                           (int64_t)b.hb.pairs.size(), b.prm.q_chunk, b.prm.min_prob, b.prm.target_total};
  int64_t kept = 0, decided = 0;
  for (int ps = 0; ps < total; ++ps) {
    const int64_t lo = off[(size_t)ps], hi = off[(size_t)ps + 1];
    off[(size_t)ps] = kept;
    for (int64_t i = lo; i < hi; ++i) {
      decided += rec[(size_t)i].pad_[0] == SXC_REC_NEAR;
      const int k = sxc::resolve_record(rec[(size_t)i], rc);
      if (k < 0) return fail(SXC_ERR_HIP, "a record flagged near min_prob does not fit its chunk pair");
      if (k) rec[(size_t)kept++] = rec[(size_t)i];
    }
  }
  if (kept > 0)
    HIPCHK(hipMemcpyAsync(b.final_.p, rec.data(), (size_t)kept * sizeof(sxc_result_dev), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(b.offsets.p, off.data(), (size_t)total * sizeof(int64_t), hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  b.n_results = kept;
  b.resolved = decided;
  return SXC_OK;
}

int batch_end(sxc_ctx* ctx, sxc_batch_dev& b) {
  hipStream_t s = b.s;
  const int total = (int)b.hb.pairs.size() * b.prm.strands;
  if (total == 0) {
    HIPCHK(hipEventSynchronize(b.ev[1]));
    b.n_results = 0;
    b.spectra_ms = elapsed(b.ev[0], b.ev[1]);
    b.correlate_ms = b.scan_ms = b.finalize_ms = 0;
    b.device_ms = b.spectra_ms;
    return SXC_OK;
  }
  HIPCHK(hipStreamSynchronize(s));
  if (b.last_graph) {
    b.spectra_ms = b.correlate_ms = b.scan_ms = b.finalize_ms = 0;
  } else {
    b.spectra_ms = elapsed(b.ev[0], b.ev[1]);
    b.correlate_ms = elapsed(b.ev[1], b.ev[2]);
    b.scan_ms = elapsed(b.ev[2], b.ev[3]);
    b.finalize_ms = elapsed(b.ev[3], b.ev[4]);
  }
  b.device_ms = elapsed(b.ev[0], b.ev[4]);
  const int32_t maxc = (int32_t)b.h_ctl[0];
  int64_t tot = b.h_ctl[1];
  bool order = false;
  if (maxc > b.slot_cap) {
    // some (pair, strand) kept more records than its slot holds: enlarge and recompute
    b.slot_cap = maxc;
    int rc = batch_alloc(ctx, b);
    if (rc) return rc;
    rc = launch_pairs(ctx, b);
    if (rc) return rc;
    HIPCHK(sxc::launch_count_offsets(b.counts.as<int32_t>(), total, b.slot_cap, b.offsets.as<int64_t>(),
                                     b.total.as<int64_t>(), b.maxc.as<int32_t>(), s, b.nflag.as<int32_t>()));
    HIPCHK(hipMemcpyAsync(&b.h_ctl[1], b.total.p, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    tot = b.h_ctl[1];
    order = true;
  }
  if (tot > b.out_cap) {
    b.out_cap = tot + tot / 4;
    HIPCHK(b.final_.ensure((size_t)b.out_cap * sizeof(sxc_result_dev)));
    order = true;
  }
  if (order) {
    HIPCHK(sxc::launch_order_records(b.counts.as<int32_t>(), total, b.slot_cap, b.offsets.as<int64_t>(),
                                     b.slots.as<sxc_result_dev>(), b.keys.as<uint64_t>(),
                                     b.final_.as<sxc_result_dev>(), b.out_cap, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  b.n_results = tot;
  b.resolved = 0;
  if (b.h_ctl[2] > 0) return batch_resolve(b);
  return SXC_OK;
}

int batch_download(sxc_ctx* ctx, sxc_batch_dev& b, sxc_result* out, int64_t cap, int64_t* n_out) {
  *n_out = b.n_results;
  if (b.n_results > cap) return fail(SXC_ERR_CAPACITY, "output capacity too small");
  if (b.n_results > 0)
    HIPCHK(hipMemcpyAsync(out, b.final_.p, (size_t)b.n_results * sizeof(sxc_result), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return SXC_OK;
}

int batch_prepare(sxc_batch_dev& b, const sxc_pair* pairs, int32_t n, const sxc_params* prm,
                  const sxc_signal_tail* t_tail = nullptr, const sxc_signal_tail* q_tail = nullptr) {
  if (!prm || n < 0 || (n > 0 && !pairs)) return fail(SXC_ERR_ARG, "null argument");
  if (prm->strands != 1 && prm->strands != 2) return fail(SXC_ERR_ARG, "strands must be 1 or 2");
  if (prm->q_chunk < 0) return fail(SXC_ERR_ARG, "q_chunk must be >= 0");
  int nb;
  int rc = check_N(prm->N, &nb);
  if (rc) return rc;
  std::string e = sxc::pack_batch(pairs, n, prm->N, b.hb, t_tail, q_tail);
  if (!e.empty()) return fail(SXC_ERR_ARG, e);
  b.prm = *prm;
  b.nb = nb;
  b.slot_cap = 64;  // records per (pair, strand) a run first makes room for; grows on overflow
  return SXC_OK;
}

// ---------------------------------------------------------------- length classes
//
// With prm->N == 0 every pair gets its own FFT length, the next power of two >= tlen + qlen
// (at least 512), and a batch is one sub-batch per length class, run class after class on
// the context's stream; records are merged back into input order.  With prm->N > 0 there is
// one class and no merge.

struct Group {
  int nb = 0;
  std::vector<int32_t> idx;  // input pair index of each entry (empty: the identity)
  sxc_batch_dev* b = nullptr;
};

int auto_nb(const sxc_pair& p) {
  int nb = SXC_MIN_NB;
  while ((1 << nb) < p.tlen + p.qlen && nb < 30) ++nb;
  return nb;
}

int group_pairs(const sxc_pair* pairs, int32_t n, const sxc_params* prm, std::vector<Group>& gs) {
  gs.clear();
  if (!prm || n < 0 || (n > 0 && !pairs)) return fail(SXC_ERR_ARG, "null argument");
  if (prm->N != 0) {
    gs.resize(1);
    gs[0].nb = sxc::nb_of(prm->N);
    return SXC_OK;
  }
  std::map<int, std::vector<int32_t>> by;
  for (int32_t k = 0; k < n; ++k) {
    if (pairs[k].tlen < 0 || pairs[k].qlen < 0) return fail(SXC_ERR_ARG, "negative length");
    const int nb = auto_nb(pairs[k]);
    if (nb > SXC_MAX_NB)
      return fail(SXC_ERR_ARG, "pair " + std::to_string(k) + " needs N = 2^" + std::to_string(nb) +
                                   "; FFT lengths above 32768 are not supported");
    by[nb].push_back(k);
  }
  for (auto& kv : by) {
    Group g;
    g.nb = kv.first;
    g.idx = std::move(kv.second);
    gs.push_back(std::move(g));
  }
  return SXC_OK;
}

// Prepares every group's sub-batch (its pairs and tails gathered in input order).
int groups_prepare(std::vector<Group>& gs, const sxc_pair* pairs, int32_t n, const sxc_params* prm,
                   const sxc_signal_tail* t_tail, const sxc_signal_tail* q_tail) {
  for (Group& g : gs) {
    if (g.idx.empty() && gs.size() == 1 && prm->N != 0) {
      const int rc = batch_prepare(*g.b, pairs, n, prm, t_tail, q_tail);
      if (rc) return rc;
      continue;
    }
    std::vector<sxc_pair> sub(g.idx.size());
    std::vector<sxc_signal_tail> st, sq;
    for (size_t i = 0; i < g.idx.size(); ++i) sub[i] = pairs[g.idx[i]];
    if (t_tail) {
      st.resize(g.idx.size());
      for (size_t i = 0; i < g.idx.size(); ++i) st[i] = t_tail[g.idx[i]];
    }
    if (q_tail) {
      sq.resize(g.idx.size());
      for (size_t i = 0; i < g.idx.size(); ++i) sq[i] = q_tail[g.idx[i]];
    }
    sxc_params p = *prm;
    p.N = 1 << g.nb;
    const int rc = batch_prepare(*g.b, sub.data(), (int32_t)sub.size(), &p, t_tail ? st.data() : nullptr,
                                 q_tail ? sq.data() : nullptr);
    if (rc) return rc;
  }
  return SXC_OK;
}

// Runs every group; with `upload` the group's inputs are copied first.  All groups are
// queued on their own streams before the first is waited for, so length classes overlap.
int groups_run(sxc_ctx* ctx, std::vector<Group>& gs, sxc_timing& t, bool upload = false) {
  t.spectra_ms = t.correlate_ms = t.scan_ms = t.finalize_ms = 0;
  t.near_threshold = 0;
  for (Group& g : gs) {
    const auto t0 = std::chrono::steady_clock::now();
    int rc = upload ? batch_upload(ctx, *g.b) : SXC_OK;
    t.upload_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (!rc) rc = batch_begin(ctx, *g.b);
    if (rc) return rc;
  }
  for (Group& g : gs) {
    const int rc = batch_end(ctx, *g.b);
    if (rc) return rc;
    t.device_ms += g.b->device_ms;
    t.graph |= g.b->last_graph ? 1 : 0;
    t.spectra_ms += g.b->spectra_ms;
    t.correlate_ms += g.b->correlate_ms;
    t.scan_ms += g.b->scan_ms;
    t.finalize_ms += g.b->finalize_ms;
    t.near_threshold += g.b->resolved;
  }
  t.pair_ms = t.correlate_ms + t.scan_ms;
  return SXC_OK;
}

int64_t groups_results(const std::vector<Group>& gs) {
  int64_t n = 0;
  for (const Group& g : gs) n += g.b->n_results;
  return n;
}

// Records of all groups in input pair order.
int groups_download(sxc_ctx* ctx, std::vector<Group>& gs, int32_t n, sxc_result* out, int64_t cap, int64_t* n_out) {
  if (gs.size() == 1 && gs[0].idx.empty()) return batch_download(ctx, *gs[0].b, out, cap, n_out);
  *n_out = groups_results(gs);
  if (*n_out > cap) return fail(SXC_ERR_CAPACITY, "output capacity too small");
  hipStream_t s = ctx->stream;
  std::vector<std::vector<sxc_result>> recs(gs.size());
  std::vector<std::vector<int64_t>> offs(gs.size());
  std::vector<int32_t> where((size_t)n * 2);  // (group, entry) of every input pair
  for (size_t gi = 0; gi < gs.size(); ++gi) {
    sxc_batch_dev& b = *gs[gi].b;
    const int total = (int)b.hb.pairs.size() * b.prm.strands;
    recs[gi].resize((size_t)b.n_results);
    offs[gi].resize((size_t)total + 1);
    if (b.n_results > 0)
      HIPCHK(hipMemcpyAsync(recs[gi].data(), b.final_.p, (size_t)b.n_results * sizeof(sxc_result),
                            hipMemcpyDeviceToHost, s));
    if (total > 0 && b.n_results > 0)
      HIPCHK(hipMemcpyAsync(offs[gi].data(), b.offsets.p, (size_t)total * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    else
      std::fill(offs[gi].begin(), offs[gi].end(), 0);
    offs[gi][(size_t)total] = b.n_results;
    for (size_t i = 0; i < gs[gi].idx.size(); ++i) {
      where[2 * (size_t)gs[gi].idx[i]] = (int32_t)gi;
      where[2 * (size_t)gs[gi].idx[i] + 1] = (int32_t)i;
    }
  }
  HIPCHK(hipStreamSynchronize(s));
  int64_t at = 0;
  for (int32_t k = 0; k < n; ++k) {
    const int gi = where[2 * (size_t)k], i = where[2 * (size_t)k + 1];
    const int strands = gs[(size_t)gi].b->prm.strands;
    const int64_t b0 = offs[(size_t)gi][(size_t)i * strands], b1 = offs[(size_t)gi][(size_t)(i + 1) * strands];
    if (b1 > b0) memcpy(out + at, recs[(size_t)gi].data() + b0, (size_t)(b1 - b0) * sizeof(sxc_result));
    at += b1 - b0;
  }
  return SXC_OK;
}

// ---------------------------------------------------------------- autotuner
//
// Geometry codes (sxc_fft.h) <-> the ABI's sxc_geometry.

sxc_geometry geometry_of(int g) {
  return sxc_geometry{sxc::geo_sa(g), sxc::geo_km(g), sxc::geo_occ(g), sxc::geo_lean(g)};
}
//
// Times every compiled geometry of the pair kernel and the target-spectra kernel for one FFT
// length on a synthetic batch shaped like the slave's (target and query chunks of N/2 bases,
// both strands, half the pairs carrying a planted homology), checks that all geometries
// accept the same lags, and makes the fastest the context's choice.  Called with ctx->mu held.

uint64_t tune_rand(uint64_t& st) {
  st = st * 6364136223846793005ull + 1442695040888963407ull;
  return st >> 33;
}

int tune_class(sxc_ctx* ctx, int nb, std::vector<sxc_tune_entry>& out) {
  const int N = 1 << nb, L = N / 2;
  const int P = std::min(16384, std::max(256, (1 << 24) >> nb));
  std::vector<std::string> T((size_t)P), Q((size_t)P);
  uint64_t st = 0x5eed0000ull + (uint64_t)nb;
  const char acgt[4] = {'A', 'C', 'G', 'T'};
  auto base = [&]() {
    const uint64_t r = tune_rand(st) % 1000;  // GC 0.41
    return r < 295 ? 'A' : r < 500 ? 'C' : r < 705 ? 'G' : 'T';
  };
  for (int k = 0; k < P; ++k) {
    T[(size_t)k].resize((size_t)L);
    Q[(size_t)k].resize((size_t)L);
    for (int i = 0; i < L; ++i) T[(size_t)k][(size_t)i] = base();
    for (int i = 0; i < L; ++i) Q[(size_t)k][(size_t)i] = base();
    if (k % 2 == 0) {  // a homology of L/4 bases at 90% identity
      const int len = L / 4, at_t = (int)(tune_rand(st) % (uint64_t)(L - len)), at_q = (int)(tune_rand(st) % (uint64_t)(L - len));
      for (int i = 0; i < len; ++i)
        Q[(size_t)k][(size_t)(at_q + i)] =
            (tune_rand(st) % 10 == 0) ? acgt[tune_rand(st) % 4] : T[(size_t)k][(size_t)(at_t + i)];
    }
  }
  std::vector<sxc_pair> pairs((size_t)P);
  for (int k = 0; k < P; ++k) {
    sxc_pair& p = pairs[(size_t)k];
    p = sxc_pair{};
    p.tseq = T[(size_t)k].data();
    p.tlen = L;
    p.qseq = Q[(size_t)k].data();
    p.qlen = L;
    p.t_record = k;
    p.q_record = k;
    p.q_record_len = L;
  }
  sxc_params prm{};
  prm.N = N;
  prm.q_chunk = L;
  prm.strands = 2;
  prm.cutoff = 1.8;
  prm.cutoff_fast = 2.9;
  prm.min_prob = 0.99;
  prm.target_total = (double)L * P;
  sxc_batch_dev b;
  int rc = batch_prepare(b, pairs.data(), P, &prm);
  if (!rc) rc = batch_upload(ctx, b);
  if (!rc) rc = batch_alloc(ctx, b);
  if (rc) {
    b.release();
    return rc;
  }
  const int total = 2 * P;
  int geos[8];
  const int ng = sxc::fft_geometries(nb, geos, 8);
  std::vector<int32_t> ref_n, got_n((size_t)total);
  std::vector<uint16_t> ref_c, got_c((size_t)total * N);
  const size_t first = out.size();
  double best_c = 1e300, best_s = 1e300;
  int pick_c = nb, pick_s = nb;
  const int reps = 3;
  for (int gi = 0; gi < ng && !rc; ++gi) {
    const int g = geos[gi];
    ctx->geo_spec[nb] = ctx->geo_corr[nb] = g;
    double ms_s = 0, ms_c = 0;
    for (int r = 0; r <= reps && !rc; ++r) {  // the first repetition warms up
      if (hipEventRecord(b.ev[0], b.s) != hipSuccess) rc = fail(SXC_ERR_HIP, "hipEventRecord");
      if (!rc) rc = launch_target_spectra(ctx, b);
      if (!rc && hipEventRecord(b.ev[1], b.s) != hipSuccess) rc = fail(SXC_ERR_HIP, "hipEventRecord");
      if (!rc) rc = launch_correlate(ctx, b);
      if (!rc && hipEventRecord(b.ev[2], b.s) != hipSuccess) rc = fail(SXC_ERR_HIP, "hipEventRecord");
      if (!rc && hipStreamSynchronize(b.s) != hipSuccess) rc = fail(SXC_ERR_HIP, "hipStreamSynchronize");
      if (!rc && r > 0) {
        ms_s += elapsed(b.ev[0], b.ev[1]);
        ms_c += elapsed(b.ev[1], b.ev[2]);
      }
    }
    if (rc) break;
    // every geometry must accept the same lags (the same FFTReal arithmetic, bit for bit)
    if (hipMemcpy(got_n.data(), b.ncand.p, got_n.size() * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(got_c.data(), b.cands.p, got_c.size() * sizeof(uint16_t), hipMemcpyDeviceToHost) != hipSuccess) {
      rc = fail(SXC_ERR_HIP, "hipMemcpy");
      break;
    }
    for (int ps = 0; ps < total; ++ps)
      std::fill(got_c.begin() + (size_t)ps * N + std::min(N, std::max(0, got_n[(size_t)ps])),
                got_c.begin() + (size_t)(ps + 1) * N, (uint16_t)0);
    if (gi == 0) {
      ref_n = got_n;
      ref_c = got_c;
    } else if (got_n != ref_n || got_c != ref_c) {
      rc = fail(SXC_ERR_HIP, "autotune: FFT geometry " + std::to_string(g) + " of N = " + std::to_string(N) +
                                 " accepts different lags than the default geometry");
      break;
    }
    sxc_tune_entry e{};
    e.N = N;
    e.geo = geometry_of(g);
    e.pairs = P;
    e.correlate_ms = ms_c / reps;
    e.spectra_ms = ms_s / reps;
    out.push_back(e);
    if (e.correlate_ms < best_c) best_c = e.correlate_ms, pick_c = g;
    if (e.spectra_ms < best_s) best_s = e.spectra_ms, pick_s = g;
  }
  b.release();
  if (rc) {
    ctx->geo_corr[nb] = ctx->geo_spec[nb] = nb;
    out.resize(first);
    return rc;
  }
  ctx->geo_corr[nb] = pick_c;
  ctx->geo_spec[nb] = pick_s;
  const sxc_geometry pc = geometry_of(pick_c), ps_ = geometry_of(pick_s);
  for (size_t i = first; i < out.size(); ++i) {
    sxc_tune_entry& e = out[i];
    e.chosen_correlate = memcmp(&e.geo, &pc, sizeof pc) == 0;
    e.chosen_spectra = memcmp(&e.geo, &ps_, sizeof ps_) == 0;
  }
  return SXC_OK;
}

// The compiled geometry of length 2^nb matching `want`.
int find_geometry(int nb, const sxc_geometry& want, int* geo) {
  int geos[8];
  const int ng = sxc::fft_geometries(nb, geos, 8);
  for (int i = 0; i < ng; ++i) {
    const sxc_geometry g = geometry_of(geos[i]);
    if (g.stage_a == want.stage_a && g.round_levels == want.round_levels && g.three_per_cu == want.three_per_cu &&
        g.lean_lds == want.lean_lds) {
      *geo = geos[i];
      return SXC_OK;
    }
  }
  return fail(SXC_ERR_ARG, "no compiled FFT geometry (stage A " + std::to_string(want.stage_a) + ", " +
                               std::to_string(want.round_levels) + "-level rounds, three per CU " +
                               std::to_string(want.three_per_cu) + ", lean LDS " + std::to_string(want.lean_lds) +
                               ") for N = " + std::to_string(1 << nb));
}

}  // namespace

extern "C" {

const char* sxc_last_error(void) { return sxc::g_err.c_str(); }

int sxc_reverse_complement(char* seq, int32_t len) {
  if (len < 0 || (len > 0 && !seq)) return fail(SXC_ERR_ARG, "null argument");
  const sxc_tables& t = sxc::host_tables();
  for (int32_t i = 0, j = len - 1; i < (len + 1) / 2; ++i, --j) {
    const char a = (char)t.rc[(unsigned char)seq[i] & 127], b = (char)t.rc[(unsigned char)seq[j] & 127];
    seq[i] = b;
    seq[j] = a;
  }
  return SXC_OK;
}

double sxc_dna_equal(char a, char b) {
  const sxc_tables& t = sxc::host_tables();
  return t.equal[t.cls[(unsigned char)a & 127]][t.cls[(unsigned char)b & 127]];
}
const char* sxc_version(void) { return "sxc 0.1 gfx950"; }

int sxc_create(sxc_ctx** out, int device) {
  if (!out) return fail(SXC_ERR_ARG, "null out");
  *out = nullptr;
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(SXC_ERR_ARG, "no such HIP device");
  sxc_ctx* c = new sxc_ctx();
  c->device = device;
  HIPCHK(hipSetDevice(device));
  const int src = sxc::static_init(device);
  if (src) {
    delete c;
    return src;
  }
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (auto& e : c->ev) HIPCHK(hipEventCreate(&e));
  sxc_tables tab;
  sxc::build_tables(&tab);
  HIPCHK(hipMalloc(&c->d_tab, sizeof tab));
  HIPCHK(hipMemcpy(c->d_tab, &tab, sizeof tab, hipMemcpyHostToDevice));
  std::vector<float> osc(24576);
  sxc::build_osc(osc.data());
  HIPCHK(hipMalloc(&c->d_osc, osc.size() * sizeof(float)));
  HIPCHK(hipMemcpy(c->d_osc, osc.data(), osc.size() * sizeof(float), hipMemcpyHostToDevice));
  for (int nb = SXC_MIN_NB; nb <= SXC_MAX_NB; ++nb) {
    const int len = std::max(1, sxc::ent_lut_len(nb));
    std::vector<double> v(len, 0.);
    if (sxc::ent_lut_len(nb) > 0) sxc::build_ent_lut(nb, v.data());
    HIPCHK(hipMalloc(&c->d_ent[nb], len * sizeof(double)));
    HIPCHK(hipMemcpy(c->d_ent[nb], v.data(), len * sizeof(double), hipMemcpyHostToDevice));
  }
  *out = c;
  return SXC_OK;
}

void sxc_destroy(sxc_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  c->scratch.release();
  for (sxc_batch_dev& b : c->scratch_nb) b.release();
  for (auto& kv : c->ptables) kv.second.release();
  for (DevBuf* b : {&c->p0, &c->p1, &c->p2, &c->p3, &c->p4, &c->p5, &c->p6}) b->release();
  (void)hipFree(c->d_tab);
  (void)hipFree(c->d_osc);
  for (auto p : c->d_ent)
    if (p) (void)hipFree(p);
  for (auto e : c->ev)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

int sxc_align_batch(sxc_ctx* ctx, const sxc_pair* pairs, int32_t n, const sxc_params* prm, sxc_result* out,
                    int64_t cap, int64_t* n_out) {
  if (!ctx || !n_out || (cap > 0 && !out)) return fail(SXC_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> g(ctx->mu);
  int rc = set_device(ctx);
  if (rc) return rc;
  std::vector<Group> gs;
  rc = group_pairs(pairs, n, prm, gs);
  if (rc) return rc;
  for (Group& gr : gs) {
    gr.b = (prm->N != 0) ? &ctx->scratch : &ctx->scratch_nb[gr.nb];
    gr.b->use_graph = false;  // new inputs every call: launched kernel by kernel
  }
  rc = groups_prepare(gs, pairs, n, prm, nullptr, nullptr);
  if (rc) return rc;
  // uploads overlap the device work of earlier length classes
  const auto t0 = std::chrono::steady_clock::now();
  sxc_timing t{};
  rc = groups_run(ctx, gs, t, true);
  if (rc) return rc;
  const auto t2 = std::chrono::steady_clock::now();
  rc = groups_download(ctx, gs, n, out, cap, n_out);
  const auto t3 = std::chrono::steady_clock::now();
  auto ms = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
  };
  t.download_ms = ms(t2, t3);
  t.total_ms = ms(t0, t3);
  ctx->last = t;
  return rc;
}

int sxc_align_pair(sxc_ctx* ctx, const sxc_pair* pair, const sxc_params* prm, sxc_result* out, int32_t cap,
                   int32_t* n_out) {
  if (!n_out) return fail(SXC_ERR_ARG, "null n_out");
  int64_t n = 0;
  const int rc = sxc_align_batch(ctx, pair, 1, prm, out, cap, &n);
  *n_out = (int32_t)n;
  return rc;
}

int sxc_autotune(sxc_ctx* ctx, int32_t N_lo, int32_t N_hi, sxc_tune_entry* out, int32_t cap, int32_t* n_out) {
  if (!ctx || !n_out || (cap > 0 && !out)) return fail(SXC_ERR_ARG, "null argument");
  *n_out = 0;
  int lo, hi;
  int rc = check_N(N_lo, &lo);
  if (!rc) rc = check_N(N_hi, &hi);
  if (rc) return rc;
  std::lock_guard<std::mutex> g(ctx->mu);
  rc = set_device(ctx);
  if (rc) return rc;
  std::vector<sxc_tune_entry> res;
  for (int nb = lo; nb <= hi && !rc; ++nb) rc = tune_class(ctx, nb, res);
  if (rc) return rc;
  *n_out = (int32_t)res.size();
  if ((int64_t)res.size() > cap) return fail(SXC_ERR_CAPACITY, "output capacity too small");
  std::copy(res.begin(), res.end(), out);
  return SXC_OK;
}

int sxc_set_geometry(sxc_ctx* ctx, int32_t N, int32_t kernel, const sxc_geometry* want) {
  if (!ctx || !want) return fail(SXC_ERR_ARG, "null argument");
  if (kernel != SXC_KERNEL_PAIRS && kernel != SXC_KERNEL_SPECTRA) return fail(SXC_ERR_ARG, "unknown kernel");
  int nb, geo = 0;
  int rc = check_N(N, &nb);
  if (!rc) rc = find_geometry(nb, *want, &geo);
  if (rc) return rc;
  std::lock_guard<std::mutex> g(ctx->mu);
  (kernel == SXC_KERNEL_PAIRS ? ctx->geo_corr : ctx->geo_spec)[nb] = geo;
  return SXC_OK;
}

int sxc_get_geometry(sxc_ctx* ctx, int32_t N, int32_t kernel, sxc_geometry* out) {
  if (!ctx || !out) return fail(SXC_ERR_ARG, "null argument");
  if (kernel != SXC_KERNEL_PAIRS && kernel != SXC_KERNEL_SPECTRA) return fail(SXC_ERR_ARG, "unknown kernel");
  int nb;
  const int rc = check_N(N, &nb);
  if (rc) return rc;
  std::lock_guard<std::mutex> g(ctx->mu);
  const int geo = (kernel == SXC_KERNEL_PAIRS ? ctx->geo_corr : ctx->geo_spec)[nb];
  *out = geometry_of(geo);
  return SXC_OK;
}

int sxc_last_timing(sxc_ctx* ctx, sxc_timing* out) {
  if (!ctx || !out) return fail(SXC_ERR_ARG, "null argument");
  *out = ctx->last;
  return SXC_OK;
}

// ---------------------------------------------------------------- resident batches

typedef struct sxc_batch {
  sxc_ctx* ctx;
  int32_t n = 0;
  std::vector<Group> groups;
  std::vector<std::unique_ptr<sxc_batch_dev>> owned;
  void release() {
    for (auto& b : owned) b->release();
  }
} sxc_batch;

int sxc_batch_create(sxc_ctx* ctx, const sxc_pair* pairs, int32_t n, const sxc_params* prm, sxc_batch** out) {
  return sxc_batch_create_tails(ctx, pairs, n, prm, nullptr, nullptr, out);
}

int sxc_batch_create_tails(sxc_ctx* ctx, const sxc_pair* pairs, int32_t n, const sxc_params* prm,
                           const sxc_signal_tail* t_tail, const sxc_signal_tail* q_tail, sxc_batch** out) {
  if (!ctx || !out) return fail(SXC_ERR_ARG, "null argument");
  *out = nullptr;
  std::lock_guard<std::mutex> g(ctx->mu);
  int rc = set_device(ctx);
  if (rc) return rc;
  sxc_batch* bt = new sxc_batch();
  bt->ctx = ctx;
  bt->n = n;
  rc = group_pairs(pairs, n, prm, bt->groups);
  if (!rc) {
    for (Group& gr : bt->groups) {
      bt->owned.push_back(std::make_unique<sxc_batch_dev>());
      gr.b = bt->owned.back().get();
    }
    rc = groups_prepare(bt->groups, pairs, n, prm, t_tail, q_tail);
  }
  for (size_t i = 0; !rc && i < bt->groups.size(); ++i) rc = batch_upload(ctx, *bt->groups[i].b);
  if (rc) {
    bt->release();
    delete bt;
    return rc;
  }
  *out = bt;
  return SXC_OK;
}

int sxc_batch_run(sxc_batch* bt) {
  if (!bt) return fail(SXC_ERR_ARG, "null batch");
  std::lock_guard<std::mutex> g(bt->ctx->mu);
  int rc = set_device(bt->ctx);
  if (rc) return rc;
  sxc_timing t{};
  const auto t0 = std::chrono::steady_clock::now();
  rc = groups_run(bt->ctx, bt->groups, t);
  t.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  bt->ctx->last = t;
  return rc;
}

int sxc_batch_set_graph(sxc_batch* bt, int32_t on) {
  if (!bt) return fail(SXC_ERR_ARG, "null batch");
  std::lock_guard<std::mutex> g(bt->ctx->mu);
  for (Group& gr : bt->groups) gr.b->use_graph = on != 0;
  return SXC_OK;
}

int sxc_batch_results(sxc_batch* bt, sxc_result* out, int64_t cap, int64_t* n_out) {
  if (!bt || !n_out) return fail(SXC_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> g(bt->ctx->mu);
  int rc = set_device(bt->ctx);
  if (rc) return rc;
  return groups_download(bt->ctx, bt->groups, bt->n, out, cap, n_out);
}

int sxc_batch_stats(sxc_batch* bt, int64_t* candidates, int64_t* runs) {
  if (!bt || !candidates || !runs) return fail(SXC_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> g(bt->ctx->mu);
  int64_t c = 0, r = 0;
  for (const Group& gr : bt->groups) {
    const int total = (int)gr.b->hb.pairs.size() * gr.b->prm.strands;
    std::vector<int32_t> st((size_t)total * 2);
    if (total > 0) {
      HIPCHK(hipMemcpyAsync(st.data(), gr.b->stats.p, st.size() * sizeof(int32_t), hipMemcpyDeviceToHost,
                            bt->ctx->stream));
      HIPCHK(hipStreamSynchronize(bt->ctx->stream));
    }
    for (int i = 0; i < total; ++i) {
      c += st[2 * i];
      r += st[2 * i + 1];
    }
  }
  *candidates = c;
  *runs = r;
  return SXC_OK;
}

int sxc_batch_lags(sxc_batch* bt, int32_t pair, int32_t strand, int32_t* out, int32_t cap, int32_t* n_out) {
  if (!bt || !n_out || (cap > 0 && !out)) return fail(SXC_ERR_ARG, "null argument");
  if (pair < 0 || pair >= bt->n) return fail(SXC_ERR_ARG, "no such pair");
  std::lock_guard<std::mutex> g(bt->ctx->mu);
  int rc = set_device(bt->ctx);
  if (rc) return rc;
  // the pair's length class and its index there
  const Group* gr = nullptr;
  int32_t idx = pair;
  for (const Group& x : bt->groups) {
    if (x.idx.empty()) {
      gr = &x;
      break;
    }
    auto it = std::lower_bound(x.idx.begin(), x.idx.end(), pair);
    if (it != x.idx.end() && *it == pair) {
      gr = &x;
      idx = (int32_t)(it - x.idx.begin());
      break;
    }
  }
  if (!gr) return fail(SXC_ERR_ARG, "no such pair");
  const sxc_batch_dev& b = *gr->b;
  if (strand < 0 || strand >= b.prm.strands) return fail(SXC_ERR_ARG, "no such strand");
  if (b.runs == 0) return fail(SXC_ERR_ARG, "the batch has not run");
  const int64_t ps = (int64_t)idx * b.prm.strands + strand;
  hipStream_t s = bt->ctx->stream;
  int32_t n = 0;
  HIPCHK(hipMemcpyAsync(&n, b.ncand.as<int32_t>() + ps, sizeof n, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  *n_out = n;
  if (n > cap) return fail(SXC_ERR_CAPACITY, "output capacity too small");
  std::vector<uint16_t> lags((size_t)n);
  if (n > 0)
    HIPCHK(hipMemcpyAsync(lags.data(), b.cands.as<uint16_t>() + ps * b.prm.N, (size_t)n * sizeof(uint16_t),
                          hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int32_t i = 0; i < n; ++i) out[i] = lags[(size_t)i];
  return SXC_OK;
}

void sxc_batch_destroy(sxc_batch* bt) {
  if (!bt) return;
  {
    std::lock_guard<std::mutex> g(bt->ctx->mu);
    (void)hipSetDevice(bt->ctx->device);
    (void)hipStreamSynchronize(bt->ctx->stream);
    bt->release();
  }
  delete bt;
}

// ---------------------------------------------------------------- plain API

int sxc_encode_signal(sxc_ctx* ctx, const char* seq, int32_t len, int32_t N, float* out) {
  if (!ctx || !out || len < 0 || (len > 0 && !seq)) return fail(SXC_ERR_ARG, "null argument");
  int nb;
  int rc = check_N(N, &nb);
  if (rc) return rc;
  if (len > N) return fail(SXC_ERR_ARG, "sequence longer than N");
  std::lock_guard<std::mutex> g(ctx->mu);
  rc = set_device(ctx);
  if (rc) return rc;
  sxc_chunk ck{0, len, 0};
  const int fl = sxc::chunk_flags((const unsigned char*)seq, len);
  if (fl < 0) return fail(SXC_ERR_ARG, "sequence byte >= 128");
  ck.flags = fl & 1;  // the encoder only asks for B/D/H/V letters
  hipStream_t s = ctx->stream;
  std::vector<float> erow;
  if (ck.flags & 1) {  // row 0 of a one-row entropy table
    erow.resize(SXC_ENT_ROW);
    sxc::window_entropy_row((const unsigned char*)seq, len, N, erow.data());
    HIPCHK(ctx->p3.ensure(erow.size() * sizeof(float)));
    HIPCHK(hipMemcpyAsync(ctx->p3.p, erow.data(), erow.size() * sizeof(float), hipMemcpyHostToDevice, s));
  }
  HIPCHK(ctx->p0.ensure((size_t)len + 16));
  HIPCHK(ctx->p1.ensure(sizeof ck));
  HIPCHK(ctx->p2.ensure((size_t)4 * N * sizeof(float)));
  if (len) HIPCHK(hipMemcpyAsync(ctx->p0.p, seq, len, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(ctx->p1.p, &ck, sizeof ck, hipMemcpyHostToDevice, s));
  sxc::EncodeArgs a{ctx->d_tab, ctx->d_ent[nb], erow.empty() ? nullptr : ctx->p3.as<float>(),
                    ctx->p0.as<unsigned char>(), ctx->p1.as<sxc_chunk>(), 0, nullptr, ctx->p2.as<float>()};
  HIPCHK(sxc::launch_encode(nb, a, 1, s));
  HIPCHK(hipMemcpyAsync(out, ctx->p2.p, (size_t)4 * N * sizeof(float), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));  // erow is a host temporary
  return SXC_OK;
}

int sxc_fft_forward(sxc_ctx* ctx, const float* in, float* out, int32_t N, int32_t batch) {
  if (!ctx || !in || !out || batch < 0) return fail(SXC_ERR_ARG, "null argument");
  int nb;
  int rc = check_N(N, &nb);
  if (rc) return rc;
  if (batch == 0) return SXC_OK;
  std::lock_guard<std::mutex> g(ctx->mu);
  rc = set_device(ctx);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  // each signal is transformed as channel A of a 4-channel item
  const size_t item = (size_t)4 * N;
  HIPCHK(ctx->p0.ensure(item * batch * sizeof(float)));
  HIPCHK(ctx->p1.ensure(item * batch * sizeof(float)));
  HIPCHK(ctx->p2.ensure((size_t)batch * 2 * sizeof(int32_t)));
  HIPCHK(ctx->p3.ensure(sizeof(sxc_chunk)));
  HIPCHK(hipMemsetAsync(ctx->p0.p, 0, item * batch * sizeof(float), s));
  HIPCHK(hipMemcpy2DAsync(ctx->p0.p, item * sizeof(float), in, (size_t)N * sizeof(float), (size_t)N * sizeof(float),
                          batch, hipMemcpyHostToDevice, s));
  std::vector<int32_t> idx(batch);
  for (int i = 0; i < batch; ++i) idx[i] = i;
  std::vector<int32_t> zero(batch, 0);
  sxc_chunk ck{0, 0, 0};
  HIPCHK(hipMemcpyAsync(ctx->p3.p, &ck, sizeof ck, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(ctx->p2.p, zero.data(), batch * sizeof(int32_t), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(ctx->p2.as<int32_t>() + batch, idx.data(), batch * sizeof(int32_t), hipMemcpyHostToDevice, s));
  sxc::SpecArgs a{};
  a.tab = ctx->d_tab;
  a.osc = ctx->d_osc;
  a.ent_lut = ctx->d_ent[nb];
  a.seq = nullptr;
  a.chunks = ctx->p3.as<sxc_chunk>();
  a.which = ctx->p2.as<int32_t>();
  a.n = batch;
  a.spectra = ctx->p1.as<float>();
  a.dst = ctx->p2.as<int32_t>() + batch;
  a.signals = ctx->p0.as<float>();
  HIPCHK(sxc::launch_spectra(nb, a, s));
  HIPCHK(hipMemcpy2DAsync(out, (size_t)N * sizeof(float), ctx->p1.p, item * sizeof(float), (size_t)N * sizeof(float),
                          batch, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return SXC_OK;
}

int sxc_xcorr(sxc_ctx* ctx, const float* tsig, const float* qsig, int32_t N, float* xc) {
  if (!ctx || !tsig || !qsig || !xc) return fail(SXC_ERR_ARG, "null argument");
  int nb;
  int rc = check_N(N, &nb);
  if (rc) return rc;
  std::lock_guard<std::mutex> g(ctx->mu);
  rc = set_device(ctx);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  const size_t sig = (size_t)4 * N * sizeof(float);
  HIPCHK(ctx->p0.ensure(sig));
  HIPCHK(ctx->p1.ensure(sig));
  HIPCHK(ctx->p2.ensure(sig));
  HIPCHK(ctx->p3.ensure((size_t)N * sizeof(float)));
  HIPCHK(ctx->p4.ensure(2 * sizeof(int32_t) + sizeof(sxc_chunk)));
  HIPCHK(hipMemcpyAsync(ctx->p0.p, tsig, sig, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(ctx->p1.p, qsig, sig, hipMemcpyHostToDevice, s));
  int32_t zz[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(ctx->p4.p, zz, sizeof zz, hipMemcpyHostToDevice, s));
  sxc_chunk ck{0, 0, 0};
  HIPCHK(hipMemcpyAsync((char*)ctx->p4.p + 8, &ck, sizeof ck, hipMemcpyHostToDevice, s));
  sxc::SpecArgs sa{};
  sa.tab = ctx->d_tab;
  sa.osc = ctx->d_osc;
  sa.ent_lut = ctx->d_ent[nb];
  sa.chunks = (const sxc_chunk*)((char*)ctx->p4.p + 8);
  sa.which = ctx->p4.as<int32_t>();
  sa.n = 1;
  sa.spectra = ctx->p2.as<float>();
  sa.dst = ctx->p4.as<int32_t>() + 1;
  sa.signals = ctx->p0.as<float>();
  HIPCHK(sxc::launch_spectra(nb, sa, s));
  sxc::XcorrArgs xa{ctx->d_tab, ctx->d_osc, ctx->p2.as<float>(), ctx->p1.as<float>(), ctx->p3.as<float>()};
  HIPCHK(sxc::launch_xcorr(nb, xa, 1, s));
  HIPCHK(hipMemcpyAsync(xc, ctx->p3.p, (size_t)N * sizeof(float), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return SXC_OK;
}

int sxc_matchup(sxc_ctx* ctx, const char* q, int32_t lq, const char* t, int32_t lt, const float* xc, int32_t N,
                double cutoff, sxc_seqmatch* out, int32_t cap, int32_t* n_out) {
  if (!ctx || !xc || !n_out || (cap > 0 && !out)) return fail(SXC_ERR_ARG, "null argument");
  int nb;
  int rc = check_N(N, &nb);
  if (rc) return rc;
  sxc_pair p{t, lt, q, lq, 0, 0, 0, 0, lq, 0};
  sxc::HostBatch hb;
  std::string e = sxc::pack_batch(&p, 1, N, hb);
  if (!e.empty()) return fail(SXC_ERR_ARG, e);
  std::lock_guard<std::mutex> g(ctx->mu);
  rc = set_device(ctx);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  // every accepted lag yields at most N/4 + 1 runs
  const int slot = std::max(64, (int)std::min<int64_t>((int64_t)N * (N / 4 + 1), 1 << 24));
  HIPCHK(ctx->p0.ensure(hb.seq.size()));
  HIPCHK(ctx->p1.ensure(2 * sizeof(sxc_chunk) + sizeof(sxc_pair_dev)));
  HIPCHK(ctx->p2.ensure((size_t)N * sizeof(float) + (size_t)N * sizeof(uint16_t) + 2 * sizeof(int32_t)));
  HIPCHK(ctx->p3.ensure((size_t)slot * 3 * sizeof(int32_t)));
  HIPCHK(ctx->p4.ensure((size_t)slot * sizeof(uint64_t)));
  HIPCHK(hipMemcpyAsync(ctx->p0.p, hb.seq.data(), hb.seq.size(), hipMemcpyHostToDevice, s));
  char stage[2 * sizeof(sxc_chunk) + sizeof(sxc_pair_dev)];
  memcpy(stage, &hb.tchunks[0], sizeof(sxc_chunk));
  memcpy(stage + sizeof(sxc_chunk), &hb.qchunks[0], sizeof(sxc_chunk));
  sxc_pair_dev pd = hb.pairs[0];
  pd.t_chunk = 0;
  pd.q_chunk = 0;
  memcpy(stage + 2 * sizeof(sxc_chunk), &pd, sizeof pd);
  HIPCHK(hipMemcpyAsync(ctx->p1.p, stage, sizeof stage, hipMemcpyHostToDevice, s));
  float* d_xc = ctx->p2.as<float>();
  uint16_t* d_cands = (uint16_t*)(d_xc + N);
  int32_t* d_misc = (int32_t*)(d_cands + N);  // {ncand, count}
  HIPCHK(hipMemcpyAsync(d_xc, xc, (size_t)N * sizeof(float), hipMemcpyHostToDevice, s));
  HIPCHK(sxc::launch_peaks(nb, d_xc, cutoff, d_cands, d_misc, s));
  sxc::ScanArgs a{};
  a.tab = ctx->d_tab;
  a.seq = ctx->p0.as<unsigned char>();
  a.tchunks = (const sxc_chunk*)ctx->p1.p;
  a.qchunks = (const sxc_chunk*)ctx->p1.p + 1;
  a.pairs = (const sxc_pair_dev*)((const char*)ctx->p1.p + 2 * sizeof(sxc_chunk));
  a.cands = d_cands;
  a.ncand = d_misc;
  a.n_fft = N;
  a.strands = 1;
  a.z_skip = HUGE_VAL;
  a.mode = 1;
  a.runs = ctx->p3.as<int32_t>();
  a.keys = ctx->p4.as<uint64_t>();
  a.counts = d_misc + 1;
  a.cap = slot;
  HIPCHK(sxc::launch_scan(nb, a, 1, s));
  int32_t cnt = 0;
  HIPCHK(hipMemcpyAsync(&cnt, d_misc + 1, sizeof cnt, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  *n_out = cnt;
  if (cnt > slot) return fail(SXC_ERR_CAPACITY, "run buffer overflow");
  if (cnt > cap) return fail(SXC_ERR_CAPACITY, "output capacity too small");
  // runs come back unordered; order them by (lag rank, start) on the GPU
  HIPCHK(ctx->p6.ensure((size_t)std::max(cnt, 1) * 3 * sizeof(int32_t)));
  HIPCHK(sxc::launch_order_runs(ctx->p3.as<int32_t>(), ctx->p4.as<uint64_t>(), cnt, ctx->p6.as<int32_t>(), s));
  if (cnt) HIPCHK(hipMemcpyAsync(out, ctx->p6.p, (size_t)cnt * 3 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return SXC_OK;
}

int sxc_findtop(sxc_ctx* ctx, const float* xc, int32_t N, double cutoff, int32_t* out, int32_t cap, int32_t* n_out) {
  if (!ctx || !xc || !n_out || (cap > 0 && !out)) return fail(SXC_ERR_ARG, "null argument");
  int nb;
  int rc = check_N(N, &nb);
  if (rc) return rc;
  std::lock_guard<std::mutex> g(ctx->mu);
  rc = set_device(ctx);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  HIPCHK(ctx->p2.ensure((size_t)N * sizeof(float) + (size_t)N * sizeof(uint16_t) + 2 * sizeof(int32_t)));
  float* d_xc = ctx->p2.as<float>();
  uint16_t* d_cands = (uint16_t*)(d_xc + N);
  int32_t* d_n = (int32_t*)(d_cands + N);
  HIPCHK(hipMemcpyAsync(d_xc, xc, (size_t)N * sizeof(float), hipMemcpyHostToDevice, s));
  HIPCHK(sxc::launch_peaks(nb, d_xc, cutoff, d_cands, d_n, s));
  int32_t n = 0;
  HIPCHK(hipMemcpyAsync(&n, d_n, sizeof n, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  *n_out = n;
  if (n > cap) return fail(SXC_ERR_CAPACITY, "output capacity too small");
  std::vector<uint16_t> lags((size_t)n);
  if (n > 0) HIPCHK(hipMemcpyAsync(lags.data(), d_cands, (size_t)n * sizeof(uint16_t), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int32_t i = 0; i < n; ++i) out[i] = lags[(size_t)i];
  return SXC_OK;
}

int sxc_match_probability(sxc_ctx* ctx, const char* t, int32_t lt, const char* q, int32_t lq, const sxc_seqmatch* m,
                          int32_t n, double target_size, double* prob, double* ident) {
  if (!ctx || n < 0 || (n > 0 && (!m || !prob || !ident || !t || !q))) return fail(SXC_ERR_ARG, "null argument");
  for (int i = 0; i < n; ++i) {
    if (m[i].len <= 0 || m[i].start_target < 0 || m[i].start_query < 0 || m[i].start_target + m[i].len > lt ||
        m[i].start_query + m[i].len > lq)
      return fail(SXC_ERR_ARG, "match outside the sequences");
  }
  for (int i = 0; i < lt; ++i)
    if ((unsigned char)t[i] >= 128) return fail(SXC_ERR_ARG, "sequence byte >= 128");
  for (int i = 0; i < lq; ++i)
    if ((unsigned char)q[i] >= 128) return fail(SXC_ERR_ARG, "sequence byte >= 128");
  if (n == 0) return SXC_OK;
  std::lock_guard<std::mutex> g(ctx->mu);
  int rc = set_device(ctx);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  HIPCHK(ctx->p0.ensure((size_t)lt + 16));
  HIPCHK(ctx->p1.ensure((size_t)lq + 16));
  HIPCHK(ctx->p2.ensure((size_t)n * 3 * sizeof(int32_t)));
  HIPCHK(ctx->p3.ensure((size_t)n * 2 * sizeof(double)));
  if (lt) HIPCHK(hipMemcpyAsync(ctx->p0.p, t, lt, hipMemcpyHostToDevice, s));
  if (lq) HIPCHK(hipMemcpyAsync(ctx->p1.p, q, lq, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(ctx->p2.p, m, (size_t)n * 3 * sizeof(int32_t), hipMemcpyHostToDevice, s));
  sxc::ProbArgs a{ctx->d_tab, ctx->p0.as<unsigned char>(), ctx->p1.as<unsigned char>(), ctx->p2.as<int32_t>(), n,
                  target_size, ctx->p3.as<double>(), ctx->p3.as<double>() + n};
  HIPCHK(sxc::launch_prob(a, s));
  HIPCHK(hipMemcpyAsync(prob, ctx->p3.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(ident, ctx->p3.as<double>() + n, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return SXC_OK;
}

}  // extern "C"
