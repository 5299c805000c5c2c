// sxc_static.hip -- sxc_align_pair_static: the one-pair drop-in without device allocation.
//
// sxc_align_pair stages its inputs, the target spectra, the accepted lags and the records in
// device buffers the context allocates (and grows) on demand.  Callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc) use this variant instead: every buffer its kernels need lives in one
// file-scope device array, `scratch`, of SXC_STATIC_BYTES bytes per GPU, which the process
// holds until it ends.  The array's device address is looked up once per device at context
// creation (sxc::static_init, hipGetSymbolAddress) and only that address is used afterwards:
// in host code the array's name is not a device address.
//
// The same kernels as sxc_align_pair run on the same data, so the records are identical.
// Limits: FFT lengths up to 32768 (as everywhere) and at most SXC_STATIC_RECORDS records kept
// per strand (more returns SXC_ERR_CAPACITY; sxc_align_pair has no such limit).  The array
// keeps the last call's contents; whatever a call needs at zero it clears itself.  Calls on
// one GPU share its array: the library serialises them with a per-device lock.
#include <hip/hip_runtime.h>
#include <string.h>

#include <mutex>

#include "sxc_api_internal.h"
#include "sxc_kernels.h"
#include "sxc_tables.h"

using sxc::fail;

namespace {

constexpr size_t SXC_STATIC_BYTES = (size_t)4 << 20;
constexpr int SXC_STATIC_RECORDS = 4096;  // records kept per (pair, strand)
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[SXC_STATIC_BYTES];

// Device address of `scratch` per device (null until looked up) and the lock that serialises
// calls on that device.
void* g_ptr[MAX_DEVICES] = {};
std::mutex g_init_mu;
std::mutex g_dev_mu[MAX_DEVICES];

constexpr size_t up256(size_t x) { return (x + 255) & ~(size_t)255; }

// Byte offsets of the regions inside `scratch` for FFT length N, S strands.
struct Layout {
  size_t tseq, qseq, tch, qch, pair, which, ent, spectra, cands, ncand, counts, stats, offsets, total, maxc, nflag,
      slots, keys, final_, end;
  Layout(int N, int S) {
    size_t at = 0;
    auto take = [&](size_t bytes) {
      const size_t o = at;
      at = up256(at + bytes);
      return o;
    };
    tseq = take((size_t)N + 16);  // + 16: the kernels' aligned word reads run past the end
    qseq = take((size_t)N + 16);
    tch = take(sizeof(sxc_chunk));
    qch = take(sizeof(sxc_chunk));
    pair = take(sizeof(sxc_pair_dev));
    which = take(sizeof(int32_t));
    ent = take(2 * SXC_ENT_ROW * sizeof(float));
    spectra = take((size_t)4 * N * sizeof(float));
    cands = take((size_t)S * N * sizeof(uint16_t));
    ncand = take((size_t)S * sizeof(int32_t));
    counts = take((size_t)S * sizeof(int32_t));
    stats = take((size_t)2 * S * sizeof(int32_t));
    offsets = take((size_t)S * sizeof(int64_t));
    total = take(2 * sizeof(int64_t));
    maxc = take(sizeof(int32_t));
    nflag = take(sizeof(int32_t));
    slots = take((size_t)S * SXC_STATIC_RECORDS * sizeof(sxc_result_dev));
    keys = take((size_t)S * SXC_STATIC_RECORDS * sizeof(uint64_t));
    final_ = take((size_t)S * SXC_STATIC_RECORDS * sizeof(sxc_result_dev));
    end = at;
  }
};

// The chunk's flags as the batch packer sets them (sxc_batch_host.h), and its entropy row
// (row 0 for the target, 1 for the query) when it has B/D/H/V letters.
int chunk_of(const char* s, int len, int N, int row, float* erow, sxc_chunk* c) {
  c->off = 0;
  c->len = len;
  c->flags = sxc::chunk_flags((const unsigned char*)s, len);
  if (c->flags < 0) return fail(SXC_ERR_ARG, "sequence byte >= 128");
  if (c->flags & 1) {
    c->flags |= row << 8;
    sxc::window_entropy_row((const unsigned char*)s, len, N, erow);
  }
  return SXC_OK;
}

}  // namespace

namespace sxc {

// Looks up the device address of `scratch` on the current device (once per device).
int static_init(int device) {
  if (device < 0 || device >= MAX_DEVICES) return fail(SXC_ERR_ARG, "device index beyond the static scratch table");
  std::lock_guard<std::mutex> g(g_init_mu);
  if (g_ptr[device]) return SXC_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(SXC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_ptr[device] = p;
  return SXC_OK;
}

}  // namespace sxc

extern "C" int sxc_align_pair_static(sxc_ctx* ctx, const sxc_pair* pair, const sxc_params* prm, sxc_result* out,
                                     int32_t cap, int32_t* n_out) {
  if (!ctx || !pair || !prm || !n_out || (cap > 0 && !out)) return fail(SXC_ERR_ARG, "null argument");
  const int N = prm->N, S = prm->strands;
  const int nb = sxc::nb_of(N);
  if (nb < SXC_MIN_NB || nb > SXC_MAX_NB) return fail(SXC_ERR_ARG, "N must be a power of two in [512, 32768]");
  if (S != 1 && S != 2) return fail(SXC_ERR_ARG, "strands must be 1 or 2");
  if (prm->q_chunk < 0) return fail(SXC_ERR_ARG, "q_chunk must be >= 0");
  if (prm->prob_table) return fail(SXC_ERR_ARG, "-prob_table is not supported without allocation: use sxc_align_pair");
  const sxc_pair& p = *pair;
  if (p.tlen < 0 || p.qlen < 0 || (p.tlen > 0 && !p.tseq) || (p.qlen > 0 && !p.qseq))
    return fail(SXC_ERR_ARG, "negative length or null sequence");
  if (p.tlen > N || p.qlen > N) return fail(SXC_ERR_ARG, "sequence longer than the FFT length N");
  if (p.tlen > N / 2 && p.qlen > N / 2) return fail(SXC_ERR_ARG, "both chunks longer than N/2");
  const Layout L(N, S);
  if (L.end > SXC_STATIC_BYTES) return fail(SXC_ERR_ARG, "pair does not fit the static scratch");
  float erow[2 * SXC_ENT_ROW] = {};
  sxc_chunk tch, qch;
  int rc = chunk_of(p.tseq, p.tlen, N, 0, erow, &tch);
  if (!rc) rc = chunk_of(p.qseq, p.qlen, N, 1, erow + SXC_ENT_ROW, &qch);
  if (rc) return rc;
  qch.off = (int64_t)(L.qseq - L.tseq);  // both chunks' bytes are addressed from the t region
  const sxc_pair_dev pd{0, 0, p.t_chunk_start, p.q_chunk_start, p.t_record, p.q_record, p.q_record_len, p.fast, -1};
  const int32_t zero = 0;

  std::lock_guard<std::mutex> g(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  if (ctx->device >= MAX_DEVICES || !g_ptr[ctx->device]) return fail(SXC_ERR_HIP, "static scratch not set up");
  std::lock_guard<std::mutex> gd(g_dev_mu[ctx->device]);
  unsigned char* base = (unsigned char*)g_ptr[ctx->device];
  hipStream_t s = ctx->stream;
  // inputs (the caller's host buffers and small host records, copied before the call returns)
  if (p.tlen) HIPCHK(hipMemcpyAsync(base + L.tseq, p.tseq, (size_t)p.tlen, hipMemcpyHostToDevice, s));
  if (p.qlen) HIPCHK(hipMemcpyAsync(base + L.qseq, p.qseq, (size_t)p.qlen, hipMemcpyHostToDevice, s));
  // zeros after each chunk's bytes, as the batch packer leaves them (word reads run past the end)
  HIPCHK(hipMemsetAsync(base + L.tseq + p.tlen, 0, 16, s));
  HIPCHK(hipMemsetAsync(base + L.qseq + p.qlen, 0, 16, s));
  HIPCHK(hipMemcpyAsync(base + L.tch, &tch, sizeof tch, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(base + L.qch, &qch, sizeof qch, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(base + L.pair, &pd, sizeof pd, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(base + L.which, &zero, sizeof zero, hipMemcpyHostToDevice, s));
  if ((tch.flags | qch.flags) & 1) HIPCHK(hipMemcpyAsync(base + L.ent, erow, sizeof erow, hipMemcpyHostToDevice, s));
  // the flagged-record counter must start at zero (the counts kernel clears it after use,
  // but an earlier call may have stopped before)
  HIPCHK(hipMemsetAsync(base + L.nflag, 0, sizeof(int32_t), s));
  const float* ent_rows = ((tch.flags | qch.flags) & 1) ? (const float*)(base + L.ent) : nullptr;

  // the target's four spectra
  sxc::SpecArgs sa{};
  sa.tab = ctx->d_tab;
  sa.osc = ctx->d_osc;
  sa.ent_lut = ctx->d_ent[nb];
  sa.ent_rows = ent_rows;
  sa.seq = base + L.tseq;
  sa.chunks = (const sxc_chunk*)(base + L.tch);
  sa.which = (const int32_t*)(base + L.which);
  sa.n = 1;
  sa.spectra = (float*)(base + L.spectra);
  sa.dst = (const int32_t*)(base + L.which);
  HIPCHK(sxc::launch_spectra(ctx->geo_spec[nb], sa, s));
  // the fused pair kernel (both strands): correlation and FindTop
  sxc::PairArgs a{};
  a.tab = ctx->d_tab;
  a.osc = ctx->d_osc;
  a.ent_lut = ctx->d_ent[nb];
  a.ent_rows = ent_rows;
  a.seq = base + L.tseq;
  a.tchunks = (const sxc_chunk*)(base + L.tch);
  a.qchunks = (const sxc_chunk*)(base + L.qch);
  a.pairs = (const sxc_pair_dev*)(base + L.pair);
  a.spectra = (const float*)(base + L.spectra);
  a.npairs = 1;
  a.strands = S;
  a.cutoff = prm->cutoff;
  a.cutoff_fast = prm->cutoff_fast;
  a.cands = (uint16_t*)(base + L.cands);
  a.ncand = (int32_t*)(base + L.ncand);
  HIPCHK(sxc::launch_correlate(ctx->geo_corr[nb], a, S, s));
  // the diagonal scan and the probability tail
  sxc::ScanArgs c{};
  c.tab = ctx->d_tab;
  c.seq = base + L.tseq;
  c.tchunks = a.tchunks;
  c.qchunks = a.qchunks;
  c.pairs = a.pairs;
  c.cands = a.cands;
  c.ncand = a.ncand;
  c.n_fft = N;
  c.strands = S;
  c.min_len = prm->min_len;
  c.q_chunk = prm->q_chunk;
  sxc::set_keep(c, prm->min_prob, prm->target_total);
  c.nflag = (int32_t*)(base + L.nflag);
  c.mode = 0;
  c.out = (sxc_result_dev*)(base + L.slots);
  c.keys = (uint64_t*)(base + L.keys);
  c.counts = (int32_t*)(base + L.counts);
  c.cap = SXC_STATIC_RECORDS;
  c.stats = (int32_t*)(base + L.stats);
  HIPCHK(sxc::launch_scan(nb, c, S, s));
  HIPCHK(sxc::launch_count_offsets(c.counts, S, SXC_STATIC_RECORDS, (int64_t*)(base + L.offsets),
                                   (int64_t*)(base + L.total), (int32_t*)(base + L.maxc), s, c.nflag));
  HIPCHK(sxc::launch_order_records(c.counts, S, SXC_STATIC_RECORDS, (const int64_t*)(base + L.offsets), c.out,
                                   c.keys, (sxc_result_dev*)(base + L.final_),
                                   (int64_t)S * SXC_STATIC_RECORDS, s));
  int64_t tot[2] = {0, 0};
  int32_t maxc = 0;
  HIPCHK(hipMemcpyAsync(tot, base + L.total, sizeof tot, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(&maxc, base + L.maxc, sizeof maxc, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (maxc > SXC_STATIC_RECORDS)
    return fail(SXC_ERR_CAPACITY, "a strand kept " + std::to_string(maxc) + " records, more than the " +
                                      std::to_string(SXC_STATIC_RECORDS) +
                                      " the static scratch holds; use sxc_align_pair");
  *n_out = (int32_t)tot[0];
  if (tot[0] > cap) return fail(SXC_ERR_CAPACITY, "output capacity too small");
  if (tot[0] > 0)
    HIPCHK(hipMemcpyAsync(out, base + L.final_, (size_t)tot[0] * sizeof(sxc_result), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (tot[1] > 0) {
    // records the scan flagged near min_prob: decided here with the host's libm
    const sxc::ResolveCtx r{nullptr, &tch, &qch, &pd, 1, prm->q_chunk, prm->min_prob, prm->target_total,
                            (const unsigned char*)p.tseq, (const unsigned char*)p.qseq};
    int32_t kept = 0;
    for (int32_t i = 0; i < (int32_t)tot[0]; ++i) {
      sxc_result_dev rec;
      memcpy(&rec, &out[i], sizeof rec);
      const int k = sxc::resolve_record(rec, r);
      if (k < 0) return fail(SXC_ERR_HIP, "a record flagged near min_prob does not fit its chunk pair");
      if (k) memcpy(&out[kept++], &rec, sizeof rec);
    }
    *n_out = kept;
  }
  ctx->last = sxc_timing{};
  ctx->last.near_threshold = tot[1];
  return SXC_OK;
}
