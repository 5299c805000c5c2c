// sxc_server.hip -- the serving loop: chunk-pair requests against device-resident genomes.
//
// What it serves is the slave's unit of work, a block of target chunks x query chunks
// (t_pair, analysis/WorkQueue.h:17-22; align_target, analysis/HomologyByXCorrSlave.cc:
// 270-300), as the master's battleship search produces them (seed lines and the blocks
// around hits, analysis/GridSearch.cc:188-411).  Records per block are the slave's, in
// the slave's order.
//
// Layout in HBM, built once at creation:
//   seq       every target and query chunk's bytes, 16-byte aligned
//   tch, qch  the chunk tables (offset, length, flags)
//   spectra   the four FFTReal spectra of every target chunk (4*N floats each), so a pair
//             costs only its query transforms, product, inverse, peaks and scan
// Work runs in launch slots: each slot owns a stream and the buffers of one launch of up
// to `max_pairs_per_launch` chunk pairs.  The number of slots is what the HBM budget
// leaves after the resident data.  A dispatcher thread fills every free slot with the
// oldest pending pairs -- across requests, so a launch mixes requests and a request may
// span launches (continuous batching) -- and completes slots as their events fire.
// Results depend only on the chunk pair, never on what shares its launch.
#include <hip/hip_runtime.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <unordered_map>
#include <vector>

#include "sxc_api_internal.h"
#include "sxc_kernels.h"
#include "sxc_tables.h"

using sxc::fail;

namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

struct ChunkMeta {
  int32_t start, record, record_len;
};

struct Request {
  uint64_t id = 0;
  std::vector<sxc_block> blocks;
  int64_t total_pairs = 0;
  // expansion cursor: block, query, target of the next pair to dispatch
  size_t cb = 0;
  int32_t cq = 0, ct = 0;
  int64_t dispatched = 0, done_pairs = 0;
  std::map<int64_t, std::vector<sxc_result>> parts;  // first pair index -> records
  std::vector<sxc_result> results;
  Clock::time_point t_submit, t_done;
  bool complete = false;
  int error = 0;
};

// A run of consecutive pairs of one request inside a launch.
struct Piece {
  Request* req;
  int64_t first;  // pair index in the request
  int32_t count;
  int32_t at;     // pair index in the launch
};

enum SlotState { FREE = 0, COUNTING = 1, PACKING = 2 };

constexpr int kDefaultCap = 64;  // records per (pair, strand) a launch first makes room for

struct Slot {
  hipStream_t s = nullptr;
  hipEvent_t ev_start = nullptr, ev_count = nullptr, ev_pack = nullptr;
  DevBuf pairs, cands, ncand, outs, keys, counts, offsets, total, maxc, final_;
  sxc_pair_dev* h_pairs = nullptr;  // pinned
  int64_t* h_offsets = nullptr;     // pinned, 2 * P + 1 (the last entry is the total)
  int32_t* h_maxc = nullptr;        // pinned
  sxc_result* h_final = nullptr;    // pinned, h_final_cap records
  int64_t h_final_cap = 0;
  int64_t out_cap = 0;              // records final_ holds
  int64_t copied = 0;               // ordered records the launch itself copied to h_final
  int cap = kDefaultCap;            // records per (pair, strand) of the current launch
  int npairs = 0;
  std::vector<Piece> pieces;
  SlotState state = FREE;
};

}  // namespace

struct sxc_server {
  sxc_ctx* ctx = nullptr;
  sxc_params prm{};
  int nb = 0;
  int geo_corr = 0, geo_spec = 0;  // FFT geometry codes of the pair and spectra kernels
  int P = 2048;
  // Adapted to the request stream by the dispatcher: the record room a launch starts with
  // per (pair, strand) (raised to the largest count an overflowing launch needed), and the
  // mean records per pair (sizes the prefix of records a launch copies back with its counts).
  int cap_hint = kDefaultCap;
  double records_per_pair = 0.;
  int cus = 256;  // compute units of the device
  DevBuf seq, tch, qch, spectra, qspectra;
  DevBuf ent;  // window entropies of chunks with B/D/H/V letters (computed on the host)
  // host copies of the chunk bytes and tables: records flagged near min_prob are decided
  // on the host from them (sxc::resolve_record)
  std::vector<unsigned char> h_seq;
  std::vector<sxc_chunk> h_tch, h_qch;
  std::vector<float> h_ent;
  const double* ptable = nullptr;  // -prob_table thresholds (owned by the context)
  std::vector<ChunkMeta> tmeta, qmeta;
  int64_t resident_bytes = 0, slot_bytes = 0, qspec_bytes = 0;
  std::vector<Slot> slots;
  std::mutex mu;
  std::condition_variable work_cv, done_cv;
  std::deque<Request*> pending;  // requests with pairs not yet dispatched
  std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
  uint64_t next_id = 1;
  bool stop = false;
  std::thread dispatcher;
  int fatal = 0;
  std::string fatal_msg;
  // statistics
  int64_t st_requests = 0, st_pairs = 0, st_launches = 0, st_results = 0, st_near = 0;
  double st_device_ms = 0;
};

namespace {

int64_t block_pairs(const sxc_block& b) {
  return (int64_t)(b.target_to - b.target_from + 1) * (int64_t)(b.query_to - b.query_from + 1);
}

int slot_alloc(sxc_server* sv, Slot& sl) {
  const int P = sv->P, N = sv->prm.N;
  const size_t total = (size_t)2 * P;
  HIPCHK(hipStreamCreateWithFlags(&sl.s, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&sl.ev_start));
  HIPCHK(hipEventCreate(&sl.ev_count));
  HIPCHK(hipEventCreate(&sl.ev_pack));
  HIPCHK(sl.pairs.ensure((size_t)P * sizeof(sxc_pair_dev)));
  HIPCHK(sl.cands.ensure(total * N * sizeof(uint16_t)));
  HIPCHK(sl.ncand.ensure(total * sizeof(int32_t)));
  HIPCHK(sl.outs.ensure(total * kDefaultCap * sizeof(sxc_result_dev)));
  HIPCHK(sl.keys.ensure(total * kDefaultCap * sizeof(uint64_t)));
  HIPCHK(sl.counts.ensure(total * sizeof(int32_t)));
  HIPCHK(sl.offsets.ensure(total * sizeof(int64_t)));
  HIPCHK(sl.total.ensure(sizeof(int64_t)));
  HIPCHK(sl.maxc.ensure(sizeof(int32_t)));
  HIPCHK(sl.final_.ensure(total * kDefaultCap * sizeof(sxc_result_dev)));
  sl.out_cap = (int64_t)total * kDefaultCap;
  HIPCHK(hipHostMalloc((void**)&sl.h_pairs, (size_t)P * sizeof(sxc_pair_dev), hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void**)&sl.h_offsets, (total + 1) * sizeof(int64_t), hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void**)&sl.h_maxc, sizeof(int32_t), hipHostMallocDefault));
  sl.h_final_cap = (int64_t)total * kDefaultCap;
  HIPCHK(hipHostMalloc((void**)&sl.h_final, (size_t)sl.h_final_cap * sizeof(sxc_result), hipHostMallocDefault));
  return SXC_OK;
}

void slot_free(Slot& sl) {
  for (DevBuf* b : {&sl.pairs, &sl.cands, &sl.ncand, &sl.outs, &sl.keys, &sl.counts, &sl.offsets, &sl.total,
                    &sl.maxc, &sl.final_})
    b->release();
  if (sl.h_pairs) (void)hipHostFree(sl.h_pairs);
  if (sl.h_offsets) (void)hipHostFree(sl.h_offsets);
  if (sl.h_maxc) (void)hipHostFree(sl.h_maxc);
  if (sl.h_final) (void)hipHostFree(sl.h_final);
  for (hipEvent_t e : {sl.ev_start, sl.ev_count, sl.ev_pack})
    if (e) (void)hipEventDestroy(e);
  if (sl.s) (void)hipStreamDestroy(sl.s);
  sl = Slot();
}

// Device bytes one slot holds at the default record capacity.
int64_t slot_bytes_for(int P, int N, int cap) {
  const int64_t total = 2 * (int64_t)P;
  return (int64_t)P * sizeof(sxc_pair_dev) + total * N * 2 + total * 4 + total * cap * (72 + 8 + 72) + total * 12 + 64;
}

// split: workgroups per (pair, strand) (counts must have been cleared by the pair kernel
// when above 1).  order: also order the records into final_ (up to out_cap) and copy the
// first of them to the host in the same pass, so a launch whose records fit needs no
// second round trip.
int launch_scan_slot(sxc_server* sv, Slot& sl, int split, bool order) {
  // Room for cap records of every (pair, strand) of this launch.  The buffers only grow;
  // a launch may have more pairs than the one that last grew them.
  const size_t nrec = (size_t)2 * sl.npairs * sl.cap;
  HIPCHK(sl.outs.ensure(nrec * sizeof(sxc_result_dev)));
  HIPCHK(sl.keys.ensure(nrec * sizeof(uint64_t)));
  sxc::ScanArgs a{};
  a.tab = sv->ctx->d_tab;
  a.seq = sv->seq.as<unsigned char>();
  a.tchunks = sv->tch.as<sxc_chunk>();
  a.qchunks = sv->qch.as<sxc_chunk>();
  a.pairs = sl.pairs.as<sxc_pair_dev>();
  a.cands = sl.cands.as<uint16_t>();
  a.ncand = sl.ncand.as<int32_t>();
  a.n_fft = sv->prm.N;
  a.strands = 2;
  a.min_len = sv->prm.min_len;
  a.q_chunk = sv->prm.q_chunk;
  sxc::set_keep(a, sv->prm.min_prob, sv->prm.target_total);
  a.nflag = nullptr;  // the host looks at every record of the slot anyway (slot_complete)
  a.mode = 0;
  a.out = sl.outs.as<sxc_result_dev>();
  a.runs = nullptr;
  a.keys = sl.keys.as<uint64_t>();
  a.counts = sl.counts.as<int32_t>();
  a.cap = sl.cap;
  a.stats = nullptr;
  a.ptable = sv->ptable;
  a.table_prob = sv->prm.table_prob;
  a.split = split;
  HIPCHK(sxc::launch_scan(sv->nb, a, 2 * sl.npairs, sl.s));
  const int total = 2 * sl.npairs;
  HIPCHK(sxc::launch_count_offsets(sl.counts.as<int32_t>(), total, sl.cap, sl.offsets.as<int64_t>(),
                                   sl.total.as<int64_t>(), sl.maxc.as<int32_t>(), sl.s));
  HIPCHK(hipMemcpyAsync(sl.h_offsets, sl.offsets.p, (size_t)total * sizeof(int64_t), hipMemcpyDeviceToHost, sl.s));
  HIPCHK(hipMemcpyAsync(sl.h_offsets + total, sl.total.p, sizeof(int64_t), hipMemcpyDeviceToHost, sl.s));
  HIPCHK(hipMemcpyAsync(sl.h_maxc, sl.maxc.p, sizeof(int32_t), hipMemcpyDeviceToHost, sl.s));
  sl.copied = 0;
  if (order) {
    HIPCHK(sxc::launch_order_records(sl.counts.as<int32_t>(), total, sl.cap, sl.offsets.as<int64_t>(),
                                     sl.outs.as<sxc_result_dev>(), sl.keys.as<uint64_t>(),
                                     sl.final_.as<sxc_result_dev>(), sl.out_cap, sl.s));
    // records are few per pair (mostly none): a prefix sized by the launch usually holds all
    const double want = 2.0 * sv->records_per_pair * sl.npairs;
    sl.copied = std::min<int64_t>(std::min(sl.out_cap, sl.h_final_cap),
                                  std::max<int64_t>(std::max(512, sl.npairs), (int64_t)want));
    HIPCHK(hipMemcpyAsync(sl.h_final, sl.final_.p, (size_t)sl.copied * sizeof(sxc_result), hipMemcpyDeviceToHost,
                          sl.s));
  }
  HIPCHK(hipEventRecord(sl.ev_count, sl.s));
  return SXC_OK;
}

// Takes up to P pending pairs (oldest requests first) into the slot and launches
// correlate -> scan -> counts.  Called with sv->mu held.
int slot_fill(sxc_server* sv, Slot& sl) {
  sl.pieces.clear();
  sl.npairs = 0;
  sl.cap = sv->cap_hint;
  while (sl.npairs < sv->P && !sv->pending.empty()) {
    Request* r = sv->pending.front();
    Piece pc{r, r->dispatched, 0, sl.npairs};
    while (sl.npairs < sv->P && r->cb < r->blocks.size()) {
      const sxc_block& b = r->blocks[r->cb];
      const int qi = b.query_from + r->cq, ti = b.target_from + r->ct;
      sxc_pair_dev& d = sl.h_pairs[sl.npairs++];
      const ChunkMeta& tm = sv->tmeta[(size_t)ti];
      const ChunkMeta& qm = sv->qmeta[(size_t)qi];
      d.t_chunk = ti;
      d.q_chunk = qi;
      d.t_chunk_start = tm.start;
      d.q_chunk_start = qm.start;
      d.t_record = tm.record;
      d.q_record = qm.record;
      d.q_record_len = qm.record_len;
      d.fast = b.fast ? 1 : 0;
      d.q_tail = -1;
      ++pc.count;
      ++r->dispatched;
      // align_target order: query chunks outer, target chunks inner
      if (++r->ct > b.target_to - b.target_from) {
        r->ct = 0;
        if (++r->cq > b.query_to - b.query_from) {
          r->cq = 0;
          ++r->cb;
        }
      }
    }
    sl.pieces.push_back(pc);
    if (r->cb >= r->blocks.size()) sv->pending.pop_front();
  }
  if (sl.npairs == 0) return SXC_OK;
  HIPCHK(hipEventRecord(sl.ev_start, sl.s));
  HIPCHK(hipMemcpyAsync(sl.pairs.p, sl.h_pairs, (size_t)sl.npairs * sizeof(sxc_pair_dev), hipMemcpyHostToDevice, sl.s));
  sxc::PairArgs a{};
  a.tab = sv->ctx->d_tab;
  a.osc = sv->ctx->d_osc;
  a.ent_lut = sv->ctx->d_ent[sv->nb];
  a.ent_rows = sv->h_ent.empty() ? nullptr : sv->ent.as<float>();
  a.seq = sv->seq.as<unsigned char>();
  a.tchunks = sv->tch.as<sxc_chunk>();
  a.qchunks = sv->qch.as<sxc_chunk>();
  a.pairs = sl.pairs.as<sxc_pair_dev>();
  a.spectra = sv->spectra.as<float>();
  a.npairs = sl.npairs;
  a.strands = 2;
  a.cutoff = sv->prm.cutoff;
  a.cutoff_fast = sv->prm.cutoff_fast;
  a.cands = sl.cands.as<uint16_t>();
  a.ncand = sl.ncand.as<int32_t>();
  a.qspectra = sv->qspec_bytes > 0 ? sv->qspectra.as<float>() : nullptr;
  a.counts = sl.counts.as<int32_t>();  // cleared per (pair, strand) for a split scan
  HIPCHK(sxc::launch_correlate(sv->geo_corr, a, 2 * sl.npairs, sl.s));
  // A launch too small to fill the device splits each (pair, strand)'s lags over up to 4
  // scan workgroups (3 resident per CU), which shortens its latency.
  const int split = std::max(1, std::min(4, 3 * sv->cus / (2 * sl.npairs)));
  int rc = launch_scan_slot(sv, sl, split, true);
  if (rc) return rc;
  sl.state = COUNTING;
  ++sv->st_launches;
  return SXC_OK;
}

void slot_complete(sxc_server* sv, Slot& sl, hipEvent_t last);

// Counts are back (and, usually, the ordered records with them).  If a (pair, strand) kept
// more records than the slot holds, redo the scan with room for all; if the records
// overflowed the output buffer, order them again into a larger one; copy whatever the
// launch did not bring back.  Completes the slot at once when nothing is missing.
int slot_pack(sxc_server* sv, Slot& sl) {
  const int total = 2 * sl.npairs;
  bool order = false;
  if (*sl.h_maxc > sl.cap) {
    // the same pass again with room for every record (the record buffers grow to total x
    // cap), without waiting here: the slot comes back through COUNTING; later launches
    // start with that much room (up to 1024 records per (pair, strand))
    sl.cap = *sl.h_maxc;
    sv->cap_hint = std::min(1024, std::max(sv->cap_hint, sl.cap));
    return launch_scan_slot(sv, sl, 1, true);
  }
  const int64_t n = sl.h_offsets[total];
  sv->records_per_pair = 0.9 * sv->records_per_pair + 0.1 * (double)n / std::max(1, sl.npairs);
  if (n > sl.out_cap) {
    HIPCHK(sl.final_.ensure((size_t)n * sizeof(sxc_result_dev)));
    sl.out_cap = n;
    order = true;
  }
  if (n > sl.h_final_cap) {
    if (sl.h_final) HIPCHK(hipHostFree(sl.h_final));
    sl.h_final = nullptr;
    sl.h_final_cap = n;
    HIPCHK(hipHostMalloc((void**)&sl.h_final, (size_t)n * sizeof(sxc_result), hipHostMallocDefault));
    sl.copied = 0;  // the records copied so far went to the freed buffer
  }
  if (order) {
    HIPCHK(sxc::launch_order_records(sl.counts.as<int32_t>(), total, sl.cap, sl.offsets.as<int64_t>(),
                                     sl.outs.as<sxc_result_dev>(), sl.keys.as<uint64_t>(),
                                     sl.final_.as<sxc_result_dev>(), sl.out_cap, sl.s));
    sl.copied = 0;
  }
  if (n <= sl.copied) {
    slot_complete(sv, sl, sl.ev_count);
    return SXC_OK;
  }
  HIPCHK(hipMemcpyAsync(sl.h_final + sl.copied, sl.final_.as<sxc_result_dev>() + sl.copied,
                        (size_t)(n - sl.copied) * sizeof(sxc_result), hipMemcpyDeviceToHost, sl.s));
  HIPCHK(hipEventRecord(sl.ev_pack, sl.s));
  sl.state = PACKING;
  return SXC_OK;
}

// Records are on the host: hand each request its pairs' records.  `last`: the slot's last
// completed event.  Called with sv->mu held.
void slot_complete(sxc_server* sv, Slot& sl, hipEvent_t last) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, sl.ev_start, last) == hipSuccess) sv->st_device_ms += ms;
  const int total = 2 * sl.npairs;
  const auto now = Clock::now();
  const sxc::ResolveCtx rc{sv->h_seq.data(), sv->h_tch.data(), sv->h_qch.data(), sl.h_pairs, sl.npairs,
                           sv->prm.q_chunk, sv->prm.min_prob, sv->prm.target_total};
  for (const Piece& pc : sl.pieces) {
    const int64_t b = sl.h_offsets[2 * pc.at];
    const int64_t e = (2 * (pc.at + pc.count) < total) ? sl.h_offsets[2 * (pc.at + pc.count)] : sl.h_offsets[total];
    Request* r = pc.req;
    std::vector<sxc_result>& part = r->parts[pc.first];
    part.clear();
    part.reserve((size_t)(e - b));
    for (int64_t i = b; i < e; ++i) {
      // records flagged near min_prob are decided here with the host's libm
      sxc_result_dev rec;
      memcpy(&rec, &sl.h_final[i], sizeof rec);
      sv->st_near += rec.pad_[0] == SXC_REC_NEAR;
      const int keep = sxc::resolve_record(rec, rc);
      if (keep < 0 && !r->error) r->error = SXC_ERR_HIP;
      if (keep > 0) {
        part.emplace_back();
        memcpy(&part.back(), &rec, sizeof rec);
      }
    }
    r->done_pairs += pc.count;
    sv->st_results += (int64_t)part.size();
    if (r->done_pairs == r->total_pairs) {
      for (auto& kv : r->parts) r->results.insert(r->results.end(), kv.second.begin(), kv.second.end());
      r->parts.clear();
      r->t_done = now;
      r->complete = true;
      ++sv->st_requests;
    }
  }
  sv->st_pairs += sl.npairs;
  sl.pieces.clear();
  sl.npairs = 0;
  sl.state = FREE;
}

void fail_all(sxc_server* sv, int rc) {
  sv->fatal = rc;
  sv->fatal_msg = sxc::g_err;
  for (auto& kv : sv->reqs)
    if (!kv.second->complete) {
      kv.second->complete = true;
      kv.second->error = rc;
    }
  sv->pending.clear();
}

void dispatch_loop(sxc_server* sv) {
  (void)hipSetDevice(sv->ctx->device);
  std::unique_lock<std::mutex> lk(sv->mu);
  while (!sv->stop) {
    bool progressed = false, busy = false;
    for (Slot& sl : sv->slots) {
      if (sl.state == COUNTING && hipEventQuery(sl.ev_count) == hipSuccess) {
        const int rc = slot_pack(sv, sl);
        if (rc) fail_all(sv, rc);
        progressed = true;
      }
      if (sl.state == PACKING && hipEventQuery(sl.ev_pack) == hipSuccess) {
        slot_complete(sv, sl, sl.ev_pack);
        progressed = true;
      }
    }
    for (Slot& sl : sv->slots) {
      if (sl.state == FREE && !sv->pending.empty() && !sv->fatal) {
        const int rc = slot_fill(sv, sl);
        if (rc) fail_all(sv, rc);
        progressed = true;
      }
      busy |= sl.state != FREE;
    }
    if (progressed) sv->done_cv.notify_all();
    if (sv->fatal) {
      sv->done_cv.notify_all();
      sv->work_cv.wait(lk, [&] { return sv->stop; });
      break;
    }
    if (progressed) continue;
    if (busy) {
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(20));
      lk.lock();
    } else {
      sv->work_cv.wait(lk, [&] { return sv->stop || !sv->pending.empty(); });
    }
  }
}

int upload_chunks(std::vector<unsigned char>& bytes, std::vector<float>& ent, std::vector<sxc_chunk>& table,
                  std::vector<ChunkMeta>& meta, const sxc_chunk_desc* in, int32_t n, int max_len, int N) {
  table.resize((size_t)n);
  meta.resize((size_t)n);
  for (int32_t k = 0; k < n; ++k) {
    const sxc_chunk_desc& c = in[k];
    if (c.len < 0 || c.len > max_len || (c.len > 0 && !c.seq))
      return fail(SXC_ERR_ARG, "target chunks must be at most N/2 long, query chunks at most N");
    sxc_chunk& t = table[(size_t)k];
    t.off = (int64_t)bytes.size();
    t.len = c.len;
    t.flags = sxc::chunk_flags((const unsigned char*)c.seq, c.len);
    if (t.flags < 0) return fail(SXC_ERR_ARG, "sequence byte >= 128");
    bytes.resize(bytes.size() + (size_t)c.len);
    if (c.len) memcpy(bytes.data() + t.off, c.seq, (size_t)c.len);
    if (t.flags & 1) {  // window entropies with the host's log
      const size_t row = ent.size() / SXC_ENT_ROW;
      if (row >= (1u << 24)) return fail(SXC_ERR_ARG, "too many chunks with B, D, H or V letters");
      t.flags |= (int32_t)(row << 8);
      ent.resize(ent.size() + SXC_ENT_ROW);
      sxc::window_entropy_row(bytes.data() + t.off, c.len, N, ent.data() + row * SXC_ENT_ROW);
    }
    while (bytes.size() % 16) bytes.push_back(0);
    meta[(size_t)k] = ChunkMeta{c.start, c.record, c.record_len};
  }
  return SXC_OK;
}

int server_init(sxc_server* sv, const sxc_chunk_desc* targets, int32_t nt, const sxc_chunk_desc* queries,
                int32_t nq, const sxc_server_config* cfg) {
  const int N = sv->prm.N;
  std::vector<unsigned char>& bytes = sv->h_seq;
  std::vector<sxc_chunk>& ttab = sv->h_tch;
  std::vector<sxc_chunk>& qtab = sv->h_qch;
  // targets at most N/2 and queries at most N: every lag of every pair fits the transform
  int rc = upload_chunks(bytes, sv->h_ent, ttab, sv->tmeta, targets, nt, N / 2, N);
  if (!rc) rc = upload_chunks(bytes, sv->h_ent, qtab, sv->qmeta, queries, nq, N, N);
  if (rc) return rc;
  bytes.resize(bytes.size() + 16, 0);
  sv->P = cfg && cfg->max_pairs_per_launch > 0 ? cfg->max_pairs_per_launch : 2048;
  HIPCHK(hipDeviceGetAttribute(&sv->cus, hipDeviceAttributeMultiprocessorCount, sv->ctx->device));
  const int max_slots = cfg && cfg->max_slots > 0 ? cfg->max_slots : 3;
  size_t free_b = 0, total_b = 0;
  HIPCHK(hipMemGetInfo(&free_b, &total_b));
  const int64_t budget = cfg && cfg->hbm_budget > 0 ? cfg->hbm_budget : (int64_t)(free_b / 4 * 3);
  sv->resident_bytes = (int64_t)bytes.size() + (int64_t)(ttab.size() + qtab.size()) * (int64_t)sizeof(sxc_chunk) +
                       (int64_t)nt * 4 * N * (int64_t)sizeof(float) + (int64_t)(sv->h_ent.size() * sizeof(float));
  sv->slot_bytes = slot_bytes_for(sv->P, N, 64);
  // both strands' spectra of every query chunk, if they leave room for at least one slot
  const int64_t qspec_bytes = (int64_t)nq * 2 * 4 * N * (int64_t)sizeof(float);
  const bool want_q = !(cfg && cfg->no_query_spectra) && nq > 0;
  if (want_q && budget - sv->resident_bytes - qspec_bytes >= sv->slot_bytes) {
    sv->resident_bytes += qspec_bytes;
    sv->qspec_bytes = qspec_bytes;
  }
  const int64_t room = budget - sv->resident_bytes;
  const int nslots = (int)std::min<int64_t>(max_slots, room > 0 ? room / sv->slot_bytes : 0);
  if (nslots < 1)
    return fail(SXC_ERR_NOMEM, "HBM budget too small: resident genome and spectra need " +
                                   std::to_string(sv->resident_bytes) + " bytes plus " +
                                   std::to_string(sv->slot_bytes) + " per launch slot");
  if (sv->prm.prob_table) {
    rc = sxc::ctx_prob_table(sv->ctx, sv->prm.target_total, &sv->ptable);
    if (rc) return rc;
  }
  hipStream_t s = sv->ctx->stream;
  HIPCHK(sv->seq.ensure(bytes.size()));
  HIPCHK(sv->tch.ensure(std::max<size_t>(1, ttab.size()) * sizeof(sxc_chunk)));
  HIPCHK(sv->qch.ensure(std::max<size_t>(1, qtab.size()) * sizeof(sxc_chunk)));
  HIPCHK(sv->spectra.ensure(std::max<size_t>(1, ttab.size()) * 4 * (size_t)N * sizeof(float)));
  HIPCHK(hipMemcpyAsync(sv->seq.p, bytes.data(), bytes.size(), hipMemcpyHostToDevice, s));
  if (!sv->h_ent.empty()) {
    HIPCHK(sv->ent.ensure(sv->h_ent.size() * sizeof(float)));
    HIPCHK(hipMemcpyAsync(sv->ent.p, sv->h_ent.data(), sv->h_ent.size() * sizeof(float), hipMemcpyHostToDevice, s));
  }
  if (!ttab.empty())
    HIPCHK(hipMemcpyAsync(sv->tch.p, ttab.data(), ttab.size() * sizeof(sxc_chunk), hipMemcpyHostToDevice, s));
  if (!qtab.empty())
    HIPCHK(hipMemcpyAsync(sv->qch.p, qtab.data(), qtab.size() * sizeof(sxc_chunk), hipMemcpyHostToDevice, s));
  // resident target spectra, computed once in launches of up to 65536 chunks
  DevBuf which, qdst;
  const int step = 65536;
  HIPCHK(which.ensure((size_t)std::min<int32_t>(std::max(nt, nq), step) * sizeof(int32_t) + 4));
  std::vector<int32_t> idx;
  for (int32_t at = 0; at < nt; at += step) {
    const int n = std::min<int32_t>(step, nt - at);
    idx.resize((size_t)n);
    for (int i = 0; i < n; ++i) idx[(size_t)i] = at + i;
    HIPCHK(hipMemcpyAsync(which.p, idx.data(), (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, s));
    sxc::SpecArgs sa{};
    sa.tab = sv->ctx->d_tab;
    sa.osc = sv->ctx->d_osc;
    sa.ent_lut = sv->ctx->d_ent[sv->nb];
    sa.ent_rows = sv->h_ent.empty() ? nullptr : sv->ent.as<float>();
    sa.seq = sv->seq.as<unsigned char>();
    sa.chunks = sv->tch.as<sxc_chunk>();
    sa.which = which.as<int32_t>();
    sa.n = n;
    sa.spectra = sv->spectra.as<float>();
    sa.dst = which.as<int32_t>();
    HIPCHK(sxc::launch_spectra(sv->geo_spec, sa, s));
    HIPCHK(hipStreamSynchronize(s));  // idx is reused
  }
  if (sv->qspec_bytes > 0) {
    // resident query spectra: slot 2 q + strand holds query chunk q's (strand 1: reverse
    // complement) four spectra, the transforms the pair kernel would otherwise compute
    HIPCHK(sv->qspectra.ensure((size_t)sv->qspec_bytes));
    std::vector<int32_t> dst;
    for (int32_t at = 0; at < nq; at += step) {
      const int n = std::min<int32_t>(step, nq - at);
      for (int strand = 0; strand < 2; ++strand) {
        idx.resize((size_t)n);
        dst.resize((size_t)n);
        for (int i = 0; i < n; ++i) {
          idx[(size_t)i] = at + i;
          dst[(size_t)i] = 2 * (at + i) + strand;
        }
        HIPCHK(qdst.ensure((size_t)n * sizeof(int32_t) + 4));
        HIPCHK(hipMemcpyAsync(which.p, idx.data(), (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(qdst.p, dst.data(), (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, s));
        sxc::SpecArgs sa{};
        sa.tab = sv->ctx->d_tab;
        sa.osc = sv->ctx->d_osc;
        sa.ent_lut = sv->ctx->d_ent[sv->nb];
        sa.ent_rows = sv->h_ent.empty() ? nullptr : sv->ent.as<float>();
        sa.seq = sv->seq.as<unsigned char>();
        sa.chunks = sv->qch.as<sxc_chunk>();
        sa.which = which.as<int32_t>();
        sa.n = n;
        sa.spectra = sv->qspectra.as<float>();
        sa.dst = qdst.as<int32_t>();
        sa.strand = strand;
        HIPCHK(sxc::launch_spectra(sv->geo_spec, sa, s));
        HIPCHK(hipStreamSynchronize(s));  // idx and dst are reused
      }
    }
    qdst.release();
  }
  which.release();
  sv->slots.resize((size_t)nslots);
  for (Slot& sl : sv->slots) {
    rc = slot_alloc(sv, sl);
    if (rc) return rc;
  }
  return SXC_OK;
}

void server_release(sxc_server* sv) {
  for (Slot& sl : sv->slots) {
    if (sl.s) (void)hipStreamSynchronize(sl.s);
    slot_free(sl);
  }
  for (DevBuf* b : {&sv->seq, &sv->tch, &sv->qch, &sv->spectra, &sv->qspectra, &sv->ent}) b->release();
}

}  // namespace

extern "C" {

int sxc_server_create(sxc_ctx* ctx, const sxc_params* prm, const sxc_chunk_desc* targets, int32_t nt,
                      const sxc_chunk_desc* queries, int32_t nq, const sxc_server_config* cfg, sxc_server** out) {
  if (!ctx || !prm || !out || nt < 0 || nq < 0 || (nt > 0 && !targets) || (nq > 0 && !queries))
    return fail(SXC_ERR_ARG, "null argument");
  *out = nullptr;
  const int nb = sxc::nb_of(prm->N);
  if (nb < SXC_MIN_NB || nb > SXC_MAX_NB) return fail(SXC_ERR_ARG, "N must be a power of two in [512, 32768]");
  if (prm->q_chunk < 0) return fail(SXC_ERR_ARG, "q_chunk must be >= 0");
  std::lock_guard<std::mutex> g(ctx->mu);
  HIPCHK(hipSetDevice(ctx->device));
  sxc_server* sv = new sxc_server();
  sv->ctx = ctx;
  sv->prm = *prm;
  sv->prm.strands = 2;
  sv->nb = nb;
  sv->geo_corr = ctx->geo_corr[nb];  // the context's geometry choice when the server starts
  sv->geo_spec = ctx->geo_spec[nb];
  const int rc = server_init(sv, targets, nt, queries, nq, cfg);
  if (rc) {
    server_release(sv);
    delete sv;
    return rc;
  }
  sv->dispatcher = std::thread(dispatch_loop, sv);
  *out = sv;
  return SXC_OK;
}

int sxc_server_submit(sxc_server* sv, const sxc_block* blocks, int32_t nblocks, uint64_t* ticket) {
  if (!sv || !ticket || nblocks < 0 || (nblocks > 0 && !blocks)) return fail(SXC_ERR_ARG, "null argument");
  auto r = std::make_unique<Request>();
  const int32_t nt = (int32_t)sv->tmeta.size(), nq = (int32_t)sv->qmeta.size();
  for (int32_t k = 0; k < nblocks; ++k) {
    const sxc_block& b = blocks[k];
    if (b.target_from < 0 || b.target_to >= nt || b.target_from > b.target_to || b.query_from < 0 ||
        b.query_to >= nq || b.query_from > b.query_to)
      return fail(SXC_ERR_ARG, "block outside the server's chunk tables");
    r->blocks.push_back(b);
    r->total_pairs += block_pairs(b);
  }
  r->t_submit = Clock::now();
  std::lock_guard<std::mutex> g(sv->mu);
  if (sv->fatal) return fail(sv->fatal, "server stopped after an error: " + sv->fatal_msg);
  r->id = sv->next_id++;
  *ticket = r->id;
  Request* rp = r.get();
  sv->reqs[r->id] = std::move(r);
  if (rp->total_pairs == 0) {
    rp->complete = true;
    rp->t_done = Clock::now();
    ++sv->st_requests;
  } else {
    sv->pending.push_back(rp);
    sv->work_cv.notify_all();
  }
  return SXC_OK;
}

int sxc_server_wait(sxc_server* sv, uint64_t ticket, int32_t timeout_ms, int64_t* n_results, double* latency_ms) {
  if (!sv) return fail(SXC_ERR_ARG, "null server");
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->reqs.find(ticket);
  if (it == sv->reqs.end()) return fail(SXC_ERR_ARG, "unknown ticket");
  Request* r = it->second.get();
  auto ready = [&] { return r->complete; };
  if (timeout_ms < 0) {
    sv->done_cv.wait(lk, ready);
  } else if (!sv->done_cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
    return SXC_PENDING;
  }
  if (r->error) return fail(r->error, "request failed: " + sv->fatal_msg);
  if (n_results) *n_results = (int64_t)r->results.size();
  if (latency_ms) *latency_ms = ms_since(r->t_submit, r->t_done);
  return SXC_OK;
}

int sxc_server_results(sxc_server* sv, uint64_t ticket, sxc_result* out, int64_t cap, int64_t* n_out) {
  if (!sv || !n_out || (cap > 0 && !out)) return fail(SXC_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> g(sv->mu);
  auto it = sv->reqs.find(ticket);
  if (it == sv->reqs.end()) return fail(SXC_ERR_ARG, "unknown ticket");
  Request* r = it->second.get();
  if (!r->complete) return SXC_PENDING;
  if (r->error) return fail(r->error, "request failed: " + sv->fatal_msg);
  *n_out = (int64_t)r->results.size();
  if (*n_out > cap) return fail(SXC_ERR_CAPACITY, "output capacity too small");
  if (*n_out) memcpy(out, r->results.data(), r->results.size() * sizeof(sxc_result));
  sv->reqs.erase(it);
  return SXC_OK;
}

int sxc_server_stats(sxc_server* sv, sxc_server_info* out) {
  if (!sv || !out) return fail(SXC_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> g(sv->mu);
  out->requests_done = sv->st_requests;
  out->pairs_done = sv->st_pairs;
  out->launches = sv->st_launches;
  out->results = sv->st_results;
  out->near_threshold = sv->st_near;
  out->resident_bytes = sv->resident_bytes;
  out->slot_bytes = sv->slot_bytes;
  out->slots = (int32_t)sv->slots.size();
  out->max_pairs_per_launch = sv->P;
  out->device_ms = sv->st_device_ms;
  out->query_spectra_bytes = sv->qspec_bytes;
  return SXC_OK;
}

void sxc_server_destroy(sxc_server* sv) {
  if (!sv) return;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    sv->stop = true;
  }
  sv->work_cv.notify_all();
  if (sv->dispatcher.joinable()) sv->dispatcher.join();
  (void)hipSetDevice(sv->ctx->device);
  server_release(sv);
  delete sv;
}

}  // extern "C"
