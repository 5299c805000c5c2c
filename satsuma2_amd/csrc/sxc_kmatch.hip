// sxc_kmatch.hip -- KMatch (kmatch/KMatch.cc) with its k-mer tables built, sorted,
// filtered and joined on the GPU, and the chaining sweep on the host (include/sxc_kmatch.h).
//
// Per genome (KMatch::kmer_array_from_fasta, KMatch.cc:21-132): every position whose K
// bases are all A/C/G/T (either case) gives its canonical k-mer -- the smaller of the
// forward 2-bit code and the reverse complement's -- with its 1-based start plus the
// sequence index times 10^10, negated when the reverse complement was the smaller.  The
// table has len + 1 - K slots per sequence; slots no k-mer fills stay {0, 0} and are sorted
// and counted with the rest.  The table is sorted by k-mer, k-mers occurring more than
// max_freq times are dropped, and the last surviving entry is dropped as well
// (kposv.resize(wi - 1), :124-129).  Joining (merge_positions, :141-177) pairs each query
// entry with the target entries of its k-mer, orientation = query sign xor target sign.
//
// Device layout: both genomes' sequences in one byte buffer per genome, each sequence
// preceded by one 'N' so no k-mer window spans two sequences; 64 positions per thread, the
// window kept rolling in registers (two passes: count per thread, exclusive scan, write in
// position order).  Sorting and scans are rocPRIM's device radix sort and scan; the
// frequency filter tests, for each entry, the max_freq + 1 windows of the sorted table that
// contain it (a group is longer than max_freq iff one such window has equal ends), so it
// reads O(max_freq) neighbours instead of the group.
#include <hip/hip_runtime.h>
#include <string.h>

#include <chrono>
#include <fstream>
#include <memory>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <string>
#include <vector>

#include "../../include/sxc_kmatch.h"
#include "sxc_api_internal.h"

using sxc::fail;

namespace {

constexpr int64_t kRecordStride = 10000000000LL;  // KMATCH_POSITION_CHR_CNST (KMatch.h:18)
constexpr int kSpan = 16;                          // positions per lane
constexpr int kChunk = 64 * kSpan;                 // positions per chunk (one wave)
constexpr int kBlock = 256;

// One genome's sequences as the table builder walks them.
struct Side {
  std::string buf;                 // each sequence preceded by one 'N'
  std::vector<int64_t> off;        // start of each sequence in buf
  std::vector<int64_t> index;      // its sequence index (headers seen when it was processed)
  std::vector<uint64_t> lengths;   // seqnames[].length, in processing order
  int64_t slots = 0;               // table size: sum of len + 1 - K
};

int add_sequence(Side* s, const char* p, int64_t len, int64_t index, int K, const std::string& what) {
  if (len == 0) return SXC_OK;  // an empty record is not processed (KMatch.cc:50)
  if (len + 1 < K)
    return fail(SXC_ERR_ARG, what + ": a record of " + std::to_string(len) + " bases is shorter than K - 1");
  s->slots += len + 1 - K;
  s->lengths.push_back((uint64_t)len);
  s->buf.push_back('N');
  s->off.push_back((int64_t)s->buf.size());
  s->buf.append(p, (size_t)len);
  s->index.push_back(index);
  return SXC_OK;
}

// kmer_array_from_fasta's reading (KMatch.cc:32-103): std::getline lines, '>' starts a
// record, and at end of file the pending sequence is processed before the last line read
// is appended, so a final line without a newline is not part of it.
int read_side(const char* path, int K, Side* s) {
  std::ifstream f(path);
  if (!f.is_open()) return fail(SXC_ERR_ARG, std::string("cannot open ") + path);
  std::string line, seq;
  int64_t index = 0;
  for (;;) {
    std::getline(f, line);
    if ((!line.empty() && line[0] == '>') || f.eof()) {
      const int rc = add_sequence(s, seq.data(), (int64_t)seq.size(), index, K, path);
      if (rc != SXC_OK) return rc;
      if (f.eof()) break;
      seq.clear();
      ++index;
    } else {
      if (f.fail()) return fail(SXC_ERR_ARG, std::string("read error in ") + path);
      seq += line;
    }
  }
  return SXC_OK;
}

__device__ inline int base2(unsigned char c) {
  switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return -1;
  }
}

// Canonical k-mers of chunk c, positions [1024 c, 1024 c + 1024): one wave per chunk, each lane
// rolling its window over 16 positions.  Pass 0 counts a chunk's k-mers; pass 1 puts each
// lane's k-mers into its 16 LDS slots and the wave then copies the chunk's k-mers, in position
// order, to base[c] on as one contiguous run.  off/index: sequence starts and indices (binary
// search).
template <bool WRITE>
__global__ __launch_bounds__(kBlock) void k_kmers(const unsigned char* buf, int64_t n, int K, int64_t chunks,
                                                  uint64_t* count, const uint64_t* base, const int64_t* off,
                                                  const int64_t* index, int nseq, uint64_t* keys, int64_t* vals) {
  constexpr int kWaves = kBlock / 64;
  static_assert(kBlock % 64 == 0, "");
  __shared__ uint64_t skey[WRITE ? kWaves * kChunk : 1];
  __shared__ int64_t sval[WRITE ? kWaves * kChunk : 1];
  __shared__ int pre[kWaves][65];
  const int wave = (int)threadIdx.x / 64, lane = (int)threadIdx.x % 64;
  const int64_t c = (int64_t)blockIdx.x * kWaves + wave;
  const bool active = c < chunks;  // every wave reaches the barriers
  int got = 0;
  if (active) {
    const int64_t start = c * kChunk + (int64_t)lane * kSpan;
    const int64_t end = start + kSpan < n ? start + kSpan : n;
    const uint64_t mask = (1ull << (2 * K)) - 1;
    const int top = 2 * (K - 1);
    const int64_t s0 = start - (K - 1) > 0 ? start - (K - 1) : 0;
    int64_t last_unknown = s0 - 1;
    uint64_t f = 0, r = 0;
    for (int64_t p = s0; p < end; ++p) {
      const int b = base2(buf[p]);
      const uint64_t code = b < 0 ? 0 : (uint64_t)b;  // any other letter acts as A and breaks the window
      f = ((f << 2) + code) & mask;
      r = (r >> 2) + ((3 - code) << top);
      if (b < 0) last_unknown = p;
      if (p >= start && last_unknown + K <= p) {
        if (WRITE) {
          int lo = 0, hi = nseq - 1;  // the last sequence starting at or before p
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (off[mid] <= p) lo = mid; else hi = mid - 1;
          }
          const int64_t pos = (p - off[lo]) - K + 2 + index[lo] * kRecordStride;  // 1-based start
          const int slot = wave * kChunk + lane * kSpan + got;
          skey[slot] = f <= r ? f : r;
          sval[slot] = f <= r ? pos : -pos;
        }
        ++got;
      }
    }
  }
  pre[wave][lane + 1] = got;
  __syncthreads();
  if (lane == 0) {
    pre[wave][0] = 0;
    for (int l = 1; l <= 64; ++l) pre[wave][l] += pre[wave][l - 1];
    if (!WRITE && active) count[c] = (uint64_t)pre[wave][64];
  }
  __syncthreads();
  if (!WRITE || !active) return;
  const int total = pre[wave][64];
  const uint64_t at = base[c];
  for (int i = lane; i < total; i += 64) {
    int lo = 0, hi = 63;  // the lane whose k-mers hold entry i: pre[lo] <= i < pre[lo + 1]
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (pre[wave][mid] <= i) lo = mid; else hi = mid - 1;
    }
    const int slot = wave * kChunk + lo * kSpan + (i - pre[wave][lo]);
    keys[at + (uint64_t)i] = skey[slot];
    vals[at + (uint64_t)i] = sval[slot];
  }
}

// keep[i] = 1 unless entry i lies in a run of more than F equal k-mers.
__global__ __launch_bounds__(kBlock) void k_keep(const uint64_t* keys, int64_t n, int F, uint64_t* keep) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  bool drop = false;
  for (int64_t j = i - F > 0 ? i - F : 0; j <= i && j + F < n; ++j)
    if (keys[j] == keys[j + F]) {
      drop = true;
      break;
    }
  keep[i] = drop ? 0 : 1;
}

__global__ __launch_bounds__(kBlock) void k_scatter(const uint64_t* keys, const int64_t* vals, const uint64_t* keep,
                                                    const uint64_t* at, int64_t n, uint64_t* okeys, int64_t* ovals) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n || !keep[i]) return;
  okeys[at[i]] = keys[i];
  ovals[at[i]] = vals[i];
}

__device__ inline int64_t lower_bound(const uint64_t* a, int64_t n, uint64_t v) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Join: per query entry, the target entries of its k-mer (pass 0 counts, pass 1 writes
// {|query position|, |target position|, orientation} in query-table order).
template <bool WRITE>
__global__ __launch_bounds__(kBlock) void k_join(const uint64_t* qk, const int64_t* qv, int64_t nq,
                                                 const uint64_t* tk, const int64_t* tv, int64_t nt, uint64_t* count,
                                                 const uint64_t* base, uint64_t* jq, int64_t* jt, uint8_t* jrev) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= nq) return;
  const uint64_t k = qk[i];
  const int64_t lb = lower_bound(tk, nt, k);
  if (!WRITE) {
    int64_t c = 0;
    while (lb + c < nt && tk[lb + c] == k) ++c;
    count[i] = (uint64_t)c;
    return;
  }
  uint64_t at = base[i];
  const int64_t q = qv[i];
  for (int64_t j = lb; j < nt && tk[j] == k; ++j, ++at) {
    // the reference tests position > 0: a {0, 0} slot counts as reverse
    bool rev = !(q > 0);
    const int64_t t = tv[j];
    if (!(t > 0)) rev = !rev;
    jq[at] = (uint64_t)(q > 0 ? q : -q);
    jt[at] = t > 0 ? t : -t;
    jrev[at] = rev ? 1 : 0;
  }
}

__global__ __launch_bounds__(kBlock) void k_iota(uint64_t* v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) v[i] = (uint64_t)i;
}

// The sweep's inputs in query order: the target position and strand of each sorted entry
// (order is a permutation of 0 .. n - 1).
__global__ __launch_bounds__(kBlock) void k_gather(const uint64_t* order, const int64_t* t, const uint8_t* r,
                                                   int64_t* t_out, uint8_t* r_out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    const uint64_t j = order[i];
    t_out[i] = t[j];
    r_out[i] = r[j];
  }
}

inline unsigned grid_of(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// Device buffers of one run, freed on return.
struct Buffers {
  std::vector<void*> ptrs;
  ~Buffers() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <class T>
  hipError_t get(T** p, int64_t n) {
    void* q = nullptr;
    const hipError_t e = hipMalloc(&q, (size_t)(n > 0 ? n : 1) * sizeof(T));
    if (e == hipSuccess) ptrs.push_back(q);
    *p = (T*)q;
    return e;
  }
};

#define KCHK(expr)                                                                                          \
  do {                                                                                                      \
    hipError_t e_ = (expr);                                                                                 \
    if (e_ != hipSuccess) return fail(SXC_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
  } while (0)

// exclusive scan of n uint64 values; returns the total in *total
int scan(Buffers& B, const uint64_t* in, uint64_t* out, int64_t n, uint64_t* total, hipStream_t s) {
  if (n == 0) {
    *total = 0;
    return SXC_OK;
  }
  size_t bytes = 0;
  KCHK(rocprim::exclusive_scan(nullptr, bytes, in, out, (uint64_t)0, (size_t)n, rocprim::plus<uint64_t>(), s));
  void* tmp = nullptr;
  KCHK(B.get((char**)&tmp, (int64_t)bytes));
  KCHK(rocprim::exclusive_scan(tmp, bytes, in, out, (uint64_t)0, (size_t)n, rocprim::plus<uint64_t>(), s));
  uint64_t last_out = 0, last_in = 0;
  KCHK(hipMemcpyAsync(&last_out, out + n - 1, 8, hipMemcpyDeviceToHost, s));
  KCHK(hipMemcpyAsync(&last_in, in + n - 1, 8, hipMemcpyDeviceToHost, s));
  KCHK(hipStreamSynchronize(s));
  *total = last_out + last_in;
  return SXC_OK;
}

// One genome's table: built, sorted by k-mer, filtered; the sorted kept entries are left in
// (*keys, *vals), *n of them.
int build_table(Buffers& B, const Side& side, int K, int F, hipStream_t s, uint64_t** keys, int64_t** vals,
                int64_t* n, int64_t* kept) {
  const int64_t len = (int64_t)side.buf.size();
  const int64_t slots = side.slots;
  const int nseq = (int)side.off.size();
  *n = 0;
  *kept = 0;
  if (slots == 0) return SXC_OK;
  unsigned char* dbuf;
  int64_t *doff, *didx;
  KCHK(B.get(&dbuf, len));
  KCHK(B.get(&doff, nseq));
  KCHK(B.get(&didx, nseq));
  KCHK(hipMemcpyAsync(dbuf, side.buf.data(), (size_t)len, hipMemcpyHostToDevice, s));
  KCHK(hipMemcpyAsync(doff, side.off.data(), sizeof(int64_t) * nseq, hipMemcpyHostToDevice, s));
  KCHK(hipMemcpyAsync(didx, side.index.data(), sizeof(int64_t) * nseq, hipMemcpyHostToDevice, s));
  const int64_t chunks = (len + kChunk - 1) / kChunk;
  uint64_t *cnt, *base;
  KCHK(B.get(&cnt, chunks));
  KCHK(B.get(&base, chunks));
  hipLaunchKernelGGL(k_kmers<false>, dim3(grid_of(chunks * 64)), dim3(kBlock), 0, s, dbuf, len, K, chunks, cnt,
                     nullptr, doff, didx, nseq, nullptr, nullptr);
  KCHK(hipGetLastError());
  uint64_t filled = 0;
  int rc = scan(B, cnt, base, chunks, &filled, s);
  if (rc != SXC_OK) return rc;
  if ((int64_t)filled > slots) return fail(SXC_ERR_ARG, "k-mer table overflow");  // cannot happen: one per slot
  uint64_t *k0, *k1;
  int64_t *v0, *v1;
  KCHK(B.get(&k0, slots));
  KCHK(B.get(&v0, slots));
  KCHK(B.get(&k1, slots));
  KCHK(B.get(&v1, slots));
  // slots no k-mer fills stay {0, 0}
  KCHK(hipMemsetAsync(k0, 0, sizeof(uint64_t) * slots, s));
  KCHK(hipMemsetAsync(v0, 0, sizeof(int64_t) * slots, s));
  hipLaunchKernelGGL(k_kmers<true>, dim3(grid_of(chunks * 64)), dim3(kBlock), 0, s, dbuf, len, K, chunks, nullptr,
                     base, doff, didx, nseq, k0, v0);
  KCHK(hipGetLastError());
  size_t bytes = 0;
  KCHK(rocprim::radix_sort_pairs(nullptr, bytes, k0, k1, v0, v1, (size_t)slots, 0, 2 * K, s));
  void* tmp;
  KCHK(B.get((char**)&tmp, (int64_t)bytes));
  KCHK(rocprim::radix_sort_pairs(tmp, bytes, k0, k1, v0, v1, (size_t)slots, 0, 2 * K, s));
  uint64_t *keep, *at;
  KCHK(B.get(&keep, slots));
  KCHK(B.get(&at, slots));
  hipLaunchKernelGGL(k_keep, dim3(grid_of(slots)), dim3(kBlock), 0, s, k1, slots, F, keep);
  KCHK(hipGetLastError());
  uint64_t nk = 0;
  rc = scan(B, keep, at, slots, &nk, s);
  if (rc != SXC_OK) return rc;
  hipLaunchKernelGGL(k_scatter, dim3(grid_of(slots)), dim3(kBlock), 0, s, k1, v1, keep, at, slots, k0, v0);
  KCHK(hipGetLastError());
  *kept = (int64_t)nk;
  *n = nk > 0 ? (int64_t)nk - 1 : 0;  // kposv.resize(wi - 1)
  *keys = k0;
  *vals = v0;
  return SXC_OK;
}

}  // namespace

struct sxc_kmatch {
  std::vector<sxc_result> out;
  sxc_kmatch_stats st{};
};

namespace {

// dump_matching_blocks (KMatch.cc:196-318) over the joined positions sorted by query
// position: a block grows while the next joined position on its diagonal (anti-diagonal
// for reverse) lies within max_jump of its end; blocks are reported when the sweep passes
// their end by more than max_jump, in the order they finish (then creation order), if they
// span at least min_length bases.  The reference starts at the second joined position and
// ends with one read past the table; the first is skipped here too and the last step only
// closes the open blocks.
void sweep(const std::vector<uint64_t>& q, const std::vector<int64_t>& t, const std::vector<uint8_t>& rev,
           const sxc_kmatch_params& p, const Side& qs, const Side& ts, std::vector<sxc_result>* out) {
  struct Block {
    int64_t q_start, t_start;
    bool reverse;
    int64_t length;
  };
  // the open blocks in creation order, kept in a vector (compacted in place each step)
  std::vector<Block> active;
  const uint64_t n = q.size();
  for (uint64_t i = 1; i <= n; ++i) {
    const bool last = i == n;
    bool used = false;
    int64_t qd = 0;
    size_t w = 0;
    for (size_t r = 0; r < active.size(); ++r) {
      Block b = active[r];
      if (!last) {
        qd = (int64_t)q[i] - b.q_start;
        const int64_t td = b.reverse ? b.t_start - t[i] : t[i] - b.t_start;
        if (qd == td && qd - b.length <= p.max_jump && b.reverse == (rev[i] != 0)) {
          b.length = qd;
          used = true;
        }
      }
      if (last || qd - b.length > p.max_jump) {
        if (b.length + p.K >= p.min_length) {
          sxc_result res{};
          res.query_id = (uint64_t)(b.q_start / kRecordStride - 1);
          res.target_id = (uint64_t)(b.t_start / kRecordStride - 1);
          // query_seqs[query_id] (a record skipped as empty shifts this lookup, as there)
          res.query_size = res.query_id < qs.lengths.size() ? qs.lengths[res.query_id] : 0;
          res.qstart = (uint64_t)(b.q_start % kRecordStride - 1);
          res.tstart = b.reverse ? (uint64_t)((b.t_start - b.length) % kRecordStride - 1)
                                 : (uint64_t)(b.t_start % kRecordStride - 1);
          res.len = (uint64_t)(b.length + p.K);
          res.reverse = b.reverse ? 1 : 0;
          res.prob = 1;
          res.ident = 1;
          out->push_back(res);
        }
      } else {
        active[w++] = b;
      }
    }
    active.resize(w);
    if (!used && !last) active.push_back(Block{(int64_t)q[i], t[i], rev[i] != 0, 0});
  }
  (void)ts;
}

int run(sxc_ctx* ctx, const Side& qs, const Side& ts, const sxc_kmatch_params& p, double read_ms, sxc_kmatch** out) {
  std::lock_guard<std::mutex> lock(ctx->mu);
  KCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  sxc_kmatch* res = new (std::nothrow) sxc_kmatch;
  if (!res) return fail(SXC_ERR_NOMEM, "sxc_kmatch: out of memory");
  std::unique_ptr<sxc_kmatch> hold(res);
  res->st.read_ms = read_ms;
  res->st.query_kmers = qs.slots;
  res->st.target_kmers = ts.slots;
  hipEvent_t e0, e1;
  KCHK(hipEventCreate(&e0));
  KCHK(hipEventCreate(&e1));
  std::vector<uint64_t> hq;
  std::vector<int64_t> ht;
  std::vector<uint8_t> hr;
  {
    Buffers B;
    KCHK(hipEventRecord(e0, s));
    uint64_t *qk = nullptr, *tk = nullptr;
    int64_t *qv = nullptr, *tv = nullptr, nq = 0, nt = 0;
    int rc = build_table(B, qs, p.K, p.max_freq, s, &qk, &qv, &nq, &res->st.query_kept);
    if (rc != SXC_OK) return rc;
    rc = build_table(B, ts, p.K, p.max_freq, s, &tk, &tv, &nt, &res->st.target_kept);
    if (rc != SXC_OK) return rc;
    uint64_t nj = 0;
    uint64_t* base = nullptr;  // each query entry's first joined slot
    if (nq > 0 && nt > 0) {
      uint64_t* cnt;
      KCHK(B.get(&cnt, nq));
      KCHK(B.get(&base, nq));
      hipLaunchKernelGGL(k_join<false>, dim3(grid_of(nq)), dim3(kBlock), 0, s, qk, qv, nq, tk, tv, nt, cnt, nullptr,
                         nullptr, nullptr, nullptr);
      KCHK(hipGetLastError());
      rc = scan(B, cnt, base, nq, &nj, s);
      if (rc != SXC_OK) return rc;
    }
    res->st.joined = (int64_t)nj;
    if (nj > 0) {
      uint64_t *jq, *jq2, *ix, *ix2;
      int64_t* jt;
      uint8_t* jr;
      KCHK(B.get(&jq, (int64_t)nj));
      KCHK(B.get(&jq2, (int64_t)nj));
      KCHK(B.get(&ix, (int64_t)nj));
      KCHK(B.get(&ix2, (int64_t)nj));
      KCHK(B.get(&jt, (int64_t)nj));
      KCHK(B.get(&jr, (int64_t)nj));
      hipLaunchKernelGGL(k_join<true>, dim3(grid_of(nq)), dim3(kBlock), 0, s, qk, qv, nq, tk, tv, nt, nullptr, base,
                         jq, jt, jr);
      KCHK(hipGetLastError());
      hipLaunchKernelGGL(k_iota, dim3(grid_of((int64_t)nj)), dim3(kBlock), 0, s, ix, (int64_t)nj);
      KCHK(hipGetLastError());
      // by query position; stable, so a query position's target entries keep table order
      int bits = 1;
      const uint64_t maxq = (uint64_t)(qs.index.empty() ? 1 : qs.index.back() + 1) * (uint64_t)kRecordStride;
      while (bits < 64 && (maxq >> bits)) ++bits;
      size_t bytes = 0;
      KCHK(rocprim::radix_sort_pairs(nullptr, bytes, jq, jq2, ix, ix2, (size_t)nj, 0, bits, s));
      void* tmp;
      KCHK(B.get((char**)&tmp, (int64_t)bytes));
      KCHK(rocprim::radix_sort_pairs(tmp, bytes, jq, jq2, ix, ix2, (size_t)nj, 0, bits, s));
      int64_t* gt;
      uint8_t* gr;
      KCHK(B.get(&gt, (int64_t)nj));
      KCHK(B.get(&gr, (int64_t)nj));
      hipLaunchKernelGGL(k_gather, dim3(grid_of((int64_t)nj)), dim3(kBlock), 0, s, ix2, jt, jr, gt, gr, (int64_t)nj);
      KCHK(hipGetLastError());
      hq.resize(nj);
      ht.resize(nj);
      hr.resize(nj);
      KCHK(hipMemcpyAsync(hq.data(), jq2, 8 * nj, hipMemcpyDeviceToHost, s));
      KCHK(hipMemcpyAsync(ht.data(), gt, 8 * nj, hipMemcpyDeviceToHost, s));
      KCHK(hipMemcpyAsync(hr.data(), gr, nj, hipMemcpyDeviceToHost, s));
      KCHK(hipStreamSynchronize(s));
    }
    KCHK(hipEventRecord(e1, s));
    KCHK(hipEventSynchronize(e1));
  }
  float ms = 0;
  KCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  res->st.device_ms = ms;
  const auto t0 = std::chrono::steady_clock::now();
  sweep(hq, ht, hr, p, qs, ts, &res->out);
  res->st.sweep_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  res->st.blocks = (int64_t)res->out.size();
  *out = hold.release();
  return SXC_OK;
}

int check_params(const sxc_kmatch_params* p) {
  if (!p || p->K < 1 || p->K > 31 || p->K % 2 == 0 || p->max_freq < 0 || p->max_jump < 0)
    return fail(SXC_ERR_ARG, "sxc_kmatch: K must be odd and in 1..31, max_freq and max_jump >= 0");
  return SXC_OK;
}

}  // namespace

extern "C" {

int sxc_kmatch_files(sxc_ctx* ctx, const char* query_fasta, const char* target_fasta, const sxc_kmatch_params* p,
                     sxc_kmatch** out) {
  if (!ctx || !query_fasta || !target_fasta || !out) return fail(SXC_ERR_ARG, "sxc_kmatch_files: null argument");
  int rc = check_params(p);
  if (rc != SXC_OK) return rc;
  const auto t0 = std::chrono::steady_clock::now();
  Side qs, ts;
  rc = read_side(query_fasta, p->K, &qs);
  if (rc != SXC_OK) return rc;
  rc = read_side(target_fasta, p->K, &ts);
  if (rc != SXC_OK) return rc;
  const double read_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return run(ctx, qs, ts, *p, read_ms, out);
}

int sxc_kmatch_records(sxc_ctx* ctx, int32_t n_queries, const char* const* query_seqs, const int64_t* query_lens,
                       int32_t n_targets, const char* const* target_seqs, const int64_t* target_lens,
                       const sxc_kmatch_params* p, sxc_kmatch** out) {
  if (!ctx || !out || n_queries < 0 || n_targets < 0 || (n_queries && (!query_seqs || !query_lens)) ||
      (n_targets && (!target_seqs || !target_lens)))
    return fail(SXC_ERR_ARG, "sxc_kmatch_records: bad arguments");
  int rc = check_params(p);
  if (rc != SXC_OK) return rc;
  Side qs, ts;
  for (int i = 0; i < n_queries && rc == SXC_OK; ++i)
    rc = (!query_seqs[i] || query_lens[i] < 0) ? fail(SXC_ERR_ARG, "sxc_kmatch_records: bad query record")
                                              : add_sequence(&qs, query_seqs[i], query_lens[i], i + 1, p->K, "query");
  for (int i = 0; i < n_targets && rc == SXC_OK; ++i)
    rc = (!target_seqs[i] || target_lens[i] < 0)
             ? fail(SXC_ERR_ARG, "sxc_kmatch_records: bad target record")
             : add_sequence(&ts, target_seqs[i], target_lens[i], i + 1, p->K, "target");
  if (rc != SXC_OK) return rc;
  return run(ctx, qs, ts, *p, 0., out);
}

int sxc_kmatch_results(const sxc_kmatch* k, sxc_result* out, int64_t cap, int64_t* n) {
  if (!k || !n || cap < 0 || (cap && !out)) return fail(SXC_ERR_ARG, "sxc_kmatch_results: bad arguments");
  *n = (int64_t)k->out.size();
  for (int64_t i = 0; i < *n && i < cap; ++i) out[i] = k->out[(size_t)i];
  return *n > cap ? fail(SXC_ERR_CAPACITY, "sxc_kmatch_results: more records than cap") : SXC_OK;
}

int sxc_kmatch_get_stats(const sxc_kmatch* k, sxc_kmatch_stats* st) {
  if (!k || !st) return fail(SXC_ERR_ARG, "sxc_kmatch_get_stats: bad arguments");
  *st = k->st;
  return SXC_OK;
}

void sxc_kmatch_destroy(sxc_kmatch* k) { delete k; }

}  // extern "C"
