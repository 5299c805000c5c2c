// HomologyByXCorr -- the standalone cross-correlation aligner with its chunk pairs on the GPU.
//
// A drop-in for tools/analysis/HomologyByXCorr.cc in its default mode: same flags
// (:508-603), same chunking (target chunks of -t_chunk every t_chunk/2, query chunks of
// -q_chunk, :636-650), the same matches per chunk pair (HomologyByXCorr::Align, :168-225:
// forward then reverse complement, -cutoff, -l, -min_prob, RCQuery with the query chunk's
// own length) in the same order (target chunks outer, query chunks inner, :712-800), and
// the same output file (MultiMatches::Write, version 3, analysis/SequenceMatch.cc:320-360).
// Every chunk pair runs in batched launches through libsxc_hip.so.
//
// The reference reuses one CCSignal object for every target chunk and one for every query
// chunk (:706-716, 780-798), and CCSignal::SetSequence writes only the first L positions
// of its N-long channels (analysis/CrossCorr.cc:118-131, 182-206).  A chunk shorter than
// an earlier one therefore correlates with that earlier chunk's signal in its tail.  This
// program replays the order of SetSequence calls on a stack per object to name, for every
// signal, the earlier sequences whose values sit in its tail, and hands them to the library
// as signal tails (sxc_batch_create_tails), which encodes them on the GPU.
//
// The reference's optional modes that change what is aligned or how results are post-
// processed (-guide, -proteins, -chain, -select, -pairs, -line) are not part of this
// program; naming one is an error.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <deque>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../../include/sxc.h"
#include "sxc_flags.h"
#include "sxc_seqio.h"

using sxc_host::ChunkSet;
using sxc_host::Records;

namespace {

constexpr int kMaxPairsPerLaunch = 16384;

// Selection of ChunkManager::ChunkItSelect (analysis/SeqChunk.cc:101-121): with nBlocks > 0
// only chunks k in [first, last) keep their bytes; the rest stay empty.
void restrict_blocks(ChunkSet* cs, int n_blocks, int my_block) {
  if (n_blocks <= 0) return;
  const int n = (int)cs->chunks.size();
  const int bsize = (n + n_blocks - 1) / n_blocks;
  int first = my_block * bsize, last = (my_block + 1) * bsize;
  if (first == last) {
    printf("Nonsensical block assignment!\n");
    first = 0;
    last = n;
  }
  printf("Will process target chunks %d through %d\n", first, last - 1);
  for (int k = 0; k < n; ++k)
    if (k < first || k >= last) cs->chunks[(size_t)k].len = 0;
}

// Replays SetSequence calls on one reused CCSignal.  The stack holds the sequences whose
// values are still visible, with strictly decreasing lengths from bottom to top: entry e
// owns positions [length of the entry above it (or of the newest call), e.len).  A call of
// length L hides every entry of length <= L; the rest form its tail, top to bottom.
class SignalHistory {
 public:
  struct Src {
    int32_t len;
    const char* seq;
    int32_t strand;
    bool operator==(const Src& o) const { return len == o.len && seq == o.seq && strand == o.strand; }
    bool operator<(const Src& o) const { return std::tie(len, seq, strand) < std::tie(o.len, o.seq, o.strand); }
  };
  // Records the call and returns its tail (segments interned in `pool`).
  sxc_signal_tail call(const Src& s, std::map<std::vector<Src>, std::vector<sxc_tail_seg>*>& pool,
                       std::deque<std::vector<sxc_tail_seg>>& store) {
    while (!stack_.empty() && stack_.back().len <= s.len) stack_.pop_back();
    std::vector<Src> below(stack_.rbegin(), stack_.rend());
    stack_.push_back(s);
    if (below.empty()) return sxc_signal_tail{nullptr, 0, 0};
    auto it = pool.find(below);
    if (it == pool.end()) {
      store.emplace_back();
      for (const Src& b : below) store.back().push_back(sxc_tail_seg{b.len, b.len, b.seq, b.strand, 0});
      it = pool.emplace(below, &store.back()).first;
    }
    return sxc_signal_tail{it->second->data(), (int32_t)it->second->size(), 0};
  }
  const std::vector<Src>& state() const { return stack_; }
  void set_state(const std::vector<Src>& s) { stack_ = s; }

 private:
  std::vector<Src> stack_;
};

// MultiMatches::Write, version 3 (analysis/SequenceMatch.cc:320-360, SingleMatch::Write
// :49-64; CMWriteFileStream writes raw host-order values and strings as int64 length
// (strlen + 1) followed by the bytes and their NUL, util/mutil.cc:1005-1049).
class MatchFile {
 public:
  explicit MatchFile(const std::string& path) : f_(fopen(path.c_str(), "wb")) {}
  ~MatchFile() {
    if (f_) fclose(f_);
  }
  bool ok() const { return f_ != nullptr && good_; }
  void i32(int32_t v) { raw(&v, 4); }
  void f64(double v) { raw(&v, 8); }
  void str(const std::string& s) {
    const std::string c = s.c_str();  // up to the first NUL, as strlen sees it
    const int64_t len = (int64_t)c.size() + 1;
    raw(&len, 8);
    raw(c.c_str(), (size_t)len);
  }
  bool close() {
    const bool r = f_ && fclose(f_) == 0 && good_;
    f_ = nullptr;
    return r;
  }

 private:
  void raw(const void* p, size_t n) {
    if (f_ && fwrite(p, 1, n, f_) != n) good_ = false;
  }
  FILE* f_;
  bool good_ = true;
};

}  // namespace

int main(int argc, char** argv) {
  sxc_host::Flags P(argc, argv, "Compares two sequences via cross-correlation");
  P.add("-q", "string", "query fasta sequence", nullptr);
  P.add("-t", "string", "target fasta sequence", nullptr);
  P.add("-cq", "int", "query sequence id", "-1");
  P.add("-ct", "int", "target sequence id", "-1");
  P.add("-l", "int", "minimum alignment length", "0");
  P.add("-t_chunk", "int", "target chunk size", "4096");
  P.add("-q_chunk", "int", "query chunk size", "4096");
  P.add("-o", "int", "output file (binary)", nullptr);
  P.add("-nblocks", "int", "total number of blocks to process (when parallel)", "0");
  P.add("-block", "int", "block # to process (when parallel)", "0");
  P.add("-nblocks_query", "int", "total number of QUERY blocks to process (when parallel)", "0");
  P.add("-block_query", "int", "QUERY block # to process (when parallel)", "0");
  P.add("-same_only", "bool", "only align sequences that have the same name.", "0");
  P.add("-min_prob", "double", "minimum probability to keep match", "0.9999");
  P.add("-guide", "string", "XCorr input file (binary) to guide alignments", "");
  P.add("-proteins", "bool", "translate seqeunces to proteins when aligning.", "0");
  P.add("-cutoff", "double", "signal selection cutoff", "1.8");
  P.add("-chain", "bool", "chain matches before writing them out", "0");
  P.add("-select", "string", "selection file", "");
  P.add("-pairs", "string", "valid block pairs to process", "");
  P.add("-line", "double", "pick one block and compare to the query", "-1.0");
  P.add("-device", "int", "HIP device to run on", "0");
  P.add("-autotune", "bool", "time the compiled FFT kernel geometries at this length and run the fastest", "1");
  P.parse();

  const std::string qfile = P.str("-q"), tfile = P.str("-t"), output = P.str("-o");
  const std::string guide = P.str("-guide");
  const int min_len = P.i32("-l");
  const int t_chunk = P.i32("-t_chunk"), q_chunk = P.i32("-q_chunk");
  const int n_blocks = P.i32("-nblocks"), my_block = P.i32("-block");
  const int n_blocks_q = P.i32("-nblocks_query"), my_block_q = P.i32("-block_query");
  const bool same_only = P.boolean("-same_only");
  const double min_prob = P.f64("-min_prob");
  const double cutoff = P.f64("-cutoff");
  const double one_line = P.f64("-line");
  const int device = P.i32("-device");
  const bool autotune = P.boolean("-autotune");
  (void)P.i32("-cq");
  (void)P.i32("-ct");

  if (output.empty()) {
    printf("You MUST specify a valid output file (option '-o')!\n");
    return -1;
  }
  printf("Command line:\n");
  for (int i = 0; i < argc; ++i) printf("%s ", argv[i]);
  printf("\n\n");
  if (!guide.empty() && n_blocks != 0) {
    printf("Guided mode, chunking not available!!\n");
    return -1;
  }
  const char* unsupported = !guide.empty()              ? "-guide"
                            : P.boolean("-proteins")    ? "-proteins"
                            : P.boolean("-chain")       ? "-chain"
                            : !P.str("-select").empty() ? "-select"
                            : !P.str("-pairs").empty()  ? "-pairs"
                            : one_line >= -0.1          ? "-line"
                                                        : nullptr;
  if (unsupported) {
    printf("ERROR %s is not supported by this build\n", unsupported);
    return 1;
  }

  Records qraw, traw;
  std::string err;
  if (!sxc_host::read_fasta(qfile, &qraw, &err) || !sxc_host::read_fasta(tfile, &traw, &err)) {
    printf("ERROR %s\n", err.c_str());
    return 1;
  }
  printf("Query sequence:  %s\nTarget sequence: %s\nDone loading.\n", qfile.c_str(), tfile.c_str());
  ChunkSet target, query;
  sxc_host::chunk_records(traw, t_chunk, t_chunk / 2, &target);
  sxc_host::chunk_records(qraw, q_chunk, 0, &query);
  restrict_blocks(&target, n_blocks, my_block);
  restrict_blocks(&query, n_blocks_q, my_block_q);
  printf("Done chunking.\nKeeping alignments more like real than %g\n", min_prob);

  // Chunk pairs in the reference's loop order (:712-800); empty chunks are skipped there.
  // The query object's inner loop is the same for every target chunk of one record name,
  // so its tails are computed once per (starting state, name) and reused.
  std::vector<sxc_pair> pairs;
  std::vector<sxc_signal_tail> t_tails, q_tails;
  std::map<std::vector<SignalHistory::Src>, std::vector<sxc_tail_seg>*> pool;
  std::deque<std::vector<sxc_tail_seg>> store;
  SignalHistory thist, qhist;
  struct Loop {
    std::vector<sxc_signal_tail> tails;  // per query chunk index (nseg = 0 where skipped)
    std::vector<SignalHistory::Src> end_state;
  };
  std::map<std::pair<std::vector<SignalHistory::Src>, std::string>, Loop> loops;
  for (const sxc_host::Chunk& t : target.chunks) {
    if (t.len == 0) continue;
    const sxc_signal_tail ttail = thist.call({t.len, t.data, 0}, pool, store);
    const std::string& tname = traw.names[(size_t)t.record];
    const auto key = std::make_pair(qhist.state(), same_only ? tname : std::string());
    auto li = loops.find(key);
    if (li == loops.end()) {
      Loop lp;
      lp.tails.assign(query.chunks.size(), sxc_signal_tail{nullptr, 0, 0});
      for (size_t qi = 0; qi < query.chunks.size(); ++qi) {
        const sxc_host::Chunk& q = query.chunks[qi];
        if (q.len == 0) continue;
        if (same_only && tname != qraw.names[(size_t)q.record]) continue;
        lp.tails[qi] = qhist.call({q.len, q.data, 0}, pool, store);  // forward (:780-782)
        qhist.call({q.len, q.data, 1}, pool, store);                  // reverse complement (:786-792)
      }
      lp.end_state = qhist.state();
      li = loops.emplace(key, std::move(lp)).first;
    }
    qhist.set_state(li->second.end_state);
    for (size_t qi = 0; qi < query.chunks.size(); ++qi) {
      const sxc_host::Chunk& q = query.chunks[qi];
      if (q.len == 0) continue;
      if (same_only && tname != qraw.names[(size_t)q.record]) continue;
      t_tails.push_back(ttail);
      q_tails.push_back(li->second.tails[qi]);
      sxc_pair s;
      s.tseq = t.data;
      s.tlen = t.len;
      s.qseq = q.data;
      s.qlen = q.len;
      s.t_chunk_start = t.start;
      s.q_chunk_start = q.start;
      s.t_record = t.record;
      s.q_record = q.record;
      s.q_record_len = query.lengths[(size_t)q.record];
      s.fast = 0;
      pairs.push_back(s);
    }
  }

  sxc_params prm;
  prm.N = 2 * t_chunk;
  prm.q_chunk = 0;  // RCQuery with each query chunk's own length (:793-798, :226-228)
  prm.min_len = min_len;
  prm.strands = 2;
  prm.cutoff = cutoff;
  prm.cutoff_fast = cutoff;
  prm.min_prob = min_prob;
  prm.target_total = target.total();

  sxc_ctx* ctx = nullptr;
  if (sxc_create(&ctx, device) != SXC_OK) {
    printf("ERROR %s\n", sxc_last_error());
    return 1;
  }
  if (autotune) {
    sxc_tune_entry tune[16];
    int32_t nt = 0;
    if (sxc_autotune(ctx, prm.N, prm.N, tune, 16, &nt) != SXC_OK) {
      printf("ERROR %s\n", sxc_last_error());
      return 1;
    }
  }
  std::vector<sxc_result> res;
  for (size_t at = 0; at < pairs.size(); at += kMaxPairsPerLaunch) {
    const int n = (int)std::min(pairs.size() - at, (size_t)kMaxPairsPerLaunch);
    sxc_batch* b = nullptr;
    int rc = sxc_batch_create_tails(ctx, pairs.data() + at, n, &prm, t_tails.data() + at, q_tails.data() + at, &b);
    if (rc == SXC_OK) rc = sxc_batch_run(b);
    int64_t got = 0;
    std::vector<sxc_result> r;
    if (rc == SXC_OK) {
      r.resize((size_t)n * 8 + 64);
      rc = sxc_batch_results(b, r.data(), (int64_t)r.size(), &got);
      if (rc == SXC_ERR_CAPACITY) {
        r.resize((size_t)got);
        rc = sxc_batch_results(b, r.data(), (int64_t)r.size(), &got);
      }
    }
    sxc_batch_destroy(b);
    if (rc != SXC_OK) {
      printf("ERROR aligning %d chunk pairs: %s\n", n, sxc_last_error());
      sxc_destroy(ctx);
      return 1;
    }
    res.insert(res.end(), r.begin(), r.begin() + got);
  }
  sxc_destroy(ctx);

  printf("Saving alignments...\n");
  MatchFile f(output);
  if (!f.ok()) {
    printf("ERROR cannot write %s\n", output.c_str());
    return 1;
  }
  f.i32(3);
  f.i32((int32_t)traw.names.size());
  for (const std::string& s : traw.names) f.str(s);
  f.i32((int32_t)qraw.names.size());
  for (const std::string& s : qraw.names) f.str(s);
  printf("MultiMatches dump: %zu matches\n", res.size());
  f.i32((int32_t)res.size());
  for (const sxc_result& r : res) {
    f.i32((int32_t)r.target_id);
    f.i32((int32_t)r.query_id);
    f.i32((int32_t)r.query_size);
    f.i32((int32_t)r.tstart);
    f.i32((int32_t)r.qstart);
    f.i32((int32_t)r.len);
    f.i32(r.reverse ? 1 : 0);
    f.f64(0. + r.ident * (double)(int32_t)r.len);  // SingleMatch::AddMatches on a fresh match
    f.f64(r.prob);
    f.f64(r.ident);
  }
  for (int l : target.lengths) f.i32(l);
  for (int l : query.lengths) f.i32(l);
  if (!f.close()) {
    printf("ERROR writing %s\n", output.c_str());
    return 1;
  }
  printf("all done!\n");
  return 0;
}
