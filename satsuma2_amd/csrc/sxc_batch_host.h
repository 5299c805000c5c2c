// sxc_batch_host.h -- packs caller chunk pairs into the flat device layout.
//
// HBM layout of a batch (uploaded once, then read by every kernel):
//   seq      all distinct chunk byte strings back to back (1 byte per base)
//   tchunks  {offset, length, flags} per distinct target chunk; the target spectra
//            buffer holds 4*N floats per entry in the same order
//   qchunks  the same for distinct query chunks
//   pairs    one sxc_pair_dev per chunk pair (indices into the chunk tables + the
//            chunk metadata FilterMatches writes into t_result)
//   ent      window entropies of the chunks with B/D/H/V letters, computed here with the
//            host's log (SXC_ENT_ROW floats per such chunk; sxc_chunk flags bits 8-31)
// Chunks are deduplicated by (pointer, length), so a target shared by many pairs, as in
// the slave's align_target blocks (analysis/HomologyByXCorrSlave.cc:290-299), is
// transformed once.
#pragma once
#include <string.h>

#include <map>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/sxc.h"
#include "sxc_tables.h"
#include "sxc_types.h"

namespace sxc {

struct HostBatch {
  std::vector<unsigned char> seq;
  std::vector<sxc_chunk> tchunks, qchunks;
  std::vector<sxc_pair_dev> pairs;
  std::vector<float> ent;
  int max_len = 0;
  // signal tails (sxc_batch_create_tails): source sequences, their strands, the tail and
  // segment tables, and the tail of each target chunk (-1: none)
  std::vector<sxc_chunk> schunks;
  std::vector<int32_t> sstrand;
  std::vector<sxc_tail_dev> tails;
  std::vector<sxc_tail_seg_dev> segs;
  std::vector<int32_t> ttail;
};

// Returns an empty string on success, else the reason the input is rejected.
// t_tail / q_tail (optional, n entries each) are the signal tails of include/sxc.h.
inline std::string pack_batch(const sxc_pair* in, int n, int N, HostBatch& hb,
                              const sxc_signal_tail* t_tail = nullptr, const sxc_signal_tail* q_tail = nullptr) {
  hb = HostBatch();
  using Key = std::tuple<const char*, int, int>;
  std::map<Key, int> tmap, qmap, smap;
  std::map<std::pair<const sxc_tail_seg*, int>, int> tailmap;
  auto bytes = [&](const char* s, int len, sxc_chunk& c) -> std::string {
    if (len < 0 || (len > 0 && !s)) return "negative length or null sequence";
    if (len > N) return "sequence longer than the FFT length N";
    c.off = (int64_t)hb.seq.size();
    c.len = len;
    // bit 1: only A, C, G, T (scores are 100 or 0); bit 2: only A, C, G, T, N
    c.flags = chunk_flags((const unsigned char*)s, len);
    if (c.flags < 0) return "sequence byte >= 128";
    hb.seq.resize(hb.seq.size() + (size_t)len);
    if (len) memcpy(hb.seq.data() + c.off, s, (size_t)len);
    if (c.flags & 1) {
      const size_t row = hb.ent.size() / SXC_ENT_ROW;
      if (row >= (1u << 24)) return "too many chunks with B, D, H or V letters";
      c.flags |= (int32_t)(row << 8);
      hb.ent.resize(hb.ent.size() + SXC_ENT_ROW);
      window_entropy_row(hb.seq.data() + c.off, len, N, hb.ent.data() + row * SXC_ENT_ROW);
    }
    while (hb.seq.size() % 16) hb.seq.push_back(0);  // next chunk 16-byte aligned
    if (len > hb.max_len) hb.max_len = len;
    return "";
  };
  auto add = [&](const char* s, int len, int tail, std::vector<sxc_chunk>& tab, std::map<Key, int>& m,
                 int& idx) -> std::string {
    const Key key(s, len, tail);
    auto it = m.find(key);
    if (it != m.end()) {
      idx = it->second;
      return "";
    }
    sxc_chunk c;
    std::string e = bytes(s, len, c);
    if (!e.empty()) return e;
    idx = (int)tab.size();
    tab.push_back(c);
    m[key] = idx;
    return "";
  };
  // A tail for a chunk of length L: segments must tile [L, end) upwards within each
  // source's length and N.  Returns the tail index (-1 for none) through idx.
  auto add_tail = [&](const sxc_signal_tail* t, int L, int& idx) -> std::string {
    idx = -1;
    if (!t || t->nseg == 0) return "";
    if (t->nseg < 0 || !t->segs) return "bad signal tail";
    int start = L;
    for (int k = 0; k < t->nseg; ++k) {
      const sxc_tail_seg& g = t->segs[k];
      if (g.end <= start || g.end > g.len || g.end > N || (g.strand != 0 && g.strand != 1))
        return "signal tail segments must tile [chunk length, end) within their sources";
      start = g.end;
    }
    const auto key = std::make_pair(t->segs, t->nseg);
    auto it = tailmap.find(key);
    if (it != tailmap.end()) {
      idx = it->second;
      return "";
    }
    sxc_tail_dev td{(int32_t)hb.segs.size(), t->nseg};
    for (int k = 0; k < t->nseg; ++k) {
      const sxc_tail_seg& g = t->segs[k];
      const Key sk(g.seq, g.len, g.strand);
      int src;
      auto si = smap.find(sk);
      if (si != smap.end()) {
        src = si->second;
      } else {
        sxc_chunk c;
        std::string e = bytes(g.seq, g.len, c);
        if (!e.empty()) return e;
        src = (int)hb.schunks.size();
        hb.schunks.push_back(c);
        hb.sstrand.push_back(g.strand);
        smap[sk] = src;
      }
      hb.segs.push_back(sxc_tail_seg_dev{g.end, src});
    }
    idx = (int)hb.tails.size();
    hb.tails.push_back(td);
    tailmap[key] = idx;
    return "";
  };
  hb.pairs.resize(n);
  for (int k = 0; k < n; ++k) {
    const sxc_pair& p = in[k];
    sxc_pair_dev& d = hb.pairs[k];
    // lags span +-N/2: the shorter chunk must fit in half the transform (the slave uses
    // N = 2 * t_chunk, so every chunk does)
    if (p.tlen > N / 2 && p.qlen > N / 2) return "both chunks longer than N/2";
    int tt, qt;
    std::string e = add_tail(t_tail ? &t_tail[k] : nullptr, p.tlen, tt);
    if (!e.empty()) return e;
    e = add_tail(q_tail ? &q_tail[k] : nullptr, p.qlen, qt);
    if (!e.empty()) return e;
    const size_t before = hb.tchunks.size();
    e = add(p.tseq, p.tlen, tt, hb.tchunks, tmap, d.t_chunk);
    if (!e.empty()) return e;
    if (hb.tchunks.size() > before) hb.ttail.push_back(tt);
    e = add(p.qseq, p.qlen, 0, hb.qchunks, qmap, d.q_chunk);
    if (!e.empty()) return e;
    d.t_chunk_start = p.t_chunk_start;
    d.q_chunk_start = p.q_chunk_start;
    d.t_record = p.t_record;
    d.q_record = p.q_record;
    d.q_record_len = p.q_record_len;
    d.fast = p.fast;
    d.q_tail = qt;
  }
  if (hb.tails.empty()) hb.ttail.clear();  // no tails at all: kernels skip the lookups
  hb.seq.resize(hb.seq.size() + 16, 0);  // read-ahead of aligned word copies
  return "";
}

inline int nb_of(int N) {
  int nb = 0;
  while ((1 << nb) < N) ++nb;
  return ((1 << nb) == N) ? nb : -1;
}

}  // namespace sxc
