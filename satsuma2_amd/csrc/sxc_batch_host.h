// sxc_batch_host.h -- packs caller chunk pairs into the flat device layout.
//
// HBM layout of a batch (uploaded once, then read by every kernel):
//   seq      all distinct chunk byte strings back to back (1 byte per base)
//   tchunks  {offset, length, flags} per distinct target chunk; the target spectra
//            buffer holds 4*N floats per entry in the same order
//   qchunks  the same for distinct query chunks
//   pairs    one sxc_pair_dev per chunk pair (indices into the chunk tables + the
//            chunk metadata FilterMatches writes into t_result)
// Chunks are deduplicated by (pointer, length), so a target shared by many pairs, as in
// the slave's align_target blocks (analysis/HomologyByXCorrSlave.cc:290-299), is
// transformed once.
#pragma once
#include <map>
#include <string>
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
  int max_len = 0;
};

// Returns an empty string on success, else the reason the input is rejected.
inline std::string pack_batch(const sxc_pair* in, int n, int N, HostBatch& hb) {
  hb = HostBatch();
  std::map<std::pair<const char*, int>, int> tmap, qmap;
  auto add = [&](const char* s, int len, std::vector<sxc_chunk>& tab,
                 std::map<std::pair<const char*, int>, int>& m, int& idx) -> std::string {
    if (len < 0 || (len > 0 && !s)) return "negative length or null sequence";
    if (len > N) return "sequence longer than the FFT length N";
    auto key = std::make_pair(s, len);
    auto it = m.find(key);
    if (it != m.end()) {
      idx = it->second;
      return "";
    }
    sxc_chunk c;
    c.off = (int64_t)hb.seq.size();
    c.len = len;
    c.flags = 2;  // bit 1: only A, C, G, T (scores are 100 or 0)
    for (int i = 0; i < len; ++i) {
      const unsigned char b = (unsigned char)s[i];
      if (b >= 128) return "sequence byte >= 128";
      const int cl = class_of(b);
      if (class_has_thirds(cl)) c.flags |= 1;
      if (cl > 3) c.flags &= ~2;
      hb.seq.push_back(b);
    }
    while (hb.seq.size() % 16) hb.seq.push_back(0);  // next chunk 16-byte aligned
    idx = (int)tab.size();
    tab.push_back(c);
    m[key] = idx;
    if (len > hb.max_len) hb.max_len = len;
    return "";
  };
  hb.pairs.resize(n);
  for (int k = 0; k < n; ++k) {
    const sxc_pair& p = in[k];
    sxc_pair_dev& d = hb.pairs[k];
    // lags span +-N/2: the shorter chunk must fit in half the transform (the slave uses
    // N = 2 * t_chunk, so every chunk does)
    if (p.tlen > N / 2 && p.qlen > N / 2) return "both chunks longer than N/2";
    std::string e = add(p.tseq, p.tlen, hb.tchunks, tmap, d.t_chunk);
    if (!e.empty()) return e;
    e = add(p.qseq, p.qlen, hb.qchunks, qmap, d.q_chunk);
    if (!e.empty()) return e;
    d.t_chunk_start = p.t_chunk_start;
    d.q_chunk_start = p.q_chunk_start;
    d.t_record = p.t_record;
    d.q_record = p.q_record;
    d.q_record_len = p.q_record_len;
    d.fast = p.fast;
  }
  hb.seq.resize(hb.seq.size() + 16, 0);  // read-ahead of aligned word copies
  return "";
}

inline int nb_of(int N) {
  int nb = 0;
  while ((1 << nb) < N) ++nb;
  return ((1 << nb) == N) ? nb : -1;
}

}  // namespace sxc
