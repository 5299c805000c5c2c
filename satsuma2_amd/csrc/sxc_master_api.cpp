// sxc_master_api.cpp -- the C ABI of include/sxc_master.h over host/sxc_master.h.
#include <string.h>

#include <new>
#include <string>
#include <vector>

#include "../../include/sxc_master.h"
#include "host/sxc_master.h"
#include "host/sxc_seqio.h"

namespace sxc {
int fail(int code, const std::string& msg);  // sxc_api.hip: sets sxc_last_error()
}

using sxc::fail;
using sxc_master::Match;
using sxc_master::MatchSet;

struct sxc_matchset {
  MatchSet s;
};
struct sxc_coverage {
  sxc_master::CoverageTracker t;
};
struct sxc_grid {
  sxc_master::Grid g;
  std::vector<sxc_master::GridTarget> last;  // the last collection
};

static_assert(sizeof(sxc_result) == sizeof(sxc_master::Result), "t_result layouts");

namespace {

Match from_c(const sxc_match& c) {
  Match m;
  m.target = c.target_id;
  m.query = c.query_id;
  m.query_len = c.query_len;
  m.start_target = c.start_target;
  m.start_query = c.start_query;
  m.length = c.length;
  m.rc = c.rc != 0;
  m.matches = c.matches;
  m.prob = c.prob;
  m.ident = c.ident;
  return m;
}

sxc_match to_c(const Match& m) {
  sxc_match c;
  c.target_id = m.target;
  c.query_id = m.query;
  c.query_len = m.query_len;
  c.start_target = m.start_target;
  c.start_query = m.start_query;
  c.length = m.length;
  c.rc = m.rc ? 1 : 0;
  c.reserved = 0;
  c.matches = m.matches;
  c.prob = m.prob;
  c.ident = m.ident;
  return c;
}

// Record ids must name records of the set: the chaining and the grid index per-record
// arrays with them.
int check_ids(const MatchSet& s, const char* what) {
  const int nt = (int)s.target_size.size(), nq = (int)s.query_size.size();
  for (const Match& m : s.m)
    if (m.target < 0 || m.target >= nt || m.query < 0 || m.query >= nq)
      return fail(SXC_ERR_ARG, std::string(what) + ": a match names a record outside the set (target " +
                                   std::to_string(m.target) + ", query " + std::to_string(m.query) + ")");
  return SXC_OK;
}

}  // namespace

extern "C" {

sxc_matchset* sxc_matchset_create(int32_t n_targets, const int32_t* target_sizes, int32_t n_queries,
                                  const int32_t* query_sizes) {
  if (n_targets < 0 || n_queries < 0 || (n_targets && !target_sizes) || (n_queries && !query_sizes)) {
    fail(SXC_ERR_ARG, "sxc_matchset_create: bad record sizes");
    return nullptr;
  }
  sxc_matchset* s = new (std::nothrow) sxc_matchset;
  if (!s) {
    fail(SXC_ERR_NOMEM, "sxc_matchset_create: out of memory");
    return nullptr;
  }
  s->s.target_size.assign(target_sizes, target_sizes + n_targets);
  s->s.query_size.assign(query_sizes, query_sizes + n_queries);
  return s;
}

void sxc_matchset_destroy(sxc_matchset* s) { delete s; }

int sxc_matchset_add(sxc_matchset* s, const sxc_match* m, int64_t n) {
  if (!s || n < 0 || (n && !m)) return fail(SXC_ERR_ARG, "sxc_matchset_add: bad arguments");
  s->s.m.reserve(s->s.m.size() + (size_t)n);
  for (int64_t i = 0; i < n; ++i) s->s.m.push_back(from_c(m[i]));
  return SXC_OK;
}

int sxc_matchset_collect(sxc_matchset* s, const sxc_result* r, int64_t n, int from_file) {
  if (!s || n < 0 || (n && !r)) return fail(SXC_ERR_ARG, "sxc_matchset_collect: bad arguments");
  s->s.collect(reinterpret_cast<const sxc_master::Result*>(r), n, from_file != 0);
  return SXC_OK;
}

int64_t sxc_matchset_count(const sxc_matchset* s) { return s ? (int64_t)s->s.m.size() : SXC_ERR_ARG; }

int64_t sxc_matchset_get(const sxc_matchset* s, sxc_match* out, int64_t cap) {
  if (!s || cap < 0 || (cap && !out)) return fail(SXC_ERR_ARG, "sxc_matchset_get: bad arguments");
  const int64_t n = (int64_t)s->s.m.size();
  for (int64_t i = 0; i < n && i < cap; ++i) out[i] = to_c(s->s.m[(size_t)i]);
  return n;
}

int sxc_matchset_sort(sxc_matchset* s) {
  if (!s) return fail(SXC_ERR_ARG, "sxc_matchset_sort: null set");
  s->s.sort();
  return SXC_OK;
}

int sxc_matchset_collapse(sxc_matchset* s) {
  if (!s) return fail(SXC_ERR_ARG, "sxc_matchset_collapse: null set");
  s->s.collapse();
  return SXC_OK;
}

int sxc_matchset_length_filter(sxc_matchset* s, int64_t min_length) {
  if (!s || min_length < 0) return fail(SXC_ERR_ARG, "sxc_matchset_length_filter: bad arguments");
  s->s.length_filter((uint64_t)min_length);
  return SXC_OK;
}

int sxc_matchset_copy(const sxc_matchset* in, sxc_matchset* out) {
  if (!in || !out) return fail(SXC_ERR_ARG, "sxc_matchset_copy: null set");
  out->s = in->s;
  return SXC_OK;
}

int sxc_matchset_chain(const sxc_matchset* in, int dups, sxc_matchset* out) {
  if (!in || !out || in == out) return fail(SXC_ERR_ARG, "sxc_matchset_chain: bad sets");
  const int rc = check_ids(in->s, "sxc_matchset_chain");
  if (rc != SXC_OK) return rc;
  try {
    if (dups)
      sxc_master::chain_dups(in->s, &out->s);
    else
      sxc_master::chain(in->s, &out->s);
  } catch (const std::bad_alloc&) {
    return fail(SXC_ERR_NOMEM, "sxc_matchset_chain: out of memory");
  }
  return SXC_OK;
}

int sxc_matchset_interpolate(sxc_matchset* s) {
  if (!s) return fail(SXC_ERR_ARG, "sxc_matchset_interpolate: null set");
  const int rc = check_ids(s->s, "sxc_matchset_interpolate");
  if (rc != SXC_OK) return rc;
  sxc_master::interpolate(&s->s);
  return SXC_OK;
}

sxc_coverage* sxc_coverage_create(void) {
  sxc_coverage* c = new (std::nothrow) sxc_coverage;
  if (!c) fail(SXC_ERR_NOMEM, "sxc_coverage_create: out of memory");
  return c;
}

void sxc_coverage_destroy(sxc_coverage* c) { delete c; }

int sxc_coverage_update(sxc_coverage* c, const sxc_matchset* s, uint64_t* added, uint64_t* removed) {
  if (!c || !s || !added || !removed) return fail(SXC_ERR_ARG, "sxc_coverage_update: bad arguments");
  c->t.update(s->s, added, removed);
  return SXC_OK;
}

sxc_grid* sxc_grid_create(int32_t chunk, int32_t blocks, int32_t n_targets, const char* const* target_names,
                          const char* const* target_seqs, const int64_t* target_lens, int32_t n_queries,
                          const char* const* query_names, const char* const* query_seqs, const int64_t* query_lens) {
  if (chunk < 8 || blocks < 1 || n_targets < 0 || n_queries < 0 ||
      (n_targets && (!target_names || !target_seqs || !target_lens)) ||
      (n_queries && (!query_names || !query_seqs || !query_lens))) {
    fail(SXC_ERR_ARG, "sxc_grid_create: bad arguments (chunk >= 8, blocks >= 1, records)");
    return nullptr;
  }
  std::vector<sxc_master::Record> t((size_t)n_targets), q((size_t)n_queries);
  for (int i = 0; i < n_targets; ++i) {
    if (!target_names[i] || !target_seqs[i] || target_lens[i] < 0 || target_lens[i] > INT32_MAX) {
      fail(SXC_ERR_ARG, "sxc_grid_create: bad target record " + std::to_string(i));
      return nullptr;
    }
    t[(size_t)i] = sxc_master::Record{target_names[i], target_seqs[i], target_lens[i]};
  }
  for (int i = 0; i < n_queries; ++i) {
    if (!query_names[i] || !query_seqs[i] || query_lens[i] < 0 || query_lens[i] > INT32_MAX) {
      fail(SXC_ERR_ARG, "sxc_grid_create: bad query record " + std::to_string(i));
      return nullptr;
    }
    q[(size_t)i] = sxc_master::Record{query_names[i], query_seqs[i], query_lens[i]};
  }
  try {
    return new sxc_grid{sxc_master::Grid(chunk, blocks, t, q), {}};
  } catch (const std::bad_alloc&) {
    fail(SXC_ERR_NOMEM, "sxc_grid_create: out of memory");
    return nullptr;
  }
}

sxc_grid* sxc_grid_create_fasta(int32_t chunk, int32_t blocks, const char* target_fasta, const char* query_fasta) {
  if (!target_fasta || !query_fasta) {
    fail(SXC_ERR_ARG, "sxc_grid_create_fasta: null file name");
    return nullptr;
  }
  sxc_host::Records t, q;
  std::string err;
  if (!sxc_host::read_fasta(target_fasta, &t, &err) || !sxc_host::read_fasta(query_fasta, &q, &err)) {
    fail(SXC_ERR_ARG, "sxc_grid_create_fasta: " + err);
    return nullptr;
  }
  auto ptrs = [](const sxc_host::Records& r, std::vector<const char*>* names, std::vector<const char*>* seqs,
                 std::vector<int64_t>* lens) {
    for (size_t i = 0; i < r.seqs.size(); ++i) {
      names->push_back(r.names[i].c_str());
      seqs->push_back(r.seqs[i].data());
      lens->push_back((int64_t)r.seqs[i].size());
    }
  };
  std::vector<const char*> tn, ts, qn, qs;
  std::vector<int64_t> tl, ql;
  ptrs(t, &tn, &ts, &tl);
  ptrs(q, &qn, &qs, &ql);
  return sxc_grid_create(chunk, blocks, (int32_t)tn.size(), tn.data(), ts.data(), tl.data(), (int32_t)qn.size(),
                         qn.data(), qs.data(), ql.data());
}

void sxc_grid_destroy(sxc_grid* g) { delete g; }

int sxc_grid_chunks(const sxc_grid* g, int32_t* n_target_chunks, int32_t* n_query_chunks) {
  if (!g || !n_target_chunks || !n_query_chunks) return fail(SXC_ERR_ARG, "sxc_grid_chunks: bad arguments");
  *n_target_chunks = g->g.target_chunks();
  *n_query_chunks = g->g.query_chunks();
  return SXC_OK;
}

int sxc_grid_chunk(const sxc_grid* g, int query, int32_t index, int32_t* record, int32_t* start) {
  if (!g || !record || !start) return fail(SXC_ERR_ARG, "sxc_grid_chunk: bad arguments");
  const int n = query ? g->g.query_chunks() : g->g.target_chunks();
  if (index < 0 || index >= n) return fail(SXC_ERR_ARG, "sxc_grid_chunk: index out of range");
  *record = query ? g->g.query_chunk_record(index) : g->g.target_chunk_record(index);
  *start = query ? g->g.query_chunk_start(index) : g->g.target_chunk_start(index);
  return SXC_OK;
}

int sxc_grid_update(sxc_grid* g, const sxc_matchset* chained) {
  if (!g || !chained) return fail(SXC_ERR_ARG, "sxc_grid_update: bad arguments");
  const int rc = check_ids(chained->s, "sxc_grid_update");
  if (rc != SXC_OK) return rc;
  if ((int)chained->s.target_size.size() > g->g.target_records() ||
      (int)chained->s.query_size.size() > g->g.query_records())
    return fail(SXC_ERR_ARG, "sxc_grid_update: the set has more records than the grid");
  g->g.update(chained->s);
  return SXC_OK;
}

int sxc_grid_last(const sxc_grid* g, sxc_grid_target* out, int32_t cap, int32_t* count) {
  if (!g || cap < 0 || (cap && !out) || !count) return fail(SXC_ERR_ARG, "sxc_grid_last: bad arguments");
  const std::vector<sxc_master::GridTarget>& t = g->last;
  *count = (int32_t)t.size();
  for (size_t i = 0; i < t.size() && (int64_t)i < cap; ++i)
    out[i] = sxc_grid_target{t[i].target_first, t[i].target_last, t[i].query_first, t[i].query_last, t[i].x,
                             t[i].y, t[i].y2, t[i].count, t[i].fast ? 1 : 0};
  return (int64_t)t.size() > cap ? fail(SXC_ERR_CAPACITY, "sxc_grid_last: more targets than cap") : SXC_OK;
}

int sxc_grid_collect(sxc_grid* g, int32_t n, sxc_grid_target* out, int32_t cap, int32_t* count,
                     int32_t* candidates) {
  if (!g || n < 0 || cap < 0 || (cap && !out) || !count || !candidates)
    return fail(SXC_ERR_ARG, "sxc_grid_collect: bad arguments");
  g->last.clear();
  *candidates = g->g.collect(&g->last, n);
  return sxc_grid_last(g, out, cap, count);
}

}  // extern "C"
