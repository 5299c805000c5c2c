// sxc_master.cpp -- match sets, chaining, interpolation, coverage and the grid search of
// Satsuma2's master (declarations and references in sxc_master.h).
#include "sxc_master.h"

#include <math.h>

#include <algorithm>
#include <atomic>
#include <exception>
#include <mutex>
#include <system_error>
#include <thread>
#include <utility>

namespace sxc_master {

namespace {

template <class A>
bool less_fields(const A& a, const A& b) {
  if (a.target != b.target) return a.target < b.target;
  if (a.query != b.query) return a.query < b.query;
  if (a.rc != b.rc) return !a.rc;
  if (a.start_target == b.start_target) return a.length < b.length;
  return a.start_target < b.start_target;
}

// std::sort of p[0 .. n) with match_less.  std::sort's permutation follows from the outcomes
// of its comparisons alone, so sorting small records of the compared fields and the index,
// then moving each match once, puts equal matches in the order sorting the matches gives.
void sort_matches(Match* p, size_t n) {
  if (n < 2) return;
  struct Key {
    int32_t target, query, start_target, length;
    bool rc;
    int32_t idx;
  };
  std::vector<Key> k(n);
  for (size_t i = 0; i < n; ++i) k[i] = Key{p[i].target, p[i].query, p[i].start_target, p[i].length, p[i].rc, (int32_t)i};
  std::sort(k.begin(), k.end(), [](const Key& a, const Key& b) { return less_fields(a, b); });
  std::vector<Match> tmp(p, p + n);
  for (size_t i = 0; i < n; ++i) p[i] = tmp[(size_t)k[i].idx];
}

}  // namespace

bool match_less(const Match& a, const Match& b) { return less_fields(a, b); }

void MatchSet::sort() { std::sort(m.begin(), m.end(), match_less); }

// Neighbouring matches (in the set's order) on the same target and strand whose target and
// query starts both differ by less than 4 merge into one: the merged match keeps its first
// member's target start and spans the members' query extents.  The query ids are not
// compared.  The group still open when the list ends is not emitted, so a set of one match
// collapses to none (analysis/SequenceMatch.cc:418-469).
//
// In place: before step i at most i - 1 groups were written, to slots below i - 1, so m[i - 1]
// and m[i] are read before step i's write (to a slot below i) can reach m[i - 1].
void MatchSet::collapse() {
  if (m.empty()) return;
  auto near = [](int a, int b) { return a - b < 4 && b - a < 4; };  // Laps (:408-416)
  Match cur = m[0];
  size_t w = 0;
  for (size_t i = 1; i < m.size(); ++i) {
    const Match& a = m[i - 1];
    const Match& b = m[i];
    if (a.target == b.target && a.rc == b.rc && near(a.start_target, b.start_target) &&
        near(a.start_query, b.start_query)) {
      if (b.start_query < cur.start_query) {
        cur.length = cur.start_query + cur.length - b.start_query;
        cur.start_query = b.start_query;
      } else {
        cur.length = b.start_query + b.length - cur.start_query;
      }
    } else {
      const Match next = b;
      m[w++] = cur;
      cur = next;
    }
  }
  m.resize(w);
}

void MatchSet::length_filter(uint64_t min_length) {
  size_t w = 0;  // in place, order kept
  for (size_t i = 0; i < m.size(); ++i)
    if ((uint64_t)(int64_t)m[i].length >= min_length) m[w++] = m[i];
  m.resize(w);
}

int64_t MatchSet::collect(const Result* r, int64_t n, bool from_file) {
  // collect_new_matches reuses one SingleMatch for every result and only adds to its
  // m_matches, so each match carries the running sum of ident * len over the results
  // collected before it in the same call.
  double acc = 0.;
  m.reserve(m.size() + (size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    Result x = r[i];
    if (from_file && x.reverse) x.qstart = x.query_size - x.len - x.qstart;
    Match y;
    y.query = (int32_t)x.query_id;
    y.target = (int32_t)x.target_id;
    y.query_len = (int32_t)x.query_size;
    y.start_query = (int32_t)x.qstart;
    y.start_target = (int32_t)x.tstart;
    y.length = (int32_t)x.len;
    y.rc = x.reverse;
    y.prob = x.prob;
    y.ident = x.ident;
    acc += x.ident * (double)x.len;
    y.matches = acc;
    m.push_back(y);
  }
  return n;
}

// ---- chaining (analysis/MatchDynProg.cc)

namespace {

constexpr double kUnreached = 999999999999999.;  // PRETTY_INFINITE (:12)
constexpr double kRcPenalty = 20.;               // RC_PENALTY (:14)
constexpr double kMaxPenalty = 40;               // MAX_PENALTY (:15)
constexpr double kOtherQueryPenalty = 150;       // GetHighMaxPenalty (:33-36)

// MatchPenalty (:22-27)
double match_penalty(double prob) { return -200 * log(prob); }

// TransPenalty (:38-86): the cost of following a match at (t1, q1) with one at (t2, q2).
double trans_penalty(int t1, int q1, bool rc1, int t2, int q2, bool rc2) {
  if (rc1 != rc2) return kRcPenalty;
  double expect = t2 - t1;
  double observe = q2 - q1;
  double sigma = 0;
  if (expect > 1000000) sigma = sqrt(expect) / 10. - 100;
  if (rc1 && rc2) observe = -observe;
  double diff = observe - expect;
  if (diff < 0.) diff = -diff;
  const double flat = diff == 1 ? 0. : 0.01;
  if (expect < 300.) expect = 300;
  double p = diff / expect + sigma + flat;
  if (p > kMaxPenalty) p = kMaxPenalty;
  return p;
}

// GetRepeatScore (:88-106): weight of a match by how many matches cover its midpoint.
double repeat_score(int v) {
  if (v <= 1) return 10.;
  if (v == 2) return 5.;
  if (v == 3) return 2.5;
  if (v == 4) return 1.2;
  if (v >= 25) return 0.000001;
  if (v >= 10) return 0.001;
  return 0.5;
}

// MatchDynProg::Chain (:194-243) and SingleMatchDP (:109-163): matches by target start; each
// one offers every later match within the look-ahead its score plus its match penalty, the
// transition cost and the repeat weights of the matches it skips; the chain is traced back
// from the last match and returned sorted.
//
// The reference pushes each source's offers forward.  Here each destination pulls them, which
// gives the same scores and back pointers bit for bit:
//  * source i reaches destination j iff j - i <= 2000 and (j - i <= 11 or their target starts
//    differ by at most 250 kbp): its push loop breaks at the first j that fails this, and with
//    the starts sorted the test fails for every later j as well;
//  * the offer from i to j is total(i) + (trans(i, j) + skip(i, j)), skip(i, j) the repeat
//    weights of i + 1 .. j - 1 added to 0 in that order, as the push loop's running sum adds
//    them after each offer;
//  * a destination keeps the first smallest offer in source order (the push loop's strict
//    comparison), starting from "unreached";
//  * total(i) is final once every earlier source offered to i, i.e. when i is pulled.
// Most sources cannot win: an offer is at least total(i) + skip(i, j), the transition cost
// being non-negative (and FP addition monotone), so only sources whose bound does not exceed
// the best offer B of the nearest few sources can, and only those are evaluated, in source
// order.  With S the prefix sums of the repeat weights, that bound is C(i) + S(j - 1),
// C(i) = total(i) - S(i): a min-tree over C finds the sources with C(i) <= B - S(j - 1)
// (widened by a margin far above the rounding of these sums), and a source's running skip
// sum is brought up to j, by the same additions in the same order, only when it is found.
namespace {

// Min-tree over the sources' C values, reporting in ascending order every index of a range
// whose value is at most a threshold.
class MinTree {
 public:
  explicit MinTree(int n) : p_(1) {
    while (p_ < n) p_ <<= 1;
    mn_.assign((size_t)(2 * p_), HUGE_VAL);
  }
  void set(int i, double x) {
    size_t v = (size_t)(i + p_);
    mn_[v] = x;
    for (v >>= 1; v >= 1; v >>= 1) mn_[v] = std::min(mn_[2 * v], mn_[2 * v + 1]);
  }
  template <class F>
  void report(int lo, int hi, double t, F&& f) const {
    visit(1, 0, p_, lo, hi, t, f);
  }

 private:
  template <class F>
  void visit(size_t v, int a, int b, int lo, int hi, double t, F& f) const {
    if (b <= lo || a >= hi || !(mn_[v] <= t)) return;
    if (b - a == 1) {
      f(a);
      return;
    }
    const int mid = (a + b) >> 1;
    visit(2 * v, a, mid, lo, hi, t, f);
    visit(2 * v + 1, mid, b, lo, hi, t, f);
  }
  int p_;
  std::vector<double> mn_;
};

}  // namespace

// Appends v[path[k]] (path: descending indices of v, which is sorted by target start) to out,
// grouped stably by (query, strand), the groups in match_less order; true iff the appended
// run is strictly increasing under match_less (target ids included: a set that was not sorted
// can put other targets' matches into a target's range, and then this fails).
bool chain_in_order(const std::vector<Match>& v, const std::vector<int>& path, std::vector<Match>* out) {
  constexpr size_t kMaxGroups = 64;
  using Group = std::pair<int32_t, bool>;  // (query, strand)
  auto group_less = [](const Group& a, const Group& b) {
    return a.first != b.first ? a.first < b.first : !a.second && b.second;
  };
  std::vector<Group> groups;
  Group last{-1, false};
  for (int i : path) {
    const Group g{v[(size_t)i].query, v[(size_t)i].rc};
    if (!groups.empty() && g == last) continue;
    last = g;
    if (std::find(groups.begin(), groups.end(), g) == groups.end()) {
      if (groups.size() == kMaxGroups) return false;
      groups.push_back(g);
    }
  }
  std::sort(groups.begin(), groups.end(), group_less);
  // counting scatter by group, each group's members in ascending index order
  const size_t n = path.size();
  std::vector<uint8_t> gid(n);
  std::vector<size_t> slot(groups.size() + 1, 0);
  for (size_t k = 0; k < n; ++k) {
    const Match& x = v[(size_t)path[k]];
    const Group g{x.query, x.rc};
    gid[k] = (uint8_t)(std::lower_bound(groups.begin(), groups.end(), g, group_less) - groups.begin());
    ++slot[(size_t)gid[k] + 1];
  }
  for (size_t g = 1; g <= groups.size(); ++g) slot[g] += slot[g - 1];
  const size_t first = out->size();
  out->resize(first + n);
  Match* o = out->data() + first;
  for (size_t k = n; k-- > 0;) o[slot[gid[k]]++] = v[(size_t)path[k]];
  for (size_t k = 1; k < n; ++k)
    if (!match_less(o[k - 1], o[k])) return false;
  return true;
}

void chain_target(std::vector<Match>& v, std::vector<double>& rep_of, std::vector<Match>* out) {
  constexpr int kLookAheadDist = 250000;  // m_laDist (:173)
  constexpr int kLookAheadCount = 2000;   // la_limit (:200)
  constexpr int kNear = 4;                // sources evaluated for the bound
  const int n = (int)v.size();
  if (n == 0) return;
  // the reference sorts its match vector by target start with std::sort; std::sort of
  // (start, index) pairs by start makes the same comparisons, so equal starts come out in
  // the same order
  {
    std::vector<std::pair<int32_t, int32_t>> ord((size_t)n);
    for (int i = 0; i < n; ++i) ord[(size_t)i] = {v[(size_t)i].start_target, i};
    std::sort(ord.begin(), ord.end(),
              [](const std::pair<int32_t, int32_t>& a, const std::pair<int32_t, int32_t>& b) { return a.first < b.first; });
    std::vector<Match> sv((size_t)n);
    std::vector<double> sr((size_t)n);
    for (int i = 0; i < n; ++i) {
      sv[(size_t)i] = v[(size_t)ord[(size_t)i].second];
      sr[(size_t)i] = rep_of[(size_t)ord[(size_t)i].second];
    }
    v.swap(sv);
    rep_of.swap(sr);
  }
  std::vector<int> back((size_t)n, -1), next((size_t)n);
  std::vector<double> total((size_t)n), skip((size_t)n, 0.), pre((size_t)n);
  double acc = 0.;
  for (int i = 0; i < n; ++i) {
    acc += rep_of[(size_t)i];
    pre[(size_t)i] = acc;
    next[(size_t)i] = i + 1;  // skip(i) holds the weights of i + 1 .. next(i) - 1
  }
  MinTree tree(n);
  auto settle = [&](int i) {
    const double c = total[(size_t)i] - pre[(size_t)i];
    tree.set(i, c == c ? c : HUGE_VAL);  // a NaN total never wins
  };
  total[0] = 0. + match_penalty(v[0].prob);  // the first match starts at score 0
  settle(0);
  int lo = 0;
  for (int j = 1; j < n; ++j) {
    const Match& two = v[(size_t)j];
    while (two.start_target - v[(size_t)lo].start_target > kLookAheadDist) ++lo;
    int w0 = std::min(lo, j - 11);
    if (w0 < j - kLookAheadCount) w0 = j - kLookAheadCount;
    if (w0 < 0) w0 = 0;
    auto offer = [&](int i) {
      double& k = skip[(size_t)i];
      for (int& m = next[(size_t)i]; m < j; ++m) k += rep_of[(size_t)m];
      const Match& one = v[(size_t)i];
      double trans = kOtherQueryPenalty;
      if (one.query == two.query)
        trans = trans_penalty(one.start_target, one.start_query, one.rc, two.start_target, two.start_query, two.rc);
      return total[(size_t)i] + (trans + k);
    };
    double bound = kUnreached;  // no offer at or above "unreached" is taken
    for (int i = std::max(w0, j - kNear); i < j; ++i) {
      const double c = offer(i);
      if (c < bound) bound = c;
    }
    const double sj = pre[(size_t)j - 1];
    const double margin = 1e-7 * (fabs(bound) + 2. * fabs(sj)) + 1e-9;
    double best = kUnreached;
    int bi = -1;
    tree.report(w0, j, bound - sj + margin, [&](int i) {
      const double c = offer(i);
      if (c < best) {
        best = c;
        bi = i;
      }
    });
    back[(size_t)j] = bi;
    total[(size_t)j] = best + match_penalty(two.prob);
    settle(j);
  }
  // The chain, traced back from the last match, sorted with match_less.  Traced back it runs
  // through descending indices, i.e. by target start; grouped stably by (query, strand) in
  // ascending (query, strand) order, it is sorted with match_less unless two neighbours in a
  // group have equal target starts.  When the result is strictly increasing under match_less
  // it is the only sorted order, the one std::sort gives; otherwise std::sort of the chain in
  // trace order decides, as in the reference.
  std::vector<int> path;
  for (int i = n - 1; i >= 0; i = back[(size_t)i]) path.push_back(i);
  const size_t first = out->size();
  if (!chain_in_order(v, path, out)) {
    out->resize(first);
    for (int i : path) out->push_back(v[(size_t)i]);
    sort_matches(out->data() + first, out->size() - first);
  }
}

// Sorts v by key(x) (unsigned 32-bit): LSD radix in three passes of 11 bits, std::sort for
// short vectors.  Equal keys come out in any order (the callers count).
template <class T, class Key>
void radix_sort(std::vector<T>& v, Key key) {
  if (v.size() < 4096) {
    std::sort(v.begin(), v.end(), [&](const T& a, const T& b) { return key(a) < key(b); });
    return;
  }
  std::vector<T> tmp(v.size());
  std::vector<size_t> count(2049);
  for (int shift = 0; shift < 33; shift += 11) {
    std::fill(count.begin(), count.end(), 0);
    for (const T& x : v) ++count[((key(x) >> shift) & 2047u) + 1];
    for (int i = 0; i < 2048; ++i) count[(size_t)i + 1] += count[(size_t)i];
    for (const T& x : v) tmp[count[(key(x) >> shift) & 2047u]++] = x;
    v.swap(tmp);
  }
}

// Capped per-base coverage of one sequence by half-open intervals of bases.
struct Cover {
  std::vector<int> lo, hi;  // interval starts and ends, sorted separately by finish()
  int size = 0;
  void add(int a, int b) {
    if (a < 1) a = 1;
    if (b > size) b = size;
    if (a < b) {
      lo.push_back(a);
      hi.push_back(b);
    }
  }
  void finish() {
    radix_sort(lo, [](int x) { return (uint32_t)x; });  // >= 1
    radix_sort(hi, [](int x) { return (uint32_t)x; });
  }
  // For each (position p, slot k) of q: the intervals holding p (started at or before p, not
  // ended at or before p), at most 100, into cnt[k]; a position outside the sequence reads
  // as uncovered (the reference reads past its array).  One sweep over q sorted by position.
  void lookup(std::vector<std::pair<int, int>>& q, std::vector<int>& cnt) const {
    // by position, signed order
    radix_sort(q, [](const std::pair<int, int>& e) { return (uint32_t)e.first ^ 0x80000000u; });
    size_t a = 0, b = 0;
    for (const auto& e : q) {
      const int p = e.first;
      if (p < 0 || p >= size) {
        cnt[(size_t)e.second] = 0;
        continue;
      }
      while (a < lo.size() && lo[a] <= p) ++a;
      while (b < hi.size() && hi[b] <= p) ++b;
      const long c = (long)a - (long)b;
      cnt[(size_t)e.second] = c < 100 ? (int)c : 100;
    }
  }
};

}  // namespace

void chain(const MatchSet& in, MatchSet* out) {
  const int nt = (int)in.target_size.size(), nq = (int)in.query_size.size();
  out->target_size = in.target_size;
  out->query_size = in.query_size;
  out->m.clear();
  // per-base coverage counts, capped at 100, with a match overlapping its predecessor on
  // the same target and query counted only past it (:444-485).  Only the counts at match
  // midpoints are read, so each sequence keeps its covered intervals' sorted ends instead of
  // a count per base: the count at p is the number of intervals holding it, capped.
  std::vector<Cover> tcov((size_t)nt), qcov((size_t)nq);
  for (int i = 0; i < nt; ++i) tcov[(size_t)i].size = std::max(0, in.target_size[(size_t)i]);
  for (int i = 0; i < nq; ++i) qcov[(size_t)i].size = std::max(0, in.query_size[(size_t)i]);
  Match last;
  for (const Match& x : in.m) {
    if (x.query < 0) continue;
    int st = x.start_target, sq = x.start_query;
    if (x.target == last.target && x.query == last.query && x.start_target <= last.start_target + last.length &&
        x.start_query <= last.start_query + last.length && x.start_target > last.start_target &&
        x.start_query > last.start_query) {
      st = last.start_target + last.length;
      sq = last.start_query + last.length;
    }
    last = x;
    // bases 1 .. size - 1 are counted (the reference skips base 0)
    tcov[(size_t)x.target].add(st, x.start_target + x.length);
    qcov[(size_t)x.query].add(sq, x.start_query + x.length);
  }
  for (Cover& c : tcov) c.finish();
  for (Cover& c : qcov) c.finish();
  // each target's matches are taken from its first to its last index (:489-505)
  const int n = (int)in.m.size();
  std::vector<int> first((size_t)nt, n + 1), lastI((size_t)nt, -1);
  for (int i = 0; i < n; ++i) {
    const int id = in.m[(size_t)i].target;
    if (i < first[(size_t)id]) first[(size_t)id] = i;
    if (i > lastI[(size_t)id]) lastI[(size_t)id] = i;
  }
  // query coverage at each match's query midpoint
  std::vector<int> cq((size_t)n);
  {
    std::vector<std::vector<std::pair<int, int>>> qq((size_t)nq);
    for (int i = 0; i < n; ++i) {
      const Match& x = in.m[(size_t)i];
      qq[(size_t)x.query].emplace_back(x.start_query + x.length / 2, i);
    }
    for (int k = 0; k < nq; ++k) qcov[(size_t)k].lookup(qq[(size_t)k], cq);
  }
  // targets are chained independently: on up to 16 host threads, each taking the next
  // target, results concatenated in target order
  std::vector<std::vector<Match>> results((size_t)nt);
  std::atomic<int> next_target{0};
  // the first exception a worker meets (out of memory) stops the others and is rethrown
  // on the calling thread once every worker has joined
  std::exception_ptr failure;
  std::mutex failure_mu;
  auto chain_targets = [&]() {
    std::vector<std::pair<int, int>> tq;
    std::vector<int> ct;
    std::vector<Match> nodes;
    std::vector<double> reps;
    for (int j = next_target++; j < nt; j = next_target++) {
      int f = first[(size_t)j];
      const int l = lastI[(size_t)j];
      if (l == -1) f = 0;
      nodes.clear();
      reps.clear();
      // target j's coverage at the target midpoints of the matches in its index range
      tq.clear();
      for (int i = f; i <= l; ++i) {
        const Match& x = in.m[(size_t)i];
        tq.emplace_back(x.start_target + x.length / 2, i - f);
      }
      ct.assign(tq.size(), 0);
      tcov[(size_t)j].lookup(tq, ct);
      for (int i = f; i <= l; ++i) {
        const Match& x = in.m[(size_t)i];
        double r = repeat_score(ct[(size_t)(i - f)]);
        const double r2 = repeat_score(cq[(size_t)i]);
        if (r2 < r) r = r2;
        if (r > 0.0001) {
          nodes.push_back(x);
          reps.push_back(r);
        }
      }
      chain_target(nodes, reps, &results[(size_t)j]);
    }
  };
  auto worker = [&]() {
    try {
      chain_targets();
    } catch (...) {
      next_target = nt;
      std::lock_guard<std::mutex> lock(failure_mu);
      if (!failure) failure = std::current_exception();
    }
  };
  int nthreads = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
  if (nthreads > nt) nthreads = nt;
  std::vector<std::thread> pool;
  pool.reserve((size_t)std::max(0, nthreads - 1));
  for (int k = 1; k < nthreads; ++k) {
    try {
      pool.emplace_back(worker);
    } catch (const std::system_error&) {
      break;  // no more threads to be had: the ones started (and this one) take every target
    }
  }
  worker();
  for (std::thread& t : pool) t.join();
  if (failure) std::rethrow_exception(failure);
  for (const std::vector<Match>& r : results) out->m.insert(out->m.end(), r.begin(), r.end());
}

void chain_dups(const MatchSet& in, MatchSet* out) {
  MatchSet first;
  chain(in, &first);
  first.sort();
  // per target, the query with the most chained matches; every match of the target to
  // another query goes into a second chaining round (:263-341)
  MatchSet rest;
  rest.target_size = in.target_size;
  rest.query_size = in.query_size;
  const int nt = (int)in.target_size.size(), nq = (int)in.query_size.size();
  std::vector<int> hits((size_t)nq);
  for (int j = 0; j < nt; ++j) {
    std::fill(hits.begin(), hits.end(), 0);
    for (const Match& x : first.m)
      if (x.target == j && x.query >= 0) ++hits[(size_t)x.query];
    int best = -1, most = 0;
    for (int i = 0; i < nq; ++i)
      if (hits[(size_t)i] > most) {
        most = hits[(size_t)i];
        best = i;
      }
    for (const Match& x : in.m)
      if (x.target == j && x.query != best) rest.m.push_back(x);
  }
  chain(rest, out);
  out->m.insert(out->m.end(), first.m.begin(), first.m.end());
  out->sort();
}

// ---- interpolation (analysis/SatsumaSynteny2.cc:18-146)

void interpolate(MatchSet* chained) {
  constexpr int kMinRun = 40;       // m_min
  constexpr int kAdded = 2;         // m_num
  constexpr int kSpace = 100000;    // m_space
  std::vector<Match>& v = chained->m;
  const int n = (int)v.size();
  // mean angle of the steps between consecutive matches first..last (Slope, :30-49)
  auto slope = [&](int first, int last, double& sx, double& sy) {
    double eps = 0, div = 0;
    for (int i = first + 1; i <= last; ++i) {
      const double x = v[(size_t)i].start_target - v[(size_t)(i - 1)].start_target;
      const double y = v[(size_t)i].start_query - v[(size_t)(i - 1)].start_query;
      double s = y / x;
      s = atan(s);
      eps += s;
      div += 1.;
    }
    eps /= div;
    sy = sin(eps);
    sx = cos(eps);
  };
  int last = 0, count = 0;
  // a run that is still open when the list ends is not extended
  for (int i = 1; i < n; ++i) {
    const Match& a = v[(size_t)i];
    const Match& b = v[(size_t)(i - 1)];
    const int x1 = a.start_target, y1 = a.start_query, x2 = b.start_target, y2 = b.start_query;
    if (a.target == b.target && a.query == b.query && a.rc == b.rc && x1 - x2 < kSpace / 2 && y1 - y2 < kSpace / 2 &&
        y1 > y2) {
      if (x1 - x2 > b.length && y1 - y2 > b.length) ++count;
      continue;
    }
    if (count >= kMinRun) {
      double sx = 0., sy = 0;
      slope(last, i - 1, sx, sy);
      const Match head = v[(size_t)last];
      const Match tail = v[(size_t)(i - 1)];
      auto add = [&](const Match& at, int x, int y) {
        Match d;
        d.prob = 0.999;
        d.ident = 0.999;
        d.query = at.query;
        d.target = at.target;
        d.query_len = chained->query_size[(size_t)at.query];
        d.start_query = y;
        d.start_target = x;
        d.length = 50 + i - last;
        d.rc = at.rc;
        v.push_back(d);
      };
      for (int j = 1; j <= kAdded; ++j)
        add(head, head.start_target - (int)((double)kSpace * (double)j * sx),
            head.start_query - (int)((double)kSpace * (double)j * sy));
      for (int j = 1; j <= kAdded; ++j)
        add(tail, tail.start_target + (int)((double)kSpace * (double)j * sx),
            tail.start_query + (int)((double)kSpace * (double)j * sy));
    }
    last = i;
    count = 0;
  }
}

// ---- coverage (analysis/SatsumaSynteny2.cc:151-214)

void CoverageTracker::update(const MatchSet& s, uint64_t* added, uint64_t* removed) {
  constexpr int64_t kRecordStride = 10000000000L;  // POSITION_CHR_CNST (:16)
  std::vector<Seg> cur(s.m.size());
  for (size_t i = 0; i < s.m.size(); ++i) {
    cur[i].start = (uint64_t)(s.m[i].target * kRecordStride + s.m[i].start_target);
    cur[i].end = cur[i].start + (uint64_t)(int64_t)s.m[i].length;
  }
  std::sort(cur.begin(), cur.end(), [](const Seg& a, const Seg& b) { return a.start < b.start; });
  // a sweep over the boundaries of both lists; each step measures the span from the last
  // boundary to the next by which of the two current segments contain it
  const std::vector<Seg>& old = prev_;
  size_t i = 0, j = 0;
  uint64_t lastp = 0, next = 0;
  uint64_t cover[4] = {0, 0, 0, 0};
  while (i < old.size() || j < cur.size()) {
    if (i < old.size()) next = old[i].end;
    if (j < cur.size() && cur[j].end > next) next = cur[j].end;
    if (i < old.size()) {
      if (old[i].start < next && old[i].start > lastp)
        next = old[i].start;
      else if (old[i].end < next && old[i].end > lastp)
        next = old[i].end;
    }
    if (j < cur.size()) {
      if (cur[j].start < next && cur[j].start > lastp)
        next = cur[j].start;
      else if (cur[j].end < next && cur[j].end > lastp)
        next = cur[j].end;
    }
    int who = 0;
    if (i < old.size() && lastp >= old[i].start && next <= old[i].end) who += 1;
    if (j < cur.size() && lastp >= cur[j].start && next <= cur[j].end) who += 2;
    cover[who] += next - lastp;
    lastp = next;
    while (i < old.size() && lastp >= old[i].end) ++i;
    while (j < cur.size() && lastp >= cur[j].end) ++j;
  }
  prev_.swap(cur);
  *added = cover[2];
  *removed = cover[1];
}

// ---- grid search (analysis/GridSearch.cc)

Grid::Grid(int size, int blocks, const std::vector<Record>& target, const std::vector<Record>& query)
    : size_(size), blocks_(blocks) {
  // ChunkManager::ChunkIt (analysis/SeqChunk.cc:72-165): chunks every size - overlap bases
  // of records of at least 6 bases; a target chunk of N/X only is remembered as such.
  auto chunk = [](const std::vector<Record>& recs, int sz, int overlap, std::vector<int>* rec, std::vector<int>* st,
                  std::vector<bool>* empty) {
    const int step = sz - overlap;
    for (size_t i = 0; i < recs.size(); ++i) {
      const int l = (int)recs[i].len;
      if (l < 6) continue;
      for (int j = 0; j < 1 + l / step; ++j) {
        const int start = j * step;
        int end = (j + 1) * step + overlap;
        if (end >= l) end = l;
        rec->push_back((int)i);
        st->push_back(start);
        if (empty) {
          int ns = 0;
          for (int k = start; k < end; ++k) ns += (recs[i].seq[k] == 'N' || recs[i].seq[k] == 'X');
          empty->push_back(ns >= end - start);
        }
      }
    }
  };
  chunk(target, size, size / 4, &t_chunk_rec_, &t_chunk_start_, &all_n_);
  chunk(query, size, 0, &q_chunk_rec_, &q_chunk_start_, nullptr);
  // each record's first chunk is the first chunk carrying its name (GridSearch::SetUp,
  // :77-107): records sharing a name share their first chunk
  auto firsts = [](const std::vector<Record>& recs, const std::vector<int>& rec, std::vector<int>* out) {
    out->assign(recs.size(), -1);
    for (size_t i = 0; i < recs.size(); ++i)
      for (size_t c = 0; c < rec.size(); ++c)
        if (recs[(size_t)rec[c]].name == recs[i].name) {
          (*out)[i] = (int)c;
          break;
        }
  };
  firsts(target, t_chunk_rec_, &t_first_);
  firsts(query, q_chunk_rec_, &q_first_);
  // SearchMatrix::Setup (analysis/GridSearch.h:99-128)
  mx_ = target_chunks() / blocks + 1;
  my_ = query_chunks() / blocks + 1;
  state_.assign((size_t)mx_ * my_, UNKNOWN);
  count_.assign((size_t)mx_ * my_, 0);
  vertical_.assign((size_t)query_chunks(), 0);
  horizontal_.assign((size_t)target_chunks(), 0);
}

// ConsiderTargets (:123-186) with CoordsToBlocks (:410-458).  The reference's repeat
// tracker only records here; its answer is never used, so it is left out.
void Grid::consider(int tid, int st, int et, int qid, int sq, int eq, double ident) {
  // a record too short for chunks has no block to weight (the reference indexes before
  // its arrays there)
  if (t_first_[(size_t)tid] < 0 || q_first_[(size_t)qid] < 0) return;
  const int tmid = (st + et) / 2, qmid = (sq + eq) / 2;
  const int qblock = q_first_[(size_t)qid] + qmid / size_;
  const int tstep = size_ - size_ / 4;  // target chunks overlap by a quarter
  const int off = tmid / tstep;
  int tblock = t_first_[(size_t)tid] + off;
  if (off > 0 && tmid - tstep * off < size_ / 4) --tblock;  // the lower of two overlapping chunks
  const int x = tblock / blocks_, y = qblock / blocks_;
  vertical_[(size_t)qblock] += et - st;
  horizontal_[(size_t)tblock] += et - st;
  const int plus = state_[(size_t)index(x, y)] != UNKNOWN ? 1 : 0;
  for (int i = x - plus; i <= x + plus; ++i)
    for (int j = y - plus; j <= y + plus; ++j) {
      if (i < 0 || j < 0 || i >= mx_ || j >= my_) continue;
      const int w = (int)(100. * ident * (double)(et - st));
      int& c = count_[(size_t)index(i, j)];
      if (w > c) c = w;
    }
}

void Grid::update(const MatchSet& chained) {
  // ClearTargetWeights (GridSearch.h:376-385): cell weights and query occupancy cleared,
  // target occupancy kept as a mark
  std::fill(count_.begin(), count_.end(), 0);
  std::fill(vertical_.begin(), vertical_.end(), 0);
  for (int& h : horizontal_)
    if (h > 0) h = 1;
  for (const Match& m : chained.m) {
    const int x1 = m.start_target, x2 = m.start_target + m.length;
    int y1 = m.start_query, y2 = m.start_query + m.length;
    if (m.rc) {
      const int s = chained.query_size[(size_t)m.query];
      const int t = y1;
      y1 = s - y2;
      y2 = s - t;
    }
    if (x1 < 0 || x2 < 0 || y1 < 0 || y2 < 0) continue;
    if (x2 >= chained.target_size[(size_t)m.target] || y1 >= chained.query_size[(size_t)m.query] ||
        y2 >= chained.query_size[(size_t)m.query])
      continue;
    consider(m.target, x1, x2, m.query, y1, y2, m.ident);
  }
}

// GridTarget ordering (analysis/GridSearch.h:258-273): heavier first, then by cell.
static bool target_less(const GridTarget& a, const GridTarget& b) {
  if (a.count != b.count) return -a.count < -b.count;
  if (a.x != b.x) return a.x < b.x;
  return a.y < b.y;
}

int Grid::collect(std::vector<GridTarget>* out, int n) {
  for (int i = 0; i < mx_; ++i)
    for (int j = 0; j < my_; ++j) {
      const size_t k = (size_t)index(i, j);
      if (state_[k] != UNKNOWN || count_[k] == 0) continue;
      const int tl = std::min((i + 1) * blocks_ - 1, target_chunks() - 1);
      const int ql = std::min((j + 1) * blocks_ - 1, query_chunks() - 1);
      out->push_back(GridTarget{i * blocks_, tl, j * blocks_, ql, i, j, -1, count_[k], false});
    }
  std::sort(out->begin(), out->end(), target_less);
  const int candidates = (int)out->size();
  if (n < (int)out->size()) out->resize((size_t)n);
  for (const GridTarget& t : *out) state_[(size_t)index(t.x, t.y)] = SUBMITTED;
  if ((int)out->size() < n) seeds(out, n - (int)out->size());
  return candidates;
}

// CollectSeeds (:189-315): lines through untouched target columns.  Every run of >= 5
// unused target chunks (none of N only) gives its middle chunk; down the query axis, every
// stretch that starts on an unused query chunk and runs until 16 used chunks in a row
// (more than 2 chunks long) becomes one line, weighted by the run's length.  The heaviest
// lines are taken, the cut extended to keep every line of the last target chunk taken.
void Grid::seeds(std::vector<GridTarget>* out, int n) {
  if (n == 0) return;
  constexpr int kSkipMax = 16;
  std::vector<GridTarget> lines;
  const int nh = (int)horizontal_.size(), nv = (int)vertical_.size();
  for (int i = 0; i < nh; ++i) {
    if (horizontal_[(size_t)i] > 0 || all_n_[(size_t)i]) continue;
    int j = i;
    while (j < nh && horizontal_[(size_t)j] <= 0 && !all_n_[(size_t)j]) ++j;
    const int len = j - i - 1;
    if (len < 4) continue;
    const int x = i + len / 2;
    i = j;
    for (int y = 0; y < nv; ++y) {
      if (vertical_[(size_t)y] > 0) continue;
      int skip = 0, k = y;
      for (; k < nv; ++k) {
        if (vertical_[(size_t)k] > 0) {
          if (++skip >= kSkipMax) break;
        } else {
          skip = 0;
        }
      }
      const int y1 = y, y2 = k - 1;
      if (y2 - y1 < 2) continue;
      lines.push_back(GridTarget{x, x, y1, y2, x, y1, y2, len, true});
      y = k;
    }
  }
  std::sort(lines.begin(), lines.end(), target_less);
  if (n < (int)lines.size()) {
    const int lastx = lines[(size_t)(n - 1)].target_first;
    size_t i = (size_t)n;
    while (i < lines.size() && lines[i].target_first == lastx) ++i;
    lines.resize(i);
  }
  for (const GridTarget& t : lines) {
    for (int j = t.target_first; j <= t.target_last; ++j) horizontal_[(size_t)j] = 1;
    out->push_back(t);
  }
}

}  // namespace sxc_master
