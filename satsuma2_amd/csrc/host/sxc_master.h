// sxc_master.h -- the master's side of Satsuma2's search, around the chunk-pair path:
// match sets (MultiMatches), the chaining dynamic programme, the synteny interpolation,
// the coverage tracker and the grid search that turns chained matches into the next chunk
// pairs to align.  Host C++, deterministic, with the reference's results.
//
// Sorting uses std::sort with the reference's comparators, so ties come out in the order
// the same libstdc++ gives the reference.  Quirks of the reference that decide outputs are
// kept and named where they occur.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace sxc_master {

// SingleMatch (analysis/SequenceMatch.h:16-171).  `matches` is SingleMatch's accumulated
// m_matches (AddMatches).
struct Match {
  int32_t target = -1, query = -1, query_len = 0;
  int32_t start_target = 0, start_query = 0, length = 0;
  bool rc = false;
  double matches = 0., prob = 0., ident = 0.;
};

// SingleMatch::operator< (analysis/SequenceMatch.h:98-113): target, query, forward before
// reverse, start on the target, then length.
bool match_less(const Match& a, const Match& b);

// t_result (analysis/WorkQueue.h:25-35), the slaves' record.
struct Result {
  uint64_t query_id, target_id, query_size, qstart, tstart, len;
  bool reverse;
  double prob, ident;
};
static_assert(sizeof(Result) == 72, "t_result layout");

// MultiMatches (analysis/SequenceMatch.h:176-257): sequence sizes and a list of matches.
struct MatchSet {
  std::vector<int32_t> target_size, query_size;
  std::vector<Match> m;

  // MultiMatches::Sort (SequenceMatch.h:211-215)
  void sort();
  // MultiMatches::Collapse (SequenceMatch.cc:418-469)
  void collapse();
  // MultiMatches::LengthFilter (SequenceMatch.cc:471-492)
  void length_filter(uint64_t min_length);
  // WorkQueue::collect_new_matches (analysis/WorkQueue.cc:315-353) after results_from_file
  // (:271-288, from_file: reverse-strand query starts mirrored) or receive_solutions
  // (:126-165, as delivered).  Returns the number added.
  int64_t collect(const Result* r, int64_t n, bool from_file);
};

// RunMatchDynProg (analysis/MatchDynProg.cc:401-561): one chain per target through its
// matches.  `in` must hold each target's matches contiguously (a sorted set), as the
// reference assumes.
void chain(const MatchSet& in, MatchSet* out);
// RunMatchDynProgMult (analysis/MatchDynProg.cc:245-398): the chain, plus a second chain
// through the matches of every target to queries other than its best one.
void chain_dups(const MatchSet& in, MatchSet* out);
// SyntenyInterpolator::Interpolate (analysis/SatsumaSynteny2.cc:62-146): extends runs of
// >= 40 collinear chained matches by two guessed matches on each side.
void interpolate(MatchSet* chained);

// TargetCoverageTracker (analysis/SatsumaSynteny2.cc:151-214): target bases covered by the
// new set and not the previous one, and the reverse.
class CoverageTracker {
 public:
  void update(const MatchSet& s, uint64_t* added, uint64_t* removed);

 private:
  struct Seg {
    uint64_t start, end;
  };
  std::vector<Seg> prev_;
};

// GridTarget (analysis/GridSearch.h:193-300): a block of target x query chunks to align.
struct GridTarget {
  int32_t target_first, target_last, query_first, query_last, x, y, y2, count;
  bool fast;
};

// One FASTA record as the master holds it: name (without '>') and bases.
struct Record {
  std::string name;
  const char* seq;
  int64_t len;
};

// GridSearch (analysis/GridSearch.h:349-424, analysis/GridSearch.cc): target chunks of
// `size` every 3/4 size and query chunks of `size`, grouped `blocks` x `blocks` into the
// search matrix.
class Grid {
 public:
  Grid(int size, int blocks, const std::vector<Record>& target, const std::vector<Record>& query);
  // GridSearch::UpdateTargetWeights (:318-352)
  void update(const MatchSet& chained);
  // GridSearch::CollectTargets (:355-407): appends to `out` and returns the number of
  // candidate cells before the cut to n.
  int collect(std::vector<GridTarget>* out, int n);
  int target_records() const { return (int)t_first_.size(); }
  int query_records() const { return (int)q_first_.size(); }
  int target_chunks() const { return (int)t_chunk_rec_.size(); }
  int query_chunks() const { return (int)q_chunk_rec_.size(); }
  // chunk i's record and start (GridSearch::m_targetChunks / m_queryChunks)
  int target_chunk_record(int i) const { return t_chunk_rec_[(size_t)i]; }
  int target_chunk_start(int i) const { return t_chunk_start_[(size_t)i]; }
  int query_chunk_record(int i) const { return q_chunk_rec_[(size_t)i]; }
  int query_chunk_start(int i) const { return q_chunk_start_[(size_t)i]; }

 private:
  enum State : uint8_t { UNKNOWN, SUBMITTED };
  void consider(int tid, int st, int et, int qid, int sq, int eq, double ident);
  void seeds(std::vector<GridTarget>* out, int n);
  int index(int x, int y) const { return y * mx_ + x; }

  int size_, blocks_;
  std::vector<int> t_chunk_rec_, t_chunk_start_, q_chunk_rec_, q_chunk_start_;
  std::vector<bool> all_n_;                      // target chunks of N only
  std::vector<int> t_first_, q_first_;           // first chunk of each record (-1: none)
  int mx_ = 0, my_ = 0;                          // matrix cells along target and query
  std::vector<State> state_;
  std::vector<int> count_;
  std::vector<int> vertical_, horizontal_;       // per query / target chunk
};

}  // namespace sxc_master
