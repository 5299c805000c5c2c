/* sxc_master.h -- C ABI of the master side of Satsuma2's search (libsxc_hip.so).
 *
 * The chunk-pair path (sxc.h) aligns the blocks of chunk pairs that Satsuma2's master asks
 * for.  The master (analysis/SatsumaSynteny2.cc:484-599) turns their matches into the next
 * requests: it sorts and collapses the match set (analysis/SequenceMatch.cc:418-469),
 * chains it per target (RunMatchDynProg, analysis/MatchDynProg.cc:245-561), extends long
 * chains (SyntenyInterpolator, analysis/SatsumaSynteny2.cc:18-146), weights the cells of
 * its search grid by the chained matches and collects the heaviest unsearched cells and,
 * when too few remain, seed lines (GridSearch, analysis/GridSearch.cc:123-407).  These
 * entry points restate those steps with the reference's results, so a caller can run the
 * whole search around sxc_server_submit (sxc.h) in one process.
 *
 * All functions are host code and thread-compatible (one object per thread at a time).
 * Status codes and sxc_last_error() are those of sxc.h.
 */
#ifndef SXC_MASTER_H_
#define SXC_MASTER_H_

#include <stdint.h>

#include "sxc.h"

#ifdef __cplusplus
extern "C" {
#endif

/* SingleMatch (analysis/SequenceMatch.h:16-171).  rc: query reverse-complemented;
 * matches: SingleMatch's accumulated m_matches. */
typedef struct sxc_match {
  int32_t target_id, query_id, query_len;
  int32_t start_target, start_query, length;
  int32_t rc;
  int32_t reserved;
  double matches, prob, ident;
} sxc_match;

/* MultiMatches (analysis/SequenceMatch.h:176-257): the record sizes and a match list. */
typedef struct sxc_matchset sxc_matchset;

/* A set with n_targets / n_queries records of the given sizes and no matches. */
sxc_matchset* sxc_matchset_create(int32_t n_targets, const int32_t* target_sizes, int32_t n_queries,
                                  const int32_t* query_sizes);
void sxc_matchset_destroy(sxc_matchset* s);
/* Appends n matches (MultiMatches::AddMatch, SequenceMatch.h:198-205). */
int sxc_matchset_add(sxc_matchset* s, const sxc_match* m, int64_t n);
/* Appends n slave records (WorkQueue::collect_new_matches, analysis/WorkQueue.cc:315-353).
 * from_file != 0 first mirrors the query start of reverse-strand records as seed files
 * are read (WorkQueue::results_from_file, :271-288). */
int sxc_matchset_collect(sxc_matchset* s, const sxc_result* r, int64_t n, int from_file);
int64_t sxc_matchset_count(const sxc_matchset* s);
/* Copies up to cap matches to out; returns the set's count (or < 0 on error). */
int64_t sxc_matchset_get(const sxc_matchset* s, sxc_match* out, int64_t cap);
/* MultiMatches::Sort, Collapse, LengthFilter (SequenceMatch.h:211-215, .cc:418-492). */
int sxc_matchset_sort(sxc_matchset* s);
int sxc_matchset_collapse(sxc_matchset* s);
int sxc_matchset_length_filter(sxc_matchset* s, int64_t min_length);
/* out := in (MultiMatches assignment). */
int sxc_matchset_copy(const sxc_matchset* in, sxc_matchset* out);
/* out := RunMatchDynProg(in), or RunMatchDynProgMult(in) with dups != 0
 * (analysis/MatchDynProg.cc:401-561, 245-398).  out's previous contents are replaced. */
int sxc_matchset_chain(const sxc_matchset* in, int dups, sxc_matchset* out);
/* SyntenyInterpolator::Interpolate (analysis/SatsumaSynteny2.cc:62-146), in place. */
int sxc_matchset_interpolate(sxc_matchset* s);

/* TargetCoverageTracker (analysis/SatsumaSynteny2.cc:151-214). */
typedef struct sxc_coverage sxc_coverage;
sxc_coverage* sxc_coverage_create(void);
void sxc_coverage_destroy(sxc_coverage* c);
/* update_and_report: target bases covered by s and not by the previous set (added) and
 * the reverse (removed); s becomes the previous set. */
int sxc_coverage_update(sxc_coverage* c, const sxc_matchset* s, uint64_t* added, uint64_t* removed);

/* GridSearch (analysis/GridSearch.h:349-424). */
typedef struct sxc_grid sxc_grid;

/* GridTarget (analysis/GridSearch.h:193-300): target chunks [target_first, target_last] x
 * query chunks [query_first, query_last]; cell (x, y); y2 the last query chunk of a seed
 * line (-1 otherwise); count its weight; fast: a seed line, aligned with the fast cutoff. */
typedef struct sxc_grid_target {
  int32_t target_first, target_last, query_first, query_last;
  int32_t x, y, y2, count, fast;
} sxc_grid_target;

/* GridSearch(chunk, blocks) + SetUp(target, query) (GridSearch.cc:15-121) over in-memory
 * records: names without '>' (records sharing a name share their first chunk, as in the
 * reference), bases and lengths. */
sxc_grid* sxc_grid_create(int32_t chunk, int32_t blocks, int32_t n_targets, const char* const* target_names,
                          const char* const* target_seqs, const int64_t* target_lens, int32_t n_queries,
                          const char* const* query_names, const char* const* query_seqs, const int64_t* query_lens);
/* The same with the records of FASTA files (comma-separated lists), read as the master
 * reads them (vecDNAVector::Read, analysis/DNAVector.cc:1176-1293). */
sxc_grid* sxc_grid_create_fasta(int32_t chunk, int32_t blocks, const char* target_fasta, const char* query_fasta);
void sxc_grid_destroy(sxc_grid* g);
/* Chunk counts of the grid and each chunk's record and start (m_targetChunks /
 * m_queryChunks): the request's chunk indices refer to these. */
int sxc_grid_chunks(const sxc_grid* g, int32_t* n_target_chunks, int32_t* n_query_chunks);
int sxc_grid_chunk(const sxc_grid* g, int query, int32_t index, int32_t* record, int32_t* start);
/* UpdateTargetWeights (GridSearch.cc:318-352) with a chained set. */
int sxc_grid_update(sxc_grid* g, const sxc_matchset* chained);
/* CollectTargets (GridSearch.cc:355-407): collects targets (at most n, more only to keep
 * a seed column whole), marks their cells searched, writes up to cap of them to out, sets
 * *count to how many were collected and *candidates to the weighted cells before the cut.
 * SXC_ERR_CAPACITY if *count > cap: the collection is kept and sxc_grid_last returns it. */
int sxc_grid_collect(sxc_grid* g, int32_t n, sxc_grid_target* out, int32_t cap, int32_t* count,
                     int32_t* candidates);
/* The targets of the last sxc_grid_collect (SXC_ERR_CAPACITY if more than cap). */
int sxc_grid_last(const sxc_grid* g, sxc_grid_target* out, int32_t cap, int32_t* count);

#ifdef __cplusplus
}
#endif

#endif /* SXC_MASTER_H_ */
