/* sxc_kmatch.h -- C ABI of the GPU KMatch seeding (libsxc_hip.so).
 *
 * KMatch (kmatch/KMatch.cc:19-318) finds the k-mers that occur at most max_freq times in
 * each of two genomes (canonical: the smaller of a k-mer and its reverse complement), joins
 * the query's with the target's and chains joined positions that continue along one
 * diagonal, with gaps of up to max_jump positions, into blocks of at least min_length
 * bases.  SatsumaSynteny2 runs it for K = 11, 13, ..., 31 (analysis/SatsumaSynteny2.cc:
 * 417-463) and reads its t_result files as the search's first matches.  Here the k-mer
 * tables are built, sorted, filtered and joined on the GPU; the chaining sweep, whose
 * output order follows the sorted positions, runs on the host.  Results equal the
 * reference's (t_result records, same order; the 7 padding bytes after `reverse`, which
 * the reference leaves uninitialized, are zero) for max_freq = 1, SatsumaSynteny2's
 * default.  With max_freq > 1 the records are the same and come in the same order of query
 * starts, but blocks sharing a query start can come in another order: there a query
 * position joins several target positions, whose order the reference leaves to std::sort
 * of equal keys; here they keep target-table order.
 *
 * Status codes and sxc_last_error() are those of sxc.h.
 */
#ifndef SXC_KMATCH_H_
#define SXC_KMATCH_H_

#include <stdint.h>

#include "sxc.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t K;           /* odd, 3 .. 31 */
  int32_t min_length;  /* blocks of at least this many bases are reported */
  int32_t max_jump;    /* largest gap (in positions) a block may bridge */
  int32_t max_freq;    /* k-mers occurring more often in a genome are dropped */
} sxc_kmatch_params;

typedef struct {
  int64_t query_kmers, target_kmers;  /* table sizes (KMatch's kposv, with its unfilled slots) */
  int64_t query_kept, target_kept;    /* after the frequency filter */
  int64_t joined;                     /* joined positions (kmatches) */
  int64_t blocks;                     /* reported t_result records */
  double read_ms, device_ms, sweep_ms;
} sxc_kmatch_stats;

typedef struct sxc_kmatch sxc_kmatch;

/* KMatch on two FASTA files, read as kmatch/KMatch.cc:21-104 reads them (std::getline
 * lines; a last line without a newline is not read; other letters than ACGTacgt break
 * k-mers).  Records shorter than K - 1 bases are an error (the reference's table size
 * underflows there). */
int sxc_kmatch_files(sxc_ctx* ctx, const char* query_fasta, const char* target_fasta, const sxc_kmatch_params* p,
                     sxc_kmatch** out);
/* The same on in-memory records (record i has sequence index i + 1). */
int sxc_kmatch_records(sxc_ctx* ctx, int32_t n_queries, const char* const* query_seqs, const int64_t* query_lens,
                       int32_t n_targets, const char* const* target_seqs, const int64_t* target_lens,
                       const sxc_kmatch_params* p, sxc_kmatch** out);
/* The t_result records (dump_matching_blocks, KMatch.cc:196-318), in the reference's order:
 * up to cap of them; *n receives the count (SXC_ERR_CAPACITY if it exceeds cap). */
int sxc_kmatch_results(const sxc_kmatch* k, sxc_result* out, int64_t cap, int64_t* n);
int sxc_kmatch_get_stats(const sxc_kmatch* k, sxc_kmatch_stats* st);
void sxc_kmatch_destroy(sxc_kmatch* k);

#ifdef __cplusplus
}
#endif

#endif /* SXC_KMATCH_H_ */
