/*
 * sxc_oracle.h -- CPU restatement of Satsuma2's chunk-pair cross-correlation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU path is compared
 * against (tests/, __graft_entry__.smoke(), and bench.py's cpu_baseline leg).  The
 * product (libsxc_hip.so) never links, loads or calls it.
 *
 * Every function restates one reference function with the same IEEE operation
 * order (compiled with -ffp-contract=off, no FMA, x86-64 SSE arithmetic), so its
 * outputs are bit-identical to the reference on the same inputs.  Each definition in
 * sxc_oracle.c cites the reference file:line it follows.
 */
#ifndef SXC_ORACLE_H
#define SXC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t start_target;
  int32_t start_query;
  int32_t len;
} sxo_seqmatch;

/* One kept match of a chunk pair, the fields of t_result (analysis/WorkQueue.h:23-33). */
typedef struct {
  uint64_t query_id;
  uint64_t target_id;
  uint64_t query_size;
  uint64_t qstart;
  uint64_t tstart;
  uint64_t len;
  uint8_t reverse;
  uint8_t pad_[7];
  double prob;
  double ident;
} sxo_result;

/* Per-chunk-pair job: chunk bytes plus the chunk metadata FilterMatches uses. */
typedef struct {
  const char* tseq;
  int32_t tlen;
  const char* qseq;
  int32_t qlen;
  int32_t t_chunk_start; /* SeqChunk::GetStart of the target chunk */
  int32_t q_chunk_start; /* SeqChunk::GetStart of the query chunk */
  int32_t t_record;      /* target record id */
  int32_t q_record;      /* query record id */
  int32_t q_record_len;  /* ChunkManager::GetSize(query record) */
  int32_t fast;          /* t_pair.fast: cutoff_fast instead of cutoff */
} sxo_pair;

typedef struct {
  int32_t N;            /* FFT length (2 * t_chunk in the slave) */
  int32_t q_chunk;      /* -q_chunk, used by RCQuery; 0 = the query chunk's own length */
  int32_t min_len;      /* -l */
  int32_t strands;      /* 1 = forward only, 2 = forward then reverse complement */
  double cutoff;        /* -cutoff */
  double cutoff_fast;   /* -cutoff_fast */
  double min_prob;      /* effective keep threshold (0.99 on the slave path) */
  double target_total;  /* sum of target record lengths */
  const double* ptable; /* -prob_table: sxo_prob_table output, or NULL for GetMatchProbabilityEx */
  double table_prob;    /* the probability a table hit reports (ProbTable's cutoff, -min_prob) */
} sxo_params;

#define SXO_PT_ROWS 512
#define SXO_PT_LEN 2048
/* ProbTable::Setup (analysis/ProbTable.cc:15-52): t[i * 2048 + j] for i = 1..511, j = 1..2047
 * is the identity found by bisection (width 1e-8) where GetMatchProbabilityRaw(j, ident,
 * i / 511, target_size) stops being 0 (the reference tests the probability for non-zero,
 * not against the cutoff); t[i * 2048] = 2; row 0 is never set (zeros here).  Rows are
 * computed on `threads` threads. */
void sxo_prob_table(double target_size, int32_t threads, double* t);
/* ProbTable::GetMatchProbability (analysis/ProbTable.cc:105-141): `cutoff` if ident reaches
 * the table's threshold, else 0; *ident = matches / len.  A GC expectation below 1/511 maps
 * to row 0, which the reference never fills (it indexes an empty vector there); it is
 * reported as 0 here. */
double sxo_prob_table_lookup(const double* t, const char* target, const char* query, int32_t st, int32_t sq,
                             int32_t len, double cutoff, double* ident);

/* DNACodec tables (analysis/DNAVector.cc:13-57). */
void sxo_codec(double* a256, double* c256, double* g256, double* t256, unsigned char* rc256);
double sxo_dna_equal(int a, int b);
double sxo_dna_equal_amb(int a, int b);
void sxo_reverse_complement(const char* in, int32_t len, char* out);

/* CCSignal::SetSequence(seq, N) (analysis/CrossCorr.cc:35-134,179-206): out = 4*N floats (A,C,G,T). */
int sxo_encode(const char* seq, int32_t len, int32_t N, float* out);

/* FFTReal<float> (extern/RealFFT/FFTReal.hpp) on one vector of length N (power of two >= 8). */
int sxo_fft_forward(int32_t N, const float* x, float* f);
int sxo_fft_inverse(int32_t N, const float* f, float* x);
void sxo_fft_rescale(int32_t N, float* x);

/* CrossCorrelation::CrossCorrelate(out, target, query) (analysis/CrossCorr.cc:386-507). */
int sxo_xcorr(const float* tsig4, const float* qsig4, int32_t N, float* xc);

/* SeqAnalyzer::FindTop (analysis/CrossCorr.cc:878-944): accepted lags ascending; returns count. */
int32_t sxo_findtop(const float* xc, int32_t N, double cutoff, int32_t* top, int32_t cap);

/* SeqAnalyzer::MatchUp (analysis/CrossCorr.cc:583-605, DoOne :667-724): appends runs; returns count
 * (may exceed cap; only the first cap are written). */
int32_t sxo_matchup(const char* q, int32_t lq, const char* t, int32_t lt, const float* xc, int32_t N,
                    double cutoff, sxo_seqmatch* out, int32_t cap);

/* GetMatchProbabilityEx (analysis/AlignProbability.cc:62-127). */
double sxo_match_probability(const char* t, const char* q, int32_t st, int32_t sq, int32_t len,
                             double target_size);
/* PrintMatch(query, target, m, silent) return value (analysis/CrossCorr.cc:967-1037). */
double sxo_print_match_ident(const char* q, const char* t, int32_t st, int32_t sq, int32_t len);

/* HomologyByXCorr::Align + FilterMatches for one chunk pair (analysis/HomologyByXCorrSlave.cc:
 * 168-253).  Returns number of results (may exceed cap). */
int32_t sxo_align_pair(const sxo_pair* p, const sxo_params* prm, sxo_result* out, int32_t cap);

/* Signal tail of a reused CCSignal (tools/analysis/HomologyByXCorr.cc:706-716, 780-798):
 * CCSignal::SetSequence writes only [0, L) (analysis/CrossCorr.cc:118-131, 182-206), so the
 * positions from L on keep an earlier, longer sequence's signal.  Segment s covers
 * [L or the previous end, end) with the signal of seq (reverse complement if strand = 1). */
typedef struct {
  int32_t end;
  int32_t len;
  const char* seq;
  int32_t strand;
  int32_t pad_;
} sxo_tail_seg;

typedef struct {
  const sxo_tail_seg* segs;
  int32_t nseg;
  int32_t pad_;
} sxo_tail;

/* Overwrites sig (4*N floats, channels A, C, G, T) at [L, last end) from the tail. */
int sxo_apply_tail(float* sig, int32_t L, int32_t N, const sxo_tail* t);

/* Many pairs on `threads` POSIX threads; results of pair k go to out + offsets[k],
 * counts[k] receives the number produced (clipped to cap_per_pair). Returns total produced. */
int64_t sxo_align_batch(const sxo_pair* pairs, int32_t n, const sxo_params* prm, int32_t threads,
                        sxo_result* out, int32_t cap_per_pair, int32_t* counts);

/* sxo_align_batch with signal tails: t_tails[k] / q_tails[k] (either may be NULL) extend
 * pair k's target signal and its query signal on both strands. */
int64_t sxo_align_batch_tails(const sxo_pair* pairs, int32_t n, const sxo_params* prm, const sxo_tail* t_tails,
                              const sxo_tail* q_tails, int32_t threads, sxo_result* out, int32_t cap_per_pair,
                              int32_t* counts);

#ifdef __cplusplus
}
#endif
#endif
