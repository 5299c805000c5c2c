/*
 * sxc.h -- C ABI of the MI355X chunk-pair cross-correlation library (libsxc_hip.so).
 *
 * This is the drop-in boundary for Satsuma2's cross-correlation path: each entry point
 * replaces one reference interface (cited below) and computes the same results on the
 * GPU.  Plain pointers and sizes only; all host buffers are caller-owned; device memory
 * is owned by the context.  Every function returns SXC_OK (0) or a negative status;
 * sxc_last_error() describes the last failure of the calling thread.
 *
 * Sequences are byte strings of uppercase ASCII letters (< 128), as the reference holds
 * them after vecDNAVector::Read (analysis/DNAVector.cc:1176-1293, ToUpper at :586).
 * FFT lengths N are powers of two in [512, 32768] (2 * t_chunk in the slave,
 * analysis/HomologyByXCorrSlave.cc:259); sequence lengths must not exceed N.
 */
#ifndef SXC_H
#define SXC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SXC_OK = 0,
  SXC_ERR_ARG = -1,       /* bad argument (size, N, byte >= 128, null pointer) */
  SXC_ERR_HIP = -2,       /* a HIP runtime call failed */
  SXC_ERR_CAPACITY = -3,  /* output capacity too small; *n_out holds the required count */
  SXC_ERR_NOMEM = -4,
  SXC_PENDING = 1         /* a server request is not finished yet */
};

typedef struct sxc_ctx sxc_ctx;

/* One diagonal run found by the scan: SeqMatch (analysis/CrossCorr.h:165-196). */
typedef struct {
  int32_t start_target;
  int32_t start_query;
  int32_t len;
} sxc_seqmatch;

/* t_result (analysis/WorkQueue.h:23-33), byte-compatible on LP64 (72 bytes). */
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
} sxc_result;

/* Slave parameters (analysis/HomologyByXCorrSlave.cc:334-384 flags). */
typedef struct {
  int32_t N;            /* FFT length = 2 * t_chunk; 0 = per pair, the next power of two >= tlen + qlen
                           (at least 512), pairs of each length run as one class */
  int32_t q_chunk;      /* RCQuery chunk size: -q_chunk on the slave path; 0 = each query chunk's own
                           length, as the standalone aligner computes it
                           (tools/analysis/HomologyByXCorr.cc:226-228, 793-798) */
  int32_t min_len;      /* -l */
  int32_t strands;      /* 2 = forward then reverse complement (Align), 1 = forward only */
  double cutoff;        /* -cutoff */
  double cutoff_fast;   /* -cutoff_fast */
  double min_prob;      /* keep threshold: 0.99 on the slave path (m_minProb, :76) */
  double target_total;  /* sum of target record lengths (:405-408) */
  int32_t prob_table;   /* -prob_table: keep by ProbTable lookup (analysis/ProbTable.cc:105-141)
                           instead of GetMatchProbability (:184-187) */
  int32_t pad_;
  double table_prob;    /* the probability a table hit reports: ProbTable's cutoff, the slave's
                           -min_prob (:413-421); a hit is kept if it is >= min_prob */
} sxc_params;

/* One chunk pair: chunk bytes and the chunk metadata FilterMatches uses (:168-219). */
typedef struct {
  const char* tseq;
  int32_t tlen;
  const char* qseq;
  int32_t qlen;
  int32_t t_chunk_start; /* SeqChunk::GetStart of the target chunk */
  int32_t q_chunk_start; /* SeqChunk::GetStart of the query chunk */
  int32_t t_record;      /* t_result.target_id */
  int32_t q_record;      /* t_result.query_id */
  int32_t q_record_len;  /* t_result.query_size; RCQuery's qLen */
  int32_t fast;          /* t_pair.fast: use cutoff_fast */
} sxc_pair;

/* Context on HIP device `device`; builds FFT twiddle, codec and entropy tables. */
int sxc_create(sxc_ctx** out, int device);
void sxc_destroy(sxc_ctx* ctx);
const char* sxc_last_error(void);
const char* sxc_version(void);

/* Host-side codec helpers (no device work; the tables the kernels use).
 * DNAVector::ReverseComplement on seq[0, len) in place (analysis/DNAVector.cc:482-521, the
 * DNACodec map of :13-57); DNA_Equal(a, b) (analysis/DNAVector.cc:390-403). */
int sxc_reverse_complement(char* seq, int32_t len);
double sxc_dna_equal(char a, char b);

/* CCSignal::SetSequence(seq, N) (analysis/CrossCorr.cc:179-206): out = 4*N floats, channels A, C, G, T. */
int sxc_encode_signal(sxc_ctx* ctx, const char* seq, int32_t len, int32_t N, float* out);

/* FFTReal<float>::do_fft (extern/RealFFT/FFTReal.hpp:146) on `batch` signals of N floats;
 * spectra in FFTReal's packed layout. */
int sxc_fft_forward(sxc_ctx* ctx, const float* in, float* out, int32_t N, int32_t batch);

/* CrossCorrelation::CrossCorrelate(out, target, query) (analysis/CrossCorr.cc:386-403):
 * tsig and qsig are 4*N encoded signals; xc receives N floats. */
int sxc_xcorr(sxc_ctx* ctx, const float* tsig, const float* qsig, int32_t N, float* xc);

/* SeqAnalyzer::MatchUp(out, query, target, xc) (analysis/CrossCorr.cc:583-605) with
 * SetTopCutoff(cutoff): runs of every accepted lag in reference order.  *n_out receives
 * the number found; SXC_ERR_CAPACITY if it exceeds cap. */
int sxc_matchup(sxc_ctx* ctx, const char* q, int32_t lq, const char* t, int32_t lt, const float* xc, int32_t N,
                double cutoff, sxc_seqmatch* out, int32_t cap, int32_t* n_out);

/* SeqAnalyzer::FindTop(xc) with SetTopCutoff(cutoff) (analysis/CrossCorr.cc:878-944): the accepted
 * lags, ascending (indices into xc; lag = index - N/2).  *n_out receives their number;
 * SXC_ERR_CAPACITY if it exceeds cap. */
int sxc_findtop(sxc_ctx* ctx, const float* xc, int32_t N, double cutoff, int32_t* out, int32_t cap, int32_t* n_out);

/* GetMatchProbability(target, query, sT, sQ, len, targetSize) (analysis/AlignProbability.cc:42-127)
 * and PrintMatch's identity (analysis/CrossCorr.cc:967-1037) for n matches. */
int sxc_match_probability(sxc_ctx* ctx, const char* t, int32_t lt, const char* q, int32_t lq,
                          const sxc_seqmatch* m, int32_t n, double target_size, double* prob, double* ident);

/* HomologyByXCorr::Align + FilterMatches for one chunk pair (analysis/HomologyByXCorrSlave.cc:
 * 221-253, 168-219): the plain drop-in, one pair per call.  Results in reference order
 * (forward strand, then reverse; lags ascending; runs ascending). */
int sxc_align_pair(sxc_ctx* ctx, const sxc_pair* pair, const sxc_params* prm, sxc_result* out, int32_t cap,
                   int32_t* n_out);

/* sxc_align_pair without device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Every device buffer it uses lives in a static
 * per-GPU array of 4 MiB that the library holds until the process ends; calls on one GPU are
 * serialised.  Same results as sxc_align_pair; limits: at most 4096 records kept per strand
 * (SXC_ERR_CAPACITY beyond that), no -prob_table. */
int sxc_align_pair_static(sxc_ctx* ctx, const sxc_pair* pair, const sxc_params* prm, sxc_result* out,
                          int32_t cap, int32_t* n_out);

/* Many chunk pairs in one fused launch (align_target over a block of t_pairs,
 * :270-300).  Results grouped by pair in input order, each pair as sxc_align_pair. */
int sxc_align_batch(sxc_ctx* ctx, const sxc_pair* pairs, int32_t n, const sxc_params* prm, sxc_result* out,
                    int64_t cap, int64_t* n_out);

/* Signal tails of the standalone aligner (tools/analysis/HomologyByXCorr.cc:706-716, 780-798).
 * That program reuses one CCSignal per side, and CCSignal::SetSequence writes only positions
 * [0, L) of its N-long channels (analysis/CrossCorr.cc:118-131, 182-206): positions from the
 * chunk length on keep what an earlier, longer sequence left there.  A tail lists those
 * earlier sequences as consecutive segments: the first covers [L, end[0]), the next
 * [end[0], end[1]), ...; each holds the signal of `seq` (reverse-complemented when
 * strand = 1) at the same positions, so end <= len.  Positions past the last end are 0. */
typedef struct {
  int32_t end;
  int32_t len;
  const char* seq;
  int32_t strand;
  int32_t pad_;
} sxc_tail_seg;

typedef struct {
  const sxc_tail_seg* segs;
  int32_t nseg;
  int32_t pad_;
} sxc_signal_tail;

/* Resident batch: the inputs of sxc_align_batch uploaded once to HBM, so the device work
 * can be run (and timed) repeatedly without host transfers.  Same results as
 * sxc_align_batch. */
typedef struct sxc_batch sxc_batch;
int sxc_batch_create(sxc_ctx* ctx, const sxc_pair* pairs, int32_t n, const sxc_params* prm, sxc_batch** out);
/* sxc_batch_create with signal tails (see sxc_tail_seg): t_tail[k] extends pair k's target
 * signal, q_tail[k] its query signal on both strands (the reverse-complement call of the
 * standalone aligner overwrites exactly the positions the forward call wrote).  Either
 * array may be NULL; nseg = 0 means no tail. */
int sxc_batch_create_tails(sxc_ctx* ctx, const sxc_pair* pairs, int32_t n, const sxc_params* prm,
                           const sxc_signal_tail* t_tail, const sxc_signal_tail* q_tail, sxc_batch** out);
/* Runs the device work.  The first run is launched kernel by kernel (with phase timing);
 * later runs replay the launch sequence captured as a HIP graph on the second run, unless
 * graphs are switched off with sxc_batch_set_graph(batch, 0). */
int sxc_batch_run(sxc_batch* batch);
int sxc_batch_set_graph(sxc_batch* batch, int32_t on);
int sxc_batch_results(sxc_batch* batch, sxc_result* out, int64_t cap, int64_t* n_out);
/* Accepted lags (FindTop) and diagonal runs (DoOne) of the last run, summed over the batch. */
int sxc_batch_stats(sxc_batch* batch, int64_t* candidates, int64_t* runs);
/* The lags the fused pair kernel accepted for one pair and strand in the last run: FindTop's
 * list (analysis/CrossCorr.cc:878-944) that DoOne then scans, ascending indices into xc. */
int sxc_batch_lags(sxc_batch* batch, int32_t pair, int32_t strand, int32_t* out, int32_t cap, int32_t* n_out);
void sxc_batch_destroy(sxc_batch* batch);

/* ---------------------------------------------------------------- serving loop
 *
 * A server keeps a target and a query genome's chunks resident in HBM, with the four
 * spectra of every target chunk computed once, and serves requests made of t_pair blocks
 * (analysis/WorkQueue.h:17-22) -- the battleship search's seed lines and hit
 * neighbourhoods (analysis/GridSearch.cc:188-411).  Requests are taken as they arrive:
 * a dispatcher thread packs the oldest pending chunk pairs, across requests, into launch
 * slots that run concurrently on their own streams, and admits more as slots free up.
 * The number of slots is what the HBM budget leaves after the resident data.  A request's
 * records are those align_target produces for its blocks, in the same order, whatever
 * else shares its launches. */
typedef struct sxc_server sxc_server;

/* One chunk of a resident genome (a ChunkManager chunk, analysis/SeqChunk.cc:72-165). */
typedef struct {
  const char* seq;
  int32_t len;          /* targets at most N/2, queries at most N */
  int32_t start;        /* SeqChunk::GetStart */
  int32_t record;       /* record id (t_result.target_id / query_id) */
  int32_t record_len;   /* ChunkManager::GetSize(record) (t_result.query_size, RCQuery) */
} sxc_chunk_desc;

/* t_pair: chunk index ranges, both inclusive, and the fast flag (cutoff_fast). */
typedef struct {
  int32_t target_from, target_to, query_from, query_to;
  int32_t fast;
} sxc_block;

typedef struct {
  int64_t hbm_budget;            /* device bytes for resident data + slots; 0 = 3/4 of free memory */
  int32_t max_pairs_per_launch;  /* 0 = 2048 */
  int32_t max_slots;             /* launches in flight; 0 = 3 */
  int32_t no_query_spectra;      /* 1: do not keep query spectra resident (0: keep both strands'
                                    spectra of every query chunk when they fit the budget, so a
                                    served pair skips the query's encode and forward transforms) */
  int32_t pad_;
} sxc_server_config;

typedef struct {
  int64_t requests_done, pairs_done, launches, results;
  int64_t resident_bytes;        /* genome bytes, chunk tables, target (and query) spectra */
  int64_t slot_bytes;            /* device bytes of one launch slot */
  int32_t slots, max_pairs_per_launch;
  double device_ms;              /* summed device time of all launches */
  int64_t query_spectra_bytes;   /* resident query spectra (0: the pair kernel transforms queries) */
  int64_t near_threshold;        /* records decided on the host because their device probability fell
                                    within a few device-libm ulps of min_prob (see sxc_timing) */
} sxc_server_info;

/* prm->strands is ignored (both strands, as Align). */
int sxc_server_create(sxc_ctx* ctx, const sxc_params* prm, const sxc_chunk_desc* targets, int32_t nt,
                      const sxc_chunk_desc* queries, int32_t nq, const sxc_server_config* cfg, sxc_server** out);
int sxc_server_submit(sxc_server* srv, const sxc_block* blocks, int32_t nblocks, uint64_t* ticket);
/* Waits up to timeout_ms (< 0: no limit) for the request; SXC_PENDING if it is not done.
 * latency_ms: submission to completion. */
int sxc_server_wait(sxc_server* srv, uint64_t ticket, int32_t timeout_ms, int64_t* n_results, double* latency_ms);
/* Copies a finished request's records and forgets the request (SXC_PENDING if not finished). */
int sxc_server_results(sxc_server* srv, uint64_t ticket, sxc_result* out, int64_t cap, int64_t* n_out);
int sxc_server_stats(sxc_server* srv, sxc_server_info* out);
void sxc_server_destroy(sxc_server* srv);

/* ---------------------------------------------------------------- autotuner
 *
 * Each FFT length has several compiled kernel geometries that compute the same FFTReal
 * dataflow bit for bit and differ in how it maps onto the GPU: the stage-A width (16 or 32
 * positions per thread: the radix of the register-resident first levels), the LDS round
 * depth (at most 2 or 3 butterfly levels per pass over LDS), a register budget for more
 * workgroups per CU, and a lean LDS plan (see sxc_geometry).  sxc_autotune times every
 * geometry of the pair kernel and of the target-spectra kernel for each length in
 * [N_lo, N_hi] on a synthetic batch (chunks of N/2 bases, both strands), checks that all
 * geometries accept identical lags, and makes the fastest the context's choice for later
 * batches and servers.  out receives one entry per (length, geometry).  The reference has
 * one CPU FFT and no counterpart; sxc_set_geometry applies a saved choice without timing. */
typedef struct {
  int32_t stage_a;          /* positions per thread in stage A (16 or 32) */
  int32_t round_levels;     /* butterfly levels per LDS round, at most (2 or 3) */
  int32_t three_per_cu;     /* 1: register budget for three workgroups per CU */
  int32_t lean_lds;         /* 1: twiddles and chunk bytes read through the caches, not LDS */
} sxc_geometry; /* one compiled FFT kernel geometry */
typedef struct {
  int32_t N;
  int32_t pairs;            /* chunk pairs of the synthetic batch */
  sxc_geometry geo;
  double correlate_ms;      /* pair kernel over the batch (both strands), mean of 3 launches */
  double spectra_ms;        /* target-spectra kernel over the batch's target chunks */
  int32_t chosen_correlate; /* 1: the context now runs the pair kernel in this geometry */
  int32_t chosen_spectra;   /* 1: ... the target-spectra kernel */
} sxc_tune_entry;
#define SXC_KERNEL_PAIRS 0
#define SXC_KERNEL_SPECTRA 1
int sxc_autotune(sxc_ctx* ctx, int32_t N_lo, int32_t N_hi, sxc_tune_entry* out, int32_t cap, int32_t* n_out);
int sxc_set_geometry(sxc_ctx* ctx, int32_t N, int32_t kernel, const sxc_geometry* geo);
int sxc_get_geometry(sxc_ctx* ctx, int32_t N, int32_t kernel, sxc_geometry* geo);

/* Time of the last sxc_align_batch / sxc_align_pair / sxc_batch_run (milliseconds): total_ms,
 * upload_ms and download_ms are wall time on the host; the phases and device_ms are device
 * time summed over the length classes of the batch, which run side by side (so they can add
 * up to more than total_ms).  A run replayed from a captured graph (graph = 1) has no phase
 * times, only device_ms. */
typedef struct {
  double total_ms;
  double upload_ms;
  double spectra_ms;
  double pair_ms;
  double finalize_ms;
  double download_ms;
  int64_t candidates;
  int64_t runs;
  double correlate_ms;  /* encode + FFT + product + iFFT + peak search (part of pair_ms) */
  double scan_ms;       /* diagonal scan + probability tail (part of pair_ms) */
  double device_ms;     /* first kernel to last copy of the run, per length class, summed */
  int32_t graph;        /* 1 if the run was replayed from a captured graph */
  int32_t pad_;
  int64_t near_threshold; /* records whose device probability fell within a few device-libm ulps
                             of min_prob, decided again on the host with its libm (glibc erf/exp,
                             analysis/AlignProbability.cc:122-124) */
} sxc_timing;
int sxc_last_timing(sxc_ctx* ctx, sxc_timing* out);

#ifdef __cplusplus
}
#endif
#endif
