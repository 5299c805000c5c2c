// sxc_tables.h -- host construction of the kernels' constant tables (sxc_tables.cpp).
#pragma once
#include "sxc_types.h"

namespace sxc {

void build_tables(sxc_tables* t);
void build_osc(float* osc);  // 24576 floats: oscillator {cos, sin} for FFT levels 13 and 14
int ent_lut_len(int nb);     // doubles in the entropy table for N = 2^nb
void build_ent_lut(int nb, double* out);
int class_of(int byte);
bool class_has_thirds(int c);
// sxc_chunk flags bits 0-2 of a chunk's bytes (bit 0: a B/D/H/V letter; bit 1: only A/C/G/T;
// bit 2: only A/C/G/T/N), or -1 if a byte is >= 128.
int chunk_flags(const unsigned char* s, int len);
// z above which a run's probability is certainly below min_prob (see sxc_scan.h tail).
double keep_z_bound(double min_prob, double target_size);
// ProbTable thresholds for a target size: SXC_PT_ROWS x SXC_PT_LEN doubles (row 0 unused).
void build_prob_table(double target_size, double* t);

// The kernels' tables built once on the host (host-side lookups use them).
const sxc_tables& host_tables();

// Window entropies of a chunk with a 1/3-valued letter (B, D, H, V) at FFT length N, both
// strands, with the host's log: SXC_ENT_ROW floats (the forward strand's windows, then the
// reverse complement's, SXC_ENT_ROW / 2 slots each).  Only read when L >= 1024.
void window_entropy_row(const unsigned char* seq, int L, int N, float* row);

// Keep band of the device's match probability.  The device computes prob with its own erf
// and exp; a run with prob >= keep_lo is emitted, and flagged SXC_REC_NEAR when
// prob < flag_hi, i.e. when the device's libm could put it on the other side of min_prob
// from the host's.  The host then decides it (resolve_record).
void keep_band(double min_prob, double target_size, double* keep_lo, double* flag_hi);

// GetMatchProbabilityEx (analysis/AlignProbability.cc:62-127) on host bytes with the host's
// libm: target bytes t from start_t; query bytes q (qlen of them, reverse-complemented when
// rc) from start_q.  *ident is PrintMatch's identity.
double host_match_probability(const unsigned char* t, const unsigned char* q, int qlen, int rc, int start_t,
                              int start_q, int len, double target_size, double* ident);

// What the host needs to re-decide flagged records of one launch: the chunk bytes and tables
// and the launch's pairs (a flagged record names its pair in pad_[1..4]).
struct ResolveCtx {
  const unsigned char* seq;
  const sxc_chunk* tch;
  const sxc_chunk* qch;
  const sxc_pair_dev* pairs;
  int64_t npairs;
  int q_chunk;
  double min_prob;
  double target_total;
  // optional: the target and query chunk bytes themselves (one pair per launch), instead of
  // seq + the chunk tables' offsets
  const unsigned char* tbytes;
  const unsigned char* qbytes;
};
// A record without the flag: 1 (kept as is).  A flagged record: its probability and identity
// recomputed on the host, the flag cleared; 1 if kept (prob >= min_prob), 0 if dropped; -1 if
// the record does not fit its pair.
int resolve_record(sxc_result_dev& r, const ResolveCtx& c);

}  // namespace sxc
