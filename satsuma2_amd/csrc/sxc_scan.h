// sxc_scan.h -- diagonal scan of every accepted lag and the FP64 match tail.
//
// One workgroup per (pair, strand).  SeqAnalyzer::DoOne (OPTIM_DOONE,
// analysis/CrossCorr.cc:667-724) slides a 46-wide window of integer scores
// (int)(100 * DNA_EqualAmb + 0.5) along the diagonal of each lag FindTop accepted and
// opens a run while the window sum exceeds 1889.  Lags are taken in epochs of 8 or 16.
// Pairs of pure A/C/G/T chunks (scan_epoch_bitmap): threads take whole 32-position blocks of
// a lag, keep the window's flags in an LDS bitmap, and a second pass turns each rising edge
// into a run.  Other pairs (scan_epoch_segments): a lag's diagonal is split into contiguous
// segments, one per thread; each thread recomputes the window at its segment start,
// records the runs that open and close inside its segment, and a second phase closes runs
// that cross segment boundaries.  Runs go to an LDS queue; the queue is drained by all
// threads at once, one run per thread, through GetMatchProbabilityEx
// (analysis/AlignProbability.cc:62-127) and FilterMatches
// (analysis/HomologyByXCorrSlave.cc:168-219).  Kept records carry the key
// (lag rank, run start) so the finalize kernel restores the reference's order.
//
// Queue capacity: the segment scans' queue holds every run one lag can produce (runs need
// a below-threshold gap, so at most npos/2 + 1 per lag, npos <= N/2), and an epoch of
// several lags that overflows is discarded and redone one lag at a time; the bitmap scan
// re-extracts an overflowing epoch's runs in rounds that each fit its queue.
#pragma once
#include <math.h>

#include "sxc_fft.h"
#include "sxc_tables.h"

namespace sxc {

struct ScanArgs {
  const sxc_tables* tab;
  const unsigned char* seq;
  const sxc_chunk* tchunks;
  const sxc_chunk* qchunks;
  const sxc_pair_dev* pairs;
  const uint16_t* cands;     // accepted lags, N slots per (pair, strand)
  const int32_t* ncand;      // accepted lags per (pair, strand)
  int32_t n_fft;             // N
  int32_t strands;
  int32_t min_len;
  int32_t q_chunk;
  double min_prob;
  double target_total;
  double z_skip;             // runs with (m - x) / s / 1.414213562 >= z_skip have prob < min_prob
  // keep band (keep_band): runs with prob >= keep_lo are written; those with prob < flag_hi
  // are flagged SXC_REC_NEAR for the host to decide, and counted in *nflag (optional)
  double keep_lo;
  double flag_hi;
  int32_t* nflag;
  int32_t mode;              // 0: filter runs into t_results; 1: write runs (MatchUp)
  sxc_result_dev* out;       // mode 0: cap records per (pair, strand)
  int32_t* runs;             // mode 1: cap x {start_target, start_query, len} per (pair, strand)
  uint64_t* keys;
  int32_t* counts;
  int32_t cap;
  int32_t* stats;            // optional {candidates, runs} per (pair, strand)
  const double* ptable;      // optional ProbTable thresholds (SXC_PT_ROWS x SXC_PT_LEN): -prob_table
  double table_prob;         // what a table hit reports (the table's cutoff)
  // workgroups per (pair, strand), each scanning a contiguous share of its accepted lags
  // (0 or 1: one).  With more than one, records take their slots through a global atomic
  // on counts[ps], which must be zero at launch (the pair kernel clears it), and stats
  // runs are added atomically.
  int32_t split;
};

// Host: the keep decision's parameters for min_prob and the target size -- the float
// pre-filter's z bound and the keep band (sxc_tables.h).
inline void set_keep(ScanArgs& a, double min_prob, double target_total) {
  a.min_prob = min_prob;
  a.target_total = target_total;
  a.z_skip = keep_z_bound(min_prob, target_total);
  keep_band(min_prob, target_total, &a.keep_lo, &a.flag_hi);
}

constexpr int SCAN_NT = 256;  // threads per scan workgroup
// smallest count c of identical A/C/G/T bases with 100 * c > 1889 (19 of 46)
constexpr int SXC_MIN_MATCHES = SXC_WIN_THRESHOLD / 100 + 1;
static_assert(100 * SXC_MIN_MATCHES > SXC_WIN_THRESHOLD && 100 * (SXC_MIN_MATCHES - 1) <= SXC_WIN_THRESHOLD, "");
// With N: a position involving an N scores 25 (N:N and N against A/C/G/T), so a window of M
// identical A/C/G/T bases and K positions with an N exceeds 1889 <=> 4 M + K >= 76.
constexpr int SXC_MIN_QUARTERS = SXC_WIN_THRESHOLD / 25 + 1;
static_assert(25 * SXC_MIN_QUARTERS > SXC_WIN_THRESHOLD && 25 * (SXC_MIN_QUARTERS - 1) <= SXC_WIN_THRESHOLD, "");

// NB: log2 N in bits 0-4; bit 5 selects the kernel for pairs whose two chunks are pure
// A/C/G/T, whose LDS plan holds only the packed bases, the run queue and the counters (the
// accepted lags are read from global memory), so twice as many of its workgroups share a
// CU; bit 6 the same kernel for pairs of A/C/G/T/N chunks with an N (HASN: N masks beside
// the packed bases).  The general kernel takes every other pair.  Each exits at once on the
// others' pairs.
template <int NB>
struct ScanSmem {
  static constexpr int LOG = NB & 31;
  static constexpr bool HASN = (NB & 64) != 0;
  static constexpr bool PURE = (NB & 96) != 0;
  static constexpr int N = 1 << LOG;
  // The segment scans' queue holds every run one lag can produce (npos / 2 + 1, npos <= N / 2);
  // the bitmap scan flushes its queue whenever it would overflow (896 entries keep its LDS
  // plan at N = 8192 within 20 KB, 8 workgroups per CU).
  static constexpr int QCAP = (PURE ? 896 : N / 4 + 64);
  static constexpr int PKW = N / 16 + 4;                // 2-bit packed words (+3 words of read-ahead)
  static constexpr int TCODE = 0;                       // uchar[N] (general kernel)
  static constexpr int QCODE = TCODE + (PURE ? 0 : N);  // uchar[N] (general kernel)
  static constexpr int TPK = QCODE + (PURE ? 0 : N);    // uint32[PKW]: 16 bases per word (pure A/C/G/T only)
  static constexpr int QPK = TPK + 4 * PKW;             // uint32[PKW]
  // The pure kernel (BMRUNS) keeps each epoch's flags in a bitmap of WPL words per lag and
  // turns its rising edges into runs afterwards; the general kernel scans contiguous
  // segments per thread and stitches runs across them.  BMRUNS, and the general kernel from
  // N = 16384, take one epoch at a time (drained after every epoch) in
  // 4-byte queue entries {lag - first lag of the epoch (RB bits), ns (NSB = LOG - 1), ne
  // (LOG)}; otherwise 16-lag epochs accumulate in 8-byte entries {lag rank, ns | ne << 16}
  // until the queue is half full.
  static constexpr bool BMRUNS = PURE;
  static constexpr bool COMPACT = BMRUNS || LOG >= 14;
  static constexpr int CE = LOG >= 14 ? 8 : 16;         // lags per epoch
  static constexpr int GSZ = SCAN_NT / CE;              // threads per lag in an epoch
  static constexpr int NSB = LOG - 1;                   // ns < N / 2
  static constexpr int RB = 32 - NSB - LOG;             // ne <= N / 2 takes LOG bits
  static_assert(!COMPACT || (RB > 0 && (1 << RB) >= CE), "compact queue entries");
  static constexpr int QE = COMPACT ? 4 : 8;            // bytes per queue entry
  static constexpr int WPL = N / 64;                    // bitmap words per lag: npos <= N / 2
  static constexpr int TNK = QPK + 4 * PKW;             // uint32[PKW]: bit 2b set where base b of a word is N (HASN)
  static constexpr int QNK = TNK + (HASN ? 4 * PKW : 0);
  static constexpr int QUEUE = QNK + (HASN ? 4 * PKW : 0);  // queue entries [QCAP]
  static constexpr int SUMM = QUEUE + QE * QCAP;        // int2[SCAN_NT] (segments)
  static constexpr int BITMAP = SUMM;                   // uint32[CE][WPL] (BMRUNS)
  static constexpr int RLIST = BITMAP + 4 * CE * WPL;   // uint16[CE * WPL]: blocks with a rising edge (BMRUNS)
  static constexpr int EQUAL = SUMM + (BMRUNS ? 6 * CE * WPL : 8 * SCAN_NT);  // double[18][18]
  static constexpr int NTAB = PURE ? 0 : 1;             // the score tables (general kernel)
  static constexpr int GC = EQUAL + NTAB * 8 * SXC_NCLASS * SXC_NCLASS;  // double[18]
  static constexpr int SCORE = GC + NTAB * 8 * SXC_NCLASS;              // uchar[18][18]
  static constexpr int CLS = SCORE + NTAB * SXC_NCLASS * SXC_NCLASS;     // uchar[128]
  static constexpr int RC = CLS + NTAB * 128;                           // uchar[128]
  static constexpr int MISC = (RC + NTAB * 128 + 15) & ~15;             // int[16 + CE]
  static constexpr int BYTES = MISC + 64 + 4 * CE;
};

enum { SM_QN = 0, SM_KEPT = 1, SM_RUNS = 2, SM_NCAND = 3, SM_LT = 4, SM_LQ = 5, SM_NRB = 7, SM_NPOS = 16 };

// Base code: class id (0..16) of a codec letter, or 128 | byte for a byte whose codec
// row is zero.  Two bases are identical bytes iff their codes are equal.
SXC_HD inline unsigned char base_code(const sxc_tables* tab, unsigned char b) {
  const unsigned char c = tab->cls[b & 127];
  return c < SXC_CLASS_ZERO ? c : (unsigned char)(128 | b);
}

// (int)(100 * DNA_EqualAmb(a, b) + 0.5): identical letters other than 'N' score 100
// (analysis/DNAVector.cc:372-387), 'N':'N' scores 25, a zero-row byte scores 0 against
// any different byte.
SXC_HD inline int code_score(const unsigned char* score, unsigned ca, unsigned cb) {
  if (ca == cb) return ca == SXC_CLASS_N ? 25 : 100;
  if ((ca | cb) & 128) return 0;
  return score[ca * SXC_NCLASS + cb];
}

SXC_HD inline int code_class(unsigned c) { return c < 128 ? (int)c : SXC_CLASS_ZERO; }

// GetMatchProbabilityEx on base codes; `ident` is also PrintMatch's identity (the same
// DNA_Equal sum in the same order).
SXC_HD inline double code_tail(const unsigned char* tc, const unsigned char* qc, const double* equal,
                               const double* gcv, int start_t, int start_q, int len, double target_size,
                               double& ident) {
  double matches = 0, gct = 0, gcq = 0;
  for (int i = 0; i < len; ++i) {
    const int ca = code_class(tc[start_t + i]), cb = code_class(qc[start_q + i]);
    matches += equal[ca * SXC_NCLASS + cb];
    gct += gcv[ca];
    gcq += gcv[cb];
  }
  ident = matches / (double)len;
  const double gc_target = gct / (double)len;
  // GCAdjustExpect, Sigma, CDF (analysis/AlignProbability.cc:11-40)
  const double at_target = 1. - gc_target;
  double r = gcq * gc_target;
  r += ((double)len - gcq) * at_target;
  const double p = r / (double)len / 2.;
  const double s = sqrt(p * (1. - p) * (double)len);
  const double m = p * (double)len;
  const double x = (double)len * ident;
  const double cdf = 0.5 * (1. + erf((m - x) / s / 1.414213562));
  const double expect = cdf * target_size;
  return exp(-expect);
}

// ---- bit-parallel window counts for pure A/C/G/T diagonals
//
// Bases are packed 2 bits each, 16 per 32-bit word.  32 bases of the target and of the
// query diagonal XOR to 32 2-bit groups; a group is zero exactly where the bases match.
// Block k of a diagonal holds the match bits of positions [32k, 32k + 32) at even bit
// positions; the 46-wide window count at position n = 32k + r reads blocks k, k-1, k-2.

SXC_HD inline uint64_t load32(const uint32_t* pk, int p) {  // 32 bases starting at base p >= 0
  const int w = p >> 4, b = (p & 15) * 2;
  const uint64_t lo = (uint64_t)pk[w] | ((uint64_t)pk[w + 1] << 32);
  if (!b) return lo;
  return (lo >> b) | ((uint64_t)pk[w + 2] << (64 - b));
}

SXC_HD inline uint64_t match_block(const uint32_t* tpk, const uint32_t* qpk, int ti, int qj, int k) {
  if (k < 0) return 0;
  const uint64_t x = load32(tpk, ti + 32 * k) ^ load32(qpk, qj + 32 * k);
  return ~(x | (x >> 1)) & 0x5555555555555555ull;
}

SXC_HD inline int popc64(uint64_t v) { return __builtin_popcountll(v); }

// bits [s, s + 32) of hi:lo, 0 <= s < 32 (a funnel shift: one v_alignbit)
SXC_HD inline uint32_t funnel32(uint32_t hi, uint32_t lo, int s) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_alignbit(hi, lo, (unsigned)s);
#else
  return (lo >> s) | ((hi << 1) << (31 - s));
#endif
}

// Bytes of hi:lo chosen by the selector bytes of sel (0-3: lo's bytes, 4-7: hi's; v_perm_b32).
SXC_HD inline uint32_t byte_perm(uint32_t hi, uint32_t lo, uint32_t sel) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_perm(hi, lo, sel);
#else
  const uint64_t v = ((uint64_t)hi << 32) | lo;
  uint32_t r = 0;
  for (int b = 0; b < 4; ++b) r |= (uint32_t)((v >> (8 * ((sel >> (8 * b)) & 7))) & 0xffu) << (8 * b);
  return r;
#endif
}

// Hides x's value from the optimizer (no instruction).
SXC_HD inline void opaque(uint32_t& x) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm("" : "+v"(x));
#else
  (void)x;
#endif
}

// a where mask is set, b elsewhere (v_bfi_b32)
SXC_HD inline uint32_t bit_select(uint32_t mask, uint32_t a, uint32_t b) { return (a & mask) | (b & ~mask); }

// Window flags of 16 consecutive positions, four at a time in the bytes of a word.  Group p
// (bits 2p, 2p + 1) of z holds m(n) + 1 - m(n - 46) in {0, 1, 2} for the half's position p;
// byte i covers positions 4i .. 4i + 3 and carries u = C - 19 + 128, C the window count.  As
// 0 <= C <= 46, every u below lies in [109, 155]: no add carries and no subtraction borrows
// across a byte, and bit 7 of u is the flag C >= 19.  ub holds u of the position before the
// half in every byte and is advanced to the half's last position.  The flag of position
// 4i + r is returned at bit 8i + 4 + r (bits 8i .. 8i + 3 are not defined).
SXC_HD inline uint32_t half_flags(uint32_t z, uint32_t& ub) {
  const uint32_t A = z & 0x33333333u;               // per byte: z0 | z2 << 4
  const uint32_t P = A + ((z >> 2) & 0x33333333u);  // z0 + z1 | (z2 + z3) << 4
  const uint32_t P01 = P & 0x0F0F0F0Fu;
  const uint32_t s = P01 + ((P >> 4) & 0x0F0F0F0Fu);  // the byte's four z, <= 8
  // the earlier bytes' s (<= 24): shift-adds, kept from being folded into a quarter-rate multiply
  uint32_t x = s + (s << 8);
  opaque(x);
  x = x + (x << 16);
  opaque(x);
  x <<= 8;
  const uint32_t u0 = ub + x - 0x0C080400u;  // u at position 4i - 1
  const uint32_t u1 = u0 + (A & 0x03030303u) - 0x01010101u;
  const uint32_t u2 = u0 + P01 - 0x02020202u;
  const uint32_t u3 = u2 + ((A >> 4) & 0x03030303u) - 0x01010101u;
  const uint32_t u4 = u0 + s - 0x04040404u;
  ub = byte_perm(u4, u4, 0x03030303u);
  return bit_select(0x80808080u, u4, bit_select(0x40404040u, u3 >> 1, bit_select(0x20202020u, u2 >> 2, u1 >> 3)));
}

// Window flags of a 32-position block (bit r: position r) from its z groups (positions 0-15
// in zl, 16-31 in zh); ub as in half_flags.
SXC_HD inline uint32_t window_flags(uint32_t zl, uint32_t zh, uint32_t& ub) {
  const uint32_t n0 = half_flags(zl, ub);
  const uint32_t n1 = half_flags(zh, ub);
  // bytes 0 and 2 of c gather the half's positions 0-7 and 8-15
  const uint32_t c0 = bit_select(0x0F0F0F0Fu, n0 >> 4, n0 >> 8);
  const uint32_t c1 = bit_select(0x0F0F0F0Fu, n1 >> 4, n1 >> 8);
  return byte_perm(c1, c0, 0x06040200u);
}

// half_flags' four u words (byte i of u[t]: u at position 4i + t) for any count that changes
// by z - 1 per position and stays within [0, 255] at every position; ub as in half_flags.
SXC_HD inline void half_values(uint32_t z, uint32_t& ub, uint32_t* u) {
  const uint32_t A = z & 0x33333333u;
  const uint32_t P = A + ((z >> 2) & 0x33333333u);
  const uint32_t P01 = P & 0x0F0F0F0Fu;
  const uint32_t s = P01 + ((P >> 4) & 0x0F0F0F0Fu);
  uint32_t x = s + (s << 8);
  opaque(x);
  x = x + (x << 16);
  opaque(x);
  x <<= 8;
  const uint32_t u0 = ub + x - 0x0C080400u;
  u[0] = u0 + (A & 0x03030303u) - 0x01010101u;
  u[1] = u0 + P01 - 0x02020202u;
  u[2] = u[1] + ((A >> 4) & 0x03030303u) - 0x01010101u;
  u[3] = u0 + s - 0x04040404u;
  ub = byte_perm(u[3], u[3], 0x03030303u);
}

// Window flags of a 32-position block of an A/C/G/T/N pair (bit r: position r).  The window
// holds M identical A/C/G/T bases and K positions with an N (M + K <= 46); its score
// 100 M + 25 K exceeds 1889 iff 4 M + K >= 76.  zm* and zk* hold m(n) + 1 - m(n - 46) and
// k(n) + 1 - k(n - 46) per position as z of half_flags (positions 0-15 in *l, 16-31 in *h);
// ubm carries M and ubk K + 128 - 76 in every byte, advanced as ub of half_flags.  Per byte
// 4 M + (K + 52) <= 3 M + 98 <= 236 (4 M < 256 keeps M's bits in its byte), so bit 7 of the
// combined byte is the flag.
SXC_HD inline uint32_t window_flags_mk(uint32_t zml, uint32_t zmh, uint32_t zkl, uint32_t zkh, uint32_t& ubm,
                                       uint32_t& ubk) {
  auto half = [&](uint32_t zm, uint32_t zk) {
    uint32_t um[4], uk[4], c[4];
    half_values(zm, ubm, um);
    half_values(zk, ubk, uk);
    for (int t = 0; t < 4; ++t) c[t] = (um[t] << 2) + uk[t];
    return bit_select(0x80808080u, c[3], bit_select(0x40404040u, c[2] >> 1, bit_select(0x20202020u, c[1] >> 2, c[0] >> 3)));
  };
  const uint32_t n0 = half(zml, zkl);
  const uint32_t n1 = half(zmh, zkh);
  const uint32_t c0 = bit_select(0x0F0F0F0Fu, n0 >> 4, n0 >> 8);
  const uint32_t c1 = bit_select(0x0F0F0F0Fu, n1 >> 4, n1 >> 8);
  return byte_perm(c1, c0, 0x06040200u);
}

// matches in the window [n - 45, n], n = 32k + r
SXC_HD inline int window_count(uint64_t e0, uint64_t e1, uint64_t e2, int r) {
  const uint64_t low = (r == 31) ? ~0ull : ((1ull << (2 * r + 2)) - 1);
  int c = popc64(e0 & low);
  if (r >= 13) {
    c += popc64(e1 & ~((1ull << (2 * (r - 13))) - 1));
  } else {
    c += popc64(e1) + popc64(e2 & ~((1ull << (2 * (r + 19))) - 1));
  }
  return c;
}

// Counts over a run of pure A/C/G/T bases: identical positions and C/G bases of each
// side.  The reference's double sums of 0/1 terms equal these counts exactly.
SXC_HD inline void packed_counts(const uint32_t* tpk, const uint32_t* qpk, int st, int sq, int len, int& match,
                                 int& gct, int& gcq) {
  match = gct = gcq = 0;
  for (int off = 0; off < len; off += 32) {
    const int m = (len - off < 32) ? len - off : 32;
    const uint64_t keep = (m == 32) ? ~0ull : ((1ull << (2 * m)) - 1);
    const uint64_t tv = load32(tpk, st + off), qv = load32(qpk, sq + off);
    const uint64_t x = tv ^ qv;
    match += popc64(~(x | (x >> 1)) & 0x5555555555555555ull & keep);
    gct += popc64((tv ^ (tv >> 1)) & 0x5555555555555555ull & keep);  // codes 1 (C), 2 (G)
    gcq += popc64((qv ^ (qv >> 1)) & 0x5555555555555555ull & keep);
  }
}

// The same counts where bases may be N (bit 2b of the N mask words set): identical A/C/G/T
// bases, positions with an N on either side, C/G bases and N bases of each side.
SXC_HD inline void packed_counts_n(const uint32_t* tpk, const uint32_t* qpk, const uint32_t* tnk,
                                   const uint32_t* qnk, int st, int sq, int len, int& match, int& nk, int& gct,
                                   int& ntn, int& gcq, int& nqn) {
  match = nk = gct = ntn = gcq = nqn = 0;
  for (int off = 0; off < len; off += 32) {
    const int m = (len - off < 32) ? len - off : 32;
    const uint64_t keep = ((m == 32) ? ~0ull : ((1ull << (2 * m)) - 1)) & 0x5555555555555555ull;
    const uint64_t tv = load32(tpk, st + off), qv = load32(qpk, sq + off);
    const uint64_t tn = load32(tnk, st + off) & keep, qn = load32(qnk, sq + off) & keep;
    const uint64_t x = tv ^ qv;
    match += popc64(~(x | (x >> 1)) & keep & ~(tn | qn));
    nk += popc64(tn | qn);
    gct += popc64((tv ^ (tv >> 1)) & keep & ~tn);
    gcq += popc64((qv ^ (qv >> 1)) & keep & ~qn);
    ntn += popc64(tn);
    nqn += popc64(qn);
  }
}

// The tail after the sums (GCAdjustExpect, Sigma, CDF; analysis/AlignProbability.cc:11-40,
// 107-126).  Returns false, without evaluating erf/exp, when z >= z_skip guarantees the
// probability is below the keep threshold.
SXC_HD inline bool tail_from_sums(double matches, double gct, double gcq, int len, double target_size, double z_skip,
                                  double& prob, double& ident) {
  ident = matches / (double)len;
  const double gc_target = gct / (double)len;
  const double at_target = 1. - gc_target;
  double r = gcq * gc_target;
  r += ((double)len - gcq) * at_target;
  const double p = r / (double)len / 2.;
  const double s = sqrt(p * (1. - p) * (double)len);
  const double m = p * (double)len;
  const double x = (double)len * ident;
  const double z = (m - x) / s / 1.414213562;
  if (z >= z_skip) return false;
  const double cdf = 0.5 * (1. + erf(z));
  const double expect = cdf * target_size;
  prob = exp(-expect);
  return true;
}

// The z of tail_from_sums estimated in float from the integer sums: its error (|m - x| off
// by at most ~5e-7 * len, s by a few ulp) is far below the margin, so a true result implies
// z >= z_skip in FP64 as well and the run can be dropped without the FP64 divisions.
SXC_HD inline bool z_clearly_skipped(int matches, int gct, int gcq, int len, double z_skip) {
  const float fl = (float)len;
  const float gt = (float)gct / fl;
  const float r = (float)gcq * gt + (float)(len - gcq) * (1.f - gt);
  const float p = r / fl * 0.5f;
  const float s = sqrtf(p * (1.f - p) * fl);
  const float z = (p * fl - (float)matches) / s * 0.70710678f;
  const float margin = 1e-5f * (fl / s + fabsf(z) + 1.f);
  return z - margin >= (float)z_skip;  // false for NaN (s = 0) and for z_skip = inf
}

// ProbTable::GetMatchProbability after the sums (analysis/ProbTable.cc:118-141): the table's
// cutoff if the identity reaches the threshold for (expected identity, length), else 0.  An
// expectation below 1/511 selects row 0, which the reference leaves empty; it reports 0.
SXC_HD inline double table_from_sums(double matches, double gct, double gcq, int len, const double* ptable,
                                     double table_prob, double& ident) {
  ident = matches / (double)len;
  const double gc_target = gct / (double)len;
  const double at_target = 1. - gc_target;
  double r = gcq * gc_target;
  r += ((double)len - gcq) * at_target;
  const double p = r / (double)len / 2.;
  const int index = (int)(p * (SXC_PT_ROWS - 1));
  if (index < 1) return 0.;
  const int l = len >= SXC_PT_LEN ? SXC_PT_LEN - 1 : len;
  return ident >= ptable[(size_t)index * SXC_PT_LEN + l] ? table_prob : 0.;
}

struct Lag {
  int shift, i0, npos;
};

SXC_HD inline Lag lag_geometry(int idx, int H, int Lq, int Lt) {
  // SeqAnalyzer::DoOne loop bounds: i from max(0, -shift) while j = i + shift < Lq and
  // i + 1 < Lt (analysis/CrossCorr.cc:685-696)
  Lag g;
  g.shift = idx - H;
  g.i0 = g.shift < 0 ? -g.shift : 0;
  int iend = Lq - g.shift;
  if (Lt - 1 < iend) iend = Lt - 1;
  g.npos = (g.i0 < Lt && iend > g.i0) ? iend - g.i0 : 0;
  return g;
}

// Queue entry of a run (both kernels): compact {lag - first lag of the epoch, ns, ne}
// or the 8-byte {lag rank, ns | ne << 16}.
template <int NB, class Ex>
SXC_HD inline void push_run(Ex& ex, char* sm, int c0, int rank, int ns, int ne) {
  using SS = ScanSmem<NB>;
  int* misc = (int*)(sm + SS::MISC);
  int* queue = (int*)(sm + SS::QUEUE);
  const int k = ex.atomic_add_shared(&misc[SM_QN], 1);
  if (k < SS::QCAP) {
    if constexpr (SS::COMPACT) {
      queue[k] = (int)((unsigned)(rank - c0) | ((unsigned)ns << SS::RB) | ((unsigned)ne << (SS::RB + SS::NSB)));
    } else {
      queue[2 * k] = rank;
      queue[2 * k + 1] = ns | (ne << 16);
    }
  }
}

// Pure A/C/G/T lags [c0, c1), gsz threads per lag.  Score = 100 * [match], and a window
// sum > 1889 <=> >= 19 matches.  C(n) = C(n - 1) + m(n) - m(n - 46) with m(< 0) = 0 equals
// the reference's running `match` (analysis/CrossCorr.cc:700-705).  Each thread takes whole
// 32-position blocks of one lag, updates every position branch-free and stores the block's
// flags (window above the threshold, positions >= 46 and < npos) in the epoch's LDS
// bitmap.  A second pass over the bitmap words, spread over all threads, turns every rising
// edge at ns and the first clear flag after it at ne (or npos, where the loop ends) into
// the run [ns, ne) -- the runs DoOne opens and closes.
template <int NB, class Ex, class Flush>
SXC_HD inline void scan_epoch_bitmap(Ex& ex, char* sm, const uint16_t* cands, int c0, int c1, int gsz, Flush&& flush) {
  using SS = ScanSmem<NB>;
  constexpr int H = SS::N / 2;
  constexpr int WPL = SS::WPL;
  int* misc = (int*)(sm + SS::MISC);
  int* npos_of = misc + SM_NPOS;
  uint32_t* bm = (uint32_t*)(sm + SS::BITMAP);
  const uint32_t* tpk = (const uint32_t*)(sm + SS::TPK);
  const uint32_t* qpk = (const uint32_t*)(sm + SS::QPK);
  const int Lt = misc[SM_LT], Lq = misc[SM_LQ];
  const int nl = c1 - c0;
  ex.each([&](int tid) {
    const int j = tid / gsz;
    const int l = tid % gsz;
    if (j >= nl) return;
    const Lag g = lag_geometry(cands[c0 + j], H, Lq, Lt);
    if (l == 0) npos_of[j] = g.npos;
    // whole blocks per thread: no block is written by two threads
    const int S = ((g.npos + gsz - 1) / gsz + 31) & ~31;
    const int na = l * S;
    const int nb = (na + S < g.npos) ? na + S : g.npos;
    if (na >= nb) return;
    uint32_t* row = bm + j * WPL;
    const int ti = g.i0, qj = g.i0 + g.shift;
    const int k0 = na >> 5, klast = (nb - 1) >> 5;
    uint16_t* rl = (uint16_t*)(sm + SS::RLIST);
    if constexpr (SS::HASN) {
      // the window test 4 M + K >= 76 over M identical A/C/G/T bases (m bits) and K positions
      // with an N (k bits), byte-parallel from the 2-bit groups m(n) + 1 - m(n - 46) and
      // k(n) + 1 - k(n - 46) (window_flags_mk)
      const uint32_t* tnk = (const uint32_t*)(sm + SS::TNK);
      const uint32_t* qnk = (const uint32_t*)(sm + SS::QNK);
      auto mk_at = [&](int kb, uint64_t& m, uint64_t& kk) {
        if (kb < 0) {
          m = kk = 0;
          return;
        }
        const uint64_t x = load32(tpk, ti + 32 * kb) ^ load32(qpk, qj + 32 * kb);
        kk = (load32(tnk, ti + 32 * kb) | load32(qnk, qj + 32 * kb)) & 0x5555555555555555ull;
        m = ~(x | (x >> 1)) & 0x5555555555555555ull & ~kk;
      };
      uint64_t m3, k3, m2, k2, m1, k1;
      mk_at(k0 - 3, m3, k3);
      mk_at(k0 - 2, m2, k2);
      mk_at(k0 - 1, m1, k1);
      const int M0 = window_count(m1, m2, m3, 31), K0 = window_count(k1, k2, k3, 31);  // at 32 k0 - 1
      uint32_t prev = (na - 1 >= 46 && 4 * M0 + K0 >= SXC_MIN_QUARTERS) ? 1u : 0u;
      uint32_t ubm = byte_perm(0u, (uint32_t)M0, 0u);
      uint32_t ubk = byte_perm(0u, (uint32_t)(K0 + 128 - SXC_MIN_QUARTERS), 0u);
      for (int k = k0; k <= klast; ++k) {
        uint64_t m0, kk0;
        mk_at(k, m0, kk0);
        const uint64_t mm = (m2 >> 36) | (m1 << 28), km = (k2 >> 36) | (k1 << 28);  // at n - 46
        const uint32_t zml = (uint32_t)m0 + (~(uint32_t)mm & 0x55555555u);
        const uint32_t zmh = (uint32_t)(m0 >> 32) + (~(uint32_t)(mm >> 32) & 0x55555555u);
        const uint32_t zkl = (uint32_t)kk0 + (~(uint32_t)km & 0x55555555u);
        const uint32_t zkh = (uint32_t)(kk0 >> 32) + (~(uint32_t)(km >> 32) & 0x55555555u);
        uint32_t f = window_flags_mk(zml, zmh, zkl, zkh, ubm, ubk);
        if (32 * k < 46) f &= (46 - 32 * k >= 32) ? 0u : ~((1u << (46 - 32 * k)) - 1);  // flags from n = 46
        if (k == klast && nb - 32 * k < 32) f &= (1u << (nb - 32 * k)) - 1;
        row[k] = f;
        if (f & ~((f << 1) | prev)) rl[ex.atomic_add_shared(&misc[SM_NRB], 1)] = (uint16_t)(j * WPL + k);
        prev = f >> 31;
        m2 = m1;
        m1 = m0;
        k2 = k1;
        k1 = kk0;
      }
      return;
    }
    uint64_t e2 = match_block(tpk, qpk, ti, qj, k0 - 2);
    uint64_t e1 = match_block(tpk, qpk, ti, qj, k0 - 1);
    const int C = window_count(e1, e2, match_block(tpk, qpk, ti, qj, k0 - 3), 31);  // C at 32 k0 - 1
    uint32_t prev = (na - 1 >= 46 && C >= SXC_MIN_MATCHES) ? 1u : 0u;          // flag at 32 k0 - 1
    uint32_t ub = byte_perm(0u, (uint32_t)(C + 128 - SXC_MIN_MATCHES), 0u);   // half_flags' u, every byte
    // flags from n = 46: block 0 keeps none, block 1 its bits from 14, later blocks all
    uint32_t keep = k0 == 0 ? 0u : (k0 == 1 ? 0xFFFFC000u : ~0u);
    uint16_t* rlist = (uint16_t*)(sm + SS::RLIST);
    // The block's 32 bases of each side are words w, w + 1, w + 2 shifted right by the
    // diagonal's bit offset; word w + 2 of one block is word w of the next.
    const int tb = (ti & 15) * 2, qb = (qj & 15) * 2;
    const uint32_t* tw = tpk + (ti >> 4) + 2 * k0;
    const uint32_t* qw = qpk + (qj >> 4) + 2 * k0;
    uint32_t t0 = tw[0], q0 = qw[0];
    uint32_t e1l = (uint32_t)e1, e1h = (uint32_t)(e1 >> 32), e2h = (uint32_t)(e2 >> 32);
    uint32_t f = 0;
    for (int k = k0; k <= klast; ++k, tw += 2, qw += 2) {
      const uint32_t t1 = tw[1], t2 = tw[2], q1 = qw[1], q2 = qw[2];
      const uint32_t xl = funnel32(t1, t0, tb) ^ funnel32(q1, q0, qb);
      const uint32_t xh = funnel32(t2, t1, tb) ^ funnel32(q2, q1, qb);
      t0 = t2;
      q0 = q2;
      const uint32_t e0l = ~(xl | (xl >> 1)) & 0x55555555u, e0h = ~(xh | (xh >> 1)) & 0x55555555u;
      // m(n - 46) for the block: offsets r < 14 come from block k-2, r >= 14 from block k-1
      const uint32_t ml = funnel32(e1l, e2h, 4), mh = funnel32(e1h, e1l, 4);
      // 2-bit group r of z holds m(n) + 1 - m(n - 46), in {0, 1, 2}: no carry crosses a
      // group, so each 32-bit half is one add
      const uint32_t zl = e0l + (~ml & 0x55555555u);
      const uint32_t zh = e0h + (~mh & 0x55555555u);
      f = window_flags(zl, zh, ub) & keep;
      keep = 0xFFFFC000u | (keep >> 18);
      row[k] = f;
      // (the same increment on every lane: one atomic per wave).  A rise past nb in the
      // lag's last block only adds a list entry whose word (cut below) yields no run.
      if (f & ~((f << 1) | prev)) rlist[ex.atomic_add_shared(&misc[SM_NRB], 1)] = (uint16_t)(j * WPL + k);
      prev = f >> 31;
      e2h = e1h;
      e1l = e0l;
      e1h = e0h;
    }
    if (nb - 32 * klast < 32) row[klast] = f & ((1u << (nb - 32 * klast)) - 1);  // the lag ends at npos
  });
  ex.sync();
  int* queue = (int*)(sm + SS::QUEUE);
  const uint16_t* rlist = (const uint16_t*)(sm + SS::RLIST);
  // All listed blocks at once; if their runs overflowed the queue, again in rounds of blocks
  // that fit it (a word holds at most 16 rising edges), the queue flushed after each round.
  const int nrb = misc[SM_NRB];
  auto extract = [&](int x0, int x1) {
    ex.each([&](int tid) {
      for (int x = x0 + tid; x < x1; x += SCAN_NT) {
        const int e = rlist[x];
        const int j = e / WPL, w = e % WPL;
        const int npos = npos_of[j];
        const int W = (npos + 31) >> 5;
        const uint32_t* row = bm + j * WPL;
        const uint32_t f = row[w];
        const uint32_t carry = w > 0 ? row[w - 1] >> 31 : 0u;
        uint32_t rises = f & ~((f << 1) | carry);
        while (rises) {
          const int r = __builtin_ctz(rises);
          rises &= rises - 1;
          const uint32_t clear = ~f & (~0u << r);  // bit r is set: the first clear bit is above it
          int ne;
          if (clear) {
            ne = 32 * w + __builtin_ctz(clear);
          } else {
            int w2 = w + 1;
            while (w2 < W && row[w2] == ~0u) ++w2;
            ne = w2 < W ? 32 * w2 + __builtin_ctz(~row[w2]) : npos;
          }
          const int at = ex.atomic_add_shared(&misc[SM_QN], 1);  // one atomic per wave
          if (at < SS::QCAP)
            queue[at] = (int)((unsigned)j | ((unsigned)(32 * w + r) << SS::RB) | ((unsigned)ne << (SS::RB + SS::NSB)));
        }
      }
    });
    ex.sync();
  };
  extract(0, nrb);
  const int qn = misc[SM_QN];
  if (qn <= SS::QCAP) {
    flush(qn);
  } else {
    ex.sync();  // every thread has read the count
    ex.each([&](int tid) {
      if (tid == 0) misc[SM_QN] = 0;
    });
    ex.sync();
    for (int x0 = 0; x0 < nrb; x0 += SS::QCAP / 16) {
      extract(x0, (x0 + SS::QCAP / 16 < nrb) ? x0 + SS::QCAP / 16 : nrb);
      flush(misc[SM_QN]);
    }
  }
  ex.each([&](int tid) {
    if (tid == 0) misc[SM_NRB] = 0;
  });
  ex.sync();
}

// Lags [c0, c1) with gsz threads per lag, each thread scanning a contiguous segment: it
// recomputes the window at its segment start, records the runs that open and close inside
// its segment, and a second phase closes runs that cross segment boundaries.
template <int NB, class Ex>
SXC_HD inline void scan_epoch_segments(Ex& ex, char* sm, const uint16_t* cands, int c0, int c1, int gsz) {
  using SS = ScanSmem<NB>;
  constexpr int INF = 0x7fffffff;
  constexpr int H = SS::N / 2;
  const unsigned char* tc = (const unsigned char*)(sm + SS::TCODE);
  const unsigned char* qc = (const unsigned char*)(sm + SS::QCODE);
  const unsigned char* score = (const unsigned char*)(sm + SS::SCORE);
  int* misc = (int*)(sm + SS::MISC);
  int* summ = (int*)(sm + SS::SUMM);
  const int Lt = misc[SM_LT], Lq = misc[SM_LQ];
  auto push = [&](int rank, int ns, int ne) { push_run<NB>(ex, sm, c0, rank, ns, ne); };
  ex.each([&](int tid) {
    const int c = c0 + tid / gsz;
    const int l = tid % gsz;
    int first_end = INF, open = -1;
    if (c < c1) {
      const Lag g = lag_geometry(cands[c], H, Lq, Lt);
      const int S = (g.npos + gsz - 1) / gsz;
      const int na = l * S;
      const int nb = (na + S < g.npos) ? na + S : g.npos;
      if (na < nb) {
        const unsigned char* tp = tc + g.i0;
        const unsigned char* qp = qc + g.i0 + g.shift;
        auto sc = [&](int n) -> int { return code_score(score, tp[n], qp[n]); };
        int W = 0;
        for (int m = (na - 46 > 0 ? na - 46 : 0); m < na; ++m) W += sc(m);
        bool prev = (na - 1 >= 46) && W > SXC_WIN_THRESHOLD;
        for (int n = na; n < nb; ++n) {
          W += sc(n);
          if (n >= 46) W -= sc(n - 46);
          const bool on = (n >= 46) && W > SXC_WIN_THRESHOLD;
          if (on && !prev) {
            open = n;
          } else if (!on && prev) {
            if (open >= 0) {
              push(c, open, n);
              open = -1;
            } else if (first_end == INF) {
              first_end = n;
            }
          }
          prev = on;
        }
      }
    }
    summ[2 * tid] = first_end;
    summ[2 * tid + 1] = open;
  });
  ex.sync();
  ex.each([&](int tid) {
    const int c = c0 + tid / gsz;
    const int l = tid % gsz;
    const int open = summ[2 * tid + 1];
    if (c >= c1 || open < 0) return;
    int end = INF;
    const int g0 = tid - l;
    for (int k = l + 1; k < gsz; ++k) {
      const int fe = summ[2 * (g0 + k)];
      if (fe < end) end = fe;
    }
    if (end == INF) end = lag_geometry(cands[c], H, Lq, Lt).npos;  // open where the loop breaks
    push(c, open, end);
  });
  ex.sync();
}

// Drains queue entries [0, n): mode 0 filters into t_results, mode 1 writes runs.
// rank_base: the first lag of the epoch the entries come from (compact queue).
template <int NB, class Ex>
SXC_HD inline void drain(Ex& ex, char* sm, const ScanArgs& a, const sxc_pair_dev& pr, const uint16_t* cands, int ps,
                         int n, int rank_base) {
  using SS = ScanSmem<NB>;
  constexpr int H = SS::N / 2;
  const unsigned char* tc = (const unsigned char*)(sm + SS::TCODE);
  const unsigned char* qc = (const unsigned char*)(sm + SS::QCODE);
  const double* equal = (const double*)(sm + SS::EQUAL);
  const double* gcv = (const double*)(sm + SS::GC);
  int* misc = (int*)(sm + SS::MISC);
  const int* queue = (const int*)(sm + SS::QUEUE);
  const int strand = ps % a.strands;
  ex.each([&](int tid) {
    for (int k = tid; k < n; k += SCAN_NT) {
      int rank, ns, ne;
      if constexpr (SS::COMPACT) {
        const unsigned v = (unsigned)queue[k];
        rank = rank_base + (int)(v & ((1u << SS::RB) - 1));
        ns = (int)((v >> SS::RB) & ((1u << SS::NSB) - 1));
        ne = (int)(v >> (SS::RB + SS::NSB));
      } else {
        rank = queue[2 * k];
        ns = queue[2 * k + 1] & 0xffff;
        ne = (queue[2 * k + 1] >> 16) & 0xffff;
      }
      const int shift = (int)cands[rank] - H;
      const int i0 = shift < 0 ? -shift : 0;
      const int start_t = i0 + ns - SXC_MIN_LEN;
      const int len = ne - ns + SXC_MIN_LEN;
      const int start_q = start_t + shift;
      const uint64_t key = ((uint64_t)(uint32_t)rank << 32) | (uint32_t)ns;
      if (a.mode == 1) {
        const int slot = a.split > 1 ? ex.atomic_add_global(&a.counts[ps], 1) : ex.atomic_add_shared(&misc[SM_KEPT], 1);
        if (slot >= a.cap) continue;
        const size_t at = (size_t)ps * a.cap + slot;
        a.runs[3 * at] = start_t;
        a.runs[3 * at + 1] = start_q;
        a.runs[3 * at + 2] = len;
        a.keys[at] = key;
        continue;
      }
      if (len < a.min_len) continue;  // FilterMatches: GetLength() < m_minLen
      double ident, prob = 0.;
      if constexpr (SS::HASN) {
        int cm, nk, ct, ntn, cq, nqn;
        packed_counts_n((const uint32_t*)(sm + SS::TPK), (const uint32_t*)(sm + SS::QPK),
                        (const uint32_t*)(sm + SS::TNK), (const uint32_t*)(sm + SS::QNK), start_t, start_q, len, cm,
                        nk, ct, ntn, cq, nqn);
        // DNA_Equal of N against any base or N, and N's GC weight, are quarter multiples, so
        // the reference's running sums equal these products exactly
        const double matches = (double)cm + a.tab->equal[SXC_CLASS_N][0] * (double)nk;
        const double gct = (double)ct + a.tab->gc[SXC_CLASS_N] * (double)ntn;
        const double gcq = (double)cq + a.tab->gc[SXC_CLASS_N] * (double)nqn;
        if (a.ptable) {
          prob = table_from_sums(matches, gct, gcq, len, a.ptable, a.table_prob, ident);
        } else if ((nk == 0 && z_clearly_skipped(cm, ct, cq, len, a.z_skip)) ||
                   !tail_from_sums(matches, gct, gcq, len, a.target_total, a.z_skip, prob, ident)) {
          continue;
        }
      } else if constexpr (SS::PURE) {
        int cm, ct, cq;
        packed_counts((const uint32_t*)(sm + SS::TPK), (const uint32_t*)(sm + SS::QPK), start_t, start_q, len, cm,
                      ct, cq);
        if (a.ptable) {
          prob = table_from_sums((double)cm, (double)ct, (double)cq, len, a.ptable, a.table_prob, ident);
        } else if (z_clearly_skipped(cm, ct, cq, len, a.z_skip) ||
                   !tail_from_sums((double)cm, (double)ct, (double)cq, len, a.target_total, a.z_skip, prob, ident)) {
          continue;
        }
      } else {
        double matches = 0, gct = 0, gcq = 0;
        for (int i = 0; i < len; ++i) {
          const int ca = code_class(tc[start_t + i]), cb = code_class(qc[start_q + i]);
          matches += equal[ca * SXC_NCLASS + cb];
          gct += gcv[ca];
          gcq += gcv[cb];
        }
        if (a.ptable) {
          prob = table_from_sums(matches, gct, gcq, len, a.ptable, a.table_prob, ident);
        } else if (!tail_from_sums(matches, gct, gcq, len, a.target_total, a.z_skip, prob, ident)) {
          continue;
        }
      }
      // FilterMatches keeps prob >= m_minProb.  The table's decision is exact; the device's
      // erf/exp may differ from glibc's by a few ulp, so a probability within the keep band
      // of min_prob is written flagged and decided on the host with glibc.
      bool near = false;
      if (a.ptable) {
        if (prob < a.min_prob) continue;
      } else {
        if (prob < a.keep_lo) continue;
        near = prob < a.flag_hi;
      }
      const int slot = a.split > 1 ? ex.atomic_add_global(&a.counts[ps], 1) : ex.atomic_add_shared(&misc[SM_KEPT], 1);
      if (slot >= a.cap) continue;
      const int tstart = pr.t_chunk_start + start_t;
      // RCQuery (analysis/HomologyByXCorrSlave.cc:56-60) on the reverse strand, with the
      // -q_chunk size (slave) or, for q_chunk == 0, the query chunk's own length (the
      // standalone aligner, tools/analysis/HomologyByXCorr.cc:226-228, 793-798)
      const int rc_chunk = a.q_chunk > 0 ? a.q_chunk : a.qchunks[pr.q_chunk].len;
      const int qstart =
          strand == 0 ? pr.q_chunk_start + start_q : start_q + pr.q_record_len - pr.q_chunk_start - rc_chunk;
      sxc_result_dev r;
      r.query_id = (uint64_t)(int64_t)pr.q_record;
      r.target_id = (uint64_t)(int64_t)pr.t_record;
      r.query_size = (uint64_t)(int64_t)pr.q_record_len;
      r.qstart = (uint64_t)(int64_t)qstart;
      r.tstart = (uint64_t)(int64_t)tstart;
      r.len = (uint64_t)(int64_t)len;
      r.reverse = (uint8_t)(strand == 1);
      for (int z = 0; z < 7; ++z) r.pad_[z] = 0;
      if (near) {
        const uint32_t pi = (uint32_t)(ps / a.strands);  // the pair, for the host's decision
        r.pad_[0] = SXC_REC_NEAR;
        r.pad_[1] = (uint8_t)pi;
        r.pad_[2] = (uint8_t)(pi >> 8);
        r.pad_[3] = (uint8_t)(pi >> 16);
        r.pad_[4] = (uint8_t)(pi >> 24);
        if (a.nflag) ex.atomic_add_global(a.nflag, 1);
      }
      r.prob = prob;
      r.ident = ident;
      const size_t at = (size_t)ps * a.cap + slot;
      a.out[at] = r;
      a.keys[at] = key;
    }
  });
  ex.sync();
}

template <int NB, class Ex>
SXC_HD inline void scan_body(Ex& ex, char* sm, const ScanArgs& a, int blk) {
  using SS = ScanSmem<NB>;
  constexpr int N = SS::N;
  const int split = a.split > 1 ? a.split : 1;
  const int ps = blk / split, part = blk % split;
  const sxc_pair_dev pr = a.pairs[ps / a.strands];
  const int strand = ps % a.strands;
  const sxc_chunk tch = a.tchunks[pr.t_chunk];
  const sxc_chunk qch = a.qchunks[pr.q_chunk];
  const bool fast = (tch.flags & 2) && (qch.flags & 2);   // both pure A/C/G/T
  const bool acgtn = (tch.flags & 4) && (qch.flags & 4);  // both A/C/G/T/N
  if (fast != (SS::PURE && !SS::HASN) || (!fast && acgtn) != SS::HASN) return;  // another kernel's pair
  const sxc_tables* tab = a.tab;
  unsigned char* tc = (unsigned char*)(sm + SS::TCODE);
  unsigned char* qc = (unsigned char*)(sm + SS::QCODE);
  double* equal = (double*)(sm + SS::EQUAL);
  double* gcv = (double*)(sm + SS::GC);
  unsigned char* score = (unsigned char*)(sm + SS::SCORE);
  int* misc = (int*)(sm + SS::MISC);
  const unsigned char* ts = a.seq + tch.off;
  const unsigned char* qs = a.seq + qch.off;
  const uint16_t* gcands = a.cands + (size_t)ps * N;
  unsigned char* cls = (unsigned char*)(sm + SS::CLS);
  unsigned char* rc = (unsigned char*)(sm + SS::RC);
  // the general kernel stages the chunk bytes (16-byte aligned in HBM) and the tables in LDS
  ex.each([&](int tid) {
    if constexpr (!SS::PURE) {
      const uint32_t* ts4 = (const uint32_t*)ts;
      const uint32_t* qs4 = (const uint32_t*)qs;
      uint32_t* tc4 = (uint32_t*)tc;
      uint32_t* qc4 = (uint32_t*)qc;
      for (int i = tid; i < (tch.len + 3) / 4; i += SCAN_NT) tc4[i] = ts4[i];
      for (int i = tid; i < (qch.len + 3) / 4; i += SCAN_NT) qc4[i] = qs4[i];
      for (int i = tid; i < SXC_NCLASS * SXC_NCLASS; i += SCAN_NT) {
        equal[i] = (&tab->equal[0][0])[i];
        score[i] = (&tab->score[0][0])[i];
      }
      for (int i = tid; i < SXC_NCLASS; i += SCAN_NT) gcv[i] = tab->gc[i];
      for (int i = tid; i < 128; i += SCAN_NT) {
        cls[i] = tab->cls[i];
        rc[i] = tab->rc[i];
      }
    }
    if (tid == 0) {
      misc[SM_QN] = 0;
      misc[SM_NRB] = 0;
      misc[SM_KEPT] = 0;
      misc[SM_RUNS] = 0;
      misc[SM_NCAND] = a.ncand[ps];
      misc[SM_LT] = tch.len;
      misc[SM_LQ] = qch.len;
    }
  });
  ex.sync();
  const int Lt = tch.len, Lq = qch.len;
  if constexpr (SS::PURE) {
    // 2-bit codes A=0, C=1, G=2, T=3 from the ASCII letters (read from HBM); complement =
    // code ^ 3
    uint32_t* tpk = (uint32_t*)(sm + SS::TPK);
    uint32_t* qpk = (uint32_t*)(sm + SS::QPK);
    ex.each([&](int tid) {
      for (int w = tid; w < SS::PKW; w += SCAN_NT) {
        uint32_t vt = 0, vq = 0, nt = 0, nq = 0;
        for (int b = 0; b < 16; ++b) {
          const int i = 16 * w + b;
          if (i < Lt) {
            const unsigned x = ts[i];
            vt |= (((x >> 1) ^ (x >> 2)) & 3u) << (2 * b);
            nt |= (x == 'N' ? 1u : 0u) << (2 * b);
          }
          if (i < Lq) {
            // DNAVector::ReverseComplement for strand 1 (analysis/DNAVector.cc:482-521);
            // the complement of N is N
            const unsigned x = strand ? qs[Lq - 1 - i] : qs[i];
            vq |= ((((x >> 1) ^ (x >> 2)) & 3u) ^ (strand ? 3u : 0u)) << (2 * b);
            nq |= (x == 'N' ? 1u : 0u) << (2 * b);
          }
        }
        tpk[w] = vt;
        qpk[w] = vq;
        if constexpr (SS::HASN) {
          ((uint32_t*)(sm + SS::TNK))[w] = nt;
          ((uint32_t*)(sm + SS::QNK))[w] = nq;
        }
      }
    });
    ex.sync();
  } else {
    // base codes in place; the reverse complement swaps symmetric positions
    ex.each([&](int tid) {
      for (int i = tid; i < Lt; i += SCAN_NT) {
        const unsigned char c = cls[tc[i] & 127];
        tc[i] = c < SXC_CLASS_ZERO ? c : (unsigned char)(128 | tc[i]);
      }
      if (strand) {
        for (int i = tid; i < (Lq + 1) / 2; i += SCAN_NT) {
          const unsigned char x = qc[i], y = qc[Lq - 1 - i];
          const unsigned char rx = rc[x & 127], ry = rc[y & 127];
          const unsigned char cx = cls[rx & 127], cy = cls[ry & 127];
          qc[i] = cy < SXC_CLASS_ZERO ? cy : (unsigned char)(128 | ry);
          qc[Lq - 1 - i] = cx < SXC_CLASS_ZERO ? cx : (unsigned char)(128 | rx);
        }
      } else {
        for (int i = tid; i < Lq; i += SCAN_NT) {
          const unsigned char c = cls[qc[i] & 127];
          qc[i] = c < SXC_CLASS_ZERO ? c : (unsigned char)(128 | qc[i]);
        }
      }
    });
    ex.sync();
  }
  const uint16_t* cands = gcands;
  const int ncand = misc[SM_NCAND];
  // this workgroup's share of the lags (all of them unless the scan is split)
  const int lo = (int)((int64_t)ncand * part / split), hi = (int)((int64_t)ncand * (part + 1) / split);
  constexpr int CE = SS::CE;
  int runs = 0;
  // Queue entries before the current epoch, kept per thread: every thread must see the same
  // count, and the shared counter may already hold the next epoch's pushes of a faster wave.
  int queued = 0;
  for (int e0 = lo; e0 < hi; e0 += CE) {
    const int qstart = queued;
    const int e1 = (e0 + CE < hi) ? e0 + CE : hi;
    if constexpr (SS::BMRUNS) {
      scan_epoch_bitmap<NB>(ex, sm, cands, e0, e1, SS::GSZ, [&](int n) {
        drain<NB>(ex, sm, a, pr, cands, ps, n, e0);
        runs += n;
        ex.each([&](int tid) {
          if (tid == 0) misc[SM_QN] = 0;
        });
        ex.sync();
      });
      continue;
    }
    scan_epoch_segments<NB>(ex, sm, cands, e0, e1, SS::GSZ);
    const int qn = misc[SM_QN];
    ex.sync();  // every thread has read the count before any thread pushes again
    queued = qn;
    if (qn > SS::QCAP) {
      // the epoch overflowed the queue: keep what came before it, redo its lags one by one
      drain<NB>(ex, sm, a, pr, cands, ps, qstart, e0);  // (compact queue: empty)
      runs += qstart;
      for (int c = e0; c < e1; ++c) {
        ex.each([&](int tid) {
          if (tid == 0) misc[SM_QN] = 0;
        });
        ex.sync();
        scan_epoch_segments<NB>(ex, sm, cands, c, c + 1, SCAN_NT);
        const int n1 = misc[SM_QN];
        drain<NB>(ex, sm, a, pr, cands, ps, n1, c);
        runs += n1;
      }
      ex.each([&](int tid) {
        if (tid == 0) misc[SM_QN] = 0;
      });
      ex.sync();
      queued = 0;
    } else if (SS::COMPACT || qn > SS::QCAP / 2 || e1 == hi) {
      drain<NB>(ex, sm, a, pr, cands, ps, qn, e0);
      runs += qn;
      ex.each([&](int tid) {
        if (tid == 0) misc[SM_QN] = 0;
      });
      ex.sync();
      queued = 0;
    }
  }
  ex.each([&](int tid) {
    if (tid == 0) {
      if (split == 1) {
        a.counts[ps] = misc[SM_KEPT];
        if (a.stats) {
          a.stats[2 * ps] = ncand;
          a.stats[2 * ps + 1] = runs;
        }
      } else if (a.stats) {
        if (part == 0) a.stats[2 * ps] = ncand;
        ex.atomic_add_global(&a.stats[2 * ps + 1], runs);
      }
    }
  });
}

}  // namespace sxc
