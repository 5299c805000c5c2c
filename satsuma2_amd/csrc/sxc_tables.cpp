// sxc_tables.cpp -- host construction of the constant tables the kernels read.
//
// Every value is computed with the same double expression the reference evaluates
// (glibc libm on the host), so the device only ever looks values up:
//   * FFTReal twiddles: (float)cos(i * (PI / 4096)) for the level-12 LUT and its strides as
//     the lower levels' LUTs
//     (extern/RealFFT/FFTReal.hpp:352-374) and the float oscillator sequence for levels
//     13, 14 (OscSinCos.hpp:61-104, reset per block, FFTReal.hpp:626, 639-641);
//   * DNACodec rows, reverse complement and DNA_Equal (analysis/DNAVector.cc:13-57, 390-403);
//   * LookupMatch scores (int)(100 * DNA_EqualAmb + 0.5) (analysis/CrossCorr.cc:547-553);
//   * Ent(a) / window entropies for windows whose codec sums are multiples of 1/4
//     (analysis/CrossCorr.cc:28-33, 74-84).
#include "sxc_tables.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <thread>
#include <vector>

namespace sxc {

static const char kLetters[] = "ACGTKMRYSWBVHDNX-";  // class ids 0..16; 17 = zero row

static void codec_row(char l, double* v, char* rc) {
  const double third = 1. / 3.;
  struct Row {
    char l;
    double a, c, g, t;
    char rc;
  };
  static const Row rows[] = {
      {'A', 1., 0., 0., 0., 'T'},       {'C', 0., 1., 0., 0., 'G'},       {'G', 0., 0., 1., 0., 'C'},
      {'T', 0., 0., 0., 1., 'A'},       {'K', 0., 0., 0.5, 0.5, 'M'},     {'M', 0.5, 0.5, 0., 0., 'K'},
      {'R', 0.5, 0., 0.5, 0., 'Y'},     {'Y', 0., 0.5, 0., 0.5, 'R'},     {'S', 0., 0.5, 0.5, 0., 'S'},
      {'W', 0.5, 0., 0., 0.5, 'W'},     {'B', 0., third, third, third, 'V'}, {'V', third, third, third, 0., 'B'},
      {'H', third, third, 0., third, 'D'}, {'D', third, 0., third, third, 'H'}, {'-', 0., 0., 0., 0., '-'},
      {'N', 0.25, 0.25, 0.25, 0.25, 'N'}, {'X', 0.25, 0.25, 0.25, 0.25, 'X'},
  };
  v[0] = v[1] = v[2] = v[3] = 0.;
  *rc = 0;
  for (const Row& r : rows)
    if (r.l == l) {
      v[0] = r.a;
      v[1] = r.c;
      v[2] = r.g;
      v[3] = r.t;
      *rc = r.rc;
    }
}

int class_of(int byte) {
  for (int c = 0; c < 17; ++c)
    if (kLetters[c] == (char)byte) return c;
  return SXC_CLASS_ZERO;
}

bool class_has_thirds(int c) { return c == 10 || c == 11 || c == 12 || c == 13; }

int chunk_flags(const unsigned char* s, int len) {
  // per byte: bit 0 a 1/3-valued letter, bit 1 not A/C/G/T, bit 2 neither A/C/G/T nor N,
  // bit 3 not a valid byte (>= 128)
  static const auto kind = [] {
    std::array<unsigned char, 256> k{};
    for (int b = 0; b < 256; ++b) {
      if (b >= 128) {
        k[b] = 8;
        continue;
      }
      const int cl = class_of(b);
      k[b] = (unsigned char)((class_has_thirds(cl) ? 1 : 0) | (cl > 3 ? 2 : 0) | (cl > 3 && cl != SXC_CLASS_N ? 4 : 0));
    }
    return k;
  }();
  unsigned any = 0;
  for (int i = 0; i < len; ++i) any |= kind[s[i]];
  if (any & 8) return -1;
  return (int)((any & 1) | (~any & 6));
}

void build_tables(sxc_tables* t) {
  memset(t, 0, sizeof *t);
  const double PI = 3.1415926535897932384626433832795;  // extern/RealFFT/def.h:42
  const double mul = PI / (double)(2048 << 1);
  float* lut12 = t->lut + SXC_LUT_OFF(12);
  for (int i = 0; i < 2048; ++i) lut12[i] = (float)cos((double)i * mul);
  // FFTReal's level-L table holds (float)cos(i * PI / 2^L), computed from the same double
  // argument as level 12's entry i << (12 - L): the same floats
  for (int lv = 3; lv < 12; ++lv)
    for (int i = 0; i < (1 << (lv - 1)); ++i) t->lut[SXC_LUT_OFF(lv) + i] = lut12[i << (12 - lv)];
  double v[4];
  char rc;
  for (int c = 0; c < SXC_NCLASS; ++c) {
    if (c < 17) {
      codec_row(kLetters[c], v, &rc);
    } else {
      v[0] = v[1] = v[2] = v[3] = 0.;
    }
    for (int k = 0; k < 4; ++k) {
      t->codec[c][k] = v[k];
      const double q = v[k] * 4.0;
      t->quarters[c][k] = (q == (double)(int)q) ? (unsigned char)(int)q : 0;
    }
    t->gc[c] = v[1] + v[2];
  }
  for (int a = 0; a < SXC_NCLASS; ++a)
    for (int b = 0; b < SXC_NCLASS; ++b) {
      // DNA_Equal: ((a+c)+g)+t, + 0.0
      const double s = t->codec[a][0] * t->codec[b][0] + t->codec[a][1] * t->codec[b][1] +
                       t->codec[a][2] * t->codec[b][2] + t->codec[a][3] * t->codec[b][3];
      t->equal[a][b] = s + 0.;
      t->score[a][b] = (unsigned char)(int)(t->equal[a][b] * (double)100 + 0.5);
    }
  for (int byte = 0; byte < 128; ++byte) {
    t->cls[byte] = (unsigned char)class_of(byte);
    codec_row((char)byte, v, &rc);
    t->rc[byte] = (unsigned char)rc;
  }
  for (int byte = 0; byte < 128; ++byte) {
    t->rcls[byte] = t->cls[t->rc[byte] & 127];
    for (int strand = 0; strand < 2; ++strand) {
      const unsigned char* q = t->quarters[strand ? t->rcls[byte] : t->cls[byte]];
      t->qpack[strand][byte] = (uint32_t)q[0] | ((uint32_t)q[1] << 8) | ((uint32_t)q[2] << 16) | ((uint32_t)q[3] << 24);
    }
  }
}

void build_osc(float* osc) {
  // level 13 at [0, 8192): 4096 {cos, sin}; level 14 at [8192, 24576): 8192 pairs
  const double PI = 3.1415926535897932384626433832795;
  for (int lv = 13; lv <= 14; ++lv) {
    const int len = 1 << (lv - 1);  // 1 << (TRIGO_BD_LIMIT + osc index)
    const double mul = (0.5 * PI) / (double)len;
    const float sc = (float)cos(mul), ss = (float)sin(mul);
    float* out = osc + (lv == 13 ? 0 : 8192);
    float pc = 1.f, ps = 0.f;
    out[0] = pc;
    out[1] = ps;
    for (int i = 1; i < len; ++i) {
      const float oc = pc, os = ps;
      pc = oc * sc - os * ss;
      ps = oc * ss + os * sc;
      out[2 * i] = pc;
      out[2 * i + 1] = ps;
    }
  }
}

static double ent_ref(double p) {
  if (p < 0.001) return 0;
  return p * log(p) / 0.69314718056;
}

int ent_lut_len(int nb) {
  const int win = (1 << nb) / 512;
  return win < 1 ? 0 : win * (4 * win + 1);
}

void build_ent_lut(int nb, double* out) {
  const int win = (1 << nb) / 512;
  for (int k = 1; k <= win; ++k)
    for (int q = 0; q <= 4 * win; ++q) {
      const double sum = (double)q * 0.25;
      out[(size_t)(k - 1) * (4 * win + 1) + q] = ent_ref(sum / (double)k);
    }
}

double keep_z_bound(double min_prob, double target_size) {
  // keep <=> exp(-0.5 * (1 + erf(z)) * T) >= min_prob <=> cdf(z) <= -ln(min_prob) / T.
  // Return z_skip with cdf(z_skip) >= 4 * that bound: every z >= z_skip is dropped.
  if (!(min_prob > 0.) || !(target_size > 0.)) return HUGE_VAL;
  const double cmax = (-log(min_prob) + 1e-15) / target_size;
  if (4. * cmax >= 1.) return HUGE_VAL;
  double lo = -40., hi = 40.;  // cdf(lo) < 4 cmax <= cdf(hi)
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (0.5 * (1. + erf(mid)) >= 4. * cmax) hi = mid; else lo = mid;
  }
  return hi;
}

// ProbTable::Setup (analysis/ProbTable.cc:15-52) with GetMatchProbabilityRaw (:143-163):
// for expected identity i / 511 and length j, the identity where exp(-cdf * size) stops
// being 0, by bisection to a width of 1e-8 (the reference tests the probability for non-zero,
// so the cutoff does not enter the thresholds).  Computed with the host's erf/exp, as the
// reference does; rows are spread over the host's threads.
void build_prob_table(double target_size, double* t) {
  memset(t, 0, sizeof(double) * SXC_PT_ROWS * SXC_PT_LEN);
  std::atomic<int> next(1);
  auto work = [&]() {
    for (;;) {
      const int i = next.fetch_add(1);
      if (i >= SXC_PT_ROWS) return;
      double* row = t + (size_t)i * SXC_PT_LEN;
      const double p = (double)i / ((double)SXC_PT_ROWS - 1);
      row[0] = 2.;
      for (int j = 1; j < SXC_PT_LEN; ++j) {
        double lo = 0, hi = 1;
        while (hi - lo > 0.00000001) {
          const double mid = (hi + lo) / 2.0;
          const double s = sqrt(p * (1. - p) * (double)j);
          const double m = p * (double)j;
          const double x = (double)j * mid;
          const double cdf = 0.5 * (1. + erf((m - x) / s / 1.414213562));
          if (exp(-(cdf * target_size)) != 0.) hi = mid; else lo = mid;
        }
        row[j] = (hi + lo) / 2.0;
      }
    }
  };
  const int nt = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (int k = 1; k < nt; ++k) th.emplace_back(work);
  work();
  for (auto& x : th) x.join();
}

const sxc_tables& host_tables() {
  static const sxc_tables t = [] {
    sxc_tables x;
    build_tables(&x);
    return x;
  }();
  return t;
}

void window_entropy_row(const unsigned char* seq, int L, int N, float* row) {
  // CCSignal::ComputeEntropy (analysis/CrossCorr.cc:35-93) for both strands of one chunk:
  // sequential double sums of the codec values per window, Ent with the host's log
  const sxc_tables& t = host_tables();
  const int win = N / 512;
  for (int strand = 0; strand < 2; ++strand) {
    float* out = row + strand * (SXC_ENT_ROW / 2);
    for (int w = 0; w < SXC_ENT_ROW / 2; ++w) out[w] = 0.f;
    if (L < 1024 || win < 1) continue;  // weight 1 everywhere (:39-44): the row is not read
    for (int i = 0, w = 0; i < L && w < SXC_ENT_ROW / 2; i += win, ++w) {
      double a = 0, c = 0, g = 0, u = 0;
      int k = 0;
      for (int j = i; j < i + win && j < L; ++j) {
        const unsigned char b = strand ? t.rc[seq[L - 1 - j] & 127] : (unsigned char)(seq[j] & 127);
        const double* cd = t.codec[t.cls[b]];
        a += cd[0];
        c += cd[1];
        g += cd[2];
        u += cd[3];
        ++k;
      }
      a /= (double)k;
      c /= (double)k;
      g /= (double)k;
      u /= (double)k;
      const double e = ent_ref(a) + ent_ref(c) + ent_ref(g) + ent_ref(u);
      float v = (float)(-e);
      if (v < 0.f) v = 0.f;
      out[w] = v;
    }
  }
}

void keep_band(double min_prob, double target_size, double* keep_lo, double* flag_hi) {
  // prob = exp(-T * cdf), cdf = 0.5 * (1 + erf(z)).  The device and the host evaluate z with
  // the same correctly rounded IEEE operations; only erf and exp differ, each by a few ulp of
  // its result: |erf_dev - erf_host| <= k * 2^-53 (|erf| < 1) moves cdf by half that and
  // prob by prob * T * k * 2^-54; exp adds k * 2^-53 of prob.  With k = 16 on each side the
  // band below covers both with a factor of 2 to spare.
  if (!(min_prob > 0.) || !(target_size >= 0.)) {
    *keep_lo = *flag_hi = min_prob;
    return;
  }
  const double rel = (16. * target_size + 64.) * ldexp(1., -52);
  *keep_lo = min_prob * (1. - rel);
  *flag_hi = min_prob * (1. + rel);
}

double host_match_probability(const unsigned char* t, const unsigned char* q, int qlen, int rc, int start_t,
                              int start_q, int len, double target_size, double* ident) {
  // GetMatchProbabilityEx (analysis/AlignProbability.cc:62-127) with GCAdjustExpect, Sigma
  // and CDF (:11-40): sums in base order, then the host's sqrt, erf and exp
  const sxc_tables& tb = host_tables();
  double matches = 0, gct = 0, gcq = 0;
  for (int i = 0; i < len; ++i) {
    const int j = start_q + i;
    const unsigned char qb = rc ? tb.rc[q[qlen - 1 - j] & 127] : (unsigned char)(q[j] & 127);
    const int ca = tb.cls[t[start_t + i] & 127], cb = tb.cls[qb];
    matches += tb.equal[ca][cb];
    gct += tb.gc[ca];
    gcq += tb.gc[cb];
  }
  *ident = matches / (double)len;
  const double gc_target = gct / (double)len;
  const double at_target = 1. - gc_target;
  double r = gcq * gc_target;
  r += ((double)len - gcq) * at_target;
  const double p = r / (double)len / 2.;
  const double s = sqrt(p * (1. - p) * (double)len);
  const double m = p * (double)len;
  const double x = (double)len * *ident;
  const double cdf = 0.5 * (1. + erf((m - x) / s / 1.414213562));
  const double expect = cdf * target_size;
  return exp(-expect);
}

int resolve_record(sxc_result_dev& r, const ResolveCtx& c) {
  if (r.pad_[0] != SXC_REC_NEAR) return 1;
  uint32_t pi = 0;
  memcpy(&pi, &r.pad_[1], 4);
  if ((int64_t)pi >= c.npairs) return -1;
  const sxc_pair_dev& pr = c.pairs[pi];
  const sxc_chunk& tc = c.tch[pr.t_chunk];
  const sxc_chunk& qc = c.qch[pr.q_chunk];
  const int64_t len = (int64_t)r.len;
  const int64_t st = (int64_t)r.tstart - pr.t_chunk_start;
  // invert RCQuery (analysis/HomologyByXCorrSlave.cc:56-60) on the reverse strand
  const int64_t rc_chunk = c.q_chunk > 0 ? c.q_chunk : qc.len;
  const int64_t sq = r.reverse ? (int64_t)r.qstart - pr.q_record_len + pr.q_chunk_start + rc_chunk
                               : (int64_t)r.qstart - pr.q_chunk_start;
  if (len <= 0 || st < 0 || st + len > tc.len || sq < 0 || sq + len > qc.len) return -1;
  double ident = 0.;
  const unsigned char* tb = c.tbytes ? c.tbytes : c.seq + tc.off;
  const unsigned char* qb = c.qbytes ? c.qbytes : c.seq + qc.off;
  const double prob = host_match_probability(tb, qb, qc.len, r.reverse ? 1 : 0, (int)st, (int)sq, (int)len,
                                             c.target_total, &ident);
  memset(r.pad_, 0, sizeof r.pad_);
  if (prob < c.min_prob) return 0;  // FilterMatches: prob < m_minProb
  r.prob = prob;
  r.ident = ident;
  return 1;
}

}  // namespace sxc
