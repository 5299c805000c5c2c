/*
 * sxc_oracle.c -- CPU restatement of Satsuma2's chunk-pair cross-correlation path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load this library; the product never does.
 *
 * Written from the reference's behaviour, function by function, keeping its IEEE
 * operation order so results are bit-identical.  Build flags: -ffp-contract=off, no
 * -march (SSE2 double/float arithmetic, no FMA), as the reference's CMakeLists.txt:4.
 *
 * Pinning (see tests/test_oracle.py): every function here is checked against
 * golden vectors produced by the reference compiled from source (oracle/_ref), and
 * additionally against closed forms / brute force where the mathematics fixes the
 * answer (naive DFT for the FFT, direct FP64 correlation for the xcorr, brute-force
 * window sums for the diagonal scan, textbook normal tail for the probability).
 */
#define _GNU_SOURCE
#include "sxc_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ codec */

/* DNACodec::DNACodec (analysis/DNAVector.cc:13-57) and DNACodec::Set (:153-160). */
static double g_A[256], g_C[256], g_G[256], g_T[256];
static unsigned char g_rc[256];
static int g_codec_ready = 0;
static pthread_once_t g_codec_once = PTHREAD_ONCE_INIT;

static void codec_set(int l, double a, double c, double g, double t, int rc) {
  g_rc[l] = (unsigned char)rc;
  g_A[l] = a;
  g_C[l] = c;
  g_G[l] = g;
  g_T[l] = t;
}

static void codec_init(void) {
  const double third = 1. / 3.; /* ONE_THIRD, analysis/DNAVector.cc:6 */
  memset(g_A, 0, sizeof g_A);
  memset(g_C, 0, sizeof g_C);
  memset(g_G, 0, sizeof g_G);
  memset(g_T, 0, sizeof g_T);
  memset(g_rc, 0, sizeof g_rc);
  codec_set('A', 1., 0., 0., 0., 'T');
  codec_set('C', 0., 1., 0., 0., 'G');
  codec_set('G', 0., 0., 1., 0., 'C');
  codec_set('T', 0., 0., 0., 1., 'A');
  codec_set('K', 0., 0., 0.5, 0.5, 'M');
  codec_set('M', 0.5, 0.5, 0., 0., 'K');
  codec_set('R', 0.5, 0., 0.5, 0., 'Y');
  codec_set('Y', 0., 0.5, 0., 0.5, 'R');
  codec_set('S', 0., 0.5, 0.5, 0., 'S');
  codec_set('W', 0.5, 0., 0., 0.5, 'W');
  codec_set('B', 0., third, third, third, 'V');
  codec_set('V', third, third, third, 0., 'B');
  codec_set('H', third, third, 0., third, 'D');
  codec_set('D', third, 0., third, third, 'H');
  codec_set('-', 0., 0., 0., 0., '-');
  codec_set('N', 0.25, 0.25, 0.25, 0.25, 'N');
  codec_set('X', 0.25, 0.25, 0.25, 0.25, 'X');
  g_codec_ready = 1;
}

static inline void codec_ready(void) {
  if (!g_codec_ready) pthread_once(&g_codec_once, codec_init);
}

static inline int ub(char c) { return (int)(unsigned char)c; }

void sxo_codec(double* a, double* c, double* g, double* t, unsigned char* rc) {
  codec_ready();
  memcpy(a, g_A, sizeof g_A);
  memcpy(c, g_C, sizeof g_C);
  memcpy(g, g_G, sizeof g_G);
  memcpy(t, g_T, sizeof g_T);
  memcpy(rc, g_rc, sizeof g_rc);
}

/* DNA_Equal (analysis/DNAVector.cc:390-403): ((a+c)+g)+t, then + 0.0. */
double sxo_dna_equal(int l1, int l2) {
  codec_ready();
  double a = g_A[l1 & 255] * g_A[l2 & 255];
  double c = g_C[l1 & 255] * g_C[l2 & 255];
  double g = g_G[l1 & 255] * g_G[l2 & 255];
  double t = g_T[l1 & 255] * g_T[l2 & 255];
  double sum = a + c + g + t;
  double plus = 0.;
  return sum + plus;
}

/* DNA_EqualAmb (analysis/DNAVector.cc:372-387): identical letters other than 'N' score 1. */
double sxo_dna_equal_amb(int l1, int l2) {
  if ((char)l1 == (char)l2 && (char)l1 != 'N') return 1.;
  return sxo_dna_equal(l1, l2);
}

/* DNAVector::ReverseComplement (analysis/DNAVector.cc:482-521). */
void sxo_reverse_complement(const char* in, int32_t len, char* out) {
  codec_ready();
  for (int32_t i = 0; i < len; ++i) out[i] = (char)g_rc[ub(in[len - 1 - i])];
}

/* ------------------------------------------------------------------ signal */

/* Ent (analysis/CrossCorr.cc:28-33). */
static double ent(double p) {
  if (p < 0.001) return 0;
  return p * log(p) / 0.69314718056;
}

static double dna_x(int ch, int l) {
  switch (ch) {
    case 0: return g_A[l];
    case 1: return g_C[l];
    case 2: return g_G[l];
    default: return g_T[l];
  }
}

/* CCSignal::SetSequence on a fresh signal (analysis/CrossCorr.cc:179-206):
 * ComputeEntropy (:35-93) then SeqToPCM per channel A,C,G,T (:97-134). */
int sxo_encode(const char* seq, int32_t len, int32_t N, float* out) {
  codec_ready();
  if (N <= 0 || len < 0 || len > N) return -1;
  float* ent_v = (float*)malloc(sizeof(float) * (size_t)N);
  if (!ent_v) return -2;
  for (int32_t i = 0; i < N; ++i) ent_v[i] = 0.f;
  if (len < 1024) {
    for (int32_t i = 0; i < N; ++i) ent_v[i] = 1.f;
  } else {
    int32_t win = N / 512;
    if (win <= 0) {
      free(ent_v);
      return -1;
    }
    for (int32_t i = 0; i < len; i += win) {
      double a = 0, c = 0, g = 0, t = 0;
      int32_t k = 0;
      for (int32_t j = i; j < i + win; ++j) {
        if (j >= len) break;
        k++;
        int l = ub(seq[j]);
        a += g_A[l];
        c += g_C[l];
        g += g_G[l];
        t += g_T[l];
      }
      a /= (double)k;
      c /= (double)k;
      g /= (double)k;
      t /= (double)k;
      double s = ent(a) + ent(c) + ent(g) + ent(t);
      for (int32_t j = i; j < i + k; ++j) {
        ent_v[j] = (float)(-s);
        if (ent_v[j] < 0.f) ent_v[j] = 0.f;
      }
    }
  }
  for (int ch = 0; ch < 4; ++ch) {
    float* o = out + (size_t)ch * N;
    double sum = 0;
    for (int32_t i = 0; i < len; ++i) sum += dna_x(ch, ub(seq[i]));
    double off = sum / (double)len;
    for (int32_t i = 0; i < N; ++i) o[i] = 0.f;
    for (int32_t i = 0; i < len; ++i) o[i] = (float)((double)ent_v[i] * (dna_x(ch, ub(seq[i])) - off));
  }
  free(ent_v);
  return 0;
}

/* ------------------------------------------------------------------ FFTReal */

#define SXO_PI 3.1415926535897932384626433832795      /* extern/RealFFT/def.h:42 */
#define SXO_SQRT2 1.41421356237309514547462185873883  /* extern/RealFFT/def.h:43 */
#define SXO_TRIGO_BD_LIMIT 12                         /* extern/RealFFT/FFTReal.h:78 */

typedef struct {
  int nbits;
  int32_t N;
  int32_t* br;     /* bit-reversal table (FFTReal.hpp:329-347) */
  float* lut;      /* level L at lut + (1<<(L-1)) - 4, (FFTReal.hpp:352-374) */
  float* osc_step; /* per oscillator level: {cos, sin} of the step (OscSinCos.hpp:61-67) */
} fft_plan;

static int ilog2(int32_t n) {
  int b = 0;
  while ((1 << b) < n) ++b;
  return b;
}

static int plan_init(fft_plan* p, int32_t N) {
  memset(p, 0, sizeof *p);
  if (N < 8 || (N & (N - 1))) return -1;
  p->N = N;
  p->nbits = ilog2(N);
  p->br = (int32_t*)malloc(sizeof(int32_t) * (size_t)N);
  for (int32_t i = 0; i < N; ++i) {
    int32_t r = 0;
    for (int b = 0; b < p->nbits; ++b)
      if (i & (1 << b)) r |= 1 << (p->nbits - 1 - b);
    p->br[i] = r;
  }
  if (p->nbits > 3) {
    int32_t total = (1 << (p->nbits - 1)) - 4;
    p->lut = (float*)malloc(sizeof(float) * (size_t)total);
    for (int level = 3; level < p->nbits; ++level) {
      int32_t level_len = 1 << (level - 1);
      float* lp = p->lut + ((1 << (level - 1)) - 4);
      const double mul = SXO_PI / (double)(level_len << 1);
      for (int32_t i = 0; i < level_len; ++i) lp[i] = (float)cos((double)i * mul);
    }
  }
  int nosc = p->nbits - SXO_TRIGO_BD_LIMIT;
  if (nosc > 0) {
    p->osc_step = (float*)malloc(sizeof(float) * 2 * (size_t)nosc);
    for (int o = 0; o < nosc; ++o) {
      int32_t len = 1 << (SXO_TRIGO_BD_LIMIT + o);
      const double mul = (0.5 * SXO_PI) / (double)len;
      p->osc_step[2 * o] = (float)cos(mul);
      p->osc_step[2 * o + 1] = (float)sin(mul);
    }
  }
  return 0;
}

static void plan_free(fft_plan* p) {
  free(p->br);
  free(p->lut);
  free(p->osc_step);
}

static const float* level_lut(const fft_plan* p, int level) { return p->lut + ((1 << (level - 1)) - 4); }

/* compute_direct_pass_1_2 (FFTReal.hpp:466-495). */
static void fwd_pass12(const fft_plan* p, float* df, const float* x) {
  for (int32_t ci = 0; ci < p->N; ci += 4) {
    int32_t r0 = p->br[ci], r1 = p->br[ci + 1], r2 = p->br[ci + 2], r3 = p->br[ci + 3];
    float* d = df + ci;
    d[1] = x[r0] - x[r1];
    d[3] = x[r2] - x[r3];
    float s0 = x[r0] + x[r1];
    float s2 = x[r2] + x[r3];
    d[0] = s0 + s2;
    d[2] = s0 - s2;
  }
}

/* compute_direct_pass_3 (FFTReal.hpp:499-528). */
static void fwd_pass3(const fft_plan* p, float* df, const float* sf) {
  const float sq = (float)(SXO_SQRT2 * 0.5);
  for (int32_t ci = 0; ci < p->N; ci += 8) {
    float v;
    df[ci] = sf[ci] + sf[ci + 4];
    df[ci + 4] = sf[ci] - sf[ci + 4];
    df[ci + 2] = sf[ci + 2];
    df[ci + 6] = sf[ci + 6];
    v = (sf[ci + 5] - sf[ci + 7]) * sq;
    df[ci + 1] = sf[ci + 1] + v;
    df[ci + 3] = sf[ci + 1] - v;
    v = (sf[ci + 5] + sf[ci + 7]) * sq;
    df[ci + 5] = v + sf[ci + 3];
    df[ci + 7] = v - sf[ci + 3];
  }
}

/* Twiddle pair for butterfly i of a level: LUT (FFTReal.hpp:585-586) or the float
 * oscillator stepped i times from (1,0) (FFTReal.hpp:626,639-641; OscSinCos.hpp:88-104). */
typedef struct {
  float pc, ps, sc, ss;
} osc_state;

static inline void osc_step(osc_state* o) {
  const float oc = o->pc, os = o->ps;
  o->pc = oc * o->sc - os * o->ss;
  o->ps = oc * o->ss + os * o->sc;
}

/* compute_direct_pass_n_lut / _osc (FFTReal.hpp:553-656). */
static void fwd_pass_n(const fft_plan* p, float* df, const float* sf, int pass) {
  const int32_t nbr = 1 << pass, h = nbr >> 1, d_nbr = nbr << 1;
  const int use_osc = pass > SXO_TRIGO_BD_LIMIT;
  const float* cp = use_osc ? NULL : level_lut(p, pass);
  osc_state o = {1.f, 0.f, 0.f, 0.f};
  if (use_osc) {
    o.sc = p->osc_step[2 * (pass - (SXO_TRIGO_BD_LIMIT + 1))];
    o.ss = p->osc_step[2 * (pass - (SXO_TRIGO_BD_LIMIT + 1)) + 1];
  }
  for (int32_t ci = 0; ci < p->N; ci += d_nbr) {
    const float* sf1r = sf + ci;
    const float* sf2r = sf1r + nbr;
    float* dfr = df + ci;
    float* dfi = dfr + nbr;
    o.pc = 1.f;
    o.ps = 0.f;
    dfr[0] = sf1r[0] + sf2r[0];
    dfi[0] = sf1r[0] - sf2r[0];
    dfr[h] = sf1r[h];
    dfi[h] = sf2r[h];
    const float* sf1i = sf1r + h;
    const float* sf2i = sf1i + nbr;
    for (int32_t i = 1; i < h; ++i) {
      float c, s, v;
      if (use_osc) {
        osc_step(&o);
        c = o.pc;
        s = o.ps;
      } else {
        c = cp[i];
        s = cp[h - i];
      }
      v = sf2r[i] * c - sf2i[i] * s;
      dfr[i] = sf1r[i] + v;
      dfi[-i] = sf1r[i] - v;
      v = sf2r[i] * s + sf2i[i] * c;
      dfi[i] = v + sf1i[i];
      dfi[nbr - i] = v - sf1i[i];
    }
  }
}

/* compute_inverse_pass_n_lut / _osc (FFTReal.hpp:729-830). */
static void inv_pass_n(const fft_plan* p, float* df, const float* sf, int pass) {
  const int32_t nbr = 1 << pass, h = nbr >> 1, d_nbr = nbr << 1;
  const int use_osc = pass > SXO_TRIGO_BD_LIMIT;
  const float* cp = use_osc ? NULL : level_lut(p, pass);
  osc_state o = {1.f, 0.f, 0.f, 0.f};
  if (use_osc) {
    o.sc = p->osc_step[2 * (pass - (SXO_TRIGO_BD_LIMIT + 1))];
    o.ss = p->osc_step[2 * (pass - (SXO_TRIGO_BD_LIMIT + 1)) + 1];
  }
  for (int32_t ci = 0; ci < p->N; ci += d_nbr) {
    const float* sfr = sf + ci;
    const float* sfi = sfr + nbr;
    float* df1r = df + ci;
    float* df2r = df1r + nbr;
    o.pc = 1.f;
    o.ps = 0.f;
    df1r[0] = sfr[0] + sfi[0];
    df2r[0] = sfr[0] - sfi[0];
    df1r[h] = sfr[h] * 2;
    df2r[h] = sfi[h] * 2;
    float* df1i = df1r + h;
    float* df2i = df1i + nbr;
    for (int32_t i = 1; i < h; ++i) {
      df1r[i] = sfr[i] + sfi[-i];
      df1i[i] = sfi[i] - sfi[nbr - i];
      float c, s;
      if (use_osc) {
        osc_step(&o);
        c = o.pc;
        s = o.ps;
      } else {
        c = cp[i];
        s = cp[h - i];
      }
      const float vr = sfr[i] - sfi[-i];
      const float vi = sfi[i] + sfi[nbr - i];
      df2r[i] = vr * c + vi * s;
      df2i[i] = vi * c - vr * s;
    }
  }
}

/* compute_inverse_pass_3 (FFTReal.hpp:834-862). */
static void inv_pass3(const fft_plan* p, float* df, const float* sf) {
  const float sq = (float)(SXO_SQRT2 * 0.5);
  for (int32_t ci = 0; ci < p->N; ci += 8) {
    df[ci] = sf[ci] + sf[ci + 4];
    df[ci + 4] = sf[ci] - sf[ci + 4];
    df[ci + 2] = sf[ci + 2] * 2;
    df[ci + 6] = sf[ci + 6] * 2;
    df[ci + 1] = sf[ci + 1] + sf[ci + 3];
    df[ci + 3] = sf[ci + 5] - sf[ci + 7];
    const float vr = sf[ci + 1] - sf[ci + 3];
    const float vi = sf[ci + 5] + sf[ci + 7];
    df[ci + 5] = (vr + vi) * sq;
    df[ci + 7] = (vi - vr) * sq;
  }
}

/* compute_inverse_pass_1_2 (FFTReal.hpp:866-906). */
static void inv_pass12(const fft_plan* p, float* x, const float* sf) {
  for (int32_t ci = 0; ci < p->N; ci += 8) {
    const float* s2 = sf + ci;
    const int32_t* br = p->br + ci;
    {
      const float b0 = s2[0] + s2[2], b2 = s2[0] - s2[2], b1 = s2[1] * 2, b3 = s2[3] * 2;
      x[br[0]] = b0 + b1;
      x[br[1]] = b0 - b1;
      x[br[2]] = b2 + b3;
      x[br[3]] = b2 - b3;
    }
    {
      const float b0 = s2[4] + s2[6], b2 = s2[4] - s2[6], b1 = s2[5] * 2, b3 = s2[7] * 2;
      x[br[4]] = b0 + b1;
      x[br[5]] = b0 - b1;
      x[br[6]] = b2 + b3;
      x[br[7]] = b2 - b3;
    }
  }
}

/* do_fft general case (FFTReal.hpp:146-157, 428-462); buffer identity does not affect values. */
static void plan_forward(const fft_plan* p, const float* x, float* f, float* tmp) {
  float* a = tmp;
  float* b = f;
  fwd_pass12(p, a, x);
  fwd_pass3(p, b, a);
  float* src = b;
  float* dst = a;
  for (int pass = 3; pass < p->nbits; ++pass) {
    fwd_pass_n(p, dst, src, pass);
    float* t = src;
    src = dst;
    dst = t;
  }
  if (src != f) memcpy(f, src, sizeof(float) * (size_t)p->N);
}

/* do_ifft general case (FFTReal.hpp:207-218, 661-704). */
static void plan_inverse(const fft_plan* p, const float* f, float* x, float* t1, float* t2) {
  const float* src = f;
  float* bufs[2] = {t1, t2};
  int k = 0;
  for (int pass = p->nbits - 1; pass >= 3; --pass) {
    inv_pass_n(p, bufs[k], src, pass);
    src = bufs[k];
    k ^= 1;
  }
  inv_pass3(p, bufs[k], src);
  inv_pass12(p, x, bufs[k]);
}

/* rescale (FFTReal.hpp:261-293): x *= float(1.0/N). */
void sxo_fft_rescale(int32_t N, float* x) {
  const float mul = (float)(1.0 / (double)N);
  for (int32_t i = 0; i < N; ++i) x[i] *= mul;
}

int sxo_fft_forward(int32_t N, const float* x, float* f) {
  fft_plan p;
  if (plan_init(&p, N)) return -1;
  float* tmp = (float*)malloc(sizeof(float) * (size_t)N);
  plan_forward(&p, x, f, tmp);
  free(tmp);
  plan_free(&p);
  return 0;
}

int sxo_fft_inverse(int32_t N, const float* f, float* x) {
  fft_plan p;
  if (plan_init(&p, N)) return -1;
  float* t1 = (float*)malloc(sizeof(float) * (size_t)N);
  float* t2 = (float*)malloc(sizeof(float) * (size_t)N);
  plan_inverse(&p, f, x, t1, t2);
  free(t1);
  free(t2);
  plan_free(&p);
  return 0;
}

/* ------------------------------------------------------------------ xcorr */

/* CrossCorrelation::DoOne (analysis/CrossCorr.cc:445-507) + ComplexMult (:20-26). */
static void xcorr_channel(const fft_plan* p, const float* in1, const float* in2, float* o, float* w) {
  const int32_t N = p->N, H = N / 2;
  float* t1 = w;
  float* t2 = w + N;
  float* s1 = w + 2 * N;
  float* s2 = w + 3 * N;
  plan_forward(p, in1, t1, s1);
  plan_forward(p, in2, t2, s1);
  t1[0] *= t2[0];
  for (int32_t i = 1; i < H - 1; ++i) {
    const float r1 = t1[i], i1 = t1[i + H], r2 = t2[i], i2 = -t2[i + H];
    const float fr = r1 * r2 - i1 * i2;
    const float fi = i1 * r2 + i2 * r1;
    t1[i] = fr;
    t1[i + H] = -fi;
  }
  plan_inverse(p, t1, t2, s1, s2);
  sxo_fft_rescale(N, t2);
  for (int32_t i = 0; i < H; ++i) o[i] = t2[i + H];
  for (int32_t i = H; i < N; ++i) o[i] = t2[i - H];
}

/* CrossCorrelation::CrossCorrelate (analysis/CrossCorr.cc:386-403). */
static int xcorr_with_plan(const fft_plan* p, const float* tsig4, const float* qsig4, float* xc, float* w) {
  const int32_t N = p->N;
  float* tmp = w + 4 * N;
  for (int32_t i = 0; i < N; ++i) xc[i] = 0.f;
  for (int ch = 0; ch < 4; ++ch) {
    xcorr_channel(p, tsig4 + (size_t)ch * N, qsig4 + (size_t)ch * N, tmp, w);
    for (int32_t i = 0; i < N; ++i) xc[i] += tmp[i];
  }
  return 0;
}

int sxo_xcorr(const float* tsig4, const float* qsig4, int32_t N, float* xc) {
  fft_plan p;
  if (plan_init(&p, N)) return -1;
  float* w = (float*)malloc(sizeof(float) * 5 * (size_t)N);
  xcorr_with_plan(&p, tsig4, qsig4, xc, w);
  free(w);
  plan_free(&p);
  return 0;
}

/* ------------------------------------------------------------------ peaks + diagonals */

/* SeqAnalyzer::FindTop (analysis/CrossCorr.cc:878-944), m_envSize = 256 (:562). */
int32_t sxo_findtop(const float* xc, int32_t N, double cutoff, int32_t* top, int32_t cap) {
  const int32_t env_size = 256;
  int32_t np = N / env_size;
  if (np == 0) np = 1;
  double* env = (double*)calloc((size_t)np, sizeof(double));
  for (int32_t i = 0; i < N; ++i) {
    double d = (double)(xc[i] * xc[i]);
    if (np > 8) env[i / env_size] += d;
  }
  for (int32_t b = 0; b < np; ++b) env[b] = sqrt(env[b] / (double)env_size);
  int32_t n = 0;
  for (int32_t i = 0; i < N; ++i) {
    if ((double)xc[i] > env[i / env_size] * cutoff + 1.) {
      if (n < cap) top[n] = i;
      ++n;
    }
  }
  free(env);
  return n;
}

/* LookupMatch::precalc_scores (analysis/CrossCorr.cc:547-553): (int)(100*EqualAmb + 0.5). */
static int g_score[128][128];
static pthread_once_t g_score_once = PTHREAD_ONCE_INIT;
static void score_init(void) {
  codec_ready();
  for (int i = 0; i < 128; ++i)
    for (int j = 0; j < 128; ++j) g_score[i][j] = (int)(sxo_dna_equal_amb(i, j) * (double)100 + 0.5);
}

/* SeqAnalyzer::DoOne, OPTIM_DOONE variant (analysis/CrossCorr.cc:667-724);
 * m_minLen = 45, m_minIdent = 0.42 (:559-560). */
static int32_t do_one(const char* q, int32_t lq, const char* t, int32_t lt, int32_t shift,
                      sxo_seqmatch* out, int32_t cap, int32_t n_out) {
  const int min_len = 45;
  const double min_ident = 0.42;
  int match = 0;
  const int min_matches = (int)((double)min_len * min_ident * (double)100);
  int32_t last_start = -1;
  int32_t n = 0;
  for (int32_t i = 0; i < lt; ++i) {
    int32_t j = i + shift;
    if (j < 0) continue;
    if (j >= lq || i + 1 >= lt) {
      if (last_start != -1) {
        if (n_out < cap) {
          out[n_out].start_target = last_start;
          out[n_out].start_query = last_start + shift;
          out[n_out].len = i - last_start;
        }
        ++n_out;
      }
      break;
    }
    match += g_score[ub(t[i]) & 127][ub(q[j]) & 127];
    if (n > min_len) {
      match -= g_score[ub(t[i - min_len - 1]) & 127][ub(q[j - min_len - 1]) & 127];
      if (match > min_matches) {
        if (last_start == -1) last_start = i - min_len;
      } else {
        if (last_start != -1) {
          if (n_out < cap) {
            out[n_out].start_target = last_start;
            out[n_out].start_query = last_start + shift;
            out[n_out].len = i - last_start;
          }
          ++n_out;
        }
        last_start = -1;
      }
    }
    n++;
  }
  return n_out;
}

/* SeqAnalyzer::MatchUp (analysis/CrossCorr.cc:583-605). */
int32_t sxo_matchup(const char* q, int32_t lq, const char* t, int32_t lt, const float* xc, int32_t N,
                    double cutoff, sxo_seqmatch* out, int32_t cap) {
  pthread_once(&g_score_once, score_init);
  int32_t* top = (int32_t*)malloc(sizeof(int32_t) * (size_t)N);
  int32_t ntop = sxo_findtop(xc, N, cutoff, top, N);
  int32_t n = 0;
  for (int32_t k = 0; k < ntop; ++k) n = do_one(q, lq, t, lt, top[k] - N / 2, out, cap, n);
  free(top);
  return n;
}

/* ------------------------------------------------------------------ probability tail */

/* CDF / Sigma / GCAdjustExpect (analysis/AlignProbability.cc:11-40). */
static double cdf(double x, double m, double s) { return 0.5 * (1. + erf((x - m) / s / 1.414213562)); }
static double sigma(double p, int n) { return sqrt(p * (1. - p) * (double)n); }
static double gc_adjust_expect(double gc, int n, double gc_target) {
  double at_target = 1. - gc_target;
  double r = gc * gc_target;
  r += (double)(n - gc) * at_target;
  return r / (double)n / 2.;
}

/* ProbTable::GetMatchProbabilityRaw (analysis/ProbTable.cc:143-163). */
static double prob_raw(int length, double ident, double ident_expect, double target_size) {
  double p_match = ident_expect;
  double s = sigma(p_match, length);
  double m = p_match * (double)length;
  double x = (double)length * ident;
  double c = cdf(m, x, s);
  double expect = c * target_size;
  return exp(-expect);
}

typedef struct {
  double target_size;
  double* t;
  int next;
  pthread_mutex_t mu;
} pt_job;

static void* pt_worker(void* arg) {
  pt_job* jb = (pt_job*)arg;
  for (;;) {
    pthread_mutex_lock(&jb->mu);
    int i = jb->next++;
    pthread_mutex_unlock(&jb->mu);
    if (i >= SXO_PT_ROWS) break;
    double* row = jb->t + (size_t)i * SXO_PT_LEN;
    const double ident_expect = (double)i / ((double)SXO_PT_ROWS - 1);
    row[0] = 2.;
    for (int j = 1; j < SXO_PT_LEN; ++j) {
      double min = 0, max = 1, mid;
      while (max - min > 0.00000001) {
        mid = (max + min) / 2.0;
        if (prob_raw(j, mid, ident_expect, jb->target_size) != 0.) max = mid; else min = mid;
      }
      row[j] = (max + min) / 2.0;
    }
  }
  return NULL;
}

void sxo_prob_table(double target_size, int32_t threads, double* t) {
  memset(t, 0, sizeof(double) * SXO_PT_ROWS * SXO_PT_LEN);
  pt_job jb;
  jb.target_size = target_size;
  jb.t = t;
  jb.next = 1;
  pthread_mutex_init(&jb.mu, NULL);
  if (threads < 1) threads = 1;
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
  for (int k = 0; k < threads; ++k) pthread_create(&th[k], NULL, pt_worker, &jb);
  for (int k = 0; k < threads; ++k) pthread_join(th[k], NULL);
  free(th);
  pthread_mutex_destroy(&jb.mu);
}

double sxo_prob_table_lookup(const double* t, const char* target, const char* query, int32_t st, int32_t sq,
                             int32_t len, double cutoff, double* ident) {
  codec_ready();
  double gct = 0, gcq = 0, matches = 0;
  for (int32_t i = 0; i < len; ++i) {
    int a = ub(target[i + st]), b = ub(query[i + sq]);
    matches += sxo_dna_equal(a, b);
    gct += g_C[a] + g_G[a];
    gcq += g_C[b] + g_G[b];
  }
  *ident = matches / (double)len;
  double p_match = gc_adjust_expect(gcq, len, gct / (double)len);
  int index = (int)(p_match * (SXO_PT_ROWS - 1)); /* ExpectToIndex */
  if (index < 1) return 0.;
  int l = len >= SXO_PT_LEN ? SXO_PT_LEN - 1 : len;
  return (*ident >= t[(size_t)index * SXO_PT_LEN + l]) ? cutoff : 0.;
}

/* GetMatchProbabilityEx (analysis/AlignProbability.cc:62-127). */
double sxo_match_probability(const char* t, const char* q, int32_t st, int32_t sq, int32_t length,
                             double target_size) {
  codec_ready();
  double gct = 0, gcq = 0, matches = 0;
  for (int32_t i = 0; i < length; ++i) {
    int a = ub(t[i + st]), b = ub(q[i + sq]);
    matches += sxo_dna_equal(a, b);
    gct += g_C[a] + g_G[a];
    gcq += g_C[b] + g_G[b];
  }
  double ident = matches / (double)length;
  double gc = gcq;
  double p_match = gc_adjust_expect(gc, length, gct / (double)length);
  double s = sigma(p_match, length);
  double m = p_match * (double)length;
  double x = (double)length * ident;
  double c = cdf(m, x, s);
  double expect = c * target_size;
  return exp(-expect);
}

/* PrintMatch return value (analysis/CrossCorr.cc:967-1037) with bNuc = true. */
double sxo_print_match_ident(const char* q, const char* t, int32_t st, int32_t sq, int32_t len) {
  double matches = 0;
  for (int32_t i = st; i < st + len; ++i) {
    int32_t j = sq + i - st;
    matches += sxo_dna_equal(ub(q[j]), ub(t[i]));
  }
  return matches / (double)len;
}

/* ------------------------------------------------------------------ slave Align */

/* RCQuery (analysis/HomologyByXCorrSlave.cc:56-60). */
static int rc_query(int offset, int start, int chunk_size, int qlen) { return start + qlen - offset - chunk_size; }

typedef struct {
  fft_plan plan;
  float* sigT;
  float* sigQ;
  float* xc;
  float* w;
  char* qrc;
  sxo_seqmatch* m;
  int32_t mcap;
} align_ws;

int sxo_apply_tail(float* sig, int32_t L, int32_t N, const sxo_tail* t) {
  if (!t || t->nseg <= 0) return 0;
  float* tmp = (float*)malloc(sizeof(float) * 4 * (size_t)N);
  char* rc = (char*)malloc((size_t)N + 1);
  int32_t start = L;
  int err = 0;
  for (int32_t s = 0; s < t->nseg && !err; ++s) {
    const sxo_tail_seg* g = &t->segs[s];
    if (g->end <= start || g->end > g->len || g->len > N) {
      err = -1;
      break;
    }
    const char* src = g->seq;
    if (g->strand) {
      sxo_reverse_complement(g->seq, g->len, rc);
      src = rc;
    }
    sxo_encode(src, g->len, N, tmp);
    for (int ch = 0; ch < 4; ++ch)
      memcpy(sig + (size_t)ch * N + start, tmp + (size_t)ch * N + start, sizeof(float) * (size_t)(g->end - start));
    start = g->end;
  }
  free(tmp);
  free(rc);
  return err;
}

static int ws_init(align_ws* ws, int32_t N) {
  memset(ws, 0, sizeof *ws);
  if (plan_init(&ws->plan, N)) return -1;
  ws->sigT = (float*)malloc(sizeof(float) * 4 * (size_t)N);
  ws->sigQ = (float*)malloc(sizeof(float) * 4 * (size_t)N);
  ws->xc = (float*)malloc(sizeof(float) * (size_t)N);
  ws->w = (float*)malloc(sizeof(float) * 5 * (size_t)N);
  ws->qrc = (char*)malloc((size_t)N + 1);
  ws->mcap = 1 << 16;
  ws->m = (sxo_seqmatch*)malloc(sizeof(sxo_seqmatch) * (size_t)ws->mcap);
  return 0;
}

static void ws_free(align_ws* ws) {
  plan_free(&ws->plan);
  free(ws->sigT);
  free(ws->sigQ);
  free(ws->xc);
  free(ws->w);
  free(ws->qrc);
  free(ws->m);
}

/* HomologyByXCorr::Align (analysis/HomologyByXCorrSlave.cc:221-253) and
 * FilterMatches (:168-219) with prob_table off. */
static int32_t align_pair_ws(align_ws* ws, const sxo_pair* p, const sxo_params* prm, const sxo_tail* t_tail,
                             const sxo_tail* q_tail, sxo_result* out, int32_t cap) {
  pthread_once(&g_score_once, score_init);
  const int32_t N = prm->N;
  const double cutoff = p->fast ? prm->cutoff_fast : prm->cutoff;
  int32_t nres = 0;
  if (p->tlen > N || p->qlen > N) return -1;
  sxo_encode(p->tseq, p->tlen, N, ws->sigT);
  if (sxo_apply_tail(ws->sigT, p->tlen, N, t_tail)) return -1;
  for (int strand = 0; strand < prm->strands; ++strand) {
    const char* qs = p->qseq;
    if (strand == 1) {
      sxo_reverse_complement(p->qseq, p->qlen, ws->qrc);
      qs = ws->qrc;
    }
    sxo_encode(qs, p->qlen, N, ws->sigQ);
    if (sxo_apply_tail(ws->sigQ, p->qlen, N, q_tail)) return -1;
    xcorr_with_plan(&ws->plan, ws->sigT, ws->sigQ, ws->xc, ws->w);
    int32_t nm;
    for (;;) {
      nm = sxo_matchup(qs, p->qlen, p->tseq, p->tlen, ws->xc, N, cutoff, ws->m, ws->mcap);
      if (nm <= ws->mcap) break;
      free(ws->m);
      ws->mcap = nm;
      ws->m = (sxo_seqmatch*)malloc(sizeof(sxo_seqmatch) * (size_t)ws->mcap);
    }
    for (int32_t k = 0; k < nm; ++k) {
      const sxo_seqmatch* m = &ws->m[k];
      if (m->len < prm->min_len) continue;
      int32_t len = m->len;
      int32_t tstart = p->t_chunk_start + m->start_target;
      int32_t qstart = strand == 0 ? p->q_chunk_start + m->start_query
                                   : rc_query(p->q_chunk_start, m->start_query,
                                              prm->q_chunk > 0 ? prm->q_chunk : p->qlen, p->q_record_len);
      double prob, ident;
      if (prm->ptable) {
        /* FilterMatches with prob_table (analysis/HomologyByXCorrSlave.cc:184-187) */
        prob = sxo_prob_table_lookup(prm->ptable, p->tseq, qs, m->start_target, m->start_query, len,
                                     prm->table_prob, &ident);
        if (prob < prm->min_prob) continue;
      } else {
        prob = sxo_match_probability(p->tseq, qs, m->start_target, m->start_query, len, prm->target_total);
        if (prob < prm->min_prob) continue;
        ident = sxo_print_match_ident(qs, p->tseq, m->start_target, m->start_query, len);
      }
      if (nres < cap) {
        sxo_result* r = &out[nres];
        memset(r, 0, sizeof *r);
        r->query_id = (uint64_t)(int64_t)p->q_record;
        r->target_id = (uint64_t)(int64_t)p->t_record;
        r->query_size = (uint64_t)(int64_t)p->q_record_len;
        r->qstart = (uint64_t)(int64_t)qstart;
        r->tstart = (uint64_t)(int64_t)tstart;
        r->len = (uint64_t)(int64_t)len;
        r->reverse = (uint8_t)(strand == 1);
        r->prob = prob;
        r->ident = ident;
      }
      ++nres;
    }
  }
  return nres;
}

int32_t sxo_align_pair(const sxo_pair* p, const sxo_params* prm, sxo_result* out, int32_t cap) {
  align_ws ws;
  if (ws_init(&ws, prm->N)) return -1;
  int32_t n = align_pair_ws(&ws, p, prm, NULL, NULL, out, cap);
  ws_free(&ws);
  return n;
}

typedef struct {
  const sxo_pair* pairs;
  const sxo_tail* t_tails;
  const sxo_tail* q_tails;
  int32_t n;
  const sxo_params* prm;
  sxo_result* out;
  int32_t cap;
  int32_t* counts;
  int32_t next;
  pthread_mutex_t mu;
  int64_t total;
} batch_ctx;

static void* batch_worker(void* arg) {
  batch_ctx* b = (batch_ctx*)arg;
  align_ws ws;
  if (ws_init(&ws, b->prm->N)) return NULL;
  for (;;) {
    pthread_mutex_lock(&b->mu);
    int32_t k = b->next++;
    pthread_mutex_unlock(&b->mu);
    if (k >= b->n) break;
    int32_t c = align_pair_ws(&ws, &b->pairs[k], b->prm, b->t_tails ? &b->t_tails[k] : NULL,
                              b->q_tails ? &b->q_tails[k] : NULL, b->out + (size_t)k * b->cap, b->cap);
    b->counts[k] = c < b->cap ? c : b->cap;
    pthread_mutex_lock(&b->mu);
    b->total += c;
    pthread_mutex_unlock(&b->mu);
  }
  ws_free(&ws);
  return NULL;
}

int64_t sxo_align_batch(const sxo_pair* pairs, int32_t n, const sxo_params* prm, int32_t threads,
                        sxo_result* out, int32_t cap_per_pair, int32_t* counts) {
  return sxo_align_batch_tails(pairs, n, prm, NULL, NULL, threads, out, cap_per_pair, counts);
}

int64_t sxo_align_batch_tails(const sxo_pair* pairs, int32_t n, const sxo_params* prm, const sxo_tail* t_tails,
                              const sxo_tail* q_tails, int32_t threads, sxo_result* out, int32_t cap_per_pair,
                              int32_t* counts) {
  batch_ctx b;
  b.pairs = pairs;
  b.t_tails = t_tails;
  b.q_tails = q_tails;
  b.n = n;
  b.prm = prm;
  b.out = out;
  b.cap = cap_per_pair;
  b.counts = counts;
  b.next = 0;
  b.total = 0;
  pthread_mutex_init(&b.mu, NULL);
  if (threads < 1) threads = 1;
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
  for (int i = 0; i < threads; ++i) pthread_create(&th[i], NULL, batch_worker, &b);
  for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
  free(th);
  pthread_mutex_destroy(&b.mu);
  return b.total;
}
