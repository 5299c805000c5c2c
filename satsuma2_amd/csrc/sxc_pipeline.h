// sxc_pipeline.h -- the chunk-pair path as workgroup phases.
//
// Each kernel body is written once against an executor `Ex` that provides
//   ex.each(f)   run f(tid) for every thread of the workgroup
//   ex.sync()    workgroup barrier
//   ex.reg(tid)  per-thread persistent state (registers on the GPU)
//   ex.atomic_add_shared / ex.atomic_add_global
// The HIP kernels instantiate it with the device executor (sxc_kernels.hip); the CPU
// workgroup emulator used by the tests instantiates it with a sequential one.
//
// Phases and the reference functions they restate:
//   prepare     CCSignal::ComputeEntropy + SeqToPCM means (analysis/CrossCorr.cc:35-134)
//   spectrum    encode (SetSequence, :179-206) fused into FFTReal do_fft
//   correlate   DoOne product (:477-492) + do_ifft + rescale + fftshift (:493-505), summed
//               over A, C, G, T in that order (:395-403)
//   peaks       SeqAnalyzer::FindTop (:878-944)
// The diagonal scan and the FP64 tail live in sxc_scan.h.
#pragma once
#include <math.h>

#include "sxc_fft.h"

namespace sxc {

struct PairArgs {
  const sxc_tables* tab;
  const float* osc;          // oscillator twiddles for levels 13, 14
  const double* ent_lut;     // entropy of quarter-exact windows for this N, [k-1][q]
  const float* ent_rows;     // host-computed window entropies of chunks with B/D/H/V letters
                             // (sxc_chunk flags bits 8-31 index them)
  const unsigned char* seq;  // all chunk bytes
  const sxc_chunk* tchunks;
  const sxc_chunk* qchunks;
  const sxc_pair_dev* pairs;
  const float* spectra;      // resident target spectra, 4*N floats per target chunk
  int32_t npairs;
  int32_t strands;           // 1 = forward only, 2 = forward then reverse complement
  double cutoff;
  double cutoff_fast;
  uint16_t* cands;           // accepted lags, N slots per (pair, strand)
  int32_t* ncand;            // accepted lags per (pair, strand)
  float* xc_out;             // optional: N floats per (pair, strand)
  const sxc_tail_dev* tails;     // signal tails (pairs[].q_tail indexes them), may be null
  const sxc_tail_seg_dev* segs;
  const float* src_sig;          // tail source signals, 4*N floats each
  // optional resident query spectra: 4*N floats per (query chunk, strand), at
  // qspectra + (2 * q_chunk + strand) * 4N; the pair kernel then skips the query's encode
  // and forward transforms (the spectra are those transforms' output, bit for bit)
  const float* qspectra;
  int32_t* counts;               // optional: set to 0 per (pair, strand) for a split scan
};

struct SpecArgs {
  const sxc_tables* tab;
  const float* osc;
  const double* ent_lut;
  const float* ent_rows;
  const unsigned char* seq;
  const sxc_chunk* chunks;
  const int32_t* which;      // chunk indices to transform
  int32_t n;
  float* spectra;            // output: 4*N floats per entry of `which`, at spectra + dst[k]*4N
  const int32_t* dst;        // destination slot per entry
  const float* signals;      // optional: given 4*N signals instead of encoding chunk bytes
  const int32_t* chunk_tail;     // optional: signal tail per chunk (-1: none)
  const sxc_tail_dev* tails;
  const sxc_tail_seg_dev* segs;
  const float* src_sig;
  int32_t strand;            // 1: spectra of the reverse complement (resident query spectra)
};

struct EncodeArgs {
  const sxc_tables* tab;
  const double* ent_lut;
  const float* ent_rows;
  const unsigned char* seq;
  const sxc_chunk* chunks;
  int32_t strand;
  const int32_t* strands;    // optional: strand per chunk (overrides `strand`)
  float* out;                // 4*N floats per chunk
};

struct XcorrArgs {
  const sxc_tables* tab;
  const float* osc;
  const float* spectra;      // 4*N target spectra per item
  const float* qsig;         // 4*N query signals per item
  float* xc;                 // N floats per item
};

struct ProbArgs {
  const sxc_tables* tab;
  const unsigned char* t;
  const unsigned char* q;
  const int32_t* runs;       // n x {start_target, start_query, len}
  int32_t n;
  double target_size;
  double* prob;
  double* ident;
};

// ---------------------------------------------------------------- shared memory plan

template <int NB>
struct Smem {
  using G = Geo<NB>;
  static constexpr int N = G::N;
  static constexpr int NT = G::NT;
  static constexpr int NWIN = 512;  // max entropy windows (L <= N, win = N/512)
  // byte offsets
  static constexpr int Y = 0;                                  // float[YLEN]: FFT / xc / scan area
  // lean plan: twiddles and chunk bytes are read through the caches, not staged here
  static constexpr bool LEAN = geo_lean(NB) != 0;
  static constexpr int LUT = Y + 4 * G::YLEN;                  // float[SXC_LUT_LEN] twiddles
  // the peaks phase's per-thread counts reuse the twiddle area (the transforms are done)
  static constexpr int CNT = LUT;                              // int32[NT]
  static_assert(NT <= SXC_LUT_LEN, "counts fit the twiddle area");
  static constexpr int ENT = LUT + (LEAN ? 4 * NT : 4 * SXC_LUT_LEN);  // float[NWIN]
  static constexpr int ENV = ENT + 4 * NWIN;                   // double[max(N/256, 1)]
  static constexpr int NENV = (N / 256) > 0 ? N / 256 : 1;
  static constexpr int OFF = ENV + 8 * NENV;                   // double[4] channel means
  static constexpr int DSUM = OFF + 32;                        // double[4] sequential sums
  static constexpr int MISC = DSUM + 32;                       // int32[16]
  static constexpr int CODEC = MISC + 64;                      // double[18][4]; DNA_X - off[X] after prepare
  static constexpr int CLS = CODEC + 8 * 4 * SXC_NCLASS;       // uchar[128] byte -> class
  static constexpr int QUART = CLS + 128;                      // uchar[18][4], read a row at a time
  static_assert(QUART % 4 == 0, "quarter rows are read as words");
  static constexpr int RCT = QUART + 4 * SXC_NCLASS;           // uchar[128] byte -> class of its complement
  // codec classes of the strand-applied chunk bytes: an LDS copy up to N = 16384; at 32768
  // the plan would not fit the 160 KB of a workgroup, and the bytes are read from global
  // memory (through the class tables) instead
  static constexpr bool SEQ_IN_LDS = G::LOG <= 14 && !LEAN;
  static constexpr int QSEQ = (RCT + 128 + 15) & ~15;          // uchar[N] when SEQ_IN_LDS
  static constexpr int BYTES = (QSEQ + (SEQ_IN_LDS ? N : 0) + 15) & ~15;
  // after the peaks phase Y holds the accepted lags
  static constexpr int CAND = Y;                               // uint16[N]
};

// The twiddle table a kernel reads: the LDS copy, or the constant table itself under the
// lean plan.
template <int NB>
SXC_HD inline const float* twiddle_table(const char* sm, const sxc_tables* tab) {
  if constexpr (Smem<NB>::LEAN) {
    return tab->lut;
  } else {
    return (const float*)(sm + Smem<NB>::LUT);
  }
}


enum { MISC_L = 0, MISC_THIRDS = 1, MISC_NCAND = 2, MISC_KEPT = 3, MISC_Q4 = 4 /* 4 ints */, MISC_RUNS = 8 };

template <int NB>
struct Regs {
  float acc[Geo<NB>::SA];  // correlation accumulated over channels, stage-A output order
  // codec classes of the thread's stage-A positions (byte r % 4 of word r / 4), read from the
  // chunk once for the four channels (class_cache)
  uint32_t cls[Geo<NB>::SA / 4];
};

// Whether the pair kernel keeps the class cache: up to N = 8192.  At larger N its eight live
// registers cost the pair kernel occupancy or spills, and each channel reads the chunk
// again; the target-spectra kernel, with registers to spare, always keeps it.
template <int NB>
constexpr bool pair_class_cache() {
  return Geo<NB>::LOG <= 13;
}

// ---------------------------------------------------------------- helpers

// The window entropies of a chunk with a 1/3-valued letter (B, D, H, V) for one strand, or
// null.  Such windows sum thirds in the reference's order and take glibc's log of the
// result (Ent, analysis/CrossCorr.cc:28-33), which the device cannot reproduce bit for bit,
// so the host computes them with its libm (window_entropy_row) and the kernels read them.
SXC_HD inline const float* ent_row(const float* rows, const sxc_chunk& c, int strand) {
  if (!(c.flags & 1) || !rows) return nullptr;
  return rows + (size_t)((uint32_t)c.flags >> 8) * SXC_ENT_ROW + (size_t)strand * (SXC_ENT_ROW / 2);
}

SXC_HD inline int score_of(const unsigned char* cls, const unsigned char* score, unsigned char tb, unsigned char qb) {
  // LookupMatch::precalc_scores = (int)(100 * DNA_EqualAmb + 0.5) (analysis/CrossCorr.cc:547-553);
  // DNA_EqualAmb scores identical letters other than 'N' as 1 (analysis/DNAVector.cc:372-387)
  if (tb == qb && tb != 'N') return 100;
  return score[cls[tb] * SXC_NCLASS + cls[qb]];
}

// The codec classes of the bytes being encoded, with the strand applied: the LDS copy (which
// holds classes), or the chunk's bytes in global memory mapped through a 128-entry table
// (byte -> class; for the reverse strand byte -> class of its complement, read from the end:
// DNAVector::ReverseComplement, analysis/DNAVector.cc:482-521).
struct SeqView {
  const unsigned char* p;
  const unsigned char* map;  // null: p holds classes
  int L;
  int rev;
  SXC_HD unsigned char operator[](int i) const { return map ? map[p[rev ? L - 1 - i : i] & 127] : p[i]; }
};

// Loads one chunk's bytes (reverse-complemented for strand 1) and computes the per-window
// entropy weights (CCSignal::ComputeEntropy, analysis/CrossCorr.cc:35-93) and the four
// channel means.  `erow`: the host-computed window entropies when `thirds` (ent_row).
template <int NB, class Ex>
SXC_HD inline SeqView phase_prepare(Ex& ex, char* sm, const sxc_tables* tab, const double* ent_lut,
                                    const float* erow, const unsigned char* src, int L, int thirds,
                                    int strand) {
  using SM = Smem<NB>;
  using G = Geo<NB>;
  float* lut = (float*)(sm + SM::LUT);
  unsigned char* cls = (unsigned char*)(sm + SM::CLS);
  unsigned char* rct = (unsigned char*)(sm + SM::RCT);  // byte -> class of its complement
  unsigned char* quart = (unsigned char*)(sm + SM::QUART);
  double* codec = (double*)(sm + SM::CODEC);
  int* misc = (int*)(sm + SM::MISC);
  ex.each([&](int tid) {
    if constexpr (!SM::LEAN)
      for (int i = tid; i < SXC_LUT_LEN; i += G::NT) lut[i] = tab->lut[i];
    for (int i = tid; i < 128; i += G::NT) cls[i] = tab->cls[i];
    for (int i = tid; i < 4 * SXC_NCLASS; i += G::NT) {
      quart[i] = (&tab->quarters[0][0])[i];
      codec[i] = (&tab->codec[0][0])[i];
    }
    for (int i = tid; i < 128; i += G::NT) rct[i] = tab->rcls[i];  // class of the complement
    if constexpr (!SM::SEQ_IN_LDS) {
      // byte -> the four quarter counts of its class on this strand, packed in a word; held in
      // the FFT area, which the transforms take over only after this phase
      uint32_t* qtab = (uint32_t*)(sm + SM::Y);
      for (int i = tid; i < 128; i += G::NT) qtab[i] = tab->qpack[strand ? 1 : 0][i];
    }
    if constexpr (SM::SEQ_IN_LDS) {
      // classes of the strand's bytes (DNAVector::ReverseComplement, analysis/DNAVector.cc:482-521)
      unsigned char* q = (unsigned char*)(sm + SM::QSEQ);
      for (int i = tid; i < L; i += G::NT)
        q[i] = strand ? tab->cls[tab->rc[src[L - 1 - i] & 127] & 127] : tab->cls[src[i] & 127];
    }
    if (tid < 16) misc[tid] = 0;
  });
  ex.sync();
  SeqView qs;
  if constexpr (SM::SEQ_IN_LDS) {
    qs = SeqView{(const unsigned char*)(sm + SM::QSEQ), nullptr, L, 0};
  } else {
    qs = SeqView{src, strand ? rct : cls, L, strand};
  }
  float* ent = (float*)(sm + SM::ENT);
  double* dsum = (double*)(sm + SM::DSUM);
  const int win = G::N / 512;
  // a class's four quarter counts (each <= 4) as the bytes of one word: 16 of them add
  // without a carry between bytes
  auto qpack = [&](unsigned c) { return *(const uint32_t*)(quart + 4 * c); };
  const uint32_t* qtab = (const uint32_t*)(sm + SM::Y);
  auto quarter_sums = [&](int i0, int i1, int* q) {
    for (int j0 = i0; j0 < i1; j0 += 16) {
      const int j1 = (j0 + 16 < i1) ? j0 + 16 : i1;
      uint32_t acc = 0;
      if constexpr (!SM::SEQ_IN_LDS) {
        // the positions' source bytes [a, b) (reversed on strand 1; the sum is order-free),
        // read as aligned words: chunks start 16-byte aligned and are followed by 16 bytes
        const int a = strand ? L - j1 : j0, b = strand ? L - j0 : j1;
        if (((a | b) & 3) == 0) {
          for (int x = a; x < b; x += 4) {
            const uint32_t wv = *(const uint32_t*)(src + x);
            acc += qtab[wv & 127u] + qtab[(wv >> 8) & 127u] + qtab[(wv >> 16) & 127u] + qtab[(wv >> 24) & 127u];
          }
        } else {
          for (int x = a & ~3; x < b; x += 4) {
            const uint32_t wv = *(const uint32_t*)(src + x);
            static_for<0, 4>([&](auto T) {
              constexpr int t = decltype(T)::value;
              const int y = x + t;
              const uint32_t v = qtab[(wv >> (8 * t)) & 127u];
              acc += (y >= a && y < b) ? v : 0u;
            });
          }
        }
      } else {
        for (int j = j0; j < j1; ++j) acc += qpack(qs[j]);
      }
      q[0] += (int)(acc & 255u);
      q[1] += (int)((acc >> 8) & 255u);
      q[2] += (int)((acc >> 16) & 255u);
      q[3] += (int)(acc >> 24);
    }
  };
  ex.each([&](int tid) {
    if (L < 1024) return;  // entropy weight is 1 (analysis/CrossCorr.cc:39-44)
    const int nwin = (L + win - 1) / win;
    int tq[4] = {0, 0, 0, 0};  // this thread's windows' sums: its share of the channel sums
    for (int w = tid; w < nwin; w += G::NT) {
      if (thirds) {
        ent[w] = erow[w];  // computed on the host with glibc's log
        continue;
      }
      // quarter-exact window: the four sums are integers of quarters, Ent of each is a
      // table entry computed on the host
      const int i = w * win;
      const int k = (L - i < win) ? (L - i) : win;
      int q[4] = {0, 0, 0, 0};
      quarter_sums(i, i + k, q);
      const double* row = ent_lut + (size_t)(k - 1) * (4 * win + 1);
      const double s = row[q[0]] + row[q[1]] + row[q[2]] + row[q[3]];
      float e = (float)(-s);
      if (e < 0.f) e = 0.f;
      ent[w] = e;
      for (int ch = 0; ch < 4; ++ch) tq[ch] += q[ch];
    }
    // the windows partition the chunk, so their integer sums add up to the channel sums
    if (!thirds)
      for (int ch = 0; ch < 4; ++ch)
        if (tq[ch]) ex.atomic_add_shared(&misc[MISC_Q4 + ch], tq[ch]);
  });
  // channel sums: exact in quarters unless a 1/3-valued letter is present, in which case
  // the reference's sequential double order is replayed (analysis/CrossCorr.cc:108-116)
  ex.each([&](int tid) {
    if (!thirds) {
      if (L >= 1024) return;  // summed with the entropy windows above
      int q[4] = {0, 0, 0, 0};
      for (int i = tid; i < L; i += G::NT) {
        const uint32_t x = qpack(qs[i]);
        q[0] += (int)(x & 255u);
        q[1] += (int)((x >> 8) & 255u);
        q[2] += (int)((x >> 16) & 255u);
        q[3] += (int)(x >> 24);
      }
      for (int ch = 0; ch < 4; ++ch)
        if (q[ch]) ex.atomic_add_shared(&misc[MISC_Q4 + ch], q[ch]);
    } else if (tid < 4) {
      double s = 0;
      for (int i = 0; i < L; ++i) s += codec[4 * qs[i] + tid];
      dsum[tid] = s;
    }
  });
  ex.sync();
  double* off = (double*)(sm + SM::OFF);
  ex.each([&](int tid) {
    if (tid < 4) {
      const double sum = thirds ? dsum[tid] : (double)misc[MISC_Q4 + tid] * 0.25;
      off[tid] = sum / (double)L;
    }
  });
  ex.sync();
  // the encoder's (DNA_X(b) - off) per class and channel: the same double subtraction
  // SeqToPCM evaluates per base (analysis/CrossCorr.cc:127), done once per class
  ex.each([&](int tid) {
    for (int i = tid; i < 4 * SXC_NCLASS; i += G::NT) codec[i] = codec[i] - off[i & 3];
  });
  ex.sync();
  return qs;
}

// Positions at or past a chunk's length: 0, or the signal an earlier sequence left there
// when the caller reuses one CCSignal object (include/sxc.h, sxc_tail_seg).  Segments are
// ordered by end; a binary search finds the one holding p.
struct TailRef {
  const sxc_tail_seg_dev* seg;
  int n;
  const float* sig;  // source signals, 4*N floats each
};

template <int NB>
SXC_HD inline TailRef tail_ref(const sxc_tail_dev* tails, const sxc_tail_seg_dev* segs, const float* sig, int idx) {
  if (!tails || idx < 0) return TailRef{nullptr, 0, nullptr};
  const sxc_tail_dev t = tails[idx];
  return TailRef{segs + t.first, t.n, sig};
}

template <int NB>
SXC_HD inline float tail_at(const TailRef& t, int p, int ch) {
  int lo = 0, hi = t.n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (p < t.seg[mid].end) hi = mid; else lo = mid + 1;
  }
  if (lo >= t.n) return 0.f;
  return t.sig[((size_t)t.seg[lo].src * 4 + ch) * (size_t)Geo<NB>::N + p];
}

// Signal value of channel ch at time position p of a chunk of length L whose base at p has
// codec class cls (CCSignal::SeqToPCM, :118-131).
template <int NB>
SXC_HD inline float encode_class(const char* sm, int L, unsigned cls, int p, int ch, const TailRef& tail) {
  using SM = Smem<NB>;
  if (p >= L) return tail.n ? tail_at<NB>(tail, p, ch) : 0.f;
  const double* diff = (const double*)(sm + SM::CODEC);  // DNA_X(class) - off[X] (phase_prepare)
  const float e = (L < 1024) ? 1.f : ((const float*)(sm + SM::ENT))[p / (Geo<NB>::N / 512)];
  return (float)((double)e * diff[4 * cls + ch]);
}

template <int NB>
SXC_HD inline float encode_at(const char* sm, const SeqView& qs, int p, int ch, const TailRef& tail) {
  return encode_class<NB>(sm, qs.L, p < qs.L ? qs[p] : 0u, p, ch, tail);
}

// x, copied where it is used: the compiler can neither hoist nor share the copy between
// loop iterations (no instruction), so one packed word stays live instead of its fields.
SXC_HD inline uint32_t fresh(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(x));
#endif
  return x;
}

// Fills each thread's class cache (Regs::cls) for its stage-A positions: the chunk's bytes
// and the class tables are read once per transform set instead of once per channel.
template <int NB, class Ex>
SXC_HD inline void class_cache(Ex& ex, const SeqView& qs) {
  using G = Geo<NB>;
  ex.each([&](int g) {
    if (g >= G::GROUPS) return;
    Regs<NB>& rg = ex.template reg<Regs<NB>>(g);
    static_for<0, G::SA / 4>([&](auto W) {
      uint32_t w = 0;
      static_for<0, 4>([&](auto B) {
        const int p = stageA_pos<NB>(g, 4 * decltype(W)::value + decltype(B)::value);
        if (p < qs.L) w |= (uint32_t)qs[p] << (8 * decltype(B)::value);
      });
      rg.cls[decltype(W)::value] = w;
    });
  });
}

// Forward FFT of one channel into the padded LDS array: stage A (encode fused into
// the bit-reversed load, from the class cache when CACHED, or loaded from a given signal)
// then the rounds.
template <int NB, int R, class Ex>
SXC_HD inline void fwd_rounds(Ex& ex, float* Y, const Twiddle& tw, int upto = Schedule<NB>::NR) {
  using S = Schedule<NB>;
  if constexpr (R < S::NR) {
    if (R >= upto) return;
    constexpr int K = S::k(R), M = S::m(R);
    using RG = RoundGeo<NB, K, M>;
    ex.each([&](int tid) {
      for (int u = tid; u < RG::UNITS; u += Geo<NB>::NT) round_unit<NB, K, M, false>(Y, u, tw);
    });
    ex.sync();
    fwd_rounds<NB, R + 1>(ex, Y, tw, upto);
  }
}

// The last round with the product (round_unit_product): forward levels when FWD, product with
// T, inverse levels; the inverse transform then continues with round NR - 2.
template <int NB, bool FWD, class Ex>
SXC_HD inline void product_round(Ex& ex, float* Y, const Twiddle& tw, const float* T) {
  using S = Schedule<NB>;
  constexpr int R = S::NR - 1, K = S::k(R), M = S::m(R);
  using RG = RoundGeo<NB, K, M>;
  ex.each([&](int tid) {
    for (int u = tid; u < RG::UNITS; u += Geo<NB>::NT) round_unit_product<NB, K, M, FWD>(Y, u, tw, T);
  });
  ex.sync();
}

template <int NB, int R, class Ex>
SXC_HD inline void inv_rounds(Ex& ex, float* Y, const Twiddle& tw) {
  using S = Schedule<NB>;
  if constexpr (R >= 0) {
    constexpr int K = S::k(R), M = S::m(R);
    using RG = RoundGeo<NB, K, M>;
    ex.each([&](int tid) {
      for (int u = tid; u < RG::UNITS; u += Geo<NB>::NT) round_unit<NB, K, M, true>(Y, u, tw);
    });
    ex.sync();
    inv_rounds<NB, R - 1>(ex, Y, tw);
  }
}

template <int NB, bool CACHED, class Ex>
SXC_HD inline void forward_channel(Ex& ex, char* sm, const sxc_tables* tab, const Twiddle& tw, const SeqView& qs,
                                   int ch, const float* given, const TailRef& tail,
                                   int rounds = Schedule<NB>::NR) {
  using SM = Smem<NB>;
  using G = Geo<NB>;
  float* Y = (float*)(sm + SM::Y);
  ex.each([&](int g) {
    if (g >= G::GROUPS) return;
    const Regs<NB>& rg = ex.template reg<Regs<NB>>(g);
    float v[G::SA];
    if constexpr (CACHED) {
      static_for<0, G::SA / 4>([&](auto W) {
        const uint32_t w = fresh(rg.cls[decltype(W)::value]);
        static_for<0, 4>([&](auto B) {
          constexpr int r = 4 * decltype(W)::value + decltype(B)::value;
          const int p = stageA_pos<NB>(g, r);
          v[r] = given ? given[p] : encode_class<NB>(sm, qs.L, (w >> (8 * (r % 4))) & 255u, p, ch, tail);
        });
      });
    } else {
      static_for<0, G::SA>([&](auto I) {
        constexpr int r = decltype(I)::value;
        const int p = stageA_pos<NB>(g, r);
        v[r] = given ? given[p] : encode_at<NB>(sm, qs, p, ch, tail);
      });
    }
    stageA_forward<G::SA>(v, tw);
    float* ys = Y + pad(G::SA * g);  // pad(SA*g + r) = pad(SA*g) + r for r < SA (SA divides the pad row)
    static_for<0, G::SA>([&](auto I) { ys[decltype(I)::value] = v[decltype(I)::value]; });
  });
  ex.sync();
  fwd_rounds<NB, 0>(ex, Y, tw, rounds);
}

// One workgroup per chunk: CCSignal::SetSequence(chunk, N) (reverse-complemented for
// strand 1) into 4*N floats of global memory.
template <int NB, class Ex>
SXC_HD inline void encode_body(Ex& ex, char* sm, const EncodeArgs& a, int k) {
  using G = Geo<NB>;
  const sxc_chunk ck = a.chunks[k];
  const int strand = a.strands ? a.strands[k] : a.strand;
  const SeqView qs = phase_prepare<NB>(ex, sm, a.tab, a.ent_lut, ent_row(a.ent_rows, ck, strand), a.seq + ck.off,
                                       ck.len, ck.flags & 1, strand);
  const TailRef none{nullptr, 0, nullptr};
  ex.each([&](int tid) {
    for (int ch = 0; ch < 4; ++ch)
      for (int p = tid; p < G::N; p += G::NT) a.out[((size_t)k * 4 + ch) * G::N + p] = encode_at<NB>(sm, qs, p, ch, none);
  });
}

// ---------------------------------------------------------------- target spectra

// One workgroup per chunk: the four channel spectra (FFTReal layout) to global memory.
template <int NB, class Ex>
SXC_HD inline void spectra_body(Ex& ex, char* sm, const SpecArgs& a, int k) {
  using SM = Smem<NB>;
  using G = Geo<NB>;
  const sxc_chunk ck = a.chunks[a.which[k]];
  const Twiddle tw{twiddle_table<NB>(sm, a.tab), a.osc};
  SeqView qs{nullptr, nullptr, ck.len, 0};
  if (!a.signals) {
    qs = phase_prepare<NB>(ex, sm, a.tab, a.ent_lut, ent_row(a.ent_rows, ck, a.strand), a.seq + ck.off, ck.len,
                           ck.flags & 1, a.strand);
  } else if constexpr (!SM::LEAN) {
    float* lut = (float*)(sm + SM::LUT);
    ex.each([&](int tid) {
      for (int i = tid; i < SXC_LUT_LEN; i += G::NT) lut[i] = a.tab->lut[i];
    });
    ex.sync();
  }
  float* Y = (float*)(sm + SM::Y);
  float* out = a.spectra + (size_t)a.dst[k] * 4 * G::N;
  const TailRef tail =
      tail_ref<NB>(a.tails, a.segs, a.src_sig, a.chunk_tail ? a.chunk_tail[a.which[k]] : -1);
  if (!a.signals) class_cache<NB>(ex, qs);
  for (int ch = 0; ch < 4; ++ch) {
    const float* given = a.signals ? a.signals + ((size_t)k * 4 + ch) * G::N : nullptr;
    forward_channel<NB, true>(ex, sm, a.tab, tw, qs, ch, given, tail);
    ex.each([&](int tid) {
      static_assert(G::NT % PAD_ROW == 0, "");
      const float* yt = Y + pad(tid);  // pad(tid + k*NT) = pad(tid) + pad_step(k*NT)
      static_for<0, G::N / G::NT>([&](auto K) {
        constexpr int k = decltype(K)::value;
        out[(size_t)ch * G::N + tid + k * G::NT] = yt[pad_step(k * G::NT)];
      });
    });
    ex.sync();
  }
}

// ---------------------------------------------------------------- correlate

// xc = (((0 + A) + C) + G) + T of rescaled, fftshifted channel correlations, left in
// natural order in Y.  Per-thread accumulators stay in registers across channels.
template <int NB, class Ex>
SXC_HD inline void correlate(Ex& ex, char* sm, const sxc_tables* tab, const Twiddle& tw, const SeqView& qs,
                             const float* T4, const float* qsig4, const TailRef& tail,
                             const float* qspec4 = nullptr) {
  using SM = Smem<NB>;
  using G = Geo<NB>;
  float* Y = (float*)(sm + SM::Y);
  ex.each([&](int g) {
    Regs<NB>& r = ex.template reg<Regs<NB>>(g);
    static_for<0, G::SA>([&](auto I) { r.acc[decltype(I)::value] = 0.f; });
  });
  const float mul = (float)(1.0 / (double)G::N);  // FFTReal::rescale (FFTReal.hpp:264)
  if constexpr (pair_class_cache<NB>()) {
    if (!qspec4 && !qsig4) class_cache<NB>(ex, qs);
  }
  for (int ch = 0; ch < 4; ++ch) {
    if (qspec4) {
      // the load below writes Y with another thread mapping than the previous channel's
      // inverse stage A read it: every thread must have read before any thread writes
      if (ch > 0) ex.sync();
      const float* qch = qspec4 + (size_t)ch * G::N;
      ex.each([&](int tid) {
        float* yt = Y + pad(tid);
        static_for<0, G::N / G::NT>([&](auto K) {
          constexpr int k = decltype(K)::value;
          yt[pad_step(k * G::NT)] = qch[tid + k * G::NT];
        });
      });
      ex.sync();
      // the resident spectrum is final: product and inverse from the last round on
      product_round<NB, false>(ex, Y, tw, T4 + (size_t)ch * G::N);
    } else {
      // all forward rounds but the last, which runs with the product and the first inverse round
      forward_channel<NB, pair_class_cache<NB>()>(ex, sm, tab, tw, qs, ch,
                                                  qsig4 ? qsig4 + (size_t)ch * G::N : nullptr, tail,
                                                  Schedule<NB>::NR - 1);
      product_round<NB, true>(ex, Y, tw, T4 + (size_t)ch * G::N);
    }
    inv_rounds<NB, Schedule<NB>::NR - 2>(ex, Y, tw);
    ex.each([&](int g) {
      if (g >= G::GROUPS) return;
      Regs<NB>& r = ex.template reg<Regs<NB>>(g);
      float v[G::SA], o[G::SA];
      const float* ys = Y + pad(G::SA * g);
      static_for<0, G::SA>([&](auto I) { v[decltype(I)::value] = ys[decltype(I)::value]; });
      stageA_inverse<G::SA>(v, o, tw);
      static_for<0, G::SA>([&](auto I) {
        constexpr int k = decltype(I)::value;
        const float x = o[k] * mul;
        r.acc[k] = r.acc[k] + x;
      });
    });
    // the next channel's forward stage A writes the same positions from the same thread (no
    // barrier needed); loading a resident query spectrum does not (barrier above)
  }
  ex.sync();
  ex.each([&](int g) {
    if (g >= G::GROUPS) return;
    Regs<NB>& r = ex.template reg<Regs<NB>>(g);
    // fftshift by N/2 (analysis/CrossCorr.cc:500-505): slot k of group g goes to
    // stageA_pos(g, k) ^ N/2 = X_k | bitrev(g), X_k = stageA_hi(k) ^ N/2 a multiple of N/SA;
    // when N/SA is a multiple of the pad row, pad(X_k | rg) = pad(rg) + pad_step(X_k)
    constexpr int LOW = G::LOG - G::LSA;
    const int rg = (int)bitrev((unsigned)g, LOW);
    float* yg = Y + pad(rg);
    static_for<0, G::SA>([&](auto I) {
      constexpr int k = decltype(I)::value;
      if constexpr (LOW >= PAD_SHIFT) {
        yg[pad_step(stageA_hi<NB>(k) ^ G::H)] = r.acc[k];
      } else {
        Y[pad(stageA_pos<NB>(g, k) ^ G::H)] = r.acc[k];
      }
    });
  });
  ex.sync();
}

// ---------------------------------------------------------------- peaks

template <int NB, class Ex>
SXC_HD inline void peaks(Ex& ex, char* sm, double cutoff) {
  using SM = Smem<NB>;
  using G = Geo<NB>;
  constexpr int N = G::N;
  constexpr int NP = SM::NENV;
  constexpr int CS = N / G::NT;  // elements per thread, <= 32
  static_assert(CS <= 32, "flag mask is 32 bits");
  float* Y = (float*)(sm + SM::Y);
  double* env = (double*)(sm + SM::ENV);
  int* cnt = (int*)(sm + SM::CNT);
  int* misc = (int*)(sm + SM::MISC);
  ex.each([&](int tid) {
    if (tid >= NP) return;
    double e = 0.;
    if (NP > 8) {
      // the reference's sequential sum; the loads are issued 16 at a time ahead of it
      static_assert(SXC_ENV_SIZE % 16 == 0, "");
      const float* yb = Y + pad(tid * SXC_ENV_SIZE);  // + pad(i) for i < 256 (a multiple of 32 before)
      for (int i0 = 0; i0 < SXC_ENV_SIZE; i0 += 16) {
        float sq[16];
        static_for<0, 16>([&](auto K) {
          const float x = yb[pad(i0 + decltype(K)::value)];
          sq[decltype(K)::value] = x * x;
        });
        static_for<0, 16>([&](auto K) { e += (double)sq[decltype(K)::value]; });
      }
    }
    env[tid] = sqrt(e / (double)SXC_ENV_SIZE);
  });
  ex.sync();
  unsigned mask_reg[1];
  (void)mask_reg;
  ex.each([&](int tid) {
    unsigned m = 0;
    int c = 0;
    // the CS elements of this thread lie in one envelope block and one 32-float pad row
    static_assert(SXC_ENV_SIZE % CS == 0 && (PAD_ROW % CS == 0), "");
    const float* yt = Y + pad(tid * CS);
    const double thr = env[(tid * CS) / SXC_ENV_SIZE] * cutoff + 1.;
    static_for<0, CS>([&](auto K) {
      constexpr int k = decltype(K)::value;
      if ((double)yt[k] > thr) {
        m |= 1u << k;
        ++c;
      }
    });
    cnt[tid] = c;
    ex.template reg<Regs<NB>>(tid).acc[0] = __builtin_bit_cast(float, m);
  });
  ex.sync();
  // exclusive prefix of the per-thread counts in two levels over blocks of 16 threads (the
  // block sums use the entropy windows' area: the encodes are done)
  constexpr int SB = 16, NBK = (G::NT + SB - 1) / SB;
  static_assert(NBK <= SM::NWIN, "block sums fit the entropy area");
  int* bsum = (int*)(sm + SM::ENT);
  ex.each([&](int tid) {
    if (tid >= NBK) return;
    int s = 0;
    for (int t = SB * tid; t < SB * tid + SB && t < G::NT; ++t) s += cnt[t];
    bsum[tid] = s;
  });
  ex.sync();
  ex.each([&](int tid) {
    if (tid >= NBK) return;
    int s = 0;
    for (int b = 0; b < tid; ++b) s += bsum[b];
    for (int t = SB * tid; t < SB * tid + SB && t < G::NT; ++t) {
      const int c = cnt[t];
      cnt[t] = s;
      s += c;
    }
    if (tid == NBK - 1) misc[MISC_NCAND] = s;
  });
  ex.sync();
  uint16_t* cand = (uint16_t*)(sm + SM::CAND);
  ex.each([&](int tid) {
    unsigned m = __builtin_bit_cast(unsigned, ex.template reg<Regs<NB>>(tid).acc[0]);
    int o = cnt[tid];
    for (int k = 0; k < CS; ++k)
      if (m & (1u << k)) cand[o++] = (uint16_t)(tid * CS + k);
  });
  ex.sync();
}

// ---------------------------------------------------------------- fused correlate body

// One workgroup per (pair, strand): encode -> FFT -> product -> iFFT -> channel sum ->
// FindTop; the accepted lags (ascending) go to global memory for the scan kernel.
template <int NB, class Ex>
SXC_HD inline void correlate_body(Ex& ex, char* sm, const PairArgs& a, int ps) {
  using SM = Smem<NB>;
  using G = Geo<NB>;
  const int pair = ps / a.strands;
  const int strand = ps % a.strands;
  const sxc_pair_dev pr = a.pairs[pair];
  const sxc_chunk qc = a.qchunks[pr.q_chunk];
  const Twiddle tw{twiddle_table<NB>(sm, a.tab), a.osc};
  if (a.qspectra) {
    // resident query spectra: no encode; only the twiddle table and the counters to set up
    float* lut = (float*)(sm + SM::LUT);
    int* misc0 = (int*)(sm + SM::MISC);
    ex.each([&](int tid) {
      if constexpr (!SM::LEAN)
        for (int i = tid; i < SXC_LUT_LEN; i += G::NT) lut[i] = a.tab->lut[i];
      if (tid < 16) misc0[tid] = 0;
    });
    ex.sync();
    correlate<NB>(ex, sm, a.tab, tw, SeqView{nullptr, nullptr, qc.len, 0}, a.spectra + (size_t)pr.t_chunk * 4 * G::N,
                  nullptr, TailRef{nullptr, 0, nullptr},
                  a.qspectra + ((size_t)pr.q_chunk * 2 + strand) * 4 * G::N);
  } else {
    const SeqView qs = phase_prepare<NB>(ex, sm, a.tab, a.ent_lut, ent_row(a.ent_rows, qc, strand), a.seq + qc.off,
                                         qc.len, qc.flags & 1, strand);
    correlate<NB>(ex, sm, a.tab, tw, qs, a.spectra + (size_t)pr.t_chunk * 4 * G::N, nullptr,
                  tail_ref<NB>(a.tails, a.segs, a.src_sig, pr.q_tail));
  }
  if (a.xc_out) {
    const float* Y = (const float*)(sm + SM::Y);
    ex.each([&](int tid) {
      for (int p = tid; p < G::N; p += G::NT) a.xc_out[(size_t)ps * G::N + p] = Y[pad(p)];
    });
  }
  peaks<NB>(ex, sm, pr.fast ? a.cutoff_fast : a.cutoff);
  const int* misc = (const int*)(sm + SM::MISC);
  const uint16_t* cand = (const uint16_t*)(sm + SM::CAND);
  const int n = misc[MISC_NCAND];
  uint16_t* dst = a.cands + (size_t)ps * G::N;
  ex.each([&](int tid) {
    for (int i = tid; i < n; i += G::NT) dst[i] = cand[i];
    if (tid == 0) {
      a.ncand[ps] = n;
      if (a.counts) a.counts[ps] = 0;
    }
  });
}

// FindTop on a given xc (the first half of SeqAnalyzer::MatchUp, analysis/CrossCorr.cc:583-592).
template <int NB, class Ex>
SXC_HD inline void peaks_body(Ex& ex, char* sm, const float* xc, double cutoff, uint16_t* cands, int32_t* ncand) {
  using SM = Smem<NB>;
  using G = Geo<NB>;
  float* Y = (float*)(sm + SM::Y);
  int* misc0 = (int*)(sm + SM::MISC);
  ex.each([&](int tid) {
    for (int p = tid; p < G::N; p += G::NT) Y[pad(p)] = xc[p];
    if (tid < 16) misc0[tid] = 0;
  });
  ex.sync();
  peaks<NB>(ex, sm, cutoff);
  const int* misc = (const int*)(sm + SM::MISC);
  const uint16_t* cand = (const uint16_t*)(sm + SM::CAND);
  const int n = misc[MISC_NCAND];
  ex.each([&](int tid) {
    for (int i = tid; i < n; i += G::NT) cands[i] = cand[i];
    if (tid == 0) *ncand = n;
  });
}

}  // namespace sxc
