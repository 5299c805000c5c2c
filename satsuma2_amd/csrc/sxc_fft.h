// sxc_fft.h -- FFTReal<float>'s exact dataflow, blocked for a CDNA4 workgroup.
//
// FFTReal (extern/RealFFT/FFTReal.hpp) computes a length-N real DFT with a bit-reversed
// radix-4 first pass (pass 1+2, :466-495), an 8-point pass (pass 3, :499-528) and one
// radix-2 level per bit above that (pass_n, n = 3..nbits-1, :553-656).  Every output
// element is produced by a fixed sequence of IEEE float ops; any schedule that applies
// the same ops to the same operands gives the same bits.  This file regroups the
// butterflies so one thread owns a closed set of positions for several levels:
//
//   stage A : levels pass1+2, pass3, pass_3 (and pass_4 when SA = 32) on SA contiguous
//             positions (the bit-reversed gather of pass 1+2 is done while loading);
//   rounds  : levels pass_m .. pass_{m+K-1} (K <= KMAX) on "units" of 2^(K+2) positions.
//
// A unit at level m is butterfly index j and h_m - j (h_m = 2^(m-1)) in 2^(K-1)
// neighbouring blocks; level m+1 butterflies j' and 2h_m - j' read exactly what level-m
// butterfly j wrote in two blocks, so the set closes over K levels and is in place.
// Index j = 0 and j = h_m/2 form one more unit of the same size.  The inverse
// (FFTReal.hpp:661-906) runs the same units with levels in reverse order.
//
// Twiddles: FFTReal's per-level LUT (float)cos(i*PI/2^L) equals the level-12 table at
// stride 2^(12-L) (same double argument); the host lays the levels out back to back so a
// unit's lanes read consecutive entries of one level.  Levels > 12 use FFTReal's float oscillator (OscSinCos.hpp:88-104); its values
// depend only on the butterfly index and are tabulated by the host.
//
// All arithmetic here must be compiled without FP contraction (no FMA): FFTReal is built
// with plain -O3 and no -march, so it never fuses a multiply into an add.
#pragma once
#include <type_traits>

#include "sxc_types.h"

#if defined(__HIP__)
#define SXC_HD __host__ __device__
#else
#define SXC_HD
#endif

#pragma clang fp contract(off)

namespace sxc {

template <int B, int E, class F>
SXC_HD inline void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// LDS index of FFT position p: one pad word per row of PAD_ROW floats keeps the 32-float rows
// of stage A on distinct banks.
#ifndef SXC_PAD_SHIFT
#define SXC_PAD_SHIFT 5
#endif
constexpr int PAD_SHIFT = SXC_PAD_SHIFT;
constexpr int PAD_ROW = 1 << PAD_SHIFT;
SXC_HD constexpr int pad(int p) { return p + (p >> PAD_SHIFT); }
// pad(p + k*PAD_ROW*m) = pad(p) + k*(PAD_ROW + 1)*m: the index step of `d` positions, d a
// multiple of PAD_ROW
SXC_HD constexpr int pad_step(int d) { return d + d / PAD_ROW; }
// Stage-A group size (positions per thread in the first levels) and the largest number
// of levels one LDS round carries.
#ifndef SXC_STAGE_A
#define SXC_STAGE_A 16
#endif
#ifndef SXC_KMAX
#define SXC_KMAX 3
#endif

// A kernel's FFT geometry packed in one template int ("geometry code"): log2 of the
// transform length in bits 0-4, bit 5 set for a 32-wide stage A, the LDS round depth in
// bits 6-7, bit 8 asking the compiler for a register budget that fits as many workgroups
// per CU as the LDS plan allows (up to three), bit 9 for a lean LDS plan (twiddles and
// chunk bytes read through the caches instead of staged in LDS, so more workgroups fit a
// CU).  A plain log2 length (bits 5-9 clear) is the default geometry for that length.
// Every geometry computes the same FFTReal dataflow; they differ in how positions map to
// threads, how many levels one LDS round carries and how many workgroups share a CU.
SXC_HD constexpr int geo_code(int nb, int sa, int km, int occ = 0, int lean = 0) {
  return nb | (sa == 32 ? 32 : 0) | (km << 6) | (occ ? 256 : 0) | (lean ? 512 : 0);
}
SXC_HD constexpr int geo_lean(int g) { return (g >> 9) & 1; }
SXC_HD constexpr int geo_nb(int g) { return g & 31; }
SXC_HD constexpr int geo_sa(int g) {
  return ((g & 32) || geo_nb(g) >= 15) ? 32 : (((g >> 6) & 3) ? 16 : SXC_STAGE_A);
}
SXC_HD constexpr int geo_km(int g) { return ((g >> 6) & 3) ? ((g >> 6) & 3) : SXC_KMAX; }
SXC_HD constexpr int geo_occ(int g) { return (g >> 8) & 1; }

template <int NB>
struct Geo {
  static constexpr int LOG = geo_nb(NB);                   // log2(N)
  static constexpr int N = 1 << LOG;
  static constexpr int H = N / 2;
  // 16 or 32 positions per thread in stage A; N = 32768 needs 32 to stay at 1024 threads
  static constexpr int SA = geo_sa(NB);
  static constexpr int KM = geo_km(NB);                    // levels per LDS round at most
  static constexpr int LSA = SA == 32 ? 5 : 4;             // log2(SA)
  static constexpr int NT = (N / SA < 64) ? 64 : N / SA;   // threads per workgroup
  static constexpr int YLEN = N + N / PAD_ROW;            // padded LDS floats
  static constexpr int GROUPS = N / SA;                    // stage-A units
  static_assert(NT <= 1024, "a workgroup has at most 1024 threads");
};

SXC_HD inline unsigned bitrev(unsigned x, int bits) {
  unsigned r = 0;
  for (int b = 0; b < bits; ++b) r |= ((x >> b) & 1u) << (bits - 1 - b);
  return r;
}

// FFTReal's sqrt2_2 = DataType(SQRT2 * 0.5) (FFTReal.hpp:506, def.h:43)
#define SXC_SQRT2_2 ((float)(1.41421356237309514547462185873883 * 0.5))

struct Twiddle {
  const float* lut;  // per-level cosine tables (SXC_LUT_OFF), the LDS copy or the constant table
  const float* osc;  // {cos, sin} after i oscillator steps: level 13 at [0, 8192), level 14 at [8192, 24576)
  template <int LEVEL>
  SXC_HD inline void get(int beta, int hbeta, float& c, float& s) const {
    if constexpr (LEVEL <= 12) {
      c = lut[SXC_LUT_OFF(LEVEL) + beta];
      s = lut[SXC_LUT_OFF(LEVEL) + hbeta];
    } else {
      const float* t = osc + (LEVEL == 13 ? 0 : 8192) + 2 * beta;
      c = t[0];
      s = t[1];
    }
  }
};

// ---------------------------------------------------------------- butterflies

// compute_direct_pass_n body (FFTReal.hpp:589-595): inputs at block positions
// i, h+i, nbr+i, nbr+h+i; outputs at i, nbr-i, nbr+i, 2nbr-i.
SXC_HD inline void bfly_fwd(float sf1r, float sf1i, float sf2r, float sf2i, float c, float s, float& o0,
                            float& o1, float& o2, float& o3) {
  float v = sf2r * c - sf2i * s;
  o0 = sf1r + v;
  o1 = sf1r - v;
  v = sf2r * s + sf2i * c;
  o2 = v + sf1i;
  o3 = v - sf1i;
}

// compute_inverse_pass_n body (FFTReal.hpp:761-770): inputs at i, nbr-i, nbr+i, 2nbr-i;
// outputs at i, h+i, nbr+i, nbr+h+i.
SXC_HD inline void bfly_inv(float xr, float xm, float xp, float x2, float c, float s, float& o0, float& o1,
                            float& o2, float& o3) {
  o0 = xr + xm;
  o1 = xp - x2;
  const float vr = xr - xm;
  const float vi = xp + x2;
  o2 = vr * c + vi * s;
  o3 = vi * c - vr * s;
}

// One radix-2 level (pass_n with nbr = 2^P) over NBLK contiguous blocks held in v.
template <int P, int NBLK, bool INV>
SXC_HD inline void level_contig(float* v, const Twiddle& tw) {
  constexpr int nbr = 1 << P, h = nbr >> 1;
  float o[NBLK * 2 * nbr];
  static_for<0, NBLK * 2 * nbr>([&](auto I) { o[decltype(I)::value] = v[decltype(I)::value]; });
  static_for<0, NBLK>([&](auto B) {
    constexpr int ci = decltype(B)::value * 2 * nbr;
    o[ci] = v[ci] + v[ci + nbr];
    o[ci + nbr] = v[ci] - v[ci + nbr];
    if constexpr (INV) {
      o[ci + h] = v[ci + h] * 2;
      o[ci + nbr + h] = v[ci + nbr + h] * 2;
    }
    static_for<1, h>([&](auto I) {
      constexpr int i = decltype(I)::value;
      float c, s;
      tw.template get<P>(i, h - i, c, s);
      if constexpr (!INV)
        bfly_fwd(v[ci + i], v[ci + h + i], v[ci + nbr + i], v[ci + nbr + h + i], c, s, o[ci + i], o[ci + nbr - i],
                 o[ci + nbr + i], o[ci + 2 * nbr - i]);
      else
        bfly_inv(v[ci + i], v[ci + nbr - i], v[ci + nbr + i], v[ci + 2 * nbr - i], c, s, o[ci + i], o[ci + h + i],
                 o[ci + nbr + i], o[ci + nbr + h + i]);
    });
  });
  static_for<0, NBLK * 2 * nbr>([&](auto I) { v[decltype(I)::value] = o[decltype(I)::value]; });
}

// ---------------------------------------------------------------- stage A

SXC_HD constexpr unsigned bitrev_c(unsigned x, int bits) {
  unsigned r = 0;
  for (int b = 0; b < bits; ++b) r |= ((x >> b) & 1u) << (bits - 1 - b);
  return r;
}

// The part of stageA_pos(g, r) that depends on r alone (its low bits come from g).
template <int NB>
SXC_HD constexpr int stageA_hi(int r) {
  return (int)(bitrev_c((unsigned)r, Geo<NB>::LSA) << (Geo<NB>::LOG - Geo<NB>::LSA));
}

// Time position read by slot r of stage-A group g (pass 1+2 reads x[br[4k+m]]).
template <int NB>
SXC_HD inline int stageA_pos(int g, int r) {
  constexpr int L = Geo<NB>::LSA, LG = Geo<NB>::LOG;
  return (int)((bitrev((unsigned)r, L) << (LG - L)) | bitrev((unsigned)g, LG - L));
}

// Forward levels pass1+2, pass3, pass_3 (and pass_4 for SA = 32) (FFTReal.hpp:466-528,
// 553-601).  v[r] holds the time sample at stageA_pos(g, r) on entry, spectrum position
// SA*g + r on exit.
template <int SA>
SXC_HD inline void stageA_forward(float* v, const Twiddle& tw) {
  float d[SA];
  static_for<0, SA / 4>([&](auto K) {
    constexpr int b = 4 * decltype(K)::value;
    d[b + 1] = v[b + 0] - v[b + 1];
    d[b + 3] = v[b + 2] - v[b + 3];
    const float s0 = v[b + 0] + v[b + 1];
    const float s2 = v[b + 2] + v[b + 3];
    d[b + 0] = s0 + s2;
    d[b + 2] = s0 - s2;
  });
  const float sq = SXC_SQRT2_2;
  static_for<0, SA / 8>([&](auto K) {
    constexpr int c = 8 * decltype(K)::value;
    float t;
    v[c + 0] = d[c + 0] + d[c + 4];
    v[c + 4] = d[c + 0] - d[c + 4];
    v[c + 2] = d[c + 2];
    v[c + 6] = d[c + 6];
    t = (d[c + 5] - d[c + 7]) * sq;
    v[c + 1] = d[c + 1] + t;
    v[c + 3] = d[c + 1] - t;
    t = (d[c + 5] + d[c + 7]) * sq;
    v[c + 5] = t + d[c + 3];
    v[c + 7] = t - d[c + 3];
  });
  level_contig<3, SA / 16, false>(v, tw);
  if constexpr (SA == 32) level_contig<4, 1, false>(v, tw);
}

// Inverse of stage A: (pass_4), pass_3 inverse (FFTReal.hpp:729-776), inverse pass 3
// (:834-862), inverse pass 1+2 (:866-906).  On exit out[r] is the time sample for
// position stageA_pos(g, r) (the bit-reversed store of pass 1+2).
template <int SA>
SXC_HD inline void stageA_inverse(float* v, float* out, const Twiddle& tw) {
  if constexpr (SA == 32) level_contig<4, 1, true>(v, tw);
  level_contig<3, SA / 16, true>(v, tw);
  const float sq = SXC_SQRT2_2;
  float d[SA];
  static_for<0, SA / 8>([&](auto K) {
    constexpr int c = 8 * decltype(K)::value;
    d[c + 0] = v[c + 0] + v[c + 4];
    d[c + 4] = v[c + 0] - v[c + 4];
    d[c + 2] = v[c + 2] * 2;
    d[c + 6] = v[c + 6] * 2;
    d[c + 1] = v[c + 1] + v[c + 3];
    d[c + 3] = v[c + 5] - v[c + 7];
    const float vr = v[c + 1] - v[c + 3];
    const float vi = v[c + 5] + v[c + 7];
    d[c + 5] = (vr + vi) * sq;
    d[c + 7] = (vi - vr) * sq;
  });
  static_for<0, SA / 4>([&](auto K) {
    constexpr int b = 4 * decltype(K)::value;
    const float b0 = d[b + 0] + d[b + 2];
    const float b2 = d[b + 0] - d[b + 2];
    const float b1 = d[b + 1] * 2;
    const float b3 = d[b + 3] * 2;
    out[b + 0] = b0 + b1;
    out[b + 1] = b0 - b1;
    out[b + 2] = b2 + b3;
    out[b + 3] = b2 - b3;
  });
}

// ---------------------------------------------------------------- rounds

SXC_HD constexpr int slot(int A, int s) { return 2 * A - (s < 0 ? 1 : 0); }

// Regular unit j (1 <= j < h_m/2) of a K-level round starting at level M.  Slot (A, s)
// holds position A*h_m + s*j of the unit's big block.
template <int K, int M, int L, bool INV>
SXC_HD inline void unit_level_regular(float* v, int j, const Twiddle& tw) {
  constexpr int S = 1 << (K + 2);
  constexpr int HM = 1 << (M - 1);
  constexpr int HL = HM << L;
  float o[S];
  static_for<0, (1 << (K - 1 - L))>([&](auto C) {
    constexpr int base = decltype(C)::value << (L + 2);
    static_for<0, (1 << L) + 1>([&](auto AL) {
      constexpr int al = decltype(AL)::value;
      static_for<0, 2>([&](auto SGI) {
        constexpr int sg = decltype(SGI)::value == 0 ? 1 : -1;
        if constexpr (!((al == 0 && sg < 0) || (al == (1 << L) && sg > 0))) {
          const int beta = al * HM + sg * j;
          float c, s;
          tw.template get<M + L>(beta, HL - beta, c, s);
          constexpr int i0 = slot(base + al, sg);
          constexpr int i1 = slot(base + al + (1 << L), sg);
          constexpr int i2 = slot(base + al + (2 << L), sg);
          constexpr int i3 = slot(base + al + 3 * (1 << L), sg);
          constexpr int o0 = slot(base + al, sg);
          constexpr int o1 = slot(base + (2 << L) - al, -sg);
          constexpr int o2 = slot(base + (2 << L) + al, sg);
          constexpr int o3 = slot(base + (4 << L) - al, -sg);
          if constexpr (!INV)
            bfly_fwd(v[i0], v[i1], v[i2], v[i3], c, s, o[o0], o[o1], o[o2], o[o3]);
          else
            bfly_inv(v[o0], v[o1], v[o2], v[o3], c, s, o[i0], o[i1], o[i2], o[i3]);
        }
      });
    });
  });
  static_for<0, S>([&](auto I) { v[decltype(I)::value] = o[decltype(I)::value]; });
}

// The unit made of butterflies 0 and h_m/2 (and their descendants): slot q holds
// position q * h_m/2 of the big block.
template <int K, int M, int L, bool INV>
SXC_HD inline void unit_level_special(float* v, const Twiddle& tw) {
  constexpr int S = 1 << (K + 2);
  constexpr int HM = 1 << (M - 1);
  constexpr int HL = HM << L;
  float o[S];
  static_for<0, S>([&](auto I) { o[decltype(I)::value] = v[decltype(I)::value]; });
  static_for<0, (1 << (K - 1 - L))>([&](auto C) {
    constexpr int base = decltype(C)::value << (L + 3);
    constexpr int q0 = base, q1 = base + (4 << L);
    // extreme coefficients (FFTReal.hpp:575-578 / 751-754)
    o[q0] = v[q0] + v[q1];
    o[q1] = v[q0] - v[q1];
    if constexpr (INV) {
      constexpr int qa = base + (2 << L), qb = base + 3 * (2 << L);
      o[qa] = v[qa] * 2;
      o[qb] = v[qb] * 2;
    }
    static_for<1, (2 << L)>([&](auto G) {
      constexpr int g = decltype(G)::value;
      constexpr int beta = g * (HM / 2);
      float c, s;
      tw.template get<M + L>(beta, HL - beta, c, s);
      constexpr int i0 = base + g, i1 = base + g + (2 << L), i2 = base + g + (4 << L), i3 = base + g + 3 * (2 << L);
      constexpr int o0 = base + g, o1 = base + (4 << L) - g, o2 = base + (4 << L) + g, o3 = base + (8 << L) - g;
      if constexpr (!INV)
        bfly_fwd(v[i0], v[i1], v[i2], v[i3], c, s, o[o0], o[o1], o[o2], o[o3]);
      else
        bfly_inv(v[o0], v[o1], v[o2], v[o3], c, s, o[i0], o[i1], o[i2], o[i3]);
    });
  });
  static_for<0, S>([&](auto I) { v[decltype(I)::value] = o[decltype(I)::value]; });
}

// Round geometry for NB with K levels from M.
template <int NB, int K, int M>
struct RoundGeo {
  static constexpr int N = Geo<NB>::N;
  static constexpr int HM = 1 << (M - 1);
  static constexpr int UPB = HM / 2;                  // units per big block
  static constexpr int BIG = (1 << (K + 1)) * HM;     // big block length
  static constexpr int NBIG = N / BIG;
  static constexpr int NREG = NBIG * (UPB - 1);       // regular units
  static constexpr int UNITS = NBIG * UPB;            // = N / 2^(K+2)
  static constexpr int S = 1 << (K + 2);
};

// LDS indices of a unit's positions split into a per-thread base and a compile-time constant,
// so every slot is one LDS access with an immediate offset.  A regular unit j (1 <= j < h_m/2)
// of the big block starting at `off` (a multiple of the pad row R) touches off + A*h_m + s*j,
// s = +-1:
//   pad(off + A*h_m + j) = pad(off) + pj + cp(A),   pad(off + A*h_m - j) = pad(off) - mj + cm(A)
// with, for h_m >= R (h_m a multiple of R): pj = pad(j), mj = j + ceil(j/R),
// cp(A) = cm(A) = pad_step(A*h_m); for h_m < R (h_m divides R, j < h_m): pj = mj = j,
// cp(A) = pad(A*h_m), cm(A) = A*h_m + floor((A*h_m - 1)/R) (the pad word of the row holding
// A*h_m - 1).
template <int HM>
SXC_HD constexpr int pad_cp(int A) {
  return HM >= PAD_ROW ? pad_step(A * HM) : A * HM + ((A * HM) >> PAD_SHIFT);
}
template <int HM>
SXC_HD constexpr int pad_cm(int A) {
  return HM >= PAD_ROW ? pad_step(A * HM) : A * HM + ((A * HM - 1) >> PAD_SHIFT);
}
template <int HM>
SXC_HD inline int pad_pj(int j) {
  return HM >= PAD_ROW ? j + (j >> PAD_SHIFT) : j;
}
template <int HM>
SXC_HD inline int pad_mj(int j) {
  return HM >= PAD_ROW ? j + ((j + PAD_ROW - 1) >> PAD_SHIFT) : j;
}

// Runs unit u of a round on the padded LDS array Y.  Units [0, NREG) are regular,
// [NREG, UNITS) special, so only one wavefront per round mixes the two code paths.
template <int NB, int K, int M, bool INV>
SXC_HD inline void round_unit(float* Y, int u, const Twiddle& tw) {
  using RG = RoundGeo<NB, K, M>;
  constexpr int HM = RG::HM;
  static_assert(RG::BIG % PAD_ROW == 0, "big blocks start on a pad boundary");
  float v[RG::S];
  if (u < RG::NREG) {
    const int b = u / (RG::UPB - 1);
    const int j = 1 + u % (RG::UPB - 1);
    const int p0 = pad(b * RG::BIG);
    float* yp = Y + (p0 + pad_pj<HM>(j));  // slots with s = +1
    float* ym = Y + (p0 - pad_mj<HM>(j));  // slots with s = -1
    static_for<0, RG::S>([&](auto I) {
      constexpr int q = decltype(I)::value;
      constexpr int A = (q + 1) >> 1;
      if constexpr (q & 1) v[q] = ym[pad_cm<HM>(A)]; else v[q] = yp[pad_cp<HM>(A)];
    });
    if constexpr (!INV) {
      static_for<0, K>([&](auto L) { unit_level_regular<K, M, decltype(L)::value, false>(v, j, tw); });
    } else {
      static_for<0, K>([&](auto L) { unit_level_regular<K, M, K - 1 - decltype(L)::value, true>(v, j, tw); });
    }
    static_for<0, RG::S>([&](auto I) {
      constexpr int q = decltype(I)::value;
      constexpr int A = (q + 1) >> 1;
      if constexpr (q & 1) ym[pad_cm<HM>(A)] = v[q]; else yp[pad_cp<HM>(A)] = v[q];
    });
  } else {
    // positions off + q*h_m/2: pad(off) + pad(q*h_m/2), off a multiple of 32
    float* y0 = Y + pad((u - RG::NREG) * RG::BIG);
    static_for<0, RG::S>([&](auto I) {
      constexpr int q = decltype(I)::value;
      v[q] = y0[pad(q * (HM / 2))];
    });
    if constexpr (!INV) {
      static_for<0, K>([&](auto L) { unit_level_special<K, M, decltype(L)::value, false>(v, tw); });
    } else {
      static_for<0, K>([&](auto L) { unit_level_special<K, M, K - 1 - decltype(L)::value, true>(v, tw); });
    }
    static_for<0, RG::S>([&](auto I) {
      constexpr int q = decltype(I)::value;
      y0[pad(q * (HM / 2))] = v[q];
    });
  }
}

// The last round's big block is the whole transform (M + K = log2 N), so a unit holds both
// halves of each of its bins: slot (A, s) and slot (A + 2^K, s) are bins i and i + N/2, the
// real and the stored imaginary part (FFTReal's packed layout).  These run the round's
// forward levels (when FWD), the conjugate product with the target spectrum T of
// CrossCorrelation::DoOne (analysis/CrossCorr.cc:477-492; ComplexMult :20-26) on the unit's
// bins, and the same round's inverse levels, without leaving registers.  Bins N/2 - 1 and
// N/2 (the imaginary part of bin 0) take T's values unchanged, as the reference leaves them.
template <int NB, int K, int M, bool FWD>
SXC_HD inline void round_unit_product(float* Y, int u, const Twiddle& tw, const float* T) {
  using RG = RoundGeo<NB, K, M>;
  constexpr int HM = RG::HM, S = RG::S, HALF = 1 << K;
  constexpr int H = Geo<NB>::H;
  static_assert(M + K == Geo<NB>::LOG && RG::NBIG == 1, "the last round spans the transform");
  float v[S];
  if (u < RG::NREG) {
    const int j = 1 + u;
    const int p0 = 0;
    float* yp = Y + (p0 + pad_pj<HM>(j));
    float* ym = Y + (p0 - pad_mj<HM>(j));
    static_for<0, S>([&](auto I) {
      constexpr int q = decltype(I)::value;
      constexpr int A = (q + 1) >> 1;
      if constexpr (q & 1) v[q] = ym[pad_cm<HM>(A)]; else v[q] = yp[pad_cp<HM>(A)];
    });
    if constexpr (FWD) static_for<0, K>([&](auto L) { unit_level_regular<K, M, decltype(L)::value, false>(v, j, tw); });
    const float* tp = T + j;
    const float* tm = T - j;
    static_for<0, S>([&](auto I) {
      constexpr int q = decltype(I)::value;
      constexpr int A = (q + 1) >> 1;
      constexpr int sg = (q & 1) ? -1 : 1;
      // real slots: A in [0, 2^K) for s = +1, [1, 2^K] for s = -1; bin i = A h_m + s j
      if constexpr ((sg > 0 && A < HALF) || (sg < 0 && A >= 1 && A <= HALF)) {
        constexpr int qi = slot(A + HALF, sg);
        const float r1 = sg > 0 ? tp[A * HM] : tm[A * HM];
        const float i1 = sg > 0 ? tp[A * HM + H] : tm[A * HM + H];
        bool quirk = false;
        if constexpr (sg < 0 && A == HALF) quirk = j == 1;  // bin N/2 - 1
        if (quirk) {
          v[q] = r1;
          v[qi] = i1;
        } else {
          const float r2 = v[q], i2 = -v[qi];
          const float fr = r1 * r2 - i1 * i2;
          const float fi = i1 * r2 + i2 * r1;
          v[q] = fr;
          v[qi] = -fi;
        }
      }
    });
    static_for<0, K>([&](auto L) { unit_level_regular<K, M, K - 1 - decltype(L)::value, true>(v, j, tw); });
    static_for<0, S>([&](auto I) {
      constexpr int q = decltype(I)::value;
      constexpr int A = (q + 1) >> 1;
      if constexpr (q & 1) ym[pad_cm<HM>(A)] = v[q]; else yp[pad_cp<HM>(A)] = v[q];
    });
  } else {
    // bins q h_m/2, q < 2^(K+1), with their imaginary parts at q + 2^(K+1)
    static_for<0, S>([&](auto I) {
      constexpr int q = decltype(I)::value;
      v[q] = Y[pad(q * (HM / 2))];
    });
    if constexpr (FWD) static_for<0, K>([&](auto L) { unit_level_special<K, M, decltype(L)::value, false>(v, tw); });
    constexpr int QH = 2 * HALF;
    static_for<0, QH>([&](auto I) {
      constexpr int q = decltype(I)::value;
      constexpr int i = q * (HM / 2);
      if constexpr (q == 0) {
        v[0] = T[0] * v[0];
        v[QH] = T[H];
      } else {
        const float r1 = T[i], i1 = T[i + H];
        const float r2 = v[q], i2 = -v[q + QH];
        const float fr = r1 * r2 - i1 * i2;
        const float fi = i1 * r2 + i2 * r1;
        v[q] = fr;
        v[q + QH] = -fi;
      }
    });
    static_for<0, K>([&](auto L) { unit_level_special<K, M, K - 1 - decltype(L)::value, true>(v, tw); });
    static_for<0, S>([&](auto I) {
      constexpr int q = decltype(I)::value;
      Y[pad(q * (HM / 2))] = v[q];
    });
  }
}

// Level schedule above stage A: levels LSA .. log2(N)-1 in rounds of KM, the remainder last.
template <int NB>
struct Schedule {
  static constexpr int M0 = Geo<NB>::LSA;
  static constexpr int KM = Geo<NB>::KM;
  static constexpr int R = Geo<NB>::LOG - M0;
  static constexpr int NR = (R + KM - 1) / KM;
  static constexpr int k(int r) { return (r < R / KM) ? KM : (R - KM * (R / KM)); }
  static constexpr int m(int r) { return M0 + KM * r; }
};

}  // namespace sxc
