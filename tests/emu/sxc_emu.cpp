// CPU workgroup emulator of the HIP kernels (test infrastructure).
//
// Runs the exact kernel bodies of satsuma2_amd/csrc/sxc_pipeline.h with a sequential
// executor: every "thread" of a workgroup runs each phase in turn, barriers are phase
// boundaries, LDS is a heap buffer.  Built with AddressSanitizer by tests/test_emu.py so
// every LDS and global index the kernels compute is bounds-checked on the CPU before the
// same code runs on a GPU.  Not part of the product.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <vector>

#include "../../satsuma2_amd/csrc/sxc_batch_host.h"
#include "../../satsuma2_amd/csrc/sxc_pipeline.h"
#include "../../satsuma2_amd/csrc/sxc_scan.h"

namespace {

// SXC_EMU_ORDER=reverse | shuffle runs the threads of every phase in that order instead of
// ascending: a kernel whose phase reads what another thread of the same phase writes (a
// missing barrier) then computes something else.
int emu_order_mode() {
  static const int m = [] {
    const char* v = getenv("SXC_EMU_ORDER");
    if (!v) return 0;
    return strcmp(v, "reverse") == 0 ? 1 : strcmp(v, "shuffle") == 0 ? 2 : 0;
  }();
  return m;
}

template <class R>
struct EmuExec {
  int nt;
  std::vector<R> regs;
  uint64_t rng = 0x9e3779b97f4a7c15ull;
  explicit EmuExec(int n) : nt(n), regs(n) {}
  template <class F>
  void each(F&& f) {
    const int mode = emu_order_mode();
    if (mode == 0) {
      for (int t = 0; t < nt; ++t) f(t);
    } else if (mode == 1) {
      for (int t = nt - 1; t >= 0; --t) f(t);
    } else {
      std::vector<int> ord(nt);
      for (int t = 0; t < nt; ++t) ord[t] = t;
      for (int t = nt - 1; t > 0; --t) {
        rng ^= rng << 13;
        rng ^= rng >> 7;
        rng ^= rng << 17;
        std::swap(ord[t], ord[(int)(rng % (uint64_t)(t + 1))]);
      }
      for (int t : ord) f(t);
    }
  }
  void sync() {}
  template <class T>
  T& reg(int tid) {
    return regs[tid];
  }
  int atomic_add_shared(int* p, int v) {
    const int o = *p;
    *p += v;
    return o;
  }
  int atomic_add_global(int* p, int v) { return atomic_add_shared(p, v); }
};

struct Tables {
  sxc_tables tab;
  std::vector<float> osc;
  std::vector<double> ent[16];
  Tables() : osc(24576) {
    sxc::build_tables(&tab);
    sxc::build_osc(osc.data());
    for (int nb = 9; nb <= 15; ++nb) {
      ent[nb].resize(std::max(1, sxc::ent_lut_len(nb)));
      sxc::build_ent_lut(nb, ent[nb].data());
    }
  }
};

int64_t g_emu_near = 0;  // records the last emulated scan flagged near min_prob

const Tables& tables() {
  static Tables t;
  return t;
}

// GEO: the geometry code of the FFT kernels (a plain log2 length: the default geometry).
template <int GEO>
int run_batch(const sxc::HostBatch& hb, const sxc_params& prm, std::vector<sxc_result_dev>& res, int cap,
              float* xc_out, int32_t* stats) {
  constexpr int NB = sxc::geo_nb(GEO);
  using G = sxc::Geo<GEO>;
  using SM = sxc::Smem<GEO>;
  const Tables& T = tables();
  std::vector<float> spectra((size_t)hb.tchunks.size() * 4 * G::N);
  std::vector<int32_t> which(hb.tchunks.size()), dst(hb.tchunks.size());
  for (size_t i = 0; i < which.size(); ++i) which[i] = dst[i] = (int)i;
  std::vector<char> sm(SM::BYTES);
  // signals of the sequences named by signal tails
  std::vector<float> src_sig((size_t)hb.schunks.size() * 4 * G::N);
  const float* erows = hb.ent.empty() ? nullptr : hb.ent.data();
  sxc::EncodeArgs ea{&T.tab, T.ent[NB].data(), erows, hb.seq.data(), hb.schunks.data(), 0, hb.sstrand.data(),
                     src_sig.data()};
  for (int k = 0; k < (int)hb.schunks.size(); ++k) {
    EmuExec<sxc::Regs<GEO>> ex(G::NT);
    memset(sm.data(), 0xCD, sm.size());
    sxc::encode_body<GEO>(ex, sm.data(), ea, k);
  }
  const bool tails = !hb.tails.empty();
  sxc::SpecArgs sa{&T.tab, T.osc.data(), T.ent[NB].data(), erows, hb.seq.data(), hb.tchunks.data(), which.data(),
                   (int)which.size(), spectra.data(), dst.data(), nullptr,
                   tails ? hb.ttail.data() : nullptr, tails ? hb.tails.data() : nullptr,
                   tails ? hb.segs.data() : nullptr, tails ? src_sig.data() : nullptr, 0};
  for (int k = 0; k < sa.n; ++k) {
    EmuExec<sxc::Regs<GEO>> ex(G::NT);
    memset(sm.data(), 0xCD, sm.size());
    sxc::spectra_body<GEO>(ex, sm.data(), sa, k);
  }
  const int total = (int)hb.pairs.size() * prm.strands;
  std::vector<sxc_result_dev> slots((size_t)total * cap);
  std::vector<uint64_t> keys((size_t)total * cap);
  std::vector<int32_t> counts(total), ncand(total);
  std::vector<uint16_t> cands((size_t)total * G::N);
  sxc::PairArgs pa{};
  pa.tab = &T.tab;
  pa.osc = T.osc.data();
  pa.ent_lut = T.ent[NB].data();
  pa.ent_rows = erows;
  pa.seq = hb.seq.data();
  pa.tchunks = hb.tchunks.data();
  pa.qchunks = hb.qchunks.data();
  pa.pairs = hb.pairs.data();
  pa.spectra = spectra.data();
  pa.npairs = (int)hb.pairs.size();
  pa.strands = prm.strands;
  pa.cutoff = prm.cutoff;
  pa.cutoff_fast = prm.cutoff_fast;
  pa.cands = cands.data();
  pa.ncand = ncand.data();
  pa.xc_out = xc_out;
  if (tails) {
    pa.tails = hb.tails.data();
    pa.segs = hb.segs.data();
    pa.src_sig = src_sig.data();
  }
  // SXC_EMU_QSPEC=1: both strands' query spectra computed first (as the serving loop keeps
  // them resident) and read by the pair kernel instead of transforming the queries
  std::vector<float> qspec;
  if (getenv("SXC_EMU_QSPEC") && !tails) {
    const int nq = (int)hb.qchunks.size();
    qspec.resize((size_t)nq * 2 * 4 * G::N);
    std::vector<int32_t> qw(1), qd(1);
    for (int q = 0; q < nq; ++q)
      for (int strand = 0; strand < 2; ++strand) {
        qw[0] = q;
        qd[0] = 2 * q + strand;
        sxc::SpecArgs qa{&T.tab, T.osc.data(), T.ent[NB].data(), erows, hb.seq.data(), hb.qchunks.data(), qw.data(), 1,
                         qspec.data(), qd.data(), nullptr, nullptr, nullptr, nullptr, nullptr, strand};
        EmuExec<sxc::Regs<GEO>> ex(G::NT);
        memset(sm.data(), 0xCD, sm.size());
        sxc::spectra_body<GEO>(ex, sm.data(), qa, 0);
      }
    pa.qspectra = qspec.data();
  }
  for (int ps = 0; ps < total; ++ps) {
    EmuExec<sxc::Regs<GEO>> ex(G::NT);
    memset(sm.data(), 0xCD, sm.size());
    sxc::correlate_body<GEO>(ex, sm.data(), pa, ps);
  }
  sxc::ScanArgs sc{};
  sc.tab = &T.tab;
  sc.seq = hb.seq.data();
  sc.tchunks = hb.tchunks.data();
  sc.qchunks = hb.qchunks.data();
  sc.pairs = hb.pairs.data();
  sc.cands = cands.data();
  sc.ncand = ncand.data();
  sc.n_fft = G::N;
  sc.strands = prm.strands;
  sc.min_len = prm.min_len;
  sc.q_chunk = prm.q_chunk;
  sxc::set_keep(sc, prm.min_prob, prm.target_total);
  int32_t nflag = 0;
  sc.nflag = &nflag;
  sc.mode = 0;
  sc.out = slots.data();
  sc.keys = keys.data();
  sc.counts = counts.data();
  sc.cap = cap;
  sc.stats = stats;
  std::vector<double> ptable;
  if (prm.prob_table) {
    ptable.resize((size_t)SXC_PT_ROWS * SXC_PT_LEN);
    sxc::build_prob_table(prm.target_total, ptable.data());
    sc.ptable = ptable.data();
    sc.table_prob = prm.table_prob;
  }
  // each kernel's LDS plan at its exact size, so the sanitizer bounds every access to it
  std::vector<char> ssm_pure(sxc::ScanSmem<NB | 32>::BYTES), ssm_n(sxc::ScanSmem<NB | 64>::BYTES),
      ssm(sxc::ScanSmem<NB>::BYTES);
  // SXC_EMU_SPLIT=k: k scan workgroups per (pair, strand), counts cleared first (as the pair
  // kernel does for a split scan), run in reverse order
  if (const char* v = getenv("SXC_EMU_SPLIT")) sc.split = atoi(v);
  const int split = sc.split > 1 ? sc.split : 1;
  if (split > 1) {
    std::fill(counts.begin(), counts.end(), 0);
    if (stats)
      for (int ps = 0; ps < total; ++ps) stats[2 * ps + 1] = 0;
  }
  for (int blk = total * split - 1; blk >= 0; --blk) {
    // the three kernels, as launched on the device (each skips the others' pairs)
    EmuExec<int> ex(sxc::SCAN_NT);
    memset(ssm_pure.data(), 0xCD, ssm_pure.size());
    sxc::scan_body<NB | 32>(ex, ssm_pure.data(), sc, blk);
    memset(ssm_n.data(), 0xCD, ssm_n.size());
    sxc::scan_body<NB | 64>(ex, ssm_n.data(), sc, blk);
    memset(ssm.data(), 0xCD, ssm.size());
    sxc::scan_body<NB>(ex, ssm.data(), sc, blk);
  }
  res.clear();
  const sxc::ResolveCtx rc{hb.seq.data(), hb.tchunks.data(), hb.qchunks.data(), hb.pairs.data(),
                           (int64_t)hb.pairs.size(), prm.q_chunk, prm.min_prob, prm.target_total};
  for (int ps = 0; ps < total; ++ps) {
    const int n = std::min(counts[ps], cap);
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int a, int b) { return keys[(size_t)ps * cap + a] < keys[(size_t)ps * cap + b]; });
    for (int i : idx) {
      // records flagged near min_prob are decided on the host, as the library does
      sxc_result_dev r = slots[(size_t)ps * cap + i];
      const int keep = sxc::resolve_record(r, rc);
      if (keep < 0) return -4;
      if (keep) res.push_back(r);
    }
    if (counts[ps] > cap) return -3;
  }
  g_emu_near = nflag;
  return 0;
}

}  // namespace

extern "C" {

// Mirrors sxc_align_batch.  xc_out (optional) receives N floats per (pair, strand);
// stats (optional) receives {candidates, runs} per (pair, strand).
int emu_align_batch_tails(const sxc_pair* pairs, int n, const sxc_params* prm, const sxc_signal_tail* t_tail,
                          const sxc_signal_tail* q_tail, sxc_result* out, int64_t cap, int64_t* n_out, float* xc_out,
                          int32_t* stats) {
  sxc::HostBatch hb;
  if (!sxc::pack_batch(pairs, n, prm->N, hb, t_tail, q_tail).empty()) return -1;
  std::vector<sxc_result_dev> res;
  const int slot_cap = 4096;
  int rc;
  // SXC_EMU_GEO="<stage A>,<round levels>" runs that geometry of the FFT kernels (the
  // alternatives the library compiles for the autotuner); unset: the default geometry
  int geo = sxc::nb_of(prm->N);
  if (const char* v = getenv("SXC_EMU_GEO")) {
    int sa = 0, km = 0, lean = 0;
    if (sscanf(v, "%d,%d,%d", &sa, &km, &lean) < 2) return -1;
    geo = sxc::geo_code(geo, sa, km, 0, lean);
  }
#define SXC_EMU_CASE(G) \
  case (G): rc = run_batch<(G)>(hb, *prm, res, slot_cap, xc_out, stats); break;
  switch (geo) {
    SXC_EMU_CASE(9) SXC_EMU_CASE(10) SXC_EMU_CASE(11) SXC_EMU_CASE(12) SXC_EMU_CASE(13) SXC_EMU_CASE(14)
    SXC_EMU_CASE(15)
    SXC_EMU_CASE(sxc::geo_code(9, 16, 2)) SXC_EMU_CASE(sxc::geo_code(10, 16, 2)) SXC_EMU_CASE(sxc::geo_code(11, 16, 2))
    SXC_EMU_CASE(sxc::geo_code(12, 16, 2)) SXC_EMU_CASE(sxc::geo_code(13, 16, 2)) SXC_EMU_CASE(sxc::geo_code(14, 16, 2))
    SXC_EMU_CASE(sxc::geo_code(15, 32, 2))
    SXC_EMU_CASE(sxc::geo_code(11, 32, 3)) SXC_EMU_CASE(sxc::geo_code(12, 32, 3)) SXC_EMU_CASE(sxc::geo_code(13, 32, 3))
    SXC_EMU_CASE(sxc::geo_code(14, 32, 3))
    SXC_EMU_CASE(sxc::geo_code(12, 32, 3, 0, 1)) SXC_EMU_CASE(sxc::geo_code(12, 16, 2, 0, 1))
    SXC_EMU_CASE(sxc::geo_code(13, 32, 3, 0, 1)) SXC_EMU_CASE(sxc::geo_code(14, 32, 3, 0, 1))
    SXC_EMU_CASE(sxc::geo_code(11, 32, 3, 0, 1)) SXC_EMU_CASE(sxc::geo_code(11, 16, 2, 0, 1))
    SXC_EMU_CASE(sxc::geo_code(13, 32, 2, 0, 1))
    default: return -1;
  }
#undef SXC_EMU_CASE
  if (rc) return rc;
  *n_out = (int64_t)res.size();
  if ((int64_t)res.size() > cap) return -3;
  memcpy(out, res.data(), res.size() * sizeof(sxc_result));
  return 0;
}

int emu_align_batch(const sxc_pair* pairs, int n, const sxc_params* prm, sxc_result* out, int64_t cap,
                    int64_t* n_out, float* xc_out, int32_t* stats) {
  return emu_align_batch_tails(pairs, n, prm, nullptr, nullptr, out, cap, n_out, xc_out, stats);
}

// The library's ProbTable thresholds (sxc_tables.cpp), SXC_PT_ROWS x SXC_PT_LEN doubles.
void emu_prob_table(double target_size, double* out) { sxc::build_prob_table(target_size, out); }

// Records the last emulated batch flagged near min_prob (decided on the host).
int64_t emu_last_near(void) { return g_emu_near; }

// The host's window entropies of a chunk (sxc::window_entropy_row): SXC_ENT_ROW floats.
void emu_window_entropy(const char* seq, int len, int N, float* row) {
  sxc::window_entropy_row((const unsigned char*)seq, len, N, row);
}

// The keep band around min_prob (sxc::keep_band).
void emu_keep_band(double min_prob, double target_size, double* lo, double* hi) {
  sxc::keep_band(min_prob, target_size, lo, hi);
}

// FFT of one signal through stage A + rounds (spectra_body with a given signal).
int emu_fft_forward(const float* in, float* out, int N) {
  const Tables& T = tables();
  sxc_chunk ck{0, 0, 0};
  int32_t which = 0, dst = 0;
  std::vector<float> sig(4 * (size_t)N, 0.f), spec(4 * (size_t)N);
  memcpy(sig.data(), in, sizeof(float) * N);
  auto go = [&](auto nbc) {
    constexpr int NB = decltype(nbc)::value;
    sxc::SpecArgs sa{&T.tab, T.osc.data(), T.ent[NB].data(), nullptr, nullptr, &ck, &which, 1, spec.data(), &dst, sig.data(),
                     nullptr, nullptr, nullptr, nullptr, 0};
    std::vector<char> sm(sxc::Smem<NB>::BYTES);
    EmuExec<sxc::Regs<NB>> ex(sxc::Geo<NB>::NT);
    sxc::spectra_body<NB>(ex, sm.data(), sa, 0);
  };
  switch (sxc::nb_of(N)) {
    case 9: go(std::integral_constant<int, 9>{}); break;
    case 10: go(std::integral_constant<int, 10>{}); break;
    case 11: go(std::integral_constant<int, 11>{}); break;
    case 12: go(std::integral_constant<int, 12>{}); break;
    case 13: go(std::integral_constant<int, 13>{}); break;
    case 14: go(std::integral_constant<int, 14>{}); break;
    case 15: go(std::integral_constant<int, 15>{}); break;
    default: return -1;
  }
  memcpy(out, spec.data(), sizeof(float) * N);
  return 0;
}

// The float pre-filter of the tail (z_clearly_skipped) against the FP64 z of
// tail_from_sums on n random sums: with z_skip just above the exact z the filter must never
// drop the run (returns the number of such violations); fired[0] counts the samples where
// it dropped the run for z_skip well below z.
int64_t emu_check_z_filter(int64_t n, uint64_t seed, int64_t* fired) {
  uint64_t x = seed * 2654435761ull + 1;
  auto next = [&]() {
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    return x;
  };
  int64_t bad = 0, hit = 0;
  const double deltas[] = {1e-12, 1e-9, 1e-6, 1e-4};
  for (int64_t i = 0; i < n; ++i) {
    const int len = 46 + (int)(next() % 16339);
    const int cm = (int)(next() % (uint64_t)(len + 1));
    const int ct = (int)(next() % (uint64_t)(len + 1));
    const int cq = (int)(next() % (uint64_t)(len + 1));
    // z exactly as tail_from_sums computes it
    const double ident = (double)cm / (double)len;
    const double gc_target = (double)ct / (double)len;
    const double at_target = 1. - gc_target;
    double r = (double)cq * gc_target;
    r += ((double)len - (double)cq) * at_target;
    const double p = r / (double)len / 2.;
    const double sd = sqrt(p * (1. - p) * (double)len);
    const double m = p * (double)len;
    const double xx = (double)len * ident;
    const double z = (m - xx) / sd / 1.414213562;
    if (!std::isfinite(z)) continue;
    for (double d : deltas) {
      const double above = z + d * (1. + fabs(z));
      if (sxc::z_clearly_skipped(cm, ct, cq, len, above)) ++bad;
    }
    if (sxc::z_clearly_skipped(cm, ct, cq, len, z - 1e-3 * (1. + fabs(z)) - 0.01)) ++hit;
  }
  *fired = hit;
  return bad;
}

// window_flags (the byte-parallel window test of the pure A/C/G/T scan) against the window
// counted position by position: n diagonals of 32 * blocks positions whose match bits have a
// density drawn per diagonal from 0 to 1 (so counts reach 0 and 46); returns the number of
// wrong flag bits and, in *flagged, how many flags were set.
int64_t emu_check_window_flags(int64_t n, int blocks, uint64_t seed, int64_t* flagged) {
  uint64_t x = seed * 2654435761ull + 1;
  auto next = [&]() {
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    return x;
  };
  int64_t bad = 0, set = 0;
  std::vector<int> m(32 * blocks);
  for (int64_t i = 0; i < n; ++i) {
    const uint64_t density = next() % 65;  // out of 64
    for (auto& v : m) v = (next() % 64) < density ? 1 : 0;
    uint32_t ub = sxc::byte_perm(0u, (uint32_t)(128 - sxc::SXC_MIN_MATCHES), 0u);  // C = 0 before position 0
    int C = 0;
    for (int k = 0; k < blocks; ++k) {
      uint32_t zw[2] = {0, 0};
      for (int r = 0; r < 32; ++r) {
        const int p = 32 * k + r;
        const int z = m[p] + 1 - (p >= 46 ? m[p - 46] : 0);
        zw[r >> 4] |= (uint32_t)z << (2 * (r & 15));
      }
      const uint32_t f = sxc::window_flags(zw[0], zw[1], ub);
      for (int r = 0; r < 32; ++r) {
        const int p = 32 * k + r;
        C += m[p] - (p >= 46 ? m[p - 46] : 0);
        const uint32_t want = C >= sxc::SXC_MIN_MATCHES ? 1u : 0u;
        bad += ((f >> r) & 1u) != want;
        set += want;
      }
      bad += ub != sxc::byte_perm(0u, (uint32_t)(C + 128 - sxc::SXC_MIN_MATCHES), 0u);
    }
  }
  *flagged = set;
  return bad;
}

// window_flags_mk (the byte-parallel window test of the A/C/G/T/N scan) against the window
// scored position by position: n diagonals whose positions are identical A/C/G/T bases, N
// positions or neither, with densities drawn per diagonal; returns the number of wrong flag
// bits and carried values and, in *flagged, how many flags were set.
int64_t emu_check_window_flags_mk(int64_t n, int blocks, uint64_t seed, int64_t* flagged) {
  uint64_t x = seed * 2654435761ull + 1;
  auto next = [&]() {
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    return x;
  };
  int64_t bad = 0, set = 0;
  std::vector<int> m(32 * blocks), kk(32 * blocks);
  for (int64_t i = 0; i < n; ++i) {
    const uint64_t dm = next() % 65, dk = next() % 65;  // out of 64
    for (int p = 0; p < 32 * blocks; ++p) {
      kk[p] = (next() % 64) < dk ? 1 : 0;
      m[p] = !kk[p] && (next() % 64) < dm ? 1 : 0;  // a position with an N is never a match
    }
    uint32_t ubm = 0, ubk = sxc::byte_perm(0u, (uint32_t)(128 - sxc::SXC_MIN_QUARTERS), 0u);
    int M = 0, K = 0;
    for (int b = 0; b < blocks; ++b) {
      uint32_t zm[2] = {0, 0}, zk[2] = {0, 0};
      for (int r = 0; r < 32; ++r) {
        const int p = 32 * b + r;
        zm[r >> 4] |= (uint32_t)(m[p] + 1 - (p >= 46 ? m[p - 46] : 0)) << (2 * (r & 15));
        zk[r >> 4] |= (uint32_t)(kk[p] + 1 - (p >= 46 ? kk[p - 46] : 0)) << (2 * (r & 15));
      }
      const uint32_t f = sxc::window_flags_mk(zm[0], zm[1], zk[0], zk[1], ubm, ubk);
      for (int r = 0; r < 32; ++r) {
        const int p = 32 * b + r;
        M += m[p] - (p >= 46 ? m[p - 46] : 0);
        K += kk[p] - (p >= 46 ? kk[p - 46] : 0);
        const uint32_t want = 100 * M + 25 * K > SXC_WIN_THRESHOLD ? 1u : 0u;
        bad += ((f >> r) & 1u) != want;
        set += want;
      }
      bad += ubm != sxc::byte_perm(0u, (uint32_t)M, 0u);
      bad += ubk != sxc::byte_perm(0u, (uint32_t)(K + 128 - sxc::SXC_MIN_QUARTERS), 0u);
    }
  }
  *flagged = set;
  return bad;
}

}  // extern "C"



// The subset of the libsxc_hip.so C ABI that the host programs (slave, standalone
// aligner) call, served by the emulator, so their protocol and file handling can be
// tested on a machine without a GPU.  Test infrastructure only.

struct sxc_ctx {
  int device;
};
struct sxc_batch {
  std::vector<sxc_pair> pairs;
  sxc_params prm;
  std::vector<sxc_result> res;
  std::vector<sxc_signal_tail> t_tail, q_tail;
};

extern "C" {
static thread_local const char* g_emu_err = "";
int sxc_create(sxc_ctx** out, int device) {
  *out = new sxc_ctx{device};
  return SXC_OK;
}
void sxc_destroy(sxc_ctx* ctx) { delete ctx; }
const char* sxc_last_error(void) { return g_emu_err; }
const char* sxc_version(void) { return "emulator"; }
int sxc_batch_create_tails(sxc_ctx*, const sxc_pair* pairs, int32_t n, const sxc_params* prm,
                           const sxc_signal_tail* t_tail, const sxc_signal_tail* q_tail, sxc_batch** out) {
  *out = new sxc_batch{std::vector<sxc_pair>(pairs, pairs + n), *prm, {}, {}, {}};
  if (t_tail) (*out)->t_tail.assign(t_tail, t_tail + n);
  if (q_tail) (*out)->q_tail.assign(q_tail, q_tail + n);
  return SXC_OK;
}
int sxc_batch_create(sxc_ctx* c, const sxc_pair* pairs, int32_t n, const sxc_params* prm, sxc_batch** out) {
  return sxc_batch_create_tails(c, pairs, n, prm, nullptr, nullptr, out);
}
int sxc_batch_run(sxc_batch* b) {
  std::vector<sxc_result> r((size_t)b->pairs.size() * 64 + 1024);
  int64_t n = 0;
  const sxc_signal_tail* tt = b->t_tail.empty() ? nullptr : b->t_tail.data();
  const sxc_signal_tail* qt = b->q_tail.empty() ? nullptr : b->q_tail.data();
  int rc = emu_align_batch_tails(b->pairs.data(), (int)b->pairs.size(), &b->prm, tt, qt, r.data(), (int64_t)r.size(),
                                 &n, nullptr, nullptr);
  if (rc == SXC_ERR_CAPACITY) {
    r.resize((size_t)n);
    rc = emu_align_batch_tails(b->pairs.data(), (int)b->pairs.size(), &b->prm, tt, qt, r.data(), (int64_t)r.size(),
                               &n, nullptr, nullptr);
  }
  if (rc) {
    g_emu_err = "emulated batch failed";
    return SXC_ERR_ARG;
  }
  r.resize((size_t)n);
  b->res.swap(r);
  return SXC_OK;
}
int sxc_batch_set_graph(sxc_batch*, int32_t) { return SXC_OK; }
int sxc_batch_results(sxc_batch* b, sxc_result* out, int64_t cap, int64_t* n_out) {
  *n_out = (int64_t)b->res.size();
  if (*n_out > cap) return SXC_ERR_CAPACITY;
  memcpy(out, b->res.data(), b->res.size() * sizeof(sxc_result));
  return SXC_OK;
}
void sxc_batch_destroy(sxc_batch* b) { delete b; }
int sxc_align_batch(sxc_ctx*, const sxc_pair* pairs, int32_t n, const sxc_params* prm, sxc_result* out, int64_t cap,
                    int64_t* n_out) {
  return emu_align_batch(pairs, n, prm, out, cap, n_out, nullptr, nullptr);
}

// The serving loop, run synchronously at submission.  SXC_EMU_NO_SERVER=1 makes creation
// fail, as a device without room for the genomes does.
struct sxc_server {
  std::vector<sxc_chunk_desc> t, q;
  sxc_params prm;
  std::map<uint64_t, std::vector<sxc_result>> done;
  uint64_t next = 1;
};
int sxc_server_create(sxc_ctx*, const sxc_params* prm, const sxc_chunk_desc* targets, int32_t nt,
                      const sxc_chunk_desc* queries, int32_t nq, const sxc_server_config*, sxc_server** out) {
  const char* off = getenv("SXC_EMU_NO_SERVER");
  if (off && off[0] == '1') {
    g_emu_err = "serving loop switched off";
    return SXC_ERR_NOMEM;
  }
  *out = new sxc_server{std::vector<sxc_chunk_desc>(targets, targets + nt),
                        std::vector<sxc_chunk_desc>(queries, queries + nq), *prm, {}, 1};
  return SXC_OK;
}
int sxc_server_submit(sxc_server* sv, const sxc_block* blocks, int32_t nblocks, uint64_t* ticket) {
  std::vector<sxc_pair> pairs;
  for (int32_t k = 0; k < nblocks; ++k) {
    const sxc_block& b = blocks[k];
    if (b.target_from < 0 || b.target_to >= (int)sv->t.size() || b.query_from < 0 || b.query_to >= (int)sv->q.size())
      return SXC_ERR_ARG;
    for (int qi = b.query_from; qi <= b.query_to; ++qi)
      for (int ti = b.target_from; ti <= b.target_to; ++ti) {
        const sxc_chunk_desc &t = sv->t[(size_t)ti], &q = sv->q[(size_t)qi];
        pairs.push_back(sxc_pair{t.seq, t.len, q.seq, q.len, t.start, q.start, t.record, q.record, q.record_len,
                                 b.fast});
      }
  }
  std::vector<sxc_result> r(pairs.size() * 64 + 1024);
  int64_t n = 0;
  int rc = emu_align_batch(pairs.data(), (int)pairs.size(), &sv->prm, r.data(), (int64_t)r.size(), &n, nullptr,
                           nullptr);
  if (rc == SXC_ERR_CAPACITY) {
    r.resize((size_t)n);
    rc = emu_align_batch(pairs.data(), (int)pairs.size(), &sv->prm, r.data(), (int64_t)r.size(), &n, nullptr,
                         nullptr);
  }
  if (rc) return SXC_ERR_ARG;
  r.resize((size_t)n);
  *ticket = sv->next++;
  sv->done[*ticket].swap(r);
  return SXC_OK;
}
int sxc_server_wait(sxc_server* sv, uint64_t ticket, int32_t, int64_t* n_results, double* latency_ms) {
  auto it = sv->done.find(ticket);
  if (it == sv->done.end()) return SXC_ERR_ARG;
  if (n_results) *n_results = (int64_t)it->second.size();
  if (latency_ms) *latency_ms = 0;
  return SXC_OK;
}
int sxc_server_results(sxc_server* sv, uint64_t ticket, sxc_result* out, int64_t cap, int64_t* n_out) {
  auto it = sv->done.find(ticket);
  if (it == sv->done.end()) return SXC_ERR_ARG;
  *n_out = (int64_t)it->second.size();
  if (*n_out > cap) return SXC_ERR_CAPACITY;
  memcpy(out, it->second.data(), it->second.size() * sizeof(sxc_result));
  sv->done.erase(it);
  return SXC_OK;
}
void sxc_server_destroy(sxc_server* sv) { delete sv; }
// One geometry per length here: nothing to time.
int sxc_autotune(sxc_ctx*, int32_t, int32_t, sxc_tune_entry*, int32_t, int32_t* n_out) {
  *n_out = 0;
  return SXC_OK;
}
int sxc_set_geometry(sxc_ctx*, int32_t, int32_t, const sxc_geometry*) { return SXC_OK; }
int sxc_get_geometry(sxc_ctx*, int32_t N, int32_t, sxc_geometry* g) {
  *g = sxc_geometry{N >= 32768 ? 32 : 16, 3, 0, 0};
  return SXC_OK;
}
}  // extern "C"
