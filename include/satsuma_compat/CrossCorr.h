/* CrossCorr.h -- the reference's cross-correlation classes (analysis/CrossCorr.h:10-71,
 * 146-293) on the GPU path, for C++ code that includes the reference header today.
 *
 *   CCSignal::SetSequence        -> sxc_encode_signal  (analysis/CrossCorr.cc:179-206)
 *   CrossCorrelation::CrossCorrelate -> sxc_xcorr       (analysis/CrossCorr.cc:386-403)
 *   SeqAnalyzer::MatchUp         -> sxc_matchup        (analysis/CrossCorr.cc:583-605, 667-724)
 *   PrintMatch                   -> sxc_match_probability's identity (analysis/CrossCorr.cc:967-1037)
 *
 * Results are bit-identical to the reference's.  Signal sizes are powers of two in
 * [512, 32768] and a sequence must not be longer than its signal (the reference writes out
 * of bounds there); the window of MatchUp is the reference's default (SetMinLen(45)); the
 * protein and codon signals and MatchUp on signals (DoOneFloat) are not provided.  Calls
 * throw std::runtime_error when the library reports an error. */
#ifndef SXC_COMPAT_CROSSCORR_H
#define SXC_COMPAT_CROSSCORR_H

#include <iostream>
#include <string>
#include <vector>

#include "DNAVector.h"
#include "sxc_compat_context.h"

class CCSignal {
 public:
  CCSignal() : m_points(0) {}
  virtual ~CCSignal() {}

  virtual void SetSequence(const DNAVector& b, int size) {
    std::vector<float> all((size_t)4 * (size > 0 ? size : 0));
    sxc_compat::check(sxc_encode_signal(sxc_compat::context(), b.data(), b.size(), size, all.data()),
                      "CCSignal::SetSequence");
    for (int c = 0; c < 4; ++c) m_ch[c].assign(all.begin() + (size_t)c * size, all.begin() + (size_t)(c + 1) * size);
    m_points = size;
  }
  virtual int GetFullSize() const { return (int)m_ch[0].size(); }
  virtual int GetCount() const { return 4; }
  virtual const std::vector<float>& Get(int i) const { return m_ch[(i >= 0 && i < 4) ? i : 0]; }

 protected:
  std::vector<float> m_ch[4];  // A, C, G, T
  int m_points;
};

class CrossCorrelation {
 public:
  // out = xc of target `one` against query `two`, N floats (the signals' size)
  void CrossCorrelate(std::vector<float>& out, const CCSignal& one, const CCSignal& two) {
    const int N = one.GetFullSize();
    if (two.GetFullSize() != N) throw std::runtime_error("sxc: CrossCorrelate: signal sizes differ");
    std::vector<float> t((size_t)4 * N), q((size_t)4 * N);
    for (int c = 0; c < 4; ++c) {
      std::copy(one.Get(c).begin(), one.Get(c).end(), t.begin() + (size_t)c * N);
      std::copy(two.Get(c).begin(), two.Get(c).end(), q.begin() + (size_t)c * N);
    }
    out.resize(N);
    sxc_compat::check(sxc_xcorr(sxc_compat::context(), t.data(), q.data(), N, out.data()),
                      "CrossCorrelation::CrossCorrelate");
  }
};

class SeqMatch {
 public:
  SeqMatch() : m_startTarget(-1), m_startQuery(-1), m_len(0), m_ident(0.f) {}
  SeqMatch(int startTarget, int startQuery, int len, double ident) { Set(startTarget, startQuery, len, ident); }
  void Set(int startTarget, int startQuery, int len, double ident) {
    m_startTarget = startTarget;
    m_startQuery = startQuery;
    m_len = len;
    m_ident = (float)ident;
  }
  int GetStartTarget() const { return m_startTarget; }
  int GetStartQuery() const { return m_startQuery; }
  int GetLength() const { return m_len; }
  double GetIdentity() const { return m_ident; }

 private:
  int m_startTarget;
  int m_startQuery;
  int m_len;
  float m_ident;
};

class vecSeqMatch {
 public:
  vecSeqMatch() : m_len(0) {}
  void clear() { m_len = 0; }
  void push_back(const SeqMatch& m) {
    if (m_len >= (int)m_data.size()) m_data.resize(m_len + 4096);
    m_data[m_len] = m;
    m_len++;
  }
  int size() const { return m_len; }
  SeqMatch& operator[](int i) { return m_data[i]; }
  const SeqMatch& operator[](int i) const { return m_data[i]; }

 private:
  std::vector<SeqMatch> m_data;
  int m_len;
};

class SeqAnalyzer {
 public:
  SeqAnalyzer() : m_minIdent(0.42), m_topCutoff(1.8), m_minLen(45) {}

  // Appends the runs of every lag FindTop accepts (out is not cleared), as the reference.
  void MatchUp(vecSeqMatch& out, const DNAVector& query, const DNAVector& target, std::vector<float>& xc) {
    if (m_minLen != 45) throw std::runtime_error("sxc: SeqAnalyzer::MatchUp supports the default window (SetMinLen(45))");
    int32_t n = 0;
    std::vector<sxc_seqmatch> runs(1024);
    for (;;) {
      const int rc = sxc_matchup(sxc_compat::context(), query.data(), query.size(), target.data(), target.size(),
                                 xc.data(), (int32_t)xc.size(), m_topCutoff, runs.data(), (int32_t)runs.size(), &n);
      if (rc == SXC_ERR_CAPACITY && n > (int32_t)runs.size()) {
        runs.resize((size_t)n);
        continue;
      }
      sxc_compat::check(rc, "SeqAnalyzer::MatchUp");
      break;
    }
    for (int32_t i = 0; i < n; ++i) out.push_back(SeqMatch(runs[i].start_target, runs[i].start_query, runs[i].len, 0.));
  }
  void SetTopCutoff(double c) { m_topCutoff = c; }
  void SetMinLen(int l) { m_minLen = l; }
  void SetMinIdent(double m) { m_minIdent = m; }  // not read by MatchUp in the reference either

 private:
  double m_minIdent;
  double m_topCutoff;
  int m_minLen;
};

// PrintMatch (analysis/CrossCorr.cc:967-1037): the identity of m (the GPU's DNA_Equal sum,
// bit-identical to the reference's); unless bSilent, the reference's 80-column alignment.
inline double PrintMatch(const DNAVector& query, const DNAVector& target, const SeqMatch& m, bool bSilent = false,
                         bool bNukes = true) {
  double ident = 0.;
  if (bNukes) {
    const sxc_seqmatch sm{m.GetStartTarget(), m.GetStartQuery(), m.GetLength()};
    double prob = 0.;
    sxc_compat::check(sxc_match_probability(sxc_compat::context(), target.data(), target.size(), query.data(),
                                            query.size(), &sm, 1, 1.0, &prob, &ident),
                      "PrintMatch");
  }
  std::vector<std::string> q, t, a;
  std::string qq, tt, mm;
  double matches = 0;
  const int l = 80;
  int k = 0;
  for (int i = m.GetStartTarget(); i < m.GetStartTarget() + m.GetLength(); i++) {
    if ((k > 0 && k % l == 0) || i + 1 == m.GetStartTarget() + m.GetLength()) {
      q.push_back(qq);
      t.push_back(tt);
      a.push_back(mm);
      qq = tt = mm = "";
    }
    const int j = m.GetStartQuery() + i - m.GetStartTarget();
    qq += query[j];
    tt += target[i];
    const double dist = bNukes ? DNA_Equal(query[j], target[i]) : (query[j] == target[i] ? 1. : 0.);
    matches += dist;
    mm += dist > 0.9 ? "|" : dist > 0.49 ? "~" : dist > 0.3 ? "." : " ";
    k++;
  }
  if (!bNukes) ident = matches / (double)m.GetLength();
  if (!bSilent) {
    std::cout << "Identity = " << 100 * ident << " % ";
    std::cout << "  Alignment length = " << m.GetLength() << std::endl;
    std::cout << "Query " << m.GetStartQuery() << " - " << m.GetStartQuery() + m.GetLength();
    std::cout << "   Target " << m.GetStartTarget() << " - " << m.GetStartTarget() + m.GetLength() << std::endl;
    std::cout << "--------------------------------------------------------------------------------" << std::endl;
    for (size_t i = 0; i < q.size(); i++) {
      std::cout << q[i] << std::endl;
      std::cout << a[i] << std::endl;
      std::cout << t[i] << std::endl;
      std::cout << std::endl
                << "--------------------------------------------------------------------------------" << std::endl
                << std::endl;
    }
  }
  return ident;
}

#endif
