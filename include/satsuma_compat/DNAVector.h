/* DNAVector.h -- the part of the reference's DNAVector (analysis/DNAVector.h:300-360) that
 * the cross-correlation classes take: a string of base letters with the reference's
 * element access, sizing, sub-sequence and reverse-complement semantics
 * (analysis/DNAVector.cc:431-468, 482-571).  Quality values and protein modes are not
 * carried. */
#ifndef SXC_COMPAT_DNAVECTOR_H
#define SXC_COMPAT_DNAVECTOR_H

#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "sxc_compat_context.h"

class DNAVector {
 public:
  DNAVector() {}
  explicit DNAVector(const std::string& bases) { SetFromBases(bases); }

  void resize(int n) { m_data.resize(n); }
  void resize(int n, char c) { m_data.resize(n, c); }
  int size() const { return (int)m_data.size(); }
  int isize() const { return (int)m_data.size(); }
  const char& operator[](int i) const { return m_data[i]; }
  char& operator[](int i) { return m_data[i]; }
  void Set(int i, char c) { m_data[i] = c; }
  const char* data() const { return m_data.data(); }

  // DNAVector::SetFromBases (analysis/DNAVector.cc:562-571): bytes up to the first NUL.
  void SetFromBases(const std::string& s) {
    const size_t n = std::strlen(s.c_str());
    m_data.assign(s.c_str(), s.c_str() + n);
  }
  std::string AsString() const { return std::string(m_data.begin(), m_data.end()); }

  // DNAVector::SetToSubOf (analysis/DNAVector.cc:431-468), with its boundary adjustments.
  bool SetToSubOf(const DNAVector& v, int start, int len) {
    if (len < 0) {
      std::cout << "WARNING, negative length provided in DNAVector::SetToSubOf -> " << len << std::endl;
      return false;
    }
    if (start < 0) {
      std::cout << "WARNING, adjusting boundaries in DNAVector::SetToSubOf (1)" << std::endl;
      len += start;
      start = 0;
    }
    if (start + len > v.size()) {
      std::cout << "WARNING, adjusting boundaries in DNAVector::SetToSubOf (2)" << std::endl;
      len = v.size() - start;
    }
    m_data.assign(v.m_data.begin() + start, v.m_data.begin() + start + len);
    return true;
  }
  bool SetToSubOf(const DNAVector& v, int start) {
    if (start > v.size()) return false;
    return SetToSubOf(v, start, v.size() - start);
  }

  void ReverseComplement() {
    sxc_compat::check(sxc_reverse_complement(m_data.data(), (int32_t)m_data.size()), "ReverseComplement");
  }
  void Reverse() {
    for (int i = 0, j = size() - 1; i < (size() + 1) / 2; ++i, --j) std::swap(m_data[i], m_data[j]);
  }

 private:
  std::vector<char> m_data;
};

// DNA_Equal (analysis/DNAVector.cc:390-403): the codec rows' dot product.
inline double DNA_Equal(char a, char b) { return sxc_dna_equal(a, b); }

#endif
