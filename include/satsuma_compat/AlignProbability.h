/* AlignProbability.h -- GetMatchProbability / GetMatchProbabilityEx
 * (analysis/AlignProbability.h:10-27, analysis/AlignProbability.cc:42-127) on the GPU path:
 * the FP64 tail with the device's erf/exp (relative error vs glibc below 1e-11, below
 * 1e-14 where the probability is at least 0.5); identity bit-identical. */
#ifndef SXC_COMPAT_ALIGNPROBABILITY_H
#define SXC_COMPAT_ALIGNPROBABILITY_H

#include "DNAVector.h"
#include "sxc_compat_context.h"

inline double GetMatchProbabilityEx(double& ident, const DNAVector& target, const DNAVector& query, int startTarget,
                                    int startQuery, int length, double targetSize) {
  const sxc_seqmatch m{startTarget, startQuery, length};
  double prob = 0.;
  sxc_compat::check(sxc_match_probability(sxc_compat::context(), target.data(), target.size(), query.data(),
                                          query.size(), &m, 1, targetSize, &prob, &ident),
                    "GetMatchProbability");
  return prob;
}

inline double GetMatchProbability(const DNAVector& target, const DNAVector& query, int startTarget, int startQuery,
                                  int length, double targetSize) {
  double ident = 0.;
  return GetMatchProbabilityEx(ident, target, query, startTarget, startQuery, length, targetSize);
}

#endif
