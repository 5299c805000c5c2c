// compat_align.cpp -- a C++ caller written against the reference's own class interface
// (CCSignal, CrossCorrelation, SeqAnalyzer, GetMatchProbability, PrintMatch), compiled
// against include/satsuma_compat/ instead of the reference headers.  It performs
// HomologyByXCorr::Align for one chunk pair (analysis/HomologyByXCorrSlave.cc:221-253): the
// target against the query and against its reverse complement, and prints every run with
// its probability and identity, unfiltered.
//
//   compat_align <target file> <query file> <N> <cutoff> <target size>
//
// The files hold the bases of one chunk each (no header, no line breaks).  Output: one line
// per run, "strand start_target start_query length prob ident" (FP64 printed exactly).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>

#include "satsuma_compat/AlignProbability.h"
#include "satsuma_compat/CrossCorr.h"

static DNAVector read_bases(const char* path) {
  std::ifstream f(path, std::ios::binary);
  std::stringstream ss;
  ss << f.rdbuf();
  DNAVector d;
  d.SetFromBases(ss.str());
  return d;
}

int main(int argc, char** argv) {
  if (argc != 6) {
    std::fprintf(stderr, "usage: %s target query N cutoff target_size\n", argv[0]);
    return 2;
  }
  try {
    const DNAVector target = read_bases(argv[1]);
    const DNAVector query = read_bases(argv[2]);
    const int N = std::atoi(argv[3]);
    const double cutoff = std::atof(argv[4]), target_size = std::atof(argv[5]);
    DNAVector rcquery = query;
    rcquery.ReverseComplement();

    CCSignal tsig, qsig, rcsig;
    tsig.SetSequence(target, N);
    qsig.SetSequence(query, N);
    rcsig.SetSequence(rcquery, N);

    CrossCorrelation xc;
    SeqAnalyzer sa;
    sa.SetTopCutoff(cutoff);
    for (int strand = 0; strand < 2; ++strand) {
      const DNAVector& q = strand ? rcquery : query;
      std::vector<float> result;
      xc.CrossCorrelate(result, tsig, strand ? rcsig : qsig);
      vecSeqMatch matches;
      sa.MatchUp(matches, q, target, result);
      for (int i = 0; i < matches.size(); ++i) {
        const SeqMatch& m = matches[i];
        const double prob =
            GetMatchProbability(target, q, m.GetStartTarget(), m.GetStartQuery(), m.GetLength(), target_size);
        const double ident = PrintMatch(q, target, m, true);
        std::printf("%d %d %d %d %a %a\n", strand, m.GetStartTarget(), m.GetStartQuery(), m.GetLength(), prob, ident);
      }
    }
  } catch (const std::exception& e) {
    std::printf("ERROR %s\n", e.what());
    return 1;
  }
  return 0;
}
