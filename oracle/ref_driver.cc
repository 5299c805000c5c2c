// Test-only driver around the reference slave translation unit.
//
// This file is checker infrastructure: it is compiled by oracle/Makefile
// against the reference sources where they lie (REF, default /root/reference)
// and its binary goes to oracle/_ref/.  Nothing in the product links it.
//
// It exposes the reference's own functions on the chunk-pair path so the
// golden-vector generator (tests/golden/make_golden.py) and the CPU baseline
// calibration can call them on synthetic inputs:
//   fft     FFTReal<float> forward / inverse / rescale on raw vectors
//           (extern/RealFFT/FFTReal.hpp:146,207,262)
//   detail  CCSignal::SetSequence, CrossCorrelation::CrossCorrelate,
//           SeqAnalyzer::FindTop / MatchUp, GetMatchProbability, PrintMatch
//           for one (target, query) sequence pair, both strands
//           (analysis/CrossCorr.cc:179,386,583,878,967;
//            analysis/AlignProbability.cc:42)
//   pairs   the slave path: FASTA load, chunking, align_target per t_pair,
//           FilterMatches -> t_result list (analysis/HomologyByXCorrSlave.cc:
//           270-300, 168-219, 386-409), optionally on T worker threads.
//   chunks  vecDNAVector::Read + ChunkManager::ChunkIt on FASTA files, printed
//           record by record and chunk by chunk (analysis/DNAVector.cc:1176-1293,
//           analysis/SeqChunk.cc:72-165)
//
// The reference's main() is renamed so this driver can provide its own, and
// SeqAnalyzer::FindTop (private in analysis/CrossCorr.h:275) is opened up so
// its candidate list can be recorded.  System headers are included first so the
// access override only touches the reference's own classes.
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <fcntl.h>
#include <fstream>
#include <iostream>
#include <map>
#include <math.h>
#include <mutex>
#include <netdb.h>
#include <netinet/in.h>
#include <queue>
#include <set>
#include <sstream>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <sys/ioctl.h>
#include <sys/socket.h>
#include <sys/time.h>
#include <thread>
#include <typeinfo>
#include <unistd.h>
#include <vector>
#include <chrono>
#include <cstdint>
#define main satsuma_reference_slave_main
#define private public
#include "HomologyByXCorrSlave.cc"
#undef private
#undef main

#include "../extern/RealFFT/FFTReal.hpp"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <thread>

namespace {

void die(const char* msg) {
  fprintf(stderr, "ref_driver: %s\n", msg);
  exit(2);
}

std::vector<float> read_floats(const char* path) {
  FILE* f = fopen(path, "rb");
  if (!f) die("cannot open float input");
  fseek(f, 0, SEEK_END);
  long n = ftell(f) / (long)sizeof(float);
  fseek(f, 0, SEEK_SET);
  std::vector<float> v(n);
  if (n && fread(&v[0], sizeof(float), n, f) != (size_t)n) die("short read");
  fclose(f);
  return v;
}

void write_blob(FILE* f, const void* p, size_t bytes) {
  if (bytes && fwrite(p, 1, bytes, f) != bytes) die("short write");
}

// fft N in.bin out.bin : in = K*N floats; out = K*(fwd N, ifft N, rescaled N)
int mode_fft(int argc, char** argv) {
  if (argc < 5) die("usage: fft N in.bin out.bin");
  long N = atol(argv[2]);
  std::vector<float> in = read_floats(argv[3]);
  if (in.size() % N) die("input not a multiple of N");
  long K = in.size() / N;
  FFTReal<float> fft(N);
  FILE* f = fopen(argv[4], "wb");
  if (!f) die("cannot open output");
  std::vector<float> spec(N), back(N), res(N);
  for (long k = 0; k < K; ++k) {
    fft.do_fft(&spec[0], &in[k * N]);
    fft.do_ifft(&spec[0], &back[0]);
    res = back;
    fft.rescale(&res[0]);
    write_blob(f, &spec[0], N * sizeof(float));
    write_blob(f, &back[0], N * sizeof(float));
    write_blob(f, &res[0], N * sizeof(float));
  }
  fclose(f);
  return 0;
}

DNAVector read_seq_file(const char* path) {
  // raw sequence bytes, no header: the caller controls every byte
  std::ifstream in(path, std::ios::binary);
  if (!in) die("cannot open sequence file");
  std::string s((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  DNAVector d;
  d.resize((int)s.size());
  for (size_t i = 0; i < s.size(); ++i) d[(int)i] = s[i];
  return d;
}

// detail target.seq query.seq N cutoff targetSize out.bin
// out layout (all little-endian):
//   int32 N
//   float sigT[4][N], sigQ[4][N], sigQrc[4][N]
//   for strand in (fwd, rc):
//     float xc[N]
//     int32 ntop, int32 top[ntop]
//     int32 nm, then nm x {int32 sT, int32 sQ, int32 len, double prob, double ident}
int mode_detail(int argc, char** argv) {
  if (argc < 8) die("usage: detail target.seq query.seq N cutoff targetSize out.bin");
  DNAVector t = read_seq_file(argv[2]);
  DNAVector q = read_seq_file(argv[3]);
  int N = atoi(argv[4]);
  double cutoff = atof(argv[5]);
  double tsize = atof(argv[6]);
  FILE* f = fopen(argv[7], "wb");
  if (!f) die("cannot open output");

  DNAVector qrc = q;
  qrc.ReverseComplement();

  CCSignal sT, sQ, sQrc;
  sT.SetSequence(t, N);
  sQ.SetSequence(q, N);
  sQrc.SetSequence(qrc, N);
  int32_t n32 = N;
  write_blob(f, &n32, 4);
  const CCSignal* sigs[3] = {&sT, &sQ, &sQrc};
  for (int s = 0; s < 3; ++s)
    for (int c = 0; c < 4; ++c) write_blob(f, &sigs[s]->Get(c)[0], N * sizeof(float));

  for (int strand = 0; strand < 2; ++strand) {
    const CCSignal& qs = strand ? sQrc : sQ;
    const DNAVector& qseq = strand ? qrc : q;
    CrossCorrelation xc;
    std::vector<float> res;
    xc.CrossCorrelate(res, sT, qs);
    write_blob(f, &res[0], N * sizeof(float));

    SeqAnalyzer sa;
    sa.SetTopCutoff(cutoff);
    std::vector<int> top;
    std::vector<float> res_copy = res;
    sa.FindTop(top, res_copy, cutoff);
    int32_t ntop = (int32_t)top.size();
    write_blob(f, &ntop, 4);
    write_blob(f, top.data(), ntop * sizeof(int32_t));

    SeqAnalyzer sa2;
    sa2.SetTopCutoff(cutoff);
    vecSeqMatch m;
    sa2.MatchUp(m, qseq, t, res);
    int32_t nm = m.size();
    write_blob(f, &nm, 4);
    for (int i = 0; i < nm; ++i) {
      int32_t st = m[i].GetStartTarget(), sq = m[i].GetStartQuery(), len = m[i].GetLength();
      double prob = GetMatchProbability(t, qseq, st, sq, len, tsize);
      double ident = PrintMatch(qseq, t, m[i], true);
      write_blob(f, &st, 4);
      write_blob(f, &sq, 4);
      write_blob(f, &len, 4);
      write_blob(f, &prob, 8);
      write_blob(f, &ident, 8);
    }
  }
  fclose(f);
  return 0;
}

// pairs query.fa target.fa pairs.txt out.bin threads [cutoff cutoff_fast minlen q_chunk t_chunk]
// pairs.txt: one "targetFrom targetTo queryFrom queryTo fast" per line.
// out.bin: uint64 count, then count x t_result (raw struct bytes, 72 B each).
// stdout: one line "pairs=<chunk pairs> seconds=<wall> threads=<T>".
int mode_pairs(int argc, char** argv) {
  if (argc < 7)
    die("usage: pairs query.fa target.fa pairs.txt out.bin threads [cutoff cutoff_fast minlen q_chunk t_chunk "
        "prob_table min_prob]");
  query_filename = argv[2];
  target_filename = argv[3];
  int threads = atoi(argv[6]);
  topCutoff = argc > 7 ? atof(argv[7]) : 1.8;
  topCutoffFast = argc > 8 ? atof(argv[8]) : 2.9;
  minLen = argc > 9 ? atoi(argv[9]) : 0;
  queryChunk = argc > 10 ? atoi(argv[10]) : 4096;
  targetChunk = argc > 11 ? atoi(argv[11]) : 4096;
  prob_table = argc > 12 ? atoi(argv[12]) != 0 : false;
  minProb = argc > 13 ? atof(argv[13]) : 0.9999;

  // analysis/HomologyByXCorrSlave.cc:389-408, verbatim call sequence
  queryRaw.Read(query_filename, queryNames);
  cmQuery = new ChunkManager(queryChunk, 0);
  cmQuery->ChunkIt(query, queryInfo, queryRaw, queryNames, 0, 0);
  queryRaw.clear();
  targetRaw.Read(target_filename, targetNames);
  cmTarget = new ChunkManager(targetChunk, targetChunk / 4);
  cmTarget->ChunkIt(target, targetInfo, targetRaw, targetNames, 0, 0);
  targetRaw.clear();
  targetTotal = 0;
  for (int i = 0; i < cmTarget->GetCount(); i++) targetTotal += (double)cmTarget->GetSize(i);
  if (prob_table) probt = ProbTable(targetTotal, minProb);  // analysis/HomologyByXCorrSlave.cc:413-421

  std::vector<t_pair> plist;
  FILE* pf = fopen(argv[4], "r");
  if (!pf) die("cannot open pairs file");
  long long npairs_chunks = 0;
  while (true) {
    t_pair p;
    memset(&p, 0, sizeof(p));
    int fast = 0;
    int r = fscanf(pf, "%d %d %d %d %d", &p.targetFrom, &p.targetTo, &p.queryFrom, &p.queryTo, &fast);
    if (r != 5) break;
    p.fast = fast != 0;
    plist.push_back(p);
    npairs_chunks += (long long)(p.targetTo + 1 - p.targetFrom) * (p.queryTo + 1 - p.queryFrom);
  }
  fclose(pf);

  std::mutex qm;
  size_t next = 0;
  auto t0 = std::chrono::steady_clock::now();
  auto worker = [&]() {
    HomologyByXCorr h;
    while (true) {
      size_t k;
      {
        std::lock_guard<std::mutex> g(qm);
        if (next >= plist.size()) return;
        k = next++;
      }
      h.align_target(plist[k]);
    }
  };
  std::vector<std::thread> pool;
  for (int i = 0; i < threads; ++i) pool.emplace_back(worker);
  for (auto& th : pool) th.join();
  double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

  FILE* f = fopen(argv[5], "wb");
  if (!f) die("cannot open output");
  uint64_t cnt = results.size();
  write_blob(f, &cnt, 8);
  write_blob(f, results.data(), cnt * sizeof(t_result));
  fclose(f);
  printf("pairs=%lld seconds=%.6f threads=%d results=%llu\n", npairs_chunks, secs, threads,
         (unsigned long long)cnt);
  return 0;
}

// chunks FILES SIZE OVERLAP: "R index length name" per record, "C record start len hex"
// per chunk, "T total" (sum of ChunkManager::GetSize).
int mode_chunks(int argc, char** argv) {
  if (argc < 5) die("usage: chunks files size overlap");
  vecDNAVector raw, out;
  std::vector<string> names;
  raw.Read(argv[2], names);
  for (int i = 0; i < (int)raw.size(); ++i) printf("R %d %d %s\n", i, raw[i].size(), names[i].c_str());
  ChunkManager cm(atoi(argv[3]), atoi(argv[4]));
  std::vector<SeqChunk> info;
  cm.ChunkIt(out, info, raw, names, 0, 0);
  for (int k = 0; k < (int)out.size(); ++k) {
    printf("C %d %d %d ", info[k].GetID(), info[k].GetStart(), out[k].size());
    for (int i = 0; i < out[k].size(); ++i) printf("%02x", (unsigned char)out[k][i]);
    printf("\n");
  }
  double total = 0;
  for (int i = 0; i < cm.GetCount(); i++) total += (double)cm.GetSize(i);
  printf("T %.1f\n", total);
  return 0;
}

// probtable SIZE CUTOFF out.bin: ProbTable(SIZE, CUTOFF)'s 512 x 2048 thresholds (row 0,
// which the reference leaves empty, as zeros).
int mode_probtable(int argc, char** argv) {
  if (argc < 5) die("usage: probtable size cutoff out.bin");
  ProbTable pt(atof(argv[2]), atof(argv[3]));
  FILE* f = fopen(argv[4], "wb");
  if (!f) die("cannot open output");
  std::vector<double> row(2048);
  for (int i = 0; i < 512; ++i) {
    const std::vector<double>& t = pt.m_table[i];
    for (int j = 0; j < 2048; ++j) row[j] = j < (int)t.size() ? t[j] : 0.;
    write_blob(f, row.data(), row.size() * sizeof(double));
  }
  fclose(f);
  return 0;
}

// layout probe: prints sizeof/offsetof of the wire structs (analysis/WorkQueue.h:17-33)
int mode_layout() {
  printf("t_pair %zu tF %zu tT %zu qF %zu qT %zu fast %zu sid %zu status %zu\n", sizeof(t_pair),
         offsetof(t_pair, targetFrom), offsetof(t_pair, targetTo), offsetof(t_pair, queryFrom),
         offsetof(t_pair, queryTo), offsetof(t_pair, fast), offsetof(t_pair, slave_id),
         offsetof(t_pair, status));
  printf("t_result %zu qid %zu tid %zu qsize %zu qstart %zu tstart %zu len %zu rev %zu prob %zu ident %zu\n",
         sizeof(t_result), offsetof(t_result, query_id), offsetof(t_result, target_id),
         offsetof(t_result, query_size), offsetof(t_result, qstart), offsetof(t_result, tstart),
         offsetof(t_result, len), offsetof(t_result, reverse), offsetof(t_result, prob),
         offsetof(t_result, ident));
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) die("usage: ref_driver fft|detail|pairs|layout ...");
  std::string m = argv[1];
  if (m == "fft") return mode_fft(argc, argv);
  if (m == "detail") return mode_detail(argc, argv);
  if (m == "pairs") return mode_pairs(argc, argv);
  if (m == "layout") return mode_layout();
  if (m == "chunks") return mode_chunks(argc, argv);
  if (m == "probtable") return mode_probtable(argc, argv);
  die("unknown mode");
  return 2;
}
