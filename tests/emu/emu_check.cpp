// Standalone checker of the emulated kernels (test infrastructure), linked against the
// emulator object built with AddressSanitizer (tests/hostbuild.py).
#include <stdint.h>
#include <stdio.h>

#include <vector>

#include "../../include/sxc.h"

extern "C" int emu_align_batch(const sxc_pair* pairs, int n, const sxc_params* prm, sxc_result* out, int64_t cap,
                               int64_t* n_out, float* xc_out, int32_t* stats);

// Reads pairs from a file written by
// tests/test_emu.py, runs the emulated kernels, writes the t_result records.
// input: int32 N, strands, n; double cutoff, cutoff_fast, min_prob, target_total;
//        per pair: int32 tlen, qlen, t_chunk_start, q_chunk_start, t_record, q_record,
//        q_record_len, fast; then tlen + qlen bytes.
int main(int argc, char** argv) {
  if (argc < 3) return 2;
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  int32_t hdr[3];
  double d[4];
  if (fread(hdr, 4, 3, f) != 3 || fread(d, 8, 4, f) != 4) return 2;
  std::vector<sxc_pair> pairs(hdr[2]);
  std::vector<std::vector<char>> store;
  for (int k = 0; k < hdr[2]; ++k) {
    int32_t m[8];
    if (fread(m, 4, 8, f) != 8) return 2;
    store.emplace_back(m[0] + 1);
    char* t = store.back().data();
    if (fread(t, 1, m[0], f) != (size_t)m[0]) return 2;
    store.emplace_back(m[1] + 1);
    char* q = store.back().data();
    if (fread(q, 1, m[1], f) != (size_t)m[1]) return 2;
    pairs[k] = sxc_pair{t, m[0], q, m[1], m[2], m[3], m[4], m[5], m[6], m[7]};
  }
  fclose(f);
  sxc_params prm{hdr[0], 4096, 0, hdr[1], d[0], d[1], d[2], d[3]};
  std::vector<sxc_result> out(1 << 20);
  int64_t n = 0;
  const int rc = emu_align_batch(pairs.data(), hdr[2], &prm, out.data(), (int64_t)out.size(), &n, nullptr, nullptr);
  if (rc) return 3;
  FILE* o = fopen(argv[2], "wb");
  fwrite(&n, 8, 1, o);
  fwrite(out.data(), sizeof(sxc_result), (size_t)n, o);
  fclose(o);
  return 0;
}
