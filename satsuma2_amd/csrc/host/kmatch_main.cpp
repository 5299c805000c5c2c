// KMatch -- the k-mer seeding of Satsuma2 with its tables on the GPU.
//
// A drop-in for kmatch/KMatch.cc:320-344: the same arguments (query.fa target.fa K
// output min_length jump max_freq), the same checks and the same output file, t_result
// records (analysis/WorkQueue.h:25-35) as SatsumaSynteny2 reads them
// (analysis/SatsumaSynteny2.cc:417-463, WorkQueue::results_from_file).
#include <stdio.h>
#include <stdlib.h>

#include <iostream>
#include <vector>

#include "../../../include/sxc.h"
#include "../../../include/sxc_kmatch.h"

int main(int argc, char** argv) {
  if (argc != 8) {
    std::cout << "Usage: " << argv[0] << " query.fa target.fa K output.fa min_length jump max_freq" << std::endl;
    return -1;
  }
  if (atoi(argv[3]) % 2 == 0) {
    std::cout << "KMatch only accepts odd K values, please try again" << std::endl;
    return 1;
  }
  sxc_ctx* ctx = nullptr;
  if (sxc_create(&ctx, 0) != SXC_OK) {
    std::cerr << "KMatch: " << sxc_last_error() << std::endl;
    return 2;
  }
  sxc_kmatch_params p{atoi(argv[3]), atoi(argv[5]), atoi(argv[6]), atoi(argv[7])};
  sxc_kmatch* km = nullptr;
  if (sxc_kmatch_files(ctx, argv[1], argv[2], &p, &km) != SXC_OK) {
    std::cerr << "KMatch: " << sxc_last_error() << std::endl;
    sxc_destroy(ctx);
    return 2;
  }
  int64_t n = 0;
  sxc_kmatch_results(km, nullptr, 0, &n);
  std::vector<sxc_result> r((size_t)n);
  if (n > 0 && sxc_kmatch_results(km, r.data(), n, &n) != SXC_OK) {
    std::cerr << "KMatch: " << sxc_last_error() << std::endl;
    return 2;
  }
  FILE* f = fopen(argv[4], "wb");
  if (!f || (n > 0 && fwrite(r.data(), sizeof(sxc_result), (size_t)n, f) != (size_t)n) || fclose(f) != 0) {
    std::cerr << "KMatch: cannot write " << argv[4] << std::endl;
    return 2;
  }
  sxc_kmatch_stats st{};
  sxc_kmatch_get_stats(km, &st);
  std::cout << st.joined << " matching positions" << std::endl;
  std::cout << n << " matches dumped" << std::endl;
  sxc_kmatch_destroy(km);
  sxc_destroy(ctx);
  return 0;
}
