// Prints what satsuma2_amd/csrc/host/sxc_seqio.cpp reads and chunks (test infrastructure):
//   seqio_check FILES SIZE OVERLAP
// one line per record "R <index> <length> <name>" and, after the records, one line per
// chunk "C <record> <start> <len> <first bytes as hex...>" plus "T <total>".
#include <stdio.h>
#include <stdlib.h>

#include "../../satsuma2_amd/csrc/host/sxc_seqio.h"

int main(int argc, char** argv) {
  if (argc < 4) return 2;
  sxc_host::Records r;
  std::string err;
  if (!sxc_host::read_fasta(argv[1], &r, &err)) {
    printf("E %s\n", err.c_str());
    return 1;
  }
  for (size_t i = 0; i < r.seqs.size(); ++i) printf("R %zu %zu %s\n", i, r.seqs[i].size(), r.names[i].c_str());
  sxc_host::ChunkSet cs;
  sxc_host::chunk_records(r, atoi(argv[2]), atoi(argv[3]), &cs);
  for (const sxc_host::Chunk& c : cs.chunks) {
    printf("C %d %d %d ", c.record, c.start, c.len);
    for (int k = 0; k < c.len; ++k) printf("%02x", (unsigned char)c.data[k]);
    printf("\n");
  }
  printf("T %.1f\n", cs.total());
  return 0;
}
