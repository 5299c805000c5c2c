// sxc_seqio.h -- FASTA loading and chunking with the reference's semantics, for the
// slave executable and the standalone aligner built on libsxc_hip.so.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

namespace sxc_host {

// vecDNAVector after Read(file, names) (analysis/DNAVector.cc:1176-1293): every record
// uppercased; names are the header line without '>' with whitespace-separated tokens
// joined by '_' (shortName = false).
struct Records {
  std::vector<std::string> seqs;
  std::vector<std::string> names;
};

// Reads a comma-separated list of FASTA files and appends their records, as
// vecDNAVector::Read does.  FASTQ input (first item '@') is rejected.  Returns false and
// sets *err on failure.
bool read_fasta(const std::string& file_list, Records* out, std::string* err);

// One chunk of ChunkManager::ChunkItSelect (analysis/SeqChunk.cc:72-165): record id,
// start in the record and the chunk's bytes.  len == 0 marks a cleared (all-N/X) chunk.
struct Chunk {
  int record;
  int start;
  const char* data;
  int len;
};

struct ChunkSet {
  std::vector<Chunk> chunks;
  std::vector<int> lengths;  // ChunkManager::GetSize(record) for every record
  double total() const;      // sum of lengths (targetTotal, HomologyByXCorrSlave.cc:405-408)
};

// Chunks of `size` bytes every `size - overlap` bytes; records shorter than 6 bytes get
// none.  The chunks point into `recs`, which must outlive the set.
void chunk_records(const Records& recs, int size, int overlap, ChunkSet* out);

}  // namespace sxc_host
