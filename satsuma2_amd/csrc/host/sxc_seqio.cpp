// sxc_seqio.cpp -- FASTA reader and chunker following the reference's observable behaviour.
//
// Reader (vecDNAVector::ReadOne, analysis/DNAVector.cc:1199-1282; line tokenising
// util/mutil.cc:1493-1526, line reading util/mutil.cc:349-376):
//   * a line is split on ' ' and '\t'; lines with no token are skipped;
//   * a line whose first token starts with '>' opens a record named by all tokens joined
//     with '_'; any other line contributes its FIRST token only;
//   * bytes read before the first header are not dropped: they lead the first record
//     (the reference resets its fill position only when a record is closed);
//   * records are uppercased with toupper in the "C" locale.
// Chunker (ChunkManager::ChunkItSelect with no selection and no block split,
// analysis/SeqChunk.cc:72-165).
#include "sxc_seqio.h"

#include <stdio.h>
#include <string.h>

namespace sxc_host {

double ChunkSet::total() const {
  double t = 0;
  for (int l : lengths) t += (double)l;
  return t;
}

static void split_tokens(const std::string& line, std::vector<std::string>* toks) {
  toks->clear();
  std::string cur;
  for (char c : line) {
    if (c == ' ' || c == '\t') {
      if (!cur.empty()) toks->push_back(cur);
      cur.clear();
    } else {
      cur.push_back(c);
    }
  }
  if (!cur.empty()) toks->push_back(cur);
}

static void upper(std::string* s) {
  for (char& c : *s)
    if (c >= 'a' && c <= 'z') c = (char)(c - 'a' + 'A');
}

static bool read_one(const std::string& path, Records* out, std::string* err) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) {
    *err = "cannot open " + path;
    return false;
  }
  std::string line, pending;  // pending: bytes of the record being filled
  std::vector<std::string> toks;
  long open = -1;  // index of the record being filled, -1 before the first header
  int local = 0;
  bool ok = true;
  char buf[1 << 16];
  bool eof = false;
  while (!eof) {
    line.clear();
    bool got = false;
    for (;;) {
      if (!fgets(buf, sizeof buf, f)) {
        eof = true;
        break;
      }
      got = true;
      size_t n = strlen(buf);
      if (n && buf[n - 1] == '\n') {
        line.append(buf, n - 1);
        break;
      }
      line.append(buf, n);
    }
    if (!got && line.empty()) break;
    split_tokens(line, &toks);
    if (toks.empty()) continue;
    const std::string& p = toks[0];
    if (local == 0 && p[0] == '@') {
      *err = path + ": FASTQ input is not supported";
      ok = false;
      break;
    }
    if (p[0] == '>') {
      ++local;
      if (open >= 0) {
        upper(&pending);
        out->seqs[(size_t)open].swap(pending);
        pending.clear();
      }
      std::string name = toks[0];
      for (size_t x = 1; x < toks.size(); ++x) name += "_" + toks[x];
      out->names.push_back(name.substr(1));
      out->seqs.emplace_back();
      open = (long)out->seqs.size() - 1;
      continue;
    }
    pending += p;
  }
  if (ok && open >= 0) {
    upper(&pending);
    out->seqs[(size_t)open].swap(pending);
  }
  fclose(f);
  return ok;
}

bool read_fasta(const std::string& file_list, Records* out, std::string* err) {
  size_t pos = 0;
  bool any = false;
  while (pos <= file_list.size()) {
    size_t c = file_list.find(',', pos);
    if (c == std::string::npos) c = file_list.size();
    const std::string item = file_list.substr(pos, c - pos);
    if (!item.empty()) {
      if (!read_one(item, out, err)) return false;
      any = true;
    }
    pos = c + 1;
  }
  if (!any) {
    *err = "no input file named";
    return false;
  }
  return true;
}

void chunk_records(const Records& recs, int size, int overlap, ChunkSet* out) {
  out->chunks.clear();
  out->lengths.assign(recs.seqs.size(), 0);
  const int step = size - overlap;
  for (size_t i = 0; i < recs.seqs.size(); ++i) {
    const std::string& s = recs.seqs[i];
    const int l = (int)s.size();
    out->lengths[i] = l;
    if (l < 6) continue;
    const int n = 1 + l / step;
    for (int j = 0; j < n; ++j) {
      const int start = j * step;
      int end = (j + 1) * step + overlap;
      if (end >= l) end = l;
      Chunk c{(int)i, start, s.data() + start, end - start};
      int ns = 0;
      for (int k = 0; k < c.len; ++k) ns += (c.data[k] == 'N' || c.data[k] == 'X');
      if (ns >= c.len) c.len = 0;
      out->chunks.push_back(c);
    }
  }
}

}  // namespace sxc_host
