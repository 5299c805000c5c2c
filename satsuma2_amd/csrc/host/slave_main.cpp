// HomologyByXCorrSlave -- the Satsuma2 worker process with its chunk-pair alignment on the GPU.
//
// A drop-in for analysis/HomologyByXCorrSlave.cc: same flags (:334-384), same input
// loading and chunking (:386-409), same master exchange (:440-523) and the same t_result
// records per block (align_target / Align / FilterMatches, :168-300).  What changes is
// how the work runs: the reference gives each of its -p threads one t_pair block at a
// time.  Here both genomes and the spectra of every target chunk are put in HBM once, in a
// serving loop (sxc_server_*): each block is submitted as it arrives and the loop packs
// the pending chunk pairs of all blocks into shared launches.  If the genomes do not fit
// the device, the slave falls back to draining the queued blocks into one batch at a time.
// The master exchange stays as it is: ask for blocks while fewer than 2 * p wait besides the
// p being worked on, and hand back every finished result on the next connection.
//
// Results are appended block by block in arrival order, each block in the reference's
// order (query chunks outer, target chunks inner, forward strand then reverse), i.e. the
// order the reference slave produces with -p 1.
#include <arpa/inet.h>
#include <errno.h>
#include <netdb.h>
#include <netinet/in.h>
#include <stdio.h>
#include <string.h>
#include <sys/socket.h>
#include <time.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/sxc.h"
#include "sxc_flags.h"
#include "sxc_seqio.h"
#include "sxc_wire.h"

using sxc_host::ChunkSet;
using sxc_host::Records;
using sxc_host::WirePair;

namespace {

// Chunk pairs per GPU launch: large enough to fill the device (256 CUs, one workgroup
// per pair and strand), small enough to keep the batch's device buffers modest.
constexpr int kMaxPairsPerLaunch = 16384;

struct State {
  std::mutex mu;
  std::condition_variable work_cv;   // blocks were queued or the run is over
  std::condition_variable space_cv;  // blocks left the queue or finished
  std::condition_variable done_cv;   // a block was submitted to the serving loop
  std::deque<WirePair> blocks;
  std::deque<uint64_t> running;      // serving-loop tickets, in arrival order
  size_t busy = 0;                   // blocks taken off the queue and not finished
  std::vector<sxc_result> results;
  std::atomic<bool> finished{false};
};

bool write_all(int fd, const void* p, size_t n) {
  const char* c = (const char*)p;
  while (n > 0) {
    const ssize_t w = write(fd, c, n);
    if (w < 0 && errno == EINTR) continue;
    if (w <= 0) return false;
    c += w;
    n -= (size_t)w;
  }
  return true;
}

bool read_all(int fd, void* p, size_t n) {
  char* c = (char*)p;
  while (n > 0) {
    const ssize_t r = read(fd, c, n);
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) return false;
    c += r;
    n -= (size_t)r;
  }
  return true;
}

struct Job {
  const ChunkSet* query;
  const ChunkSet* target;
  sxc_params prm;
  sxc_ctx* ctx;
  sxc_server* srv = nullptr;  // null: batch fallback
};

// Appends the chunk pairs of one block (align_target, :270-300) in reference order.
bool append_block(const Job& job, const WirePair& p, std::vector<sxc_pair>* out) {
  const int nt = (int)job.target->chunks.size(), nq = (int)job.query->chunks.size();
  if (p.target_from < 0 || p.target_to >= nt || p.query_from < 0 || p.query_to >= nq ||
      p.target_from > p.target_to || p.query_from > p.query_to) {
    printf("ERROR block t:%d-%d q:%d-%d is outside the %d target / %d query chunks; skipped\n",
           p.target_from, p.target_to, p.query_from, p.query_to, nt, nq);
    return false;
  }
  for (int qi = p.query_from; qi <= p.query_to; ++qi) {
    const sxc_host::Chunk& q = job.query->chunks[(size_t)qi];
    for (int ti = p.target_from; ti <= p.target_to; ++ti) {
      const sxc_host::Chunk& t = job.target->chunks[(size_t)ti];
      sxc_pair s;
      s.tseq = t.data;
      s.tlen = t.len;
      s.qseq = q.data;
      s.qlen = q.len;
      s.t_chunk_start = t.start;
      s.q_chunk_start = q.start;
      s.t_record = t.record;
      s.q_record = q.record;
      s.q_record_len = job.query->lengths[(size_t)q.record];
      s.fast = p.fast ? 1 : 0;
      out->push_back(s);
    }
  }
  return true;
}

// Runs chunk pairs through the library in launches of at most kMaxPairsPerLaunch pairs
// and appends their records to *out.
bool align_pairs(const Job& job, const std::vector<sxc_pair>& pairs, std::vector<sxc_result>* out) {
  for (size_t at = 0; at < pairs.size(); at += kMaxPairsPerLaunch) {
    const int n = (int)std::min(pairs.size() - at, (size_t)kMaxPairsPerLaunch);
    sxc_batch* b = nullptr;
    int rc = sxc_batch_create(job.ctx, pairs.data() + at, n, &job.prm, &b);
    if (rc == SXC_OK) rc = sxc_batch_run(b);
    std::vector<sxc_result> r;
    int64_t got = 0;
    if (rc == SXC_OK) {
      r.resize((size_t)n * 8 + 64);
      rc = sxc_batch_results(b, r.data(), (int64_t)r.size(), &got);
      if (rc == SXC_ERR_CAPACITY) {
        r.resize((size_t)got);
        rc = sxc_batch_results(b, r.data(), (int64_t)r.size(), &got);
      }
    }
    sxc_batch_destroy(b);
    if (rc != SXC_OK) {
      printf("ERROR aligning %d chunk pairs: %s\n", n, sxc_last_error());
      return false;
    }
    out->insert(out->end(), r.begin(), r.begin() + got);
  }
  return true;
}

[[noreturn]] void device_failure(const char* what) {
  printf("ERROR %s: %s\n", what, sxc_last_error());
  fflush(stdout);
  _exit(1);  // a device failure leaves the master's blocks unanswered; fail loudly
}

// Batch fallback: drains every queued block into one batch.
void submitter(const Job& job, State* st) {
  for (;;) {
    std::vector<WirePair> take;
    {
      std::unique_lock<std::mutex> lk(st->mu);
      st->work_cv.wait(lk, [&] { return st->finished.load() || !st->blocks.empty(); });
      if (st->finished.load()) return;
      take.assign(st->blocks.begin(), st->blocks.end());
      st->blocks.clear();
      st->busy += take.size();
    }
    std::vector<sxc_pair> pairs;
    for (const WirePair& p : take) append_block(job, p, &pairs);
    std::vector<sxc_result> res;
    if (!align_pairs(job, pairs, &res)) device_failure("aligning a batch");
    {
      std::lock_guard<std::mutex> g(st->mu);
      st->results.insert(st->results.end(), res.begin(), res.end());
      st->busy -= take.size();
    }
    st->space_cv.notify_all();
  }
}

// Serving loop: every block is a request as soon as it arrives.
void serve_submitter(const Job& job, State* st) {
  const int nt = (int)job.target->chunks.size(), nq = (int)job.query->chunks.size();
  for (;;) {
    std::vector<WirePair> take;
    {
      std::unique_lock<std::mutex> lk(st->mu);
      st->work_cv.wait(lk, [&] { return st->finished.load() || !st->blocks.empty(); });
      if (st->finished.load()) return;
      take.assign(st->blocks.begin(), st->blocks.end());
      st->blocks.clear();
    }
    for (const WirePair& p : take) {
      if (p.target_from < 0 || p.target_to >= nt || p.query_from < 0 || p.query_to >= nq ||
          p.target_from > p.target_to || p.query_from > p.query_to) {
        printf("ERROR block t:%d-%d q:%d-%d is outside the %d target / %d query chunks; skipped\n", p.target_from,
               p.target_to, p.query_from, p.query_to, nt, nq);
        continue;
      }
      const sxc_block b{p.target_from, p.target_to, p.query_from, p.query_to, p.fast ? 1 : 0};
      uint64_t ticket = 0;
      if (sxc_server_submit(job.srv, &b, 1, &ticket) != SXC_OK) device_failure("submitting a block");
      std::lock_guard<std::mutex> g(st->mu);
      st->running.push_back(ticket);
      ++st->busy;
    }
    st->done_cv.notify_all();
  }
}

// Serving loop: collects blocks' records in arrival order.
void serve_collector(const Job& job, State* st) {
  std::vector<sxc_result> r;
  for (;;) {
    uint64_t ticket;
    {
      std::unique_lock<std::mutex> lk(st->mu);
      st->done_cv.wait(lk, [&] { return st->finished.load() || !st->running.empty(); });
      if (st->finished.load()) return;
      ticket = st->running.front();
    }
    int64_t n = 0;
    int rc;
    while ((rc = sxc_server_wait(job.srv, ticket, 200, &n, nullptr)) == SXC_PENDING)
      if (st->finished.load()) return;
    if (rc != SXC_OK) device_failure("serving a block");
    r.resize((size_t)std::max<int64_t>(n, 1));
    if (sxc_server_results(job.srv, ticket, r.data(), (int64_t)r.size(), &n) != SXC_OK)
      device_failure("collecting a block");
    {
      std::lock_guard<std::mutex> g(st->mu);
      st->results.insert(st->results.end(), r.begin(), r.begin() + n);
      st->running.pop_front();
      --st->busy;
    }
    st->space_cv.notify_all();
  }
}

// Both genomes and all target spectra resident, or null if they do not fit the device.
sxc_server* make_server(const Job& job) {
  auto descs = [](const ChunkSet& cs) {
    std::vector<sxc_chunk_desc> d(cs.chunks.size());
    for (size_t k = 0; k < cs.chunks.size(); ++k) {
      const sxc_host::Chunk& c = cs.chunks[k];
      d[k] = sxc_chunk_desc{c.data, c.len, c.start, c.record, cs.lengths[(size_t)c.record]};
    }
    return d;
  };
  const std::vector<sxc_chunk_desc> t = descs(*job.target), q = descs(*job.query);
  sxc_server* srv = nullptr;
  const int rc = sxc_server_create(job.ctx, &job.prm, t.data(), (int32_t)t.size(), q.data(), (int32_t)q.size(),
                                   nullptr, &srv);
  if (rc != SXC_OK) {
    printf("serving loop unavailable (%s); aligning queued blocks in batches\n", sxc_last_error());
    return nullptr;
  }
  return srv;
}

}  // namespace

int main(int argc, char** argv) {
  sxc_host::Flags P(argc, argv, "Compares two sequences via cross-correlation");
  P.add("-master", "string", "name of the submit host", nullptr);
  P.add("-port", "int", "port of the master to connect", "3491");
  P.add("-p", "int", "number of working processes (threads)", "1");
  P.add("-sid", "int", "slave_id", nullptr);
  P.add("-q", "string", "query fasta sequence", nullptr);
  P.add("-t", "string", "target fasta sequence", nullptr);
  P.add("-l", "int", "minimum alignment length", "0");
  P.add("-q_chunk", "int", "query chunk size", "4096");
  P.add("-t_chunk", "int", "target chunk size", "4096");
  P.add("-min_prob", "double", "minimum probability to keep match", "0.9999");
  P.add("-cutoff", "double", "signal selection cutoff", "1.8");
  P.add("-cutoff_fast", "double", "signal selection cutoff (fast)", "2.9");
  P.add("-prob_table", "bool", "lookup probability in table (faster, less accurate)", "0");
  P.add("-debug_targets", "int",
        "number of targets to receive before exiting (for debug purposes mostly, don't use)", "0");
  P.add("-device", "int", "HIP device to run on", "0");
  P.add("-autotune", "bool", "time the compiled FFT kernel geometries at this length and run the fastest", "1");
  P.parse();

  const std::string master = P.str("-master");
  const int port = P.i32("-port");
  int32_t slave_id = P.i32("-sid");
  const int threads = P.i32("-p") > 0 ? P.i32("-p") : 1;
  const std::string qfile = P.str("-q"), tfile = P.str("-t");
  const int min_len = P.i32("-l");
  const int q_chunk = P.i32("-q_chunk"), t_chunk = P.i32("-t_chunk");
  const double cutoff = P.f64("-cutoff"), cutoff_fast = P.f64("-cutoff_fast");
  const bool prob_table = P.boolean("-prob_table");
  const unsigned debug_max = (unsigned)P.i32("-debug_targets");
  const int device = P.i32("-device");
  const bool autotune = P.boolean("-autotune");
  // FilterMatches keeps prob >= 0.99 (m_minProb, :76); -min_prob is only the probability a
  // ProbTable hit reports (:413-421, analysis/ProbTable.cc:132-136)
  const double min_prob_flag = P.f64("-min_prob");

  time_t t0 = time(nullptr);
  Records qraw, traw;
  std::string err;
  printf("Loading query sequence:  %s\n", qfile.c_str());
  if (!sxc_host::read_fasta(qfile, &qraw, &err)) {
    printf("ERROR %s\n", err.c_str());
    return 1;
  }
  ChunkSet query, target;
  sxc_host::chunk_records(qraw, q_chunk, 0, &query);
  printf("Loading target sequence:  %s\n", tfile.c_str());
  if (!sxc_host::read_fasta(tfile, &traw, &err)) {
    printf("ERROR %s\n", err.c_str());
    return 1;
  }
  sxc_host::chunk_records(traw, t_chunk, t_chunk / 4, &target);
  printf("query chunks: %zu  target chunks: %zu\n", query.chunks.size(), target.chunks.size());
  printf("TIME SPENT ON LOADING: %ld\n", (long)(time(nullptr) - t0));

  Job job;
  job.query = &query;
  job.target = &target;
  job.prm.N = 2 * t_chunk;
  job.prm.q_chunk = q_chunk;
  job.prm.min_len = min_len;
  job.prm.strands = 2;
  job.prm.cutoff = cutoff;
  job.prm.cutoff_fast = cutoff_fast;
  job.prm.min_prob = 0.99;
  job.prm.target_total = target.total();
  job.prm.prob_table = prob_table ? 1 : 0;
  job.prm.pad_ = 0;
  job.prm.table_prob = min_prob_flag;
  if (sxc_create(&job.ctx, device) != SXC_OK) {
    printf("ERROR %s\n", sxc_last_error());
    return 1;
  }
  if (autotune) {
    sxc_tune_entry tune[16];
    int32_t nt = 0;
    if (sxc_autotune(job.ctx, job.prm.N, job.prm.N, tune, 16, &nt) != SXC_OK) {
      printf("ERROR %s\n", sxc_last_error());
      return 1;
    }
  }

  job.srv = make_server(job);
  State st;
  std::thread worker(job.srv ? serve_submitter : submitter, std::cref(job), &st);
  std::thread collector;
  if (job.srv) collector = std::thread(serve_collector, std::cref(job), &st);
  printf("== Entering communication loop ==\ncomm loop for %s %d\n", master.c_str(), port);
  fflush(stdout);

  struct hostent* server = gethostbyname(master.c_str());
  if (!server) {
    printf("ERROR resolving master hostname\n");
    st.finished = true;
    st.work_cv.notify_all();
    st.done_cv.notify_all();
    worker.join();
    if (collector.joinable()) collector.join();
    sxc_server_destroy(job.srv);
    sxc_destroy(job.ctx);
    return -1;
  }
  sockaddr_in addr;
  memset(&addr, 0, sizeof addr);
  addr.sin_family = AF_INET;
  memcpy(&addr.sin_addr.s_addr, server->h_addr, (size_t)server->h_length);
  addr.sin_port = htons((uint16_t)port);

  unsigned total_blocks = 0;
  while (!st.finished.load()) {
    {
      std::unique_lock<std::mutex> lk(st.mu);
      // the reference asks while fewer than 2 * p blocks wait, with up to p more being worked
      // on by its threads; blocks in the serving loop count as waiting or being worked on
      if (st.blocks.size() + st.busy >= (size_t)(3 * threads)) {
        st.space_cv.wait_for(lk, std::chrono::seconds(10));
        continue;
      }
    }
    int fd = -1;
    while (fd < 0) {
      fd = socket(AF_INET, SOCK_STREAM, 0);
      if (fd < 0) {
        printf("ERROR opening socket\n");
        sleep(1);
        continue;
      }
      if (connect(fd, (sockaddr*)&addr, sizeof addr) < 0) {
        printf("ERROR connecting to master: %s\n", strerror(errno));
        close(fd);
        fd = -1;
        sleep(1);
        continue;
      }
      if (!write_all(fd, &slave_id, sizeof slave_id)) {
        printf("ERROR sending id to master\n");
        close(fd);
        fd = -1;
        sleep(1);
      }
    }
    std::vector<sxc_result> out;
    {
      std::lock_guard<std::mutex> g(st.mu);
      out.swap(st.results);
    }
    const uint32_t count = (uint32_t)out.size();
    bool ok = write_all(fd, &count, sizeof count) && write_all(fd, out.data(), out.size() * sizeof(sxc_result));
    int32_t m = 0;
    ok = ok && read_all(fd, &m, sizeof m);
    if (!ok) {
      // The master closed early: keep the results for the next exchange.
      std::lock_guard<std::mutex> g(st.mu);
      st.results.insert(st.results.begin(), out.begin(), out.end());
      close(fd);
      sleep(1);
      continue;
    }
    if (m == -1) {
      close(fd);
      std::lock_guard<std::mutex> g(st.mu);
      st.finished = true;
      break;
    }
    if (m > 0) {
      std::vector<WirePair> tv((size_t)m);
      if (!read_all(fd, tv.data(), tv.size() * sizeof(WirePair))) {
        printf("ERROR reading %d blocks from master\n", m);
        close(fd);
        continue;
      }
      close(fd);
      {
        std::lock_guard<std::mutex> g(st.mu);
        for (const WirePair& p : tv) st.blocks.push_back(p);
      }
      st.work_cv.notify_all();
      total_blocks += (unsigned)m;
      if (debug_max && total_blocks >= debug_max) st.finished = true;
    } else {
      close(fd);
      // Nothing to do: poll the master again shortly.
      std::this_thread::sleep_for(std::chrono::milliseconds(50));
    }
  }
  st.work_cv.notify_all();
  st.done_cv.notify_all();
  printf("== Processing finished, waiting for the slaves to die ==\n");
  printf("TIME SPENT WORKING: %ld\n", (long)(time(nullptr) - t0));
  fflush(stdout);
  worker.join();
  if (collector.joinable()) collector.join();
  sxc_server_destroy(job.srv);
  sxc_destroy(job.ctx);
  return 0;
}
