"""Request streams for the serving loop, shaped like Satsuma2's search.

`GridSearchProducer` is the master's search (analysis/SatsumaSynteny2.cc:484-599) over two
genomes: the restated GridSearch, match-set steps and chaining of satsuma2_amd.master
(include/sxc_master.h) run on a thread of their own, as the reference's master does beside
its slaves; each collected GridTarget (a block of target x query chunks, seed lines with
the fast cutoff) is one request, and the records of completed requests feed the next
iteration.  `Battleship` is a simple stand-in with the same two request shapes (seed lines
at random, then the 24 x 24-chunk cell of every hit), kept for tests that need a request
stream independent of the search's state.
`run_stream` offers requests as a Poisson stream (a rate of requests, or of chunk pairs with
gaps scaled by each request's pairs) to a `Server` and reports per-request latency
percentiles and the chunk pairs per second completed.  Latency is counted from each
request's scheduled arrival (completion minus the arrival time the Poisson schedule gave
it), so a generator that falls behind an overloaded server, or waits for the master's next
targets, does not hide the queueing (coordinated omission); the server's own
submission-to-completion time is reported beside it as the service latency.
"""
from __future__ import annotations

import collections
import threading
import time
from typing import List, Optional, Sequence, Tuple

import numpy as np

Block = Tuple[int, int, int, int, int]


def genome_chunks(seq: bytes, size: int, step: int, record: int = 0):
    """ChunkManager::ChunkItSelect (analysis/SeqChunk.cc:72-165) for one record without N."""
    from .api import Chunk
    return [Chunk(seq[j * step: min(j * step + size, len(seq))], j * step, record, len(seq))
            for j in range(1 + len(seq) // step)]


class Battleship:
    def __init__(self, n_target: int, n_query: int, rng: np.random.Generator, line: int = 64, cell: int = 24,
                 t_step: int = 3072, q_step: int = 4096):
        self.nt, self.nq, self.rng = n_target, n_query, rng
        self.line, self.cell, self.t_step, self.q_step = line, cell, t_step, q_step
        self.queue: collections.deque = collections.deque()
        self.asked = set()
        self.lock = threading.Lock()

    def seed(self) -> List[Block]:
        x = int(self.rng.integers(0, self.nt))
        y = int(self.rng.integers(0, max(1, self.nq - self.line)))
        return [(x, x, y, min(y + self.line - 1, self.nq - 1), 1)]

    def on_results(self, recs: np.ndarray):
        """Queue the matrix cell of every hit (analysis/GridSearch.cc:318-349 maps matches to
        cells; reverse-strand query starts are mirrored back to forward coordinates)."""
        if len(recs) == 0:
            return
        ti = recs["tstart"].astype(np.int64) // self.t_step
        qs = recs["qstart"].astype(np.int64)
        rev = recs["reverse"] != 0
        qs = np.where(rev, recs["query_size"].astype(np.int64) - qs - recs["len"].astype(np.int64), qs)
        qi = np.maximum(qs, 0) // self.q_step
        keys = np.unique((ti // self.cell) * (1 << 32) + qi // self.cell)
        cells = np.stack([keys >> 32, keys & 0xFFFFFFFF], 1)
        with self.lock:
            for cx, cy in cells.tolist():
                if (cx, cy) in self.asked:
                    continue
                self.asked.add((cx, cy))
                self.queue.append([(cx * self.cell, min(cx * self.cell + self.cell - 1, self.nt - 1),
                                    cy * self.cell, min(cy * self.cell + self.cell - 1, self.nq - 1), 0)])

    def next(self) -> List[Block]:
        with self.lock:
            if self.queue:
                return self.queue.popleft()
        return self.seed()


class GridSearchProducer:
    """The master's search loop as a request source.  A background thread runs
    SearchIteration (grid update from the chained matches, CollectTargets) whenever results
    arrived or fewer than `queue` targets are pending, asking for the difference, as the
    master tops up its work queue (analysis/SatsumaSynteny2.cc:544-555).  When the grid has
    nothing left, the search starts over on the same genomes (`restarts`), so a stream can
    run for any duration.  `seeds` (RESULT_DTYPE, e.g. KMatch's) start the match set as seed
    files do (analysis/SatsumaSynteny2.cc:466-477)."""

    def __init__(self, target: bytes, query: bytes, chunk: int = 4096, blocks: int = 24, queue: int = 128,
                 seeds: Optional[np.ndarray] = None, min_seed_length: int = 24, dups: bool = False):
        self.T, self.Q, self.chunk, self.blocks, self.queue = target, query, chunk, blocks, queue
        self.seeds, self.min_seed_length, self.dups = seeds, min_seed_length, dups
        self.pending: collections.deque = collections.deque()
        self.inbox: List[np.ndarray] = []
        self.cv = threading.Condition()
        self.closed = False
        self.iterations = 0
        self.iteration_ms: List[float] = []
        self.restarts = 0
        self.targets = 0
        self.seed_lines = 0
        self._start()
        self.thread = threading.Thread(target=self._loop, daemon=True)
        self.thread.start()

    def _start(self):
        from .master import Grid, SearchIteration
        self.grid = Grid.from_records(self.chunk, self.blocks, [("target", self.T)], [("query", self.Q)])
        self.it = SearchIteration(self.grid, [len(self.T)], [len(self.Q)], dups=self.dups)
        if self.seeds is not None and len(self.seeds):
            self.it.seed(self.seeds, self.min_seed_length)

    def _loop(self):
        while True:
            with self.cv:
                while not self.closed and not self.inbox and len(self.pending) >= self.queue:
                    self.cv.wait()
                if self.closed:
                    return
                res = np.concatenate(self.inbox) if self.inbox else np.zeros(0, _lib_result_dtype())
                self.inbox = []
                want = max(0, self.queue - len(self.pending))
            t0 = time.perf_counter()
            targets, _, _ = self.it.step(res, want)
            if len(targets) == 0 and not len(res):
                with self.cv:
                    idle = not self.pending
                if idle:
                    self._start()
                    self.restarts += 1
                    targets, _, _ = self.it.step(np.zeros(0, _lib_result_dtype()), want)
            self.iteration_ms.append(1e3 * (time.perf_counter() - t0))
            self.iterations += 1
            with self.cv:
                for t in targets.tolist():
                    tf, tl, qf, ql, _x, _y, _y2, _count, fast = t
                    self.pending.append([(tf, tl, qf, ql, fast)])
                    self.seed_lines += fast
                self.targets += len(targets)
                self.cv.notify_all()
                if not len(targets) and not self.inbox:
                    # nothing new until results come back
                    self.cv.wait(timeout=0.05)

    def next(self) -> List[Block]:
        with self.cv:
            while not self.pending and not self.closed:
                self.cv.notify_all()
                self.cv.wait(timeout=0.1)
            return self.pending.popleft() if self.pending else []

    def on_results(self, recs: np.ndarray):
        with self.cv:
            self.inbox.append(np.asarray(recs))
            self.cv.notify_all()

    def close(self):
        with self.cv:
            self.closed = True
            self.cv.notify_all()
        self.thread.join()

    def stats(self) -> dict:
        ms = np.array(self.iteration_ms) if self.iteration_ms else np.zeros(1)
        return {"iterations": self.iterations, "targets": self.targets, "seed_lines": self.seed_lines,
                "restarts": self.restarts, "iteration_ms_p50": round(float(np.percentile(ms, 50)), 2),
                "iteration_ms_max": round(float(ms.max()), 2)}


def run_search(server, target: bytes, query: bytes, seeds: Optional[np.ndarray] = None, budget_s: float = 20.0,
               max_rounds: int = 1000, slaves: int = 1, threads: int = 16, per_block: int = 4, min_matches: int = 20,
               dups: bool = False, chunk: int = 4096, blocks: int = 24) -> Tuple[List[List[Block]], dict]:
    """SatsumaSynteny2's search loop (analysis/SatsumaSynteny2.cc:484-599) with `server` as
    its slaves, in synchronous rounds: the master collects targets, all of them are aligned,
    and their records are the next round's new matches (one schedule the reference's
    asynchronous queue can take).  The queue size starts at slaves * threads * 8 and grows by
    slaves * threads while the queue runs low (:544-555); the loop ends when no target is
    left, after 10 rounds in a row with fewer than per_block * min_matches new matches per
    slave request (:572-585; a slave request carries 2 * threads targets, WorkQueue.cc:95-110),
    or at the time budget.  Returns the requests in order and the run's figures."""
    from ._lib import RESULT_DTYPE
    from .master import Grid, SearchIteration
    grid = Grid.from_records(chunk, blocks, [("target", target)], [("query", query)])
    it = SearchIteration(grid, [len(target)], [len(query)], dups=dups)
    t_start = time.perf_counter()
    if seeds is not None and len(seeds):
        it.seed(seeds)
    seed_s = time.perf_counter() - t_start
    queue = slaves * threads * 8
    min_per_slave = per_block * min_matches
    exit_counter = extra = 10
    seen_match = False
    requests: List[List[Block]] = []
    new = np.zeros(0, RESULT_DTYPE)
    master_ms, align_ms, per_round = [], [], []
    pairs = records = 0
    end = "targets exhausted"
    for r in range(max_rounds):
        t0 = time.perf_counter()
        want = queue
        if want > 0.75 * queue:
            queue += slaves * threads
            want = queue
        targets, cand, cov = it.step(new, want)
        t1 = time.perf_counter()
        master_ms.append(1e3 * (t1 - t0))
        if not len(targets):
            break
        reqs = [[(int(t["target_first"]), int(t["target_last"]), int(t["query_first"]), int(t["query_last"]),
                  int(t["fast"]))] for t in targets]
        tickets = [server.submit(b) for b in reqs]
        got = []
        for tk in tickets:
            server.wait(tk)
            got.append(server.results(tk))
        new = np.concatenate(got) if got else np.zeros(0, RESULT_DTYPE)
        align_ms.append(1e3 * (time.perf_counter() - t1))
        requests.extend(reqs)
        npairs = _pairs([b for rq in reqs for b in rq])
        pairs += npairs
        records += len(new)
        slave_msgs = max(1, -(-len(reqs) // (2 * threads)))
        per_round.append({"targets": len(reqs), "pairs": npairs, "new_matches": len(new),
                          "matches": len(it.matches), "master_ms": round(master_ms[-1], 1),
                          "align_ms": round(align_ms[-1], 1)})
        if len(new):
            seen_match = True
        if seen_match:
            if len(new) / slave_msgs < min_per_slave:
                exit_counter -= 1
                if not exit_counter:
                    end = "converged"
                    break
            else:
                exit_counter = extra
        if time.perf_counter() - t_start > budget_s:
            end = "time budget"
            break
    wall = time.perf_counter() - t_start
    return requests, {"rounds": len(per_round), "end": end, "seconds": round(wall, 2), "seed_s": round(seed_s, 2),
                      "chunk_pairs": pairs, "pairs_per_s": round(pairs / max(wall, 1e-9), 1), "records": records,
                      "matches": len(it.matches), "master_s": round(sum(master_ms) / 1e3, 2),
                      "align_s": round(sum(align_ms) / 1e3, 2), "rounds_detail": per_round[:12]}


def _lib_result_dtype():
    from ._lib import RESULT_DTYPE
    return RESULT_DTYPE


class Replay:
    """Recorded requests (e.g. run_search's) offered again in order, cycling."""

    def __init__(self, requests: Sequence[List[Block]]):
        self.requests = list(requests) or [[(0, 0, 0, 0, 0)]]
        self.i = 0

    def next(self) -> List[Block]:
        r = self.requests[self.i % len(self.requests)]
        self.i += 1
        return r

    def on_results(self, recs: np.ndarray):
        pass


def _pairs(blocks: Sequence[Block]) -> int:
    return sum((b[1] - b[0] + 1) * (b[3] - b[2] + 1) for b in blocks)


def run_stream(server, producer, rate: float, duration: float, seed: int = 0,
               pairs_rate: float = 0.0) -> dict:
    """Offers requests for `duration` seconds: a Poisson stream of `rate` requests/s, or, with
    pairs_rate > 0, requests spaced by exponential gaps whose mean is the request's chunk
    pairs / pairs_rate (the offered chunk pairs per second, whatever mix of seeds and cells
    the search produces)."""
    rng = np.random.default_rng(seed)
    done: collections.deque = collections.deque()
    tickets: collections.deque = collections.deque()
    cv = threading.Condition()
    state = {"closed": False}

    def complete():
        while True:
            with cv:
                while not tickets and not state["closed"]:
                    cv.wait()
                if not tickets:
                    return
                t, t_sched, t_sub, npairs = tickets.popleft()
            n, lat = server.wait(t)
            recs = server.results(t)
            producer.on_results(recs)
            # completion = submission + the server's latency; latency from the schedule adds
            # the time the request waited for the generator
            done.append((t_sched, (t_sub - t_sched) * 1e3 + lat, lat, npairs, n))

    th = threading.Thread(target=complete, daemon=True)
    th.start()
    t0 = time.perf_counter()
    t_next = t0
    submitted = 0
    submitted_pairs = 0
    while True:
        blocks = producer.next()
        mean_gap = _pairs(blocks) / pairs_rate if pairs_rate > 0 else 1.0 / rate
        t_next += float(rng.exponential(mean_gap))
        if t_next - t0 > duration:
            break
        delay = t_next - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        t_sub = time.perf_counter()
        ticket = server.submit(blocks)
        submitted += 1
        submitted_pairs += _pairs(blocks)
        with cv:
            tickets.append((ticket, t_next, t_sub, _pairs(blocks)))
            cv.notify()
    with cv:
        state["closed"] = True
        cv.notify()
    th.join()
    lat = np.array([d[1] for d in done])
    svc = np.array([d[2] for d in done])
    end = max(d[0] - t0 + d[1] / 1e3 for d in done) if done else duration
    pairs = sum(d[3] for d in done)

    def pct(a, q):
        return round(float(np.percentile(a, q)), 3) if len(a) else None

    return {"offered_requests_per_s": round(submitted / duration, 1),
            "offered_pairs_per_s": round(submitted_pairs / duration, 1), "requests": submitted,
            "seconds": round(end, 3),
            "pairs_per_s": round(pairs / end, 1), "mean_pairs_per_request": round(pairs / max(1, len(done)), 1),
            "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
            "service_p50_ms": pct(svc, 50), "service_p99_ms": pct(svc, 99),
            "records": int(sum(d[4] for d in done))}
