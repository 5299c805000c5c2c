"""Request streams for the serving loop, shaped like Satsuma2's battleship search.

The master's grid search (analysis/GridSearch.cc:188-411, driven from
analysis/SatsumaSynteny2.cc:484-599) asks for two kinds of t_pair blocks:
  * seed lines (CollectSeeds): one target chunk against a run of query chunks, with the
    fast cutoff;
  * neighbourhoods (CollectTargets): the cells of its search matrix next to earlier hits,
    blocks of target x query chunks with the normal cutoff.
`Battleship` produces both from a pair of chunked genomes: seeds at random until hits come
back, then the 24 x 24-chunk matrix cell of every hit not asked for yet.
`run_stream` offers requests as a Poisson stream (a rate of requests, or of chunk pairs with
gaps scaled by each request's pairs) to a `Server` and reports per-request latency
percentiles and the chunk pairs per second completed.  Latency is counted from each
request's scheduled arrival (completion minus the arrival time the Poisson schedule gave
it), so a generator that falls behind an overloaded server does not hide the queueing
(coordinated omission); the server's own submission-to-completion time is reported beside
it as the service latency.
"""
from __future__ import annotations

import collections
import threading
import time
from typing import List, Sequence, Tuple

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


def _pairs(blocks: Sequence[Block]) -> int:
    return sum((b[1] - b[0] + 1) * (b[3] - b[2] + 1) for b in blocks)


def run_stream(server, producer: Battleship, rate: float, duration: float, seed: int = 0,
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
