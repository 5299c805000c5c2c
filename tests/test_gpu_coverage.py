"""Every benchmarked workload checked against the reference restatement (the oracle), on the
GPU through the C ABI.

  * the full 4,096-pair uniform batch bench.py times (not a sample of it);
  * A/C/G/T/N chunks (assembly gaps) at every FFT length, per-pair N, with and without
    -prob_table, and through the serving loop's small launches (split scans);
  * the serving loop on special chunks (B/D/H/V, N) and at a keep threshold decided on the host;
  * the large served run's genome model (GC windows, tandem repeats, transposon families,
    syntenic blocks) at 2 Mbp, served from battleship-shaped requests.
Bar: coordinates, lengths, strands, identities bit-exact; prob within 1e-11 relative (the
device's erf/exp; keep decisions near min_prob are made on the host)."""
import os

import numpy as np
import pytest

import oracle
from conftest import threshold_size
from satsuma2_amd import synth

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))
FIELDS_EXACT = ["query_id", "target_id", "query_size", "qstart", "tstart", "len", "reverse", "ident"]


@pytest.fixture(scope="module")
def ctx():
    from satsuma2_amd.api import Context

    c = Context(0)
    yield c
    c.close()


def assert_same(got, ref, what):
    from test_gpu_server import record_diff

    same = len(got) == len(ref) and all(np.array_equal(got[f], ref[f]) for f in FIELDS_EXACT)
    assert same, f"{what}: {record_diff(got, ref)}"
    if len(ref):
        rel = np.abs(got["prob"] - ref["prob"]) / np.maximum(np.abs(ref["prob"]), 1e-300)
        assert rel.max() <= 1e-11, f"{what}: prob rel err {rel.max()}"


def oracle_pairs(pairs):
    return [dict(tseq=p.tseq, qseq=p.qseq, t_chunk_start=p.t_chunk_start, q_chunk_start=p.q_chunk_start,
                 t_record=p.t_record, q_record=p.q_record,
                 q_record_len=p.q_record_len if p.q_record_len >= 0 else len(p.qseq), fast=p.fast) for p in pairs]


def test_uniform_4096_batch_equals_reference(ctx):
    """The bench's workload, all 4,096 pairs (both strands), as the graph replays it."""
    from satsuma2_amd.api import ChunkPair, ResidentBatch, SlaveParams

    T, Q = synth.uniform_pairs(4096, seed=1)
    pairs = [ChunkPair(T[r], Q[r], t_record=r, q_record=r, q_record_len=4096) for r in range(4096)]
    size = float(4096 * 4096)
    ref = oracle.align_batch(oracle_pairs(pairs), oracle.make_params(target_total=size), threads=THREADS,
                             cap_per_pair=512)
    assert max(len(r) for r in ref) < 512
    ref = np.concatenate(ref)
    rb = ResidentBatch(ctx, pairs, SlaveParams(target_total=size))
    for k in range(3):  # eager, capture, replay
        rb.run()
        assert_same(rb.results(), ref, f"run {k}")
    c, r = rb.stats()
    assert c > 0 and r >= len(ref)
    rb.close()


def acgtn_pair(rng, lt, lq):
    """A/C/G/T chunks with a planted homology and assembly gaps: N runs on both sides, a
    single N, N at the chunk ends."""
    t = synth.random_seq(rng, lt, 0.41)
    q = synth.random_seq(rng, lq, 0.41)
    L = min(lt, lq) // 2
    a, b = int(rng.integers(0, lt - L + 1)), int(rng.integers(0, lq - L + 1))
    q[b:b + L] = synth.mutate(rng, t[a:a + L], 0.9)
    for s in (t, q):
        for _ in range(3):
            at, n = int(rng.integers(0, s.size)), int(rng.integers(1, max(2, s.size // 20)))
            s[at:at + n] = ord("N")
        s[int(rng.integers(0, s.size))] = ord("N")
    t[0] = ord("N")
    q[-1] = ord("N")
    return t.tobytes(), q.tobytes()


# (target, query) lengths spanning every FFT length class 512 .. 32768 (N = nextpow2(q + t))
ACGTN_LENGTHS = [(300, 200), (600, 400), (1200, 800), (2000, 2000), (4096, 4096), (4000, 3000), (8000, 8000),
                 (6000, 9000), (16000, 16000), (16384, 12000)]


@pytest.mark.parametrize("table", [False, True])
def test_acgtn_pairs_each_fft_length(ctx, table):
    from satsuma2_amd.api import ChunkPair, HomologyByXCorr, ResidentBatch, SlaveParams

    rng = np.random.default_rng(17)
    pairs = []
    for k, (lt, lq) in enumerate(ACGTN_LENGTHS):
        for rep in range(2):
            t, q = acgtn_pair(rng, lt, lq)
            pairs.append(ChunkPair(t, q, t_record=k, q_record=k))
    size = 1e6
    prm = SlaveParams(t_chunk=0, q_chunk=0, target_total=size, prob_table=table, table_prob=0.9999)
    ptable = oracle.prob_table(size, threads=THREADS) if table else None
    # each pair at its own length, N = next power of two >= q + t (at least 512)
    ops = oracle_pairs(pairs)
    ref = [oracle.align_batch([op], oracle.make_params(N=max(512, synth.fft_len(len(op["qseq"]), len(op["tseq"]))),
                                                       q_chunk=0, target_total=size, ptable=ptable,
                                                       table_prob=0.9999))[0] for op in ops]
    assert sum(len(r) for r in ref) > 0
    assert len({max(512, synth.fft_len(len(op["qseq"]), len(op["tseq"]))) for op in ops}) == 7
    assert_same(HomologyByXCorr(ctx, prm).align_batch(pairs), np.concatenate(ref), f"table={table} one-shot")
    rb = ResidentBatch(ctx, pairs, prm)
    rb.run()
    rb.run()
    assert_same(rb.results(), np.concatenate(ref), f"table={table} resident")
    rb.close()


def _server(ctx, tch, qch, size, **kw):
    from satsuma2_amd.api import Chunk, Server, SlaveParams

    return Server(ctx, SlaveParams(target_total=size, **{k: v for k, v in kw.items() if k in ("prob_table",
                                                                                              "table_prob")}),
                  [Chunk(s, st, r, rl) for s, st, r, rl in tch], [Chunk(s, st, r, rl) for s, st, r, rl in qch],
                  max_pairs_per_launch=kw.get("P", 256), max_slots=2)


def test_server_special_chunks(ctx, golden):
    """A/C/G/T/N and B/D/H/V chunks through the serving loop: one block per chunk pair (small
    launches: each (pair, strand)'s lags split over several scan workgroups), the same with
    -prob_table, and the records equal the oracle's."""
    from test_gpu_server import oracle_records

    rng = np.random.default_rng(23)
    g = golden("bdhv_pairs.npz")
    ts, qs = [], []
    for k in range(4):
        t, q = acgtn_pair(rng, 4096, 4096)
        ts.append(t)
        qs.append(q)
    for name in g["names"][:3]:  # the N = 8192 ones
        ts.append(g[f"{name}_t"].tobytes()[:4096])
        qs.append(g[f"{name}_q"].tobytes()[:4096])
    tch = [(t, 0, k, len(t)) for k, t in enumerate(ts)]
    qch = [(q, 0, k, len(q)) for k, q in enumerate(qs)]
    size = 4096.0 * 64
    for table in (False, True):
        kw = dict(prob_table=True, table_prob=0.9999) if table else {}
        srv = _server(ctx, tch, qch, size, **kw)
        ptable = oracle.prob_table(size, threads=THREADS) if table else None
        for k in range(len(ts)):
            blocks = [(k, k, k, k, 0)]
            pairs = [dict(tseq=ts[k], qseq=qs[k], t_record=k, q_record=k, q_record_len=len(qs[k]))]
            ref = oracle.align_batch(pairs, oracle.make_params(target_total=size, ptable=ptable, table_prob=0.9999),
                                     threads=THREADS)[0]
            assert len(ref) > 0, k
            assert_same(srv.results(srv.submit(blocks)), ref, f"table={table} pair {k}")
        # all of them in one request, with a neighbour block
        blocks = [(0, len(ts) - 1, 0, len(qs) - 1, 0)]
        if not table:
            assert_same(srv.results(srv.submit(blocks)), oracle_records(tch, qch, blocks, size), "all pairs")
        srv.close()


def test_server_keep_band(ctx, golden):
    """The serving loop decides runs near min_prob on the host: a target size that puts a run
    of the single pair exactly on the threshold gives the oracle's records."""
    g = golden("single_pair.npz")
    t, q = g["t"].tobytes(), g["q"].tobytes()
    m = g["m0"]
    T = None
    for r in m[m["len"] >= 40]:
        T = threshold_size(t, q, int(r["st"]), int(r["sq"]), int(r["len"]))
        if T is not None and 1e3 < T < 1e8:
            break
    assert T is not None and 1e3 < T < 1e8
    for size in (T, np.nextafter(T, np.inf)):
        srv = _server(ctx, [(t, 0, 0, len(t))], [(q, 0, 0, len(q))], float(size))
        ref = oracle.align_batch([dict(tseq=t, qseq=q)], oracle.make_params(target_total=float(size)))[0]
        assert_same(srv.results(srv.submit([(0, 0, 0, 0, 0)])), ref, f"T={size!r}")
        assert srv.stats()["near_threshold"] >= 1
        srv.close()


def test_large_run_model_served(ctx):
    """The large served run's genome model at 2 Mbp (GC per 100-kbp window, 10% tandem
    repeats, five transposon families of 10 copies, 50 syntenic blocks): battleship-shaped
    requests (seed lines, then the matrix cells of hits) served and compared with the oracle."""
    from satsuma2_amd.api import Server, SlaveParams
    from satsuma2_amd.serve_load import Battleship, genome_chunks
    from test_gpu_server import oracle_records

    T, Q, _ = synth.large_genomes(length=2_000_000, n_blocks=50, seed=5, copies=10)
    tch = genome_chunks(T, 4096, 3072)
    qch = genome_chunks(Q, 4096, 4096)
    size = float(len(T))
    srv = Server(ctx, SlaveParams(target_total=size), tch, qch, max_pairs_per_launch=512)
    prod = Battleship(len(tch), len(qch), np.random.default_rng(8), line=32)
    reqs, tickets = [], []
    for k in range(60):
        blocks = prod.next()
        reqs.append(blocks)
        tickets.append(srv.submit(blocks))
        if k % 10 == 9:  # let hits come back so cells get asked for
            for tk in tickets[-10:]:
                srv.wait(tk)
    got = [srv.results(tk) for tk in tickets]
    for recs in got:
        prod.on_results(recs)
    # and the cells of the hits, all at once
    while prod.queue and len(reqs) < 80:
        blocks = prod.next()
        reqs.append(blocks)
        got.append(srv.results(srv.submit(blocks)))
    srv.close()
    tl = [(c.seq, c.start, c.record, c.record_len if c.record_len >= 0 else len(c.seq)) for c in tch]
    ql = [(c.seq, c.start, c.record, c.record_len if c.record_len >= 0 else len(c.seq)) for c in qch]
    total = 0
    for i, blocks in enumerate(reqs):
        ref = oracle_records(tl, ql, blocks, size)
        assert_same(got[i], ref, f"request {i} {blocks}")
        total += len(ref)
    assert total > 100


def test_search_loop_served(ctx):
    """SatsumaSynteny2's request stream end to end at 2 Mbp: KMatch seeds on the GPU
    (K = 11 .. 31), the master's search loop (restated grid search and chaining) with the
    server as its slaves for a few rounds; the records of the first requests of each kind
    (cells, seed lines) equal the oracle's."""
    from satsuma2_amd.api import Server, SlaveParams
    from satsuma2_amd.kmatch import satsuma_seeds
    from satsuma2_amd.serve_load import genome_chunks, run_search
    from test_gpu_server import oracle_records

    T, Q, _ = synth.syntenic_genomes(length=2_000_000, n_blocks=80, seed=12)
    tch = genome_chunks(T, 4096, 3072)
    qch = genome_chunks(Q, 4096, 4096)
    seeds, st = satsuma_seeds(ctx, [Q], [T])
    assert len(seeds) > 100 and st["blocks"] == len(seeds)
    size = float(len(T))
    srv = Server(ctx, SlaveParams(target_total=size), tch, qch, max_pairs_per_launch=512)
    reqs, stats = run_search(srv, T, Q, seeds, budget_s=20.0, max_rounds=4, threads=2)
    assert stats["rounds"] >= 2 and stats["records"] > 0, stats
    cells = [r for r in reqs if r[0][4] == 0][:3]
    lines = [r for r in reqs if r[0][4] == 1][:2]
    assert cells
    tl = [(c.seq, c.start, c.record, c.record_len if c.record_len >= 0 else len(c.seq)) for c in tch]
    ql = [(c.seq, c.start, c.record, c.record_len if c.record_len >= 0 else len(c.seq)) for c in qch]
    for i, blocks in enumerate(cells + lines):
        print(f"[search] request {i} {blocks}", flush=True)
        got = srv.results(srv.submit(blocks))
        assert_same(got, oracle_records(tl, ql, blocks, size), f"request {i} {blocks}")
    srv.close()
