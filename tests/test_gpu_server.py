"""The serving loop (sxc_server_*) on the GPU.

Requests of t_pair blocks against resident genomes must return exactly the records the
reference slave's align_target produces for the same chunk pairs
(analysis/HomologyByXCorrSlave.cc:270-300; the oracle restates it, pinned against the
reference's own slave records) -- alone, and in a mix where launches carry pairs of many
requests and a request spans several launches."""
import collections
import os
import threading

import numpy as np
import pytest

import oracle
from satsuma2_amd import synth
from satsuma2_amd._lib import SxcError

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def _chunks(seq: bytes, size: int, step: int, record: int):
    """ChunkManager::ChunkItSelect (analysis/SeqChunk.cc:72-165) for one pure-ACGT record."""
    out = []
    for j in range(1 + len(seq) // step):
        s = j * step
        out.append((seq[s:min(s + size, len(seq))], s, record, len(seq)))
    return out


@pytest.fixture(scope="module")
def genomes():
    T, Q, _ = synth.syntenic_genomes(length=600_000, n_blocks=12, seed=9)
    # a record pair carrying 1500 bp of a period-3 tandem repeat: the reverse strand keeps
    # 151 records (oracle), more than a launch slot holds at first
    rng = np.random.default_rng(3)

    def embedded_repeat(period=3, n=1500):
        r = synth.random_seq(rng, 4096, 0.41).tobytes()
        unit = synth.random_seq(rng, period, 0.5).tobytes()
        return r[:1000] + (unit * (n // period + 1))[:n] + r[1000 + n:]

    trep, qrep = embedded_repeat(), embedded_repeat()
    tch = _chunks(T, 4096, 3072, 0) + _chunks(trep, 4096, 3072, 1)
    qch = _chunks(Q, 4096, 4096, 0) + _chunks(qrep, 4096, 4096, 1)
    return T, Q, tch, qch


@pytest.fixture(scope="module")
def served(genomes):
    from satsuma2_amd.api import Chunk, Context, Server, SlaveParams
    T, Q, tch, qch = genomes
    ctx = Context(0)
    params = SlaveParams(target_total=float(len(T) + 4096))
    srv = Server(ctx, params, [Chunk(s, st, r, rl) for s, st, r, rl in tch],
                 [Chunk(s, st, r, rl) for s, st, r, rl in qch], max_pairs_per_launch=256, max_slots=3)
    yield ctx, params, srv
    srv.close()
    ctx.close()


def _requests(tch, qch, rng, n=24):
    nt, nq = len(tch), len(qch)
    reqs = []
    for k in range(n):
        if k % 3 == 0:  # seed line: one target chunk against a run of query chunks (fast)
            x = int(rng.integers(0, nt - 2))
            y = int(rng.integers(0, nq - 40))
            reqs.append([(x, x, y, y + 39, 1)])
        else:  # neighbourhood block around the diagonal
            x = int(rng.integers(0, nt - 24))
            y = int(min(max(x * 3 // 4 + rng.integers(-8, 8), 0), nq - 24))
            reqs.append([(x, x + 23, y, y + 23, 0)])
    reqs.append([(nt - 2, nt - 1, nq - 2, nq - 1, 0)])  # the repeat records
    reqs.append([(0, 1, 0, 1, 0), (5, 6, 3, 3, 1)])    # two blocks in one request
    return reqs


def oracle_records(tch, qch, blocks, target_total, threads=THREADS):
    """align_target over the blocks' chunk pairs (query chunks outer, target chunks inner),
    both strands, by the oracle."""
    pairs = [dict(tseq=tch[ti][0], qseq=qch[qi][0], t_chunk_start=tch[ti][1], q_chunk_start=qch[qi][1],
                  t_record=tch[ti][2], q_record=qch[qi][2], q_record_len=qch[qi][3], fast=f)
             for tf, tt, qf, qt, f in blocks for qi in range(qf, qt + 1) for ti in range(tf, tt + 1)]
    if not pairs:
        return np.zeros(0, oracle.RESULT_DTYPE)
    recs = oracle.align_batch(pairs, oracle.make_params(target_total=target_total), threads=threads,
                              cap_per_pair=2048)
    assert max(len(r) for r in recs) < 2048
    return np.concatenate(recs)


def record_diff(got, ref):
    """A short description of how two record lists differ (for assertion messages)."""
    key = lambda r: (int(r["reverse"]), int(r["tstart"]), int(r["qstart"]), int(r["len"]))
    g, r = [key(x) for x in got], [key(x) for x in ref]
    cg, cr = collections.Counter(g), collections.Counter(r)  # overlapping chunks repeat records
    miss, extra = sorted((cr - cg).elements()), sorted((cg - cr).elements())
    first = next((i for i in range(min(len(g), len(r))) if g[i] != r[i]), None)
    return (f"{len(got)} records vs {len(ref)}; missing {miss[:5]} ({len(miss)}), extra {extra[:5]} "
            f"({len(extra)}); first difference at {first}: got {g[first] if first is not None else None}, "
            f"want {r[first] if first is not None else None}")


def assert_records(got, ref, what):
    same = len(got) == len(ref)
    if same:
        for f in ("query_id", "target_id", "query_size", "qstart", "tstart", "len", "reverse", "ident"):
            same &= bool(np.array_equal(got[f], ref[f]))
    assert same, f"{what}: {record_diff(got, ref)}"
    if len(ref):
        rel = np.abs(got["prob"] - ref["prob"]) / np.maximum(np.abs(ref["prob"]), 1e-300)
        assert rel.max() <= 1e-11, f"{what}: prob rel err {rel.max()}"


def test_server_alone_and_mix_equal_reference(genomes, served):
    T, Q, tch, qch = genomes
    ctx, params, srv = served
    reqs = _requests(tch, qch, np.random.default_rng(3))
    ref = [oracle_records(tch, qch, blocks, params.target_total) for blocks in reqs]
    assert sum(len(a) for a in ref) > 0
    assert len(ref[-2]) > 64  # the repeat pair overflows the first slot capacity
    for i, blocks in enumerate(reqs):
        assert_records(srv.results(srv.submit(blocks)), ref[i], f"request {i} alone")
    # the mix: every request at once, in shuffled order, from two submitting threads; twice
    for rnd in range(2):
        order = np.random.default_rng(5 + rnd).permutation(len(reqs))
        tickets = {}
        lock = threading.Lock()

        def submit(ids):
            for i in ids:
                t = srv.submit(reqs[i])
                with lock:
                    tickets[i] = t

        th = [threading.Thread(target=submit, args=(order[k::2],)) for k in range(2)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        bad = []
        for i in range(len(reqs)):
            got = srv.results(tickets[i])
            try:
                assert_records(got, ref[i], f"mix {rnd} request {i}")
            except AssertionError as e:
                bad.append(str(e))
        assert not bad, "\n".join(bad)
    st = srv.stats()
    assert st["slots"] == 3 and st["max_pairs_per_launch"] == 256
    # served pairs read both strands' query spectra from HBM instead of transforming queries
    assert st["query_spectra_bytes"] == len(qch) * 2 * 4 * 8192 * 4
    assert st["pairs_done"] >= 3 * sum((b[1] - b[0] + 1) * (b[3] - b[2] + 1) for r in reqs for b in r)


def test_server_without_query_spectra(genomes):
    """The same records when the pair kernel transforms each query itself."""
    from satsuma2_amd.api import Chunk, Context, Server, SlaveParams
    T, Q, tch, qch = genomes
    ctx = Context(0)
    params = SlaveParams(target_total=float(len(T) + 4096))
    srv = Server(ctx, params, [Chunk(s, st, r, rl) for s, st, r, rl in tch],
                 [Chunk(s, st, r, rl) for s, st, r, rl in qch], max_pairs_per_launch=256, max_slots=2,
                 query_spectra=False)
    assert srv.stats()["query_spectra_bytes"] == 0
    reqs = _requests(tch, qch, np.random.default_rng(11), n=6)
    for i, blocks in enumerate(reqs):
        assert_records(srv.results(srv.submit(blocks)), oracle_records(tch, qch, blocks, params.target_total),
                       f"request {i}")
    srv.close()
    ctx.close()


def test_server_latency_and_edge_requests(served, genomes):
    T, Q, tch, qch = genomes
    ctx, params, srv = served
    t = srv.submit([])
    n, lat = srv.wait(t)
    assert n == 0 and lat >= 0
    assert len(srv.results(t)) == 0
    with pytest.raises(SxcError):
        srv.submit([(0, len(tch), 0, 0, 0)])
    with pytest.raises(SxcError):
        srv.submit([(3, 2, 0, 0, 0)])
    t = srv.submit([(0, 23, 0, 23, 0)])
    n, lat = srv.wait(t)
    assert lat > 0
    assert len(srv.results(t)) == n


def test_server_budget_rejects_small_hbm(genomes):
    from satsuma2_amd.api import Chunk, Context, Server, SlaveParams
    T, Q, tch, qch = genomes
    ctx = Context(0)
    with pytest.raises(SxcError) as e:
        Server(ctx, SlaveParams(target_total=1e6), [Chunk(s, st, r, rl) for s, st, r, rl in tch],
               [Chunk(s, st, r, rl) for s, st, r, rl in qch], hbm_budget=64 << 20)
    assert e.value.code == -4
    ctx.close()
