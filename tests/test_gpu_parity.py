"""GPU parity: libsxc_hip.so (through its C ABI) against the reference's golden vectors and
the oracle, on the same inputs.

Bar: bit-exact for the FFT spectra, encoded signals, xc, accepted lags, diagonal runs,
coordinates, lengths and identities.  The match probability goes through the device's
erf/exp (not glibc's), so it is compared with a stated tolerance: relative error
<= PROB_RTOL everywhere (the error grows with -log(prob), so it is largest for
vanishingly small probabilities) and <= PROB_RTOL_KEEP where prob >= 0.5, the region
where keep decisions are made.  Keep/drop decisions at min_prob are compared exactly (a
flip would change the result count).
"""
import hashlib

import numpy as np
import pytest

import oracle
from conftest import chunk_records, dense_run_pair, lcg_noise, n_pairs, unpack2

pytestmark = pytest.mark.gpu

PROB_RTOL = 1e-11
PROB_RTOL_KEEP = 1e-14


def prob_rel_err(got, ref):
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
    return rel


FIELDS_EXACT = ["query_id", "target_id", "query_size", "qstart", "tstart", "len", "reverse", "ident"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


@pytest.fixture(scope="module")
def ctx():
    from satsuma2_amd.api import Context

    c = Context(0)
    yield c
    c.close()


def assert_results_equal(got, ref, what=""):
    assert len(got) == len(ref), f"{what}: {len(got)} results vs {len(ref)}"
    for f in FIELDS_EXACT:
        assert np.array_equal(got[f], ref[f]), f"{what}: field {f} differs"
    if len(ref):
        rel = prob_rel_err(got["prob"], ref["prob"])
        assert rel.max() <= PROB_RTOL, f"{what}: prob rel err {rel.max()}"
        hi = ref["prob"] >= 0.5
        assert rel[hi].max(initial=0.0) <= PROB_RTOL_KEEP, f"{what}: prob rel err near 1 {rel[hi].max()}"


@pytest.mark.parametrize("nb", range(9, 16))
def test_fft_forward_golden(ctx, golden, nb):
    g = golden("fft.npz")
    N = 1 << nb
    x = lcg_noise(N, N)
    f = ctx.fft_forward(x)[0]
    assert sha(f) == str(g[f"sha_{N}"][0])


def test_fft_forward_batch(ctx):
    N = 8192
    x = np.stack([lcg_noise(N, 100 + k) for k in range(5)])
    f = ctx.fft_forward(x)
    for k in range(5):
        assert np.array_equal(bits(f[k]), bits(oracle.fft_forward(x[k])))


def test_single_pair_plain_api(ctx, golden):
    g = golden("single_pair.npz")
    t, q = g["t"].tobytes(), g["q"].tobytes()
    qrc = oracle.reverse_complement(q)
    sigs = [ctx.encode(s, 8192) for s in (t, q, qrc)]
    for k in range(3):
        assert np.array_equal(bits(sigs[k]), bits(g["sig"][k])), f"signal {k}"
    for s, qs in ((0, q), (1, qrc)):
        xc = ctx.xcorr(sigs[0], sigs[1 + s])
        assert np.array_equal(bits(xc), bits(g[f"xc{s}"])), f"xc strand {s}"
        m = ctx.matchup(qs, t, xc, 1.8)
        gm = g[f"m{s}"]
        assert np.array_equal(m, np.stack([gm["st"], gm["sq"], gm["len"]], 1)), f"runs strand {s}"
        prob, ident = ctx.match_probability(t, qs, m, 4096.0)
        assert np.array_equal(bits(ident), bits(gm["ident"]))
        rel = prob_rel_err(prob, gm["prob"])
        print(f"strand {s}: prob max rel err {rel.max():.3e}, near 1 {rel[gm['prob'] >= 0.5].max(initial=0):.3e}")
        assert rel.max() <= PROB_RTOL
        assert rel[gm["prob"] >= 0.5].max(initial=0.0) <= PROB_RTOL_KEEP
        assert np.array_equal(prob >= 0.99, gm["prob"] >= 0.99)


def test_single_pair_align_forward_only(ctx, golden):
    """The single-pair workload: forward strand only, cutoff 1.8; min_prob 0.99999 is a
    no-op on the slave path, whose keep threshold is fixed at 0.99."""
    from satsuma2_amd.api import ChunkPair, HomologyByXCorr, SlaveParams

    g = golden("single_pair.npz")
    t, q = g["t"].tobytes(), g["q"].tobytes()
    h = HomologyByXCorr(ctx, SlaveParams(strands=1, target_total=4096.0))
    got = h.align_pair(ChunkPair(t, q))
    ref = oracle.align_batch([dict(tseq=t, qseq=q)], oracle.make_params(strands=1, target_total=4096.0))[0]
    assert_results_equal(got, ref, "single pair")
    assert len(ref) >= 1


def test_uniform_pairs_batch_golden(ctx, golden):
    from satsuma2_amd.api import ChunkPair, HomologyByXCorr, SlaveParams

    g = golden("uniform_pairs.npz")
    P = g["t"].shape[0]
    T = [unpack2(g["t"][r], 4096) for r in range(P)]
    Q = [unpack2(g["q"][r], 4096) for r in range(P)]
    h = HomologyByXCorr(ctx, SlaveParams(target_total=float(4096 * P)))
    pairs = [ChunkPair(T[r], Q[r], t_record=r, q_record=r, q_record_len=4096, fast=int(g["fast"][r]))
             for r in range(P)]
    got = h.align_batch(pairs)
    assert_results_equal(got, g["results"], "uniform batch")
    # the plain one-pair call gives the same records
    per = np.concatenate([h.align_pair(p) for p in pairs[:8]])
    ref8 = g["results"][np.isin(g["results"]["query_id"], np.arange(8))]
    assert_results_equal(per, ref8, "plain calls")


def test_resident_batch_repeatable(ctx, golden):
    from satsuma2_amd.api import ChunkPair, ResidentBatch, SlaveParams

    g = golden("uniform_pairs.npz")
    P = 16
    T = [unpack2(g["t"][r], 4096) for r in range(P)]
    Q = [unpack2(g["q"][r], 4096) for r in range(P)]
    pairs = [ChunkPair(T[r], Q[r], t_record=r, q_record=r, q_record_len=4096) for r in range(P)]
    ref = np.concatenate(oracle.align_batch(
        [dict(tseq=T[r], qseq=Q[r], t_record=r, q_record=r, q_record_len=4096) for r in range(P)],
        oracle.make_params(target_total=float(4096 * 64))))
    rb = ResidentBatch(ctx, pairs, SlaveParams(target_total=float(4096 * 64)))
    for _ in range(3):
        rb.run()
        assert_results_equal(rb.results(), ref, "resident")
    cands, runs = rb.stats()
    assert cands > 0 and runs > 0
    rb.close()


def test_dense_runs(ctx):
    """A pair whose scan holds more runs in one epoch than the run queue (1,024 entries):
    the runs are extracted in rounds.  Records and the total run count equal the oracle's."""
    from satsuma2_amd.api import ChunkPair, ResidentBatch, SlaveParams

    t = dense_run_pair()
    ref = oracle.align_batch([dict(tseq=t, qseq=t, q_record_len=4096)], oracle.make_params(target_total=4096.0 * 64))[0]
    nruns = 0
    for strand in range(2):
        qs = t if strand == 0 else oracle.reverse_complement(t)
        oxc = oracle.xcorr(oracle.encode(t, 8192), oracle.encode(qs, 8192))
        nruns += len(oracle.matchup(qs, t, oxc, 1.8))
    rb = ResidentBatch(ctx, [ChunkPair(t, t, q_record_len=4096)], SlaveParams(target_total=4096.0 * 64))
    rb.run()
    assert_results_equal(rb.results(), ref, "dense runs")
    assert rb.stats()[1] == nruns > 1024
    rb.close()


def test_n_pairs(ctx, golden):
    """Pairs of A/C/G/T/N chunks (the bitmap scan with N masks) and a pure pair in one batch:
    records and the total run count equal the oracle's."""
    from satsuma2_amd.api import ChunkPair, ResidentBatch, SlaveParams

    P = n_pairs(golden("uniform_pairs.npz"))
    ref = np.concatenate(oracle.align_batch(
        [dict(tseq=t, qseq=q, t_record=r, q_record=r, q_record_len=4096) for r, (t, q) in enumerate(P)],
        oracle.make_params(target_total=4096.0 * 64)))
    nruns = 0
    for t, q in P:
        for strand in range(2):
            qs = q if strand == 0 else oracle.reverse_complement(q)
            oxc = oracle.xcorr(oracle.encode(t, 8192), oracle.encode(qs, 8192))
            nruns += len(oracle.matchup(qs, t, oxc, 1.8))
    rb = ResidentBatch(ctx, [ChunkPair(t, q, t_record=r, q_record=r, q_record_len=4096) for r, (t, q) in enumerate(P)],
                       SlaveParams(target_total=4096.0 * 64))
    rb.run()
    assert_results_equal(rb.results(), ref, "N pairs")
    assert rb.stats()[1] == nruns
    rb.close()


def test_resident_batch_graph_replays(ctx, golden):
    """Runs after the first replay a captured HIP graph.  Every replay, with other eager work
    and synchronisations between them, equals the oracle; so does the first (eager) run, in
    which tandem-repeat pairs overflow both the per-(pair, strand) record slots (> 64) and
    the first output buffer (> 1024 records), and an eager run after graphs are switched off."""
    from satsuma2_amd.api import ChunkPair, HomologyByXCorr, ResidentBatch, SlaveParams

    rng = np.random.default_rng(5)
    T, Q = [], []
    for period in (40, 60):
        unit = rng.choice(list(b"ACGT"), period).astype(np.uint8)
        t = np.resize(unit, 4096)
        q = t.copy()
        m = rng.random(4096) < 0.12
        q[m] = rng.choice(list(b"ACGT"), int(m.sum()))
        T.append(t.tobytes())
        Q.append(q.tobytes())
    g = golden("uniform_pairs.npz")
    for r in range(6):
        T.append(unpack2(g["t"][r], 4096))
        Q.append(unpack2(g["q"][r], 4096))
    P = len(T)
    prm = SlaveParams(target_total=float(4096 * 64))
    pairs = [ChunkPair(T[r], Q[r], t_record=r, q_record=r, q_record_len=4096) for r in range(P)]
    ref = np.concatenate(oracle.align_batch(
        [dict(tseq=T[r], qseq=Q[r], t_record=r, q_record=r, q_record_len=4096) for r in range(P)],
        oracle.make_params(target_total=float(4096 * 64))))
    assert len(ref) > 1024 and np.bincount(ref["query_id"] * 2 + ref["reverse"]).max() > 64
    other = HomologyByXCorr(ctx, prm)
    rb = ResidentBatch(ctx, pairs, prm)
    for k in range(5):
        rb.run()
        tm = ctx.timing()
        assert tm.graph == (1 if k > 0 else 0), k
        assert tm.device_ms > 0, k
        assert_results_equal(rb.results(), ref, f"run {k}")
        # eager work of the program between replays
        assert_results_equal(other.align_pair(pairs[2]), ref[ref["query_id"] == 2], f"eager after run {k}")
    rb.set_graph(False)
    rb.run()
    tm = ctx.timing()
    assert tm.graph == 0 and tm.correlate_ms > 0 and tm.scan_ms > 0
    assert_results_equal(rb.results(), ref, "eager again")
    rb.close()


# Alternative FFT geometries compiled for the autotuner: (log2 N, stage-A width, levels per
# LDS round).  Each must reproduce the oracle bit for bit.
# (log2 N, stage-A width, levels per LDS round, three workgroups per CU, lean LDS)
GEOMETRIES = ([(nb, 16, 2, 0, 0) for nb in range(9, 15)] + [(15, 32, 2, 0, 0)]
              + [(nb, 32, 3, 0, 0) for nb in range(11, 15)] + [(13, 16, 3, 1, 0), (13, 16, 2, 1, 0)]
              + [(12, 32, 3, 0, 1), (12, 16, 2, 0, 1), (13, 32, 3, 0, 1), (14, 32, 3, 0, 1), (11, 32, 3, 0, 1),
                 (11, 16, 2, 0, 1), (13, 32, 2, 0, 1), (14, 32, 3, 1, 1)])


def _homology_pairs(N, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for r in range(n):
        t = rng.choice(np.frombuffer(b"ACGT", np.uint8), N // 2)
        q = rng.choice(np.frombuffer(b"ACGT", np.uint8), N // 2)
        ln = N // 8
        q[N // 5: N // 5 + ln] = t[N // 7: N // 7 + ln]
        out.append((t.tobytes(), q.tobytes()))
    return out


@pytest.mark.parametrize("nb,sa,km,occ,lean", GEOMETRIES)
def test_fft_geometry_parity(ctx, nb, sa, km, occ, lean):
    from satsuma2_amd.api import ChunkPair, Geometry, HomologyByXCorr, SlaveParams

    N = 1 << nb
    tq = _homology_pairs(N, 8, nb * 7 + sa + km)
    pairs = [ChunkPair(t, q, t_record=r, q_record=r, q_record_len=N // 2) for r, (t, q) in enumerate(tq)]
    ref = np.concatenate(oracle.align_batch(
        [dict(tseq=t, qseq=q, t_record=r, q_record=r, q_record_len=N // 2) for r, (t, q) in enumerate(tq)],
        oracle.make_params(N=N, target_total=1e6)))
    assert len(ref) > 0
    dflt = ctx.geometry(N, "pairs"), ctx.geometry(N, "spectra")
    try:
        ctx.set_geometry(N, "pairs", Geometry(sa, km, occ, lean))
        ctx.set_geometry(N, "spectra", Geometry(sa, km, occ, lean))
        assert ctx.geometry(N, "pairs") == (sa, km, occ, lean)
        got = HomologyByXCorr(ctx, SlaveParams(t_chunk=N // 2, target_total=1e6)).align_batch(pairs)
        assert_results_equal(got, ref, f"geometry {nb},{sa},{km},{occ},{lean}")
    finally:
        ctx.set_geometry(N, "pairs", dflt[0])
        ctx.set_geometry(N, "spectra", dflt[1])


def test_autotune(ctx, golden):
    """The autotuner times every geometry of every length, keeps one per kernel and length,
    and batches run in the chosen geometries give the golden records."""
    from satsuma2_amd.api import ChunkPair, Geometry, HomologyByXCorr, SlaveParams

    entries = ctx.autotune(512, 32768)
    by_n = {}
    for e in entries:
        by_n.setdefault(e["N"], []).append(e)
        assert e["correlate_ms"] > 0 and e["spectra_ms"] > 0 and e["pairs"] >= 256
    assert sorted(by_n) == [1 << nb for nb in range(9, 16)]
    for N, es in by_n.items():
        assert len(es) >= 2
        assert sum(e["chosen_correlate"] for e in es) == 1 and sum(e["chosen_spectra"] for e in es) == 1
        pick = [e for e in es if e["chosen_correlate"]][0]
        assert ctx.geometry(N, "pairs") == pick["geo"]
        assert pick["correlate_ms"] == min(e["correlate_ms"] for e in es)
    g = golden("uniform_pairs.npz")
    P = 64
    T = [unpack2(g["t"][r], 4096) for r in range(P)]
    Q = [unpack2(g["q"][r], 4096) for r in range(P)]
    pairs = [ChunkPair(T[r], Q[r], t_record=r, q_record=r, q_record_len=4096, fast=int(g["fast"][r]))
             for r in range(P)]
    got = HomologyByXCorr(ctx, SlaveParams(target_total=float(4096 * g["t"].shape[0]))).align_batch(pairs)
    assert_results_equal(got, g["results"][g["results"]["query_id"] < P], "after autotune")
    for N in by_n:  # back to the defaults for the tests that follow
        ctx.set_geometry(N, "pairs", Geometry(32 if N == 32768 else 16, 3, 0, 0))
        ctx.set_geometry(N, "spectra", Geometry(32 if N == 32768 else 16, 3, 0, 0))


def test_edge_pairs(ctx, golden):
    g = golden("edge_pairs.npz")
    for name in g["names"]:
        t, q, N = g[f"{name}_t"].tobytes(), g[f"{name}_q"].tobytes(), int(g[f"{name}_N"][0])
        qrc = oracle.reverse_complement(q)
        sigs = [ctx.encode(s, N) for s in (t, q, qrc)]
        assert [sha(s) for s in sigs] == list(g[f"{name}_sigsha"]), name
        for cut in (1.8, 2.9):
            for s, qs in ((0, q), (1, qrc)):
                pre = f"{name}_{cut}_s{s}_"
                xc = ctx.xcorr(sigs[0], sigs[1 + s])
                assert sha(xc) == str(g[pre + "xc_sha"]), pre
                m = ctx.matchup(qs, t, xc, cut)
                assert m.shape[0] == int(g[pre + "m_n"][0]), pre
                assert sha(m.astype(np.int32)) == str(g[pre + "m_sha"]), pre
                if m.shape[0]:
                    prob, ident = ctx.match_probability(t, qs, m, 4096.0 * 64)
                    assert sha(ident) == str(g[pre + "ident_sha"]), pre
        # and the fused path on the same pair, both strands, against the oracle
        from satsuma2_amd.api import ChunkPair, HomologyByXCorr, SlaveParams

        h = HomologyByXCorr(ctx, SlaveParams(target_total=4096.0 * 64, t_chunk=N // 2))
        got = h.align_pair(ChunkPair(t, q))
        ref = oracle.align_batch([dict(tseq=t, qseq=q)],
                                 oracle.make_params(N=N, target_total=4096.0 * 64))[0]
        assert_results_equal(got, ref, name)


def test_uneven_pairs(ctx, golden):
    g = golden("uneven_pairs.npz")
    for k in range(int(g["n"][0])):
        N = int(g[f"N{k}"][0])
        lt, lq = g[f"len{k}"]
        t, q = unpack2(g[f"t{k}"], lt), unpack2(g[f"q{k}"], lq)
        qrc = oracle.reverse_complement(q)
        sigs = [ctx.encode(s, N) for s in (t, q, qrc)]
        assert [sha(s) for s in sigs] == list(g[f"sigsha{k}"]), k
        for s, qs in ((0, q), (1, qrc)):
            xc = ctx.xcorr(sigs[0], sigs[1 + s])
            assert sha(xc) == str(g[f"p{k}_s{s}_xc_sha"]), (k, N)
            m = ctx.matchup(qs, t, xc, 1.8)
            assert sha(m.astype(np.int32)) == str(g[f"p{k}_s{s}_m_sha"]), (k, N)


def test_samples_slice_slave_path(ctx, golden):
    from satsuma2_amd.api import ChunkPair, HomologyByXCorr, SlaveParams

    g = golden("samples_slice.npz")
    t = g["t"].tobytes().upper()
    q = g["q"].tobytes().upper()
    tch = chunk_records([t], 4096, 1024)
    qch = chunk_records([q], 4096, 0)
    res = []
    for tf, tt, qf, qt, fast in g["pairs"]:
        pairs = []
        for qi in range(qf, qt + 1):
            for ti in range(tf, tt + 1):
                trec, tstart, tseq = tch[ti]
                qrec, qstart, qseq = qch[qi]
                pairs.append(ChunkPair(tseq, qseq, tstart, qstart, trec, qrec, len(q), int(fast)))
        h = HomologyByXCorr(ctx, SlaveParams(target_total=float(len(t))))
        res.append(h.align_batch(pairs))
    assert_results_equal(np.concatenate(res), g["results"], "samples slice")


def test_errors_are_reported(ctx):
    from satsuma2_amd._lib import SxcError

    with pytest.raises(SxcError):
        ctx.encode(b"ACGT", 1000)  # N not a power of two
    with pytest.raises(SxcError):
        ctx.encode(b"AC\xffT", 1024)  # byte >= 128
    with pytest.raises(SxcError):
        ctx.encode(b"A" * 2048, 1024)  # longer than N


def test_uneven_batch_per_pair_fft_length(ctx, golden):
    """N = 0: every pair at its own FFT length (next power of two >= q + t), all lengths in
    one call.  Records equal the oracle's at the same per-pair N, in input order, and the
    same pairs through fixed-N calls."""
    from satsuma2_amd.api import ChunkPair, HomologyByXCorr, SlaveParams

    g = golden("uneven_pairs.npz")
    T, Q, Ns = [], [], []
    for k in range(int(g["n"][0])):
        N = int(g[f"N{k}"][0])
        lt, lq = g[f"len{k}"]
        T.append(unpack2(g[f"t{k}"], lt))
        Q.append(unpack2(g[f"q{k}"], lq))
        Ns.append(N)
    assert len(set(Ns)) >= 4
    pairs = [ChunkPair(t, q, t_record=k, q_record=k, q_record_len=len(q)) for k, (t, q) in enumerate(zip(T, Q))]
    got = HomologyByXCorr(ctx, SlaveParams(t_chunk=0, target_total=1e6)).align_batch(pairs)
    exp = []
    for k, (t, q, N) in enumerate(zip(T, Q, Ns)):
        exp.append(oracle.align_batch([dict(tseq=t, qseq=q, t_record=k, q_record=k, q_record_len=len(q))],
                                      oracle.make_params(N=N, target_total=1e6))[0])
        fixed = HomologyByXCorr(ctx, SlaveParams(t_chunk=N // 2, target_total=1e6)).align_pair(pairs[k])
        assert fixed.tobytes() == got[got["query_id"] == k].tobytes(), k
    assert_results_equal(got, np.concatenate(exp), "uneven batch")
