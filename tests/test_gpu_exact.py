"""Exactness by construction on the GPU, through the C ABI.

Two results depend on a libm call the device cannot reproduce bit for bit, and both are
decided on the host with the host's libm, as the reference computes them:
  * the keep decision prob >= min_prob (glibc erf/exp, analysis/AlignProbability.cc:122-124):
    runs whose device probability lies within a few device-libm ulps of min_prob are flagged
    by the scan and decided on the host;
  * window entropies of chunks with B, D, H or V (glibc log of sums of thirds,
    analysis/CrossCorr.cc:28-93): computed on the host and read by the kernels.
FindTop's accepted-lag lists (analysis/CrossCorr.cc:878-944) of the fused pair kernel and
of the plain FindTop entry point are compared with the reference's lists directly."""
import hashlib
import math

import numpy as np
import pytest

import oracle
from conftest import threshold_size, unpack2

pytestmark = pytest.mark.gpu

FIELDS_EXACT = ["query_id", "target_id", "query_size", "qstart", "tstart", "len", "reverse", "ident"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def ctx():
    from satsuma2_amd.api import Context

    c = Context(0)
    yield c
    c.close()


def assert_same(got, ref, what, prob_exact=False):
    assert len(got) == len(ref), f"{what}: {len(got)} results vs {len(ref)}"
    for f in FIELDS_EXACT:
        assert np.array_equal(got[f], ref[f]), f"{what}: field {f} differs"
    if len(ref):
        rel = np.abs(got["prob"] - ref["prob"]) / np.maximum(np.abs(ref["prob"]), 1e-300)
        assert rel.max() <= (0.0 if prob_exact else 1e-11), f"{what}: prob rel err {rel.max()}"


def test_keep_band_near_threshold(ctx, golden):
    """Target sizes chosen so that one run's reference probability is the last double at or
    above 0.99 (kept) and the first below it (dropped): the records equal the oracle's, the
    run is decided on the host (near_threshold), and its probability is the host's, bit for bit."""
    from satsuma2_amd.api import ChunkPair, HomologyByXCorr, ResidentBatch, SlaveParams

    g = golden("single_pair.npz")
    t, q = g["t"].tobytes(), g["q"].tobytes()
    m = g["m0"]  # the reference's forward-strand runs (MatchUp)
    # runs whose threshold size is moderate (the band is relative to the size)
    picked = []
    for r in m[m["len"] >= 40]:
        T = threshold_size(t, q, int(r["st"]), int(r["sq"]), int(r["len"]))
        if T is not None and 1e3 < T < 1e8:
            picked.append((r, T))
        if len(picked) >= 3:
            break
    assert picked
    for r, T in picked:
        for size in (T, np.nextafter(T, math.inf)):
            ref = oracle.align_batch([dict(tseq=t, qseq=q)], oracle.make_params(strands=1, target_total=size))[0]
            sel = (ref["tstart"] == r["st"]) & (ref["qstart"] == r["sq"]) & (ref["len"] == r["len"])
            assert sel.sum() == (1 if size == T else 0)
            prm = SlaveParams(strands=1, target_total=float(size))
            got = HomologyByXCorr(ctx, prm).align_pair(ChunkPair(t, q))
            assert ctx.timing().near_threshold >= 1
            assert_same(got, ref, f"T={size!r}")
            # the host-decided record carries the host's probability exactly
            gsel = (got["tstart"] == r["st"]) & (got["qstart"] == r["sq"]) & (got["len"] == r["len"])
            assert np.array_equal(got["prob"][gsel], ref["prob"][sel])
            assert not got["pad_"].any()
            # the resident batch (graph replays) decides the same way
            rb = ResidentBatch(ctx, [ChunkPair(t, q)], prm)
            for _ in range(3):
                rb.run()
                assert ctx.timing().near_threshold >= 1
                assert_same(rb.results(), ref, f"resident T={size!r}")
            rb.close()


def test_bdhv_signals_lags_and_runs(ctx, golden):
    """Chunks dense in B/D/H/V at N = 2048 .. 32768: signal bits, xc, FindTop lags and MatchUp
    runs equal the reference's; the fused path (both strands) equals the oracle."""
    from satsuma2_amd.api import ChunkPair, HomologyByXCorr, ResidentBatch, SlaveParams

    g = golden("bdhv_pairs.npz")
    for name in g["names"]:
        t, q, N = g[f"{name}_t"].tobytes(), g[f"{name}_q"].tobytes(), int(g[f"{name}_N"][0])
        qrc = oracle.reverse_complement(q)
        sigs = [ctx.encode(s, N) for s in (t, q, qrc)]
        assert [sha(s) for s in sigs] == list(g[f"{name}_sigsha"]), name
        for s, qs in ((0, q), (1, qrc)):
            pre = f"{name}_s{s}_"
            xc = ctx.xcorr(sigs[0], sigs[1 + s])
            assert sha(xc) == str(g[pre + "xc_sha"]), pre
            assert np.array_equal(ctx.findtop(xc, 1.8), g[pre + "top"]), pre
            mm = ctx.matchup(qs, t, xc, 1.8)
            assert mm.shape[0] == int(g[pre + "m_n"][0]), pre
            assert sha(mm.astype(np.int32)) == str(g[pre + "m_sha"]), pre
        prm = SlaveParams(target_total=4096.0 * 64, t_chunk=N // 2)
        ref = oracle.align_batch([dict(tseq=t, qseq=q)], oracle.make_params(N=N, target_total=4096.0 * 64))[0]
        assert_same(HomologyByXCorr(ctx, prm).align_pair(ChunkPair(t, q)), ref, name)
        rb = ResidentBatch(ctx, [ChunkPair(t, q)], prm)
        rb.run()
        for s in (0, 1):
            assert np.array_equal(rb.lags(0, s), g[f"{name}_s{s}_top"]), (name, s)
        rb.close()


def test_findtop_lags_golden(ctx, golden):
    """The fused pair kernel's accepted lags (sxc_batch_lags) and the plain FindTop
    (sxc_findtop) against the reference's lists: the single pair, the edge cases at both
    cutoffs, the uneven pairs at their own N."""
    from satsuma2_amd.api import ChunkPair, ResidentBatch, SlaveParams

    g = golden("single_pair.npz")
    t, q = g["t"].tobytes(), g["q"].tobytes()
    rb = ResidentBatch(ctx, [ChunkPair(t, q)], SlaveParams(target_total=4096.0))
    rb.run()
    assert np.array_equal(rb.lags(0, 0), g["top0"])
    assert np.array_equal(rb.lags(0, 1), g["top1"])
    assert np.array_equal(ctx.findtop(g["xc0"], 1.8), g["top0"])
    assert np.array_equal(ctx.findtop(g["xc1"], 1.8), g["top1"])
    c, _ = rb.stats()
    assert c == len(g["top0"]) + len(g["top1"])
    rb.close()

    e = golden("edge_pairs.npz")
    for name in e["names"]:
        t, q, N = e[f"{name}_t"].tobytes(), e[f"{name}_q"].tobytes(), int(e[f"{name}_N"][0])
        # pair 0 at the slow cutoff (1.8), pair 1 flagged fast (cutoff_fast 2.9)
        rb = ResidentBatch(ctx, [ChunkPair(t, q), ChunkPair(t, q, fast=1)],
                           SlaveParams(target_total=4096.0 * 64, t_chunk=N // 2))
        rb.run()
        for k, cut in enumerate((1.8, 2.9)):
            for s in (0, 1):
                assert np.array_equal(rb.lags(k, s), e[f"{name}_{cut}_s{s}_top"]), (name, cut, s)
        rb.close()

    u = golden("uneven_pairs.npz")
    n = int(u["n"][0])
    pairs, want = [], []
    for k in range(n):
        lt, lq = u[f"len{k}"]
        pairs.append(ChunkPair(unpack2(u[f"t{k}"], lt), unpack2(u[f"q{k}"], lq)))
        want.append((u[f"p{k}_s0_top"], u[f"p{k}_s1_top"]))
    # per-pair FFT length (N = 0): the pairs run in length classes
    rb = ResidentBatch(ctx, pairs, SlaveParams(target_total=1e6, t_chunk=0))
    rb.run()
    for k in range(n):
        for s in (0, 1):
            assert np.array_equal(rb.lags(k, s), want[k][s]), (k, s)
    rb.close()
