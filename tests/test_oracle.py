"""Pins the CPU oracle (oracle/sxc_oracle.c) before anything is compared against it.

Two kinds of checks:
  * golden vectors written by the reference compiled from source (tests/golden/*.npz),
    compared bit for bit;
  * closed forms that the mathematics fixes independently of the reference: a naive
    FP64 DFT for FFTReal's transform and layout, a direct FP64 correlation for the
    cross-correlation, brute-force window sums for the diagonal scan, the normal
    tail for the match probability, and the entropy formula.
"""
import hashlib
import math

import numpy as np
import pytest

import oracle
from conftest import chunk_records, lcg_noise, unpack2


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


# ---------------------------------------------------------------- FFT

@pytest.mark.parametrize("nb", range(3, 16))
def test_fft_golden(golden, nb):
    g = golden("fft.npz")
    N = 1 << nb
    x = lcg_noise(N, N)
    f = oracle.fft_forward(x)
    inv = oracle.fft_inverse(f)
    res = oracle.fft_rescale(inv)
    assert [sha(f), sha(inv), sha(res)] == list(g[f"sha_{N}"])
    if N <= 2048:
        assert np.array_equal(bits(f), bits(g[f"fwd_{N}"]))


def naive_packed_dft(x):
    """FFTReal's convention: f(k) = sum_p x(p) exp(+2 pi j k p / N); packed layout
    f[0..N/2] = Re(bin 0..N/2), f[N/2+k] = Im(bin k) (extern/RealFFT/readme.txt:127-165)."""
    N = x.size
    p = np.arange(N)
    k = np.arange(N // 2 + 1)
    X = (x.astype(np.float64)[None, :] * np.exp(2j * np.pi * np.outer(k, p) / N)).sum(1)
    out = np.empty(N)
    out[: N // 2 + 1] = X.real
    out[N // 2 + 1:] = X.imag[1: N // 2]
    return out


@pytest.mark.parametrize("N", [8, 16, 64, 512, 2048])
def test_fft_matches_naive_dft(N):
    x = lcg_noise(N, 7 * N + 1)
    f = oracle.fft_forward(x).astype(np.float64)
    ref = naive_packed_dft(x)
    assert np.max(np.abs(f - ref)) < 1e-5 * math.sqrt(N) * np.max(np.abs(ref))
    # inverse is unscaled: ifft(fft(x)) = N x, rescale divides by N
    back = oracle.fft_rescale(oracle.fft_inverse(oracle.fft_forward(x)))
    assert np.max(np.abs(back - x)) < 1e-5


def test_fft_linearity_and_parseval():
    N = 4096
    a = lcg_noise(N, 3)
    b = lcg_noise(N, 4)
    fa, fb, fab = (oracle.fft_forward(v).astype(np.float64) for v in (a, b, (a + 2 * b).astype(np.float32)))
    assert np.max(np.abs(fab - (fa + 2 * fb))) < 1e-3
    re = fa[: N // 2 + 1]
    im = fa[N // 2 + 1:]
    energy = re[0] ** 2 + re[N // 2] ** 2 + 2 * (np.sum(re[1: N // 2] ** 2) + np.sum(im ** 2))
    assert abs(energy / N - np.sum(a.astype(np.float64) ** 2)) < 1e-4 * np.sum(a.astype(np.float64) ** 2)


# ---------------------------------------------------------------- single pair, bit for bit

def test_single_pair_golden(golden):
    g = golden("single_pair.npz")
    t, q = g["t"].tobytes(), g["q"].tobytes()
    qrc = oracle.reverse_complement(q)
    sigs = [oracle.encode(s, 8192) for s in (t, q, qrc)]
    for k in range(3):
        assert np.array_equal(bits(sigs[k]), bits(g["sig"][k]))
    for s, qs in ((0, q), (1, qrc)):
        xc = oracle.xcorr(sigs[0], sigs[1 + s])
        assert np.array_equal(bits(xc), bits(g[f"xc{s}"]))
        assert np.array_equal(oracle.findtop(xc, 1.8), g[f"top{s}"])
        m = oracle.matchup(qs, t, xc, 1.8)
        gm = g[f"m{s}"]
        assert np.array_equal(m, np.stack([gm["st"], gm["sq"], gm["len"]], 1))
        prob = np.array([oracle.match_probability(t, qs, a, b, c, 4096.0) for a, b, c in m])
        ident = np.array([oracle.print_match_ident(qs, t, a, b, c) for a, b, c in m])
        assert np.array_equal(bits(prob), bits(gm["prob"]))
        assert np.array_equal(bits(ident), bits(gm["ident"]))
    # the planted homology is found on the forward strand: peak lag = query pos - target pos
    qo, to = g["plant"]
    xc0 = g["xc0"]
    assert int(np.argmax(xc0)) - 4096 == qo - to


def direct_xcorr64(tsig, qsig):
    """FP64 sum_ch sum_i t[i] q[i + lag] on the same float signals, with the reference's
    product quirk (bins N/2-1 and N/2 keep the target spectrum unmultiplied,
    analysis/CrossCorr.cc:480-492) added as its exact iDFT term; lag of out[k] = k - N/2."""
    N = tsig.shape[1]
    H = N // 2
    out = np.zeros(N)
    for ch in range(4):
        T = np.fft.fft(tsig[ch].astype(np.float64))
        Q = np.fft.fft(qsig[ch].astype(np.float64))
        # in numpy's sign convention the stored product is T * conj(Q); the two quirk
        # bins keep T (and their Hermitian mirrors conj(T))
        P = T * np.conj(Q)
        for b in (H - 1, H):
            P[b] = T[b]
            P[N - b] = np.conj(T[b])
        # FFTReal's inverse sums with exp(-j): numpy's forward transform, then 1/N
        x = np.fft.fft(P).real / N
        out += np.roll(x, H)
    return out


def test_xcorr_matches_direct_fp64(golden):
    g = golden("single_pair.npz")
    tsig, qsig = g["sig"][0], g["sig"][1]
    xc64 = direct_xcorr64(tsig, qsig)
    xc32 = oracle.xcorr(tsig, qsig).astype(np.float64)
    assert np.max(np.abs(xc32 - xc64)) < 2e-3
    # without the quirk term the correlation at the planted lag is the plain dot product
    N = 8192
    k = int(np.argmax(xc32))
    lag = k - N // 2
    dot = sum(np.dot(tsig[c].astype(np.float64)[max(0, -lag):N - max(0, lag)],
                     qsig[c].astype(np.float64)[max(0, lag):N - max(0, -lag)]) for c in range(4))
    assert abs(dot - xc64[k]) / abs(dot) < 0.05


# ---------------------------------------------------------------- encode closed form

def test_encode_closed_form():
    rng = np.random.default_rng(3)
    s = bytes(rng.choice(np.frombuffer(b"ACGTN", np.uint8), 3000))
    N = 8192
    sig = oracle.encode(s, N).astype(np.float64)
    arr = np.frombuffer(s, np.uint8)
    ind = {c: (arr == ord(c)).astype(np.float64) + 0.25 * (arr == ord("N")) for c in "ACGT"}
    win = N // 512
    ent = np.zeros(len(s))
    for i in range(0, len(s), win):
        seg = slice(i, min(i + win, len(s)))
        h = 0.0
        for c in "ACGT":
            p = ind[c][seg].mean()
            if p >= 0.001:
                h += p * math.log(p) / 0.69314718056
        ent[seg] = max(-h, 0.0)
    for ch, c in enumerate("ACGT"):
        exp = ent * (ind[c] - ind[c].mean())
        assert np.max(np.abs(sig[ch, : len(s)] - exp)) < 1e-6
        assert not sig[ch, len(s):].any()
    # below 1024 bp the entropy weight is 1
    short = oracle.encode(s[:500], N).astype(np.float64)
    a = ind["A"][:500]
    assert np.max(np.abs(short[0, :500] - (a - a.mean()))) < 1e-6


# ---------------------------------------------------------------- diagonal scan brute force

def brute_runs(q, t, shift):
    """Brute-force restatement of the diagonal scan's contract: windows of 46 scores
    (round(100*EqualAmb)) ending at i, checked from the 47th overlapping position on,
    sum > 1889 opens/continues a run starting 45 before; the scan stops before the
    last target base or the end of the query."""
    score = np.array([[int(oracle.lib().sxo_dna_equal_amb(a, b) * 100 + 0.5) for b in range(128)]
                      for a in range(128)])
    lt, lq = len(t), len(q)
    i0 = max(0, -shift)
    iend = min(lq - shift, lt - 1)
    runs = []
    if i0 >= lt or iend < i0:
        return runs
    s = np.array([score[t[i]][q[i + shift]] for i in range(i0, iend)], np.int64)
    start = None
    for n in range(46, len(s)):
        w = s[n - 45: n + 1].sum()
        i = i0 + n
        if w > 1889:
            if start is None:
                start = i - 45
        elif start is not None:
            runs.append((start, start + shift, i - start))
            start = None
    if start is not None:
        runs.append((start, start + shift, iend - start))
    return runs


def test_matchup_brute_force():
    rng = np.random.default_rng(9)
    base = rng.choice(np.frombuffer(b"ACGT", np.uint8), 2000)
    t = bytearray(base.tobytes())
    q = bytearray(rng.choice(np.frombuffer(b"ACGT", np.uint8), 2400).tobytes())
    q[300:1300] = bytes(t[700:1700])  # exact copy: lag 300 - 700 = -400
    for p in rng.integers(300, 1300, 120):
        q[p] = ord("ACGT"[rng.integers(4)])
    q[900:950] = b"N" * 50
    t, q = bytes(t), bytes(q)
    N = 8192
    lags = [-400, -399, -1500, 0, 350, 2399, -1999]
    xc = np.zeros(N, np.float32)
    for lag in lags:
        xc[lag + N // 2] = 1000.0
    got = oracle.matchup(q, t, xc, 1.8)
    exp = []
    for lag in sorted(lags):
        exp += brute_runs(q, t, lag)
    assert [tuple(r) for r in got] == exp
    assert len(exp) > 3


def test_window_threshold_is_1889():
    # (int)(45 * 0.42 * 100) == 1889 because 45*0.42*100 = 1889.9999999999998
    assert int(45 * 0.42 * 100) == 1889
    # 19 matches of 46 (1900) pass, 18 (1800) do not; N:N scores 25
    assert int(oracle.lib().sxo_dna_equal_amb(ord("N"), ord("N")) * 100 + 0.5) == 25
    assert int(oracle.lib().sxo_dna_equal_amb(ord("X"), ord("X")) * 100 + 0.5) == 100
    assert int(oracle.lib().sxo_dna_equal_amb(ord("."), ord(".")) * 100 + 0.5) == 100


# ---------------------------------------------------------------- probability closed form

def test_probability_closed_form():
    from scipy.special import ndtr

    rng = np.random.default_rng(4)
    t = bytes(rng.choice(np.frombuffer(b"ACGT", np.uint8), 500))
    q = bytearray(t)
    for p in rng.integers(0, 500, 150):
        q[p] = ord("ACGT"[rng.integers(4)])
    q = bytes(q)
    for st, length in ((0, 46), (10, 100), (37, 300), (0, 500)):
        prob = oracle.match_probability(t, q, st, st, length, 1e6)
        ta = np.frombuffer(t[st:st + length], np.uint8)
        qa = np.frombuffer(q[st:st + length], np.uint8)
        ident = float((ta == qa).mean())
        gct = float(np.isin(ta, [ord("C"), ord("G")]).mean())
        gcq = float(np.isin(qa, [ord("C"), ord("G")]).sum())
        pm = (gcq * gct + (length - gcq) * (1 - gct)) / length / 2
        s = math.sqrt(pm * (1 - pm) * length)
        z = (pm * length - length * ident) / s
        exp = math.exp(-1e6 * float(ndtr(z * math.sqrt(2) / 1.414213562)))
        assert prob == pytest.approx(exp, rel=1e-9, abs=1e-300)
        assert oracle.print_match_ident(q, t, st, st, length) == pytest.approx(ident, rel=1e-15)


# ---------------------------------------------------------------- edge + uneven digests

def _check_digest(g, prefix, t, qs, xc, cutoff, tsize):
    assert sha(xc) == str(g[prefix + "xc_sha"])
    assert np.array_equal(oracle.findtop(xc, cutoff), g[prefix + "top"])
    m = oracle.matchup(qs, t, xc, cutoff)
    assert m.shape[0] == int(g[prefix + "m_n"][0])
    assert sha(m.astype(np.int32)) == str(g[prefix + "m_sha"])
    prob = np.array([oracle.match_probability(t, qs, a, b, c, tsize) for a, b, c in m], np.float64)
    ident = np.array([oracle.print_match_ident(qs, t, a, b, c) for a, b, c in m], np.float64)
    assert sha(prob) == str(g[prefix + "prob_sha"])
    assert sha(ident) == str(g[prefix + "ident_sha"])


def test_edge_pairs_golden(golden):
    g = golden("edge_pairs.npz")
    for name in g["names"]:
        t, q, N = g[f"{name}_t"].tobytes(), g[f"{name}_q"].tobytes(), int(g[f"{name}_N"][0])
        qrc = oracle.reverse_complement(q)
        sigs = [oracle.encode(s, N) for s in (t, q, qrc)]
        assert [sha(s) for s in sigs] == list(g[f"{name}_sigsha"]), name
        for cut in (1.8, 2.9):
            for s, qs in ((0, q), (1, qrc)):
                xc = oracle.xcorr(sigs[0], sigs[1 + s])
                _check_digest(g, f"{name}_{cut}_s{s}_", t, qs, xc, cut, 4096.0 * 64)


def test_bdhv_pairs_golden(golden):
    """Chunks dense in B/D/H/V: every entropy window sums thirds and takes glibc's log."""
    g = golden("bdhv_pairs.npz")
    for name in g["names"]:
        t, q, N = g[f"{name}_t"].tobytes(), g[f"{name}_q"].tobytes(), int(g[f"{name}_N"][0])
        qrc = oracle.reverse_complement(q)
        sigs = [oracle.encode(s, N) for s in (t, q, qrc)]
        assert [sha(s) for s in sigs] == list(g[f"{name}_sigsha"]), name
        for s, qs in ((0, q), (1, qrc)):
            xc = oracle.xcorr(sigs[0], sigs[1 + s])
            _check_digest(g, f"{name}_s{s}_", t, qs, xc, 1.8, 4096.0 * 64)


def test_uneven_pairs_golden(golden):
    g = golden("uneven_pairs.npz")
    for k in range(int(g["n"][0])):
        lt, lq = g[f"len{k}"]
        t, q = unpack2(g[f"t{k}"], lt), unpack2(g[f"q{k}"], lq)
        N = int(g[f"N{k}"][0])
        qrc = oracle.reverse_complement(q)
        sigs = [oracle.encode(s, N) for s in (t, q, qrc)]
        assert [sha(s) for s in sigs] == list(g[f"sigsha{k}"])
        for s, qs in ((0, q), (1, qrc)):
            xc = oracle.xcorr(sigs[0], sigs[1 + s])
            _check_digest(g, f"p{k}_s{s}_", t, qs, xc, 1.8, 1e6)


# ---------------------------------------------------------------- slave path

def test_uniform_pairs_slave_golden(golden):
    g = golden("uniform_pairs.npz")
    P = g["t"].shape[0]
    T = [unpack2(g["t"][r], 4096) for r in range(P)]
    Q = [unpack2(g["q"][r], 4096) for r in range(P)]
    prm = oracle.make_params(target_total=float(4096 * P))
    pairs = [dict(tseq=T[r], qseq=Q[r], t_record=r, q_record=r, q_record_len=4096, fast=int(g["fast"][r]))
             for r in range(P)]
    got = np.concatenate(oracle.align_batch(pairs, prm, threads=4))
    ref = g["results"]
    assert got.tobytes() == ref.tobytes()


def test_samples_slice_slave_golden(golden):
    g = golden("samples_slice.npz")
    t = g["t"].tobytes().upper()
    q = g["q"].tobytes().upper()
    tch = chunk_records([t], 4096, 1024)
    qch = chunk_records([q], 4096, 0)
    prm = oracle.make_params(target_total=float(len(t)))
    pairs = []
    for tf, tt, qf, qt, fast in g["pairs"]:
        for qi in range(qf, qt + 1):
            for ti in range(tf, tt + 1):
                trec, tstart, tseq = tch[ti]
                qrec, qstart, qseq = qch[qi]
                pairs.append(dict(tseq=tseq, qseq=qseq, t_chunk_start=tstart, q_chunk_start=qstart,
                                  t_record=trec, q_record=qrec, q_record_len=len(q), fast=int(fast)))
    got = np.concatenate(oracle.align_batch(pairs, prm, threads=4))
    assert got.tobytes() == g["results"].tobytes()


def _read_fasta(b: bytes):
    recs, names = [], []
    for line in b.split(b"\n"):
        if line.startswith(b">"):
            names.append(line[1:].decode())
            recs.append([])
        elif line:
            recs[-1].append(line)
    return [b"".join(r).upper() for r in recs], names


def _mm_records(b: bytes):
    """(ints, matches, prob, ident) of a MultiMatches v3 file (analysis/SequenceMatch.cc:320-360)."""
    o = 4
    for _ in range(2):
        n = int(np.frombuffer(b, "<i4", 1, o)[0])
        o += 4
        for _ in range(n):
            o += 8 + int(np.frombuffer(b, "<i8", 1, o)[0])
    n = int(np.frombuffer(b, "<i4", 1, o)[0])
    dt = np.dtype([("ints", "<i4", (7,)), ("matches", "<f8"), ("prob", "<f8"), ("ident", "<f8")])
    return np.frombuffer(b, dt, n, o + 4)


def test_cli_signal_tails_golden(golden):
    """The standalone CLI reuses one CCSignal per side, so a chunk shorter than an earlier
    one correlates with that chunk's signal in its tail (tools/analysis/HomologyByXCorr.cc:
    706-716, 780-798; analysis/CrossCorr.cc:118-131).  The oracle with tails built by
    replaying the CLI's SetSequence calls must reproduce the reference's match file; without
    the tails it must not (the record set has chunks shorter than their predecessors)."""
    g = golden("cli_records.npz")
    T, _ = _read_fasta(g["t_fasta"].tobytes())
    Q, _ = _read_fasta(g["q_fasta"].tobytes())
    tch = [c for c in chunk_records(T, 4096, 2048) if c[2]]
    qch = [c for c in chunk_records(Q, 4096, 0) if c[2]]

    def replay(stack, seq, strand):
        while stack and len(stack[-1][0]) <= len(seq):
            stack.pop()
        tail = [(len(s), s, st) for s, st in reversed(stack)]
        stack.append((seq, strand))
        return tail

    pairs, tt, qt = [], [], []
    ts, qs = [], []
    for trec, tstart, tseq in tch:
        ttail = replay(ts, tseq, 0)
        for qrec, qstart, qseq in qch:
            qtail = replay(qs, qseq, 0)
            assert replay(qs, qseq, 1) == qtail
            pairs.append(dict(tseq=tseq, qseq=qseq, t_chunk_start=tstart, q_chunk_start=qstart, t_record=trec,
                              q_record=qrec, q_record_len=len(Q[qrec])))
            tt.append(ttail)
            qt.append(qtail)
    assert sum(1 for x in tt if x) > 0 and sum(1 for x in qt if x) > 0
    prm = oracle.make_params(q_chunk=0, min_prob=0.9999, target_total=float(sum(len(t) for t in T)))
    exp = _mm_records(g["out_default"].tobytes())
    got = np.concatenate(oracle.align_batch(pairs, prm, threads=4, t_tails=tt, q_tails=qt))
    ints = np.stack([got["target_id"], got["query_id"], got["query_size"], got["tstart"], got["qstart"], got["len"],
                     got["reverse"]], 1).astype(np.int64).astype(np.int32)
    assert np.array_equal(ints, exp["ints"])
    assert np.array_equal(got["prob"], exp["prob"]) and np.array_equal(got["ident"], exp["ident"])
    assert np.array_equal(got["ident"] * got["len"].astype(np.float64), exp["matches"])
    plain = np.concatenate(oracle.align_batch(pairs, prm, threads=4))
    assert len(plain) != len(got) or plain.tobytes() != got.tobytes()


def test_prob_table_golden(golden):
    """ProbTable::Setup thresholds (analysis/ProbTable.cc:15-52) and the slave path with
    -prob_table (analysis/HomologyByXCorrSlave.cc:184-187) against the reference's."""
    g = golden("prob_table.npz")
    size = float(g["target_size"][0])
    t = oracle.prob_table(size)
    assert sha(t) == str(g["table_sha"])
    assert np.array_equal(t[g["row_idx"]], g["rows"])
    s = golden("samples_slice.npz")
    T, Q = s["t"].tobytes().upper(), s["q"].tobytes().upper()
    tch, qch = chunk_records([T], 4096, 1024), chunk_records([Q], 4096, 0)
    pairs = [dict(tseq=tch[ti][2], qseq=qch[qi][2], t_chunk_start=tch[ti][1], q_chunk_start=qch[qi][1],
                  t_record=tch[ti][0], q_record=qch[qi][0], q_record_len=len(Q), fast=int(fast))
             for tf, tt, qf, qt, fast in s["pairs"] for qi in range(qf, qt + 1) for ti in range(tf, tt + 1)]
    for name, cut in (("hi", 0.9999), ("lo", 0.9)):
        prm = oracle.make_params(target_total=size, ptable=t, table_prob=cut)
        got = np.concatenate(oracle.align_batch(pairs, prm, threads=8, cap_per_pair=1 << 15))
        assert len(got) == int(g[f"results_{name}_n"][0])
        assert sha(got) == str(g[f"results_{name}_sha"]), name
        assert got[:256].tobytes() == g[f"results_{name}_head"].tobytes()
