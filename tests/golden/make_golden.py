"""Writes the golden vectors under tests/golden/ from the reference compiled from source.

Run in a container that has the reference sources (it calls ``make -C oracle ref``,
which compiles them where they lie into oracle/_ref/).  The fixtures are data only:
inputs (sequences, 2-bit packed where pure ACGT) and the reference's outputs.

  python tests/golden/make_golden.py

Fixtures:
  fft.npz            FFTReal forward / inverse / rescale on white noise, N = 8 .. 32768
  single_pair.npz    one 4096 x 4096 pair, N = 8192: the three encoded signals, xc for both
                     strands, FindTop lags, MatchUp runs and their prob / ident
  edge_pairs.npz     short, empty, N-rich, IUPAC and repeat sequences through the same calls
  uneven_pairs.npz   query/target lengths 256..16384 at N = nextpow2(q + t) (512 .. 32768)
  uniform_pairs.npz  the slave path (chunking, both strands, FilterMatches) on 64 pairs
  samples_slice.npz  40 kbp x 40 kbp slice of the reference's samples/ (soft-masked, with N)
                     through the slave path and through the standalone CLI's match file
  cli_records.npz    12 + 12 records of uneven lengths through the standalone CLI under four
                     flag sets (match files, MultiMatches v3)
  prob_table.npz     ProbTable thresholds (digest + rows) and the slave path with -prob_table
  bdhv_pairs.npz     chunks dense in B/D/H/V (entropy windows of thirds, glibc log) at N = 2048 ..
                     32768: signals, xc, FindTop lags and MatchUp runs
  kmatch_cases.npz   KMatch (kmatch/KMatch.cc, oracle/_ref/KMatch) on a FASTA pair with repeats,
                     soft-masked and N stretches: its t_result files at several K, lengths,
                     jumps and max_freq
  master_cases.npz   the master's match-set steps (Sort, Collapse, LengthFilter, RunMatchDynProg,
                     RunMatchDynProgMult, SyntenyInterpolator) on random match sets, through
                     oracle/_ref/ref_master

  python tests/golden/make_golden.py bdhv   (one generator only)
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from satsuma2_amd import synth  # noqa: E402

REF = os.environ.get("SATSUMA2_REF", "/root/reference")
DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
CLI = os.path.join(ROOT, "oracle", "_ref", "HomologyByXCorr")

RESULT_DTYPE = np.dtype([("query_id", "<u8"), ("target_id", "<u8"), ("query_size", "<u8"), ("qstart", "<u8"),
                         ("tstart", "<u8"), ("len", "<u8"), ("reverse", "u1"), ("pad_", "u1", (7,)),
                         ("prob", "<f8"), ("ident", "<f8")])
MATCH_DTYPE = np.dtype([("st", "<i4"), ("sq", "<i4"), ("len", "<i4"), ("prob", "<f8"), ("ident", "<f8")])


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


CODE = np.full(256, 255, np.uint8)
CODE[np.frombuffer(b"ACGT", np.uint8)] = np.arange(4, dtype=np.uint8)


def pack2(seq: bytes) -> np.ndarray:
    """2-bit pack (A=0, C=1, G=2, T=3, first base in the low bits) of a pure ACGT sequence."""
    c = CODE[np.frombuffer(seq, np.uint8)]
    assert (c < 4).all()
    c = np.concatenate([c, np.zeros((-c.size) % 4, np.uint8)])
    c = c.reshape(-1, 4)
    return (c[:, 0] | (c[:, 1] << 2) | (c[:, 2] << 4) | (c[:, 3] << 6)).astype(np.uint8)


def lcg_noise(n: int, seed: int) -> np.ndarray:
    """Deterministic white noise in [-0.5, 0.5) from a 64-bit LCG (Knuth MMIX constants)."""
    out = np.empty(n, np.float32)
    x = seed & 0xFFFFFFFFFFFFFFFF
    for i in range(n):
        x = (x * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        out[i] = np.float32((x >> 40) / float(1 << 24) - 0.5)
    return out


def digest_strand(prefix, d, xc, top, m):
    d[prefix + "xc_sha"] = np.array(sha(xc))
    d[prefix + "top"] = top
    d[prefix + "m_n"] = np.array([m.size], np.int64)
    d[prefix + "m_sha"] = np.array(sha(np.stack([m["st"], m["sq"], m["len"]], 1).astype(np.int32)))
    d[prefix + "prob_sha"] = np.array(sha(m["prob"]))
    d[prefix + "ident_sha"] = np.array(sha(m["ident"]))


def run(args, **kw):
    return subprocess.run(args, check=True, capture_output=True, text=True, **kw).stdout


def detail(tmp, t: bytes, q: bytes, N: int, cutoff: float, tsize: float):
    tp, qp, op = (os.path.join(tmp, x) for x in ("t.seq", "q.seq", "d.bin"))
    open(tp, "wb").write(t)
    open(qp, "wb").write(q)
    run([DRIVER, "detail", tp, qp, str(N), repr(cutoff), repr(tsize), op])
    b = open(op, "rb").read()
    off = 4
    sig = np.frombuffer(b, np.float32, 12 * N, off).reshape(3, 4, N).copy()
    off += 12 * N * 4
    strands = []
    for _ in range(2):
        xc = np.frombuffer(b, np.float32, N, off).copy()
        off += 4 * N
        nt = int(np.frombuffer(b, np.int32, 1, off)[0])
        off += 4
        top = np.frombuffer(b, np.int32, nt, off).copy()
        off += 4 * nt
        nm = int(np.frombuffer(b, np.int32, 1, off)[0])
        off += 4
        m = np.frombuffer(b, MATCH_DTYPE, nm, off).copy()
        off += MATCH_DTYPE.itemsize * nm
        strands.append((xc, top, m))
    return sig, strands


def gen_fft(tmp):
    """Inputs: lcg_noise(N, seed=N) (restated in tests/test_oracle.py)."""
    out = {}
    for nb in range(3, 16):
        N = 1 << nb
        x = lcg_noise(N, N)
        ip, op = os.path.join(tmp, "x.bin"), os.path.join(tmp, "y.bin")
        x.tofile(ip)
        run([DRIVER, "fft", str(N), ip, op])
        y = np.fromfile(op, np.float32).reshape(3, N)
        if N <= 2048:
            out[f"fwd_{N}"] = y[0]
            out[f"inv_{N}"] = y[1]
            out[f"res_{N}"] = y[2]
        out[f"sha_{N}"] = np.array([sha(y[0]), sha(y[1]), sha(y[2])])
    np.savez_compressed(os.path.join(HERE, "fft.npz"), **out)


def gen_single(tmp):
    t, q, meta = synth.single_pair(0)
    sig, st = detail(tmp, t, q, 8192, 1.8, 4096.0)
    d = {"t": np.frombuffer(t, np.uint8), "q": np.frombuffer(q, np.uint8), "sig": sig,
         "plant": np.array([meta["qo"], meta["to"]], np.int32)}
    for s, (xc, top, m) in enumerate(st):
        d[f"xc{s}"] = xc
        d[f"top{s}"] = top
        d[f"m{s}"] = m
    np.savez_compressed(os.path.join(HERE, "single_pair.npz"), **d)


def edge_cases():
    rng = np.random.default_rng(11)
    base = synth.random_seq(rng, 4096, 0.41)
    cases = []
    # short query (< 1024: entropy = 1), same target
    cases.append(("short_query", base.tobytes(), base[100:900].tobytes(), 8192))
    # empty query
    cases.append(("empty_query", base.tobytes(), b"", 8192))
    # all-N query
    cases.append(("all_n_query", base.tobytes(), b"N" * 4096, 8192))
    # identical sequences
    cases.append(("identical", base.tobytes(), base.tobytes(), 8192))
    # N runs inside both
    t = base.copy()
    q = base.copy()
    t[500:900] = ord("N")
    q[2000:2300] = ord("N")
    q = synth.mutate(rng, q, 0.9)
    cases.append(("n_runs", t.tobytes(), q.tobytes(), 8192))
    # IUPAC codes (fractional codec rows, incl. 1/3 entries)
    iu = np.frombuffer(b"KMRYSWBVHDNX-", np.uint8)
    t = base.copy()
    q = synth.mutate(rng, base, 0.8)
    pos = rng.integers(0, 4096, size=300)
    t[pos] = iu[rng.integers(0, iu.size, size=300)]
    pos = rng.integers(0, 4096, size=300)
    q[pos] = iu[rng.integers(0, iu.size, size=300)]
    cases.append(("iupac", t.tobytes(), q.tobytes(), 8192))
    # unknown bytes: codec rows are zero but identical letters still score 100
    t = base.copy()
    q = base.copy()
    t[1000:1100] = ord(".")
    q[1000:1100] = ord(".")
    q[3000:3050] = ord("u")
    cases.append(("unknown_bytes", t.tobytes(), q.tobytes(), 8192))
    # tandem repeats (low complexity: many peaks)
    unit = synth.random_seq(rng, 3, 0.41)
    t = base.copy()
    t[1000:1900] = np.resize(unit, 900)
    q = synth.random_seq(rng, 4096, 0.41)
    q[200:800] = np.resize(unit, 600)
    cases.append(("tandem", t.tobytes(), q.tobytes(), 8192))
    # homopolymer: zero entropy windows
    t = base.copy()
    t[0:2048] = ord("A")
    q = base.copy()
    q[1024:3072] = ord("A")
    cases.append(("homopolymer", t.tobytes(), q.tobytes(), 8192))
    # unequal lengths, target shorter than query
    cases.append(("short_target", base[:1500].tobytes(), synth.random_seq(rng, 4096, 0.41).tobytes(), 8192))
    # fast cutoff on a planted pair
    tt, qq, _ = synth.single_pair(3)
    cases.append(("planted_n8192", tt, qq, 8192))
    return cases


def gen_edges(tmp):
    d = {}
    names = []
    for name, t, q, N in edge_cases():
        names.append(name)
        for cut in (1.8, 2.9):
            sig, st = detail(tmp, t, q, N, cut, 4096.0 * 64)
            key = f"{name}_{cut}"
            if cut == 1.8:
                d[f"{name}_t"] = np.frombuffer(t, np.uint8)
                d[f"{name}_q"] = np.frombuffer(q, np.uint8)
                d[f"{name}_N"] = np.array([N], np.int32)
                d[f"{name}_sigsha"] = np.array([sha(sig[i]) for i in range(3)])
            for s, (xc, top, m) in enumerate(st):
                digest_strand(f"{key}_s{s}_", d, xc, top, m)
    d["names"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, "edge_pairs.npz"), **d)


def gen_uneven(tmp, n=32):
    T, Q = synth.uneven_pairs(n, seed=2)
    d = {}
    for k in range(n):
        t, q = T[k], Q[k]
        N = synth.fft_len(len(q), len(t))
        sig, st = detail(tmp, t, q, N, 1.8, 1e6)
        d[f"t{k}"] = pack2(t)
        d[f"q{k}"] = pack2(q)
        d[f"len{k}"] = np.array([len(t), len(q)], np.int32)
        d[f"N{k}"] = np.array([N], np.int32)
        d[f"sigsha{k}"] = np.array([sha(sig[i]) for i in range(3)])
        for s, (xc, top, m) in enumerate(st):
            digest_strand(f"p{k}_s{s}_", d, xc, top, m)
    d["n"] = np.array([n], np.int32)
    np.savez_compressed(os.path.join(HERE, "uneven_pairs.npz"), **d)


def bdhv_cases(seed=21):
    """Chunks dense in B, D, H and V (codec rows of thirds) mixed with A/C/G/T, the other IUPAC
    codes and N, with a planted homology: every entropy window sums thirds."""
    rng = np.random.default_rng(seed)
    thirds = np.frombuffer(b"BDHV", np.uint8)
    other = np.frombuffer(b"KMRYSWNX", np.uint8)
    cases = []
    for k, (lt, lq, N) in enumerate([(4096, 4096, 8192), (4096, 4096, 8192), (3000, 4096, 8192), (1024, 1500, 4096),
                                     (1100, 900, 2048), (8192, 8192, 16384), (16384, 9000, 32768)]):
        t = synth.random_seq(rng, lt, 0.45)
        q = synth.random_seq(rng, lq, 0.45)
        L = min(lt, lq) // 2
        a, b = int(rng.integers(0, lt - L)), int(rng.integers(0, lq - L))
        q[b:b + L] = synth.mutate(rng, t[a:a + L], 0.9)
        for s, dens in ((t, 0.3 if k % 2 == 0 else 0.05), (q, 0.2)):
            pos = rng.random(s.size) < dens
            s[pos] = thirds[rng.integers(0, 4, size=int(pos.sum()))]
            pos = rng.random(s.size) < 0.03
            s[pos] = other[rng.integers(0, other.size, size=int(pos.sum()))]
        cases.append((f"bdhv{k}", t.tobytes(), q.tobytes(), N))
    return cases


def gen_bdhv(tmp):
    d = {}
    names = []
    for name, t, q, N in bdhv_cases():
        names.append(name)
        sig, st = detail(tmp, t, q, N, 1.8, 4096.0 * 64)
        d[f"{name}_t"] = np.frombuffer(t, np.uint8)
        d[f"{name}_q"] = np.frombuffer(q, np.uint8)
        d[f"{name}_N"] = np.array([N], np.int32)
        d[f"{name}_sigsha"] = np.array([sha(sig[i]) for i in range(3)])
        for s, (xc, top, m) in enumerate(st):
            digest_strand(f"{name}_s{s}_", d, xc, top, m)
    d["names"] = np.array(names)
    np.savez_compressed(os.path.join(HERE, "bdhv_pairs.npz"), **d)


def read_results(path):
    b = open(path, "rb").read()
    n = int(np.frombuffer(b, np.uint64, 1)[0])
    return np.frombuffer(b, RESULT_DTYPE, n, 8).copy()


def gen_uniform(tmp, P=64):
    T, Q = synth.uniform_pairs(P, seed=1)
    qf, tf, pf, of = (os.path.join(tmp, x) for x in ("q.fa", "t.fa", "p.txt", "r.bin"))
    synth.write_fasta(qf, Q)
    synth.write_fasta(tf, T)
    with open(pf, "w") as f:
        for r in range(P):
            f.write(f"{2 * r} {2 * r} {2 * r} {2 * r} {1 if r % 8 == 7 else 0}\n")
    run([DRIVER, "pairs", qf, tf, pf, of, "1"])
    res = read_results(of)
    tq = np.stack([pack2(t) for t in T])
    qq = np.stack([pack2(q) for q in Q])
    np.savez_compressed(os.path.join(HERE, "uniform_pairs.npz"), t=tq, q=qq, results=res,
                        fast=np.array([1 if r % 8 == 7 else 0 for r in range(P)], np.int32))


def fasta_slice(path, n):
    seq = []
    name = None
    with open(path, "rb") as f:
        for line in f:
            line = line.strip()
            if line.startswith(b">"):
                name = line[1:].split()[0]
                continue
            seq.append(line)
            if sum(map(len, seq)) >= n:
                break
    return name, b"".join(seq)[:n]


def gen_samples(tmp, n=40000):
    _, h = fasta_slice(os.path.join(REF, "samples", "human.X.part.fasta"), n)
    _, g = fasta_slice(os.path.join(REF, "samples", "dog.X.part.fasta"), n)
    qf, tf, pf, of = (os.path.join(tmp, x) for x in ("q.fa", "t.fa", "p.txt", "r.bin"))
    synth.write_fasta(qf, [g], ["dogX"])
    synth.write_fasta(tf, [h], ["humanX"])
    # slave path: every target chunk against every query chunk, 2 blocks, one fast
    with open(pf, "w") as f:
        f.write("0 6 0 4 0\n7 13 5 9 1\n")
    run([DRIVER, "pairs", qf, tf, pf, of, "1"])
    res = read_results(of)
    # standalone CLI match file (tools/analysis/HomologyByXCorr.cc:850)
    mf = os.path.join(tmp, "xcorr.out")
    run([CLI, "-q", qf, "-t", tf, "-o", mf, "-min_prob", "0.99999"])
    mfile = np.frombuffer(open(mf, "rb").read(), np.uint8)
    np.savez_compressed(os.path.join(HERE, "samples_slice.npz"), t=np.frombuffer(h, np.uint8),
                        q=np.frombuffer(g, np.uint8), results=res, match_file=mfile,
                        pairs=np.array([[0, 6, 0, 4, 0], [7, 13, 5, 9, 1]], np.int32))


CLI_RUNS = [
    ("default", []),
    ("small_chunks", ["-t_chunk", "2048", "-q_chunk", "1024", "-l", "60", "-min_prob", "0.999"]),
    ("same_only", ["-same_only", "true", "-cutoff", "1.6"]),
    ("blocks", ["-nblocks", "3", "-block", "1", "-nblocks_query", "2", "-block_query", "0"]),
]


def cli_records(seed=5):
    """Query and target record sets of uneven lengths (many shorter than a chunk, some in
    descending order, some under 1024 bp) cut from one random genome, the query side
    mutated (80-95% identity, indels), soft-masked runs and N runs, shared record names."""
    rng = np.random.default_rng(seed)
    base = synth.random_seq(rng, 120_000, 0.41)
    tl = [12000, 4096, 3000, 9000, 2500, 2000, 700, 6000, 1500, 900, 5000, 3500]
    ql = [7000, 3100, 2900, 8192, 2600, 1800, 640, 5000, 4000, 1200, 950, 4500]
    T, Q, tn, qn = [], [], [], []
    pos = 0
    for k, (a, b) in enumerate(zip(tl, ql)):
        t = base[pos:pos + a].copy()
        off = int(rng.integers(0, max(1, a // 3)))
        q = synth.mutate(rng, base[pos + off:pos + off + b], float(rng.uniform(0.8, 0.95)), 0.01)[:b]
        if k % 3 == 1:  # inverted homology
            q = synth.ALPHA[3 - np.searchsorted(synth.ALPHA, q[::-1])]
        tb, qb = bytearray(t.tobytes()), bytearray(q.tobytes())
        if k % 4 == 0 and len(tb) > 600:  # soft mask
            i = int(rng.integers(0, len(tb) - 500))
            tb[i:i + 400] = tb[i:i + 400].lower()
        if k % 5 == 2 and len(qb) > 300:  # N run
            i = int(rng.integers(0, len(qb) - 200))
            qb[i:i + 150] = b"N" * 150
        T.append(bytes(tb))
        Q.append(bytes(qb))
        tn.append(f"chr{k % 5}")
        qn.append(f"chr{(k + 1) % 5}")
        pos += a + 500
    return T, Q, tn, qn


def gen_cli(tmp):
    T, Q, tn, qn = cli_records()
    qf, tf = os.path.join(tmp, "cq.fa"), os.path.join(tmp, "ct.fa")
    synth.write_fasta(qf, Q, qn)
    synth.write_fasta(tf, T, tn)
    d = {"q_fasta": np.frombuffer(open(qf, "rb").read(), np.uint8),
         "t_fasta": np.frombuffer(open(tf, "rb").read(), np.uint8)}
    for name, flags in CLI_RUNS:
        out = os.path.join(tmp, f"{name}.out")
        run([CLI, "-q", qf, "-t", tf, "-o", out, *flags])
        d[f"out_{name}"] = np.frombuffer(open(out, "rb").read(), np.uint8)
        d[f"flags_{name}"] = np.array(" ".join(flags))
    np.savez_compressed(os.path.join(HERE, "cli_records.npz"), **d)


PT_ROWS = [1, 64, 127, 200, 255, 383, 511]


def gen_probtable(tmp):
    """ProbTable thresholds for the samples slice's target size and the reference slave's
    records with -prob_table on the samples slice blocks (cutoff 0.9999, and 0.9, which the
    slave's fixed 0.99 keep threshold then rejects entirely)."""
    g = np.load(os.path.join(HERE, "samples_slice.npz"))
    size = float(len(g["t"]))
    tb = os.path.join(tmp, "pt.bin")
    run([DRIVER, "probtable", repr(size), "0.9999", tb])
    t = np.fromfile(tb, np.float64).reshape(512, 2048)
    qf, tf, pf = (os.path.join(tmp, x) for x in ("q.fa", "t.fa", "p.txt"))
    synth.write_fasta(qf, [g["q"].tobytes()], ["dogX"])
    synth.write_fasta(tf, [g["t"].tobytes()], ["humanX"])
    with open(pf, "w") as f:
        for b in g["pairs"]:
            f.write(" ".join(str(int(x)) for x in b) + "\n")
    out = {}
    for name, cut in (("hi", "0.9999"), ("lo", "0.9")):
        of = os.path.join(tmp, f"r_{name}.bin")
        run([DRIVER, "pairs", qf, tf, pf, of, "1", "1.8", "2.9", "0", "4096", "4096", "1", cut])
        r = read_results(of)
        # the table keeps almost every run (171k records here): digest, count and a sample
        out[f"results_{name}_sha"] = sha(r)
        out[f"results_{name}_n"] = np.array([len(r)])
        out[f"results_{name}_head"] = r[:256]
    np.savez_compressed(os.path.join(HERE, "prob_table.npz"), target_size=np.array([size]), table_sha=sha(t),
                        rows=t[PT_ROWS], row_idx=np.array(PT_ROWS, np.int32), **out)


MASTER = os.path.join(ROOT, "oracle", "_ref", "ref_master")
M56 = np.dtype([("target_id", "<i4"), ("query_id", "<i4"), ("query_len", "<i4"), ("start_target", "<i4"),
                ("start_query", "<i4"), ("length", "<i4"), ("rc", "<i4"), ("reserved", "<i4"),
                ("matches", "<f8"), ("prob", "<f8"), ("ident", "<f8")])


def master_set(seed, nt=2, nq=3, runs=8, noise=150):
    """Collinear runs (two long enough to interpolate), near copies and noise."""
    rng = np.random.default_rng(seed)
    ts = rng.integers(300_000, 900_000, nt).astype(np.int32)
    qs = rng.integers(300_000, 900_000, nq).astype(np.int32)
    rows = []
    for k in range(runs):
        t, q, rc = int(rng.integers(nt)), int(rng.integers(nq)), int(rng.integers(2))
        n = 60 if k < 2 else int(rng.integers(5, 30))
        step = int(rng.integers(400, 3000))
        st = int(rng.integers(0, max(1, ts[t] - n * step - 500)))
        sq = int(rng.integers(0, max(1, qs[q] - n * step - 500)))
        for i in range(n):
            ln = int(rng.integers(50, 250))
            rows.append((t, q, qs[q], st + i * step, sq + i * step, ln, rc))
            if rng.random() < 0.2:
                rows.append((t, q, qs[q], st + i * step + 2, sq + i * step + 1, ln + 10, rc))
    for _ in range(noise):
        t, q = int(rng.integers(nt)), int(rng.integers(nq))
        ln = int(rng.integers(40, 300))
        rows.append((t, q, qs[q], int(rng.integers(0, ts[t] - ln)), int(rng.integers(0, qs[q] - ln)), ln,
                     int(rng.integers(2))))
    m = np.zeros(len(rows), M56)
    for i, r in enumerate(rows):
        (m[i]["target_id"], m[i]["query_id"], m[i]["query_len"], m[i]["start_target"], m[i]["start_query"],
         m[i]["length"], m[i]["rc"]) = r
    m["prob"] = rng.uniform(0.99, 1.0, len(m))
    m["ident"] = rng.uniform(0.6, 1.0, len(m))
    m["matches"] = m["ident"] * m["length"]
    rng.shuffle(m)
    return ts, qs, m


def master_run(tmp, op, ts, qs, m, *extra):
    inp, out = os.path.join(tmp, "m.in"), os.path.join(tmp, "m.out")
    with open(inp, "wb") as f:
        f.write(np.array([len(ts), len(qs)], np.int32).tobytes() + ts.tobytes() + qs.tobytes()
                + np.array([len(m)], np.int64).tobytes() + m.tobytes())
    args = [MASTER, op, inp, *[str(e) for e in extra], out]
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)
    b = open(out, "rb").read()
    nt, nq = np.frombuffer(b, np.int32, 2)
    off = 8 + 4 * int(nt + nq)
    n = int(np.frombuffer(b, np.int64, 1, off)[0])
    return np.frombuffer(b, M56, n, off + 8).copy()


def gen_master(tmp):
    out, cases = {}, []

    def case(name, op, ts, qs, m, res, *extra):
        cases.append(name)
        out.update({f"{name}_op": np.array(op), f"{name}_ts": ts, f"{name}_qs": qs, f"{name}_in": m,
                    f"{name}_out": res})
        if extra:
            out[f"{name}_arg"] = np.array(extra)

    for seed in (101, 102):
        ts, qs, m = master_set(seed)
        srt = master_run(tmp, "sort", ts, qs, m)
        case(f"sort{seed}", "sort", ts, qs, m, srt)
        col = master_run(tmp, "collapse", ts, qs, srt)
        case(f"collapse{seed}", "collapse", ts, qs, srt, col)
        for op in ("chain", "chainmult"):
            case(f"{op}{seed}", op, ts, qs, col, master_run(tmp, op, ts, qs, col))
        ch = master_run(tmp, "chain", ts, qs, col)
        case(f"interpolate{seed}", "interpolate", ts, qs, ch, master_run(tmp, "interpolate", ts, qs, ch))
    np.savez_compressed(os.path.join(HERE, "master_cases.npz"), cases=np.array(cases), **out)


def gen_kmatch(tmp):
    sys.path.insert(0, os.path.dirname(HERE))
    from kmatch_cases import fasta_text, genome_pair
    t, q = genome_pair(4, n_records=2, size=30_000)
    tf, qf, of = os.path.join(tmp, "t.fa"), os.path.join(tmp, "q.fa"), os.path.join(tmp, "o.bin")
    tt, qt = fasta_text(t, 60, False, "t"), fasta_text(q, 70, True, "q")
    open(tf, "wb").write(tt)
    open(qf, "wb").write(qt)
    args = [(11, 11, 10, 1), (21, 21, 20, 1), (31, 31, 30, 1), (13, 40, 5, 1), (15, 30, 14, 3)]
    outs = {}
    for i, (K, ml, jump, freq) in enumerate(args):
        subprocess.run([os.path.join(ROOT, "oracle", "_ref", "KMatch"), qf, tf, str(K), of, str(ml), str(jump),
                        str(freq)], check=True, stdout=subprocess.DEVNULL)
        outs[f"out{i}"] = np.frombuffer(open(of, "rb").read(), np.uint8)
    np.savez_compressed(os.path.join(HERE, "kmatch_cases.npz"), query_fasta=np.frombuffer(qt, np.uint8),
                        target_fasta=np.frombuffer(tt, np.uint8), args=np.array(args, np.int32), **outs)


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", f"REF={REF}"], check=True)
    only = sys.argv[1:]  # generator names (e.g. bdhv); none: all
    with tempfile.TemporaryDirectory() as tmp:
        if only:
            for name in only:
                globals()["gen_" + name](tmp)
            return
        gen_fft(tmp)
        gen_single(tmp)
        gen_edges(tmp)
        gen_uneven(tmp)
        gen_uniform(tmp)
        gen_samples(tmp)
        gen_cli(tmp)
        gen_probtable(tmp)
        gen_bdhv(tmp)
        gen_master(tmp)
        gen_kmatch(tmp)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
