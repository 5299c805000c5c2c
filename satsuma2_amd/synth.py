"""Synthetic chunk-pair workloads (numpy PCG64, seeded).

Workloads, named by what they are:
  * ``single_pair``      one 4096-bp target x 4096-bp query, GC 0.41, one 500-bp
                         homologous segment at 85% identity (uniform substitutions,
                         no indels), forward strand only.
  * ``uniform_pairs``    P chunk pairs of 4096 bp, GC 0.41; even-indexed pairs carry
                         one planted homology (length U{50..2000}, identity
                         U(0.70, 0.95), 1% per-base indel starts with geometric
                         lengths of mean 2); both strands.
  * ``uneven_pairs``     P pairs with query and target lengths log-uniform in
                         [256, 16384], GC U(0.35, 0.55) per pair, planted as above,
                         10% with a tandem repeat (period 1-6, length 100-1000).

  * ``syntenic_genomes``  two genomes (one record each, default 50 Mbp) at GC 0.41
                         sharing syntenic blocks: lengths log-uniform 5-200 kbp,
                         identity U(0.75, 0.95), 1% indels, 10% inverted, placed along
                         a noisy diagonal (later blocks overwrite earlier ones where
                         they overlap).

Base probabilities at GC = g: P(C) = P(G) = g/2, P(A) = P(T) = (1 - g)/2.
"""
from __future__ import annotations

import numpy as np

ALPHA = np.frombuffer(b"ACGT", dtype=np.uint8)


def random_seq(rng: np.random.Generator, n: int, gc: float) -> np.ndarray:
    p = np.array([(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2])
    return ALPHA[rng.choice(4, size=n, p=p)]


def mutate(rng: np.random.Generator, seg: np.ndarray, ident: float, indel_rate: float = 0.0) -> np.ndarray:
    """Substitute each base with probability 1-ident (uniformly among the other 3);
    with indel_rate > 0, start an insertion or deletion (50/50, geometric length,
    p = 0.5) after a base with that probability."""
    seg = seg.copy()
    sub = rng.random(seg.size) < (1.0 - ident)
    if sub.any():
        cur = np.searchsorted(ALPHA, seg[sub])
        shift = rng.integers(1, 4, size=cur.size)
        seg[sub] = ALPHA[(cur + shift) % 4]
    if indel_rate <= 0:
        return seg
    out = []
    i = 0
    n = seg.size
    while i < n:
        out.append(seg[i])
        i += 1
        if rng.random() < indel_rate:
            ln = int(rng.geometric(0.5))
            if rng.random() < 0.5:
                out.extend(ALPHA[rng.integers(0, 4, size=ln)].tolist())
            else:
                i += ln
    return np.array(out, dtype=np.uint8)


def single_pair(seed: int = 0, length: int = 4096, seg_len: int = 500, ident: float = 0.85, gc: float = 0.41):
    rng = np.random.default_rng(seed)
    q = random_seq(rng, length, gc)
    t = random_seq(rng, length, gc)
    qo = int(rng.integers(0, length - seg_len + 1))
    to = int(rng.integers(0, length - seg_len + 1))
    t[to: to + seg_len] = mutate(rng, q[qo: qo + seg_len], ident)
    return t.tobytes(), q.tobytes(), {"qo": qo, "to": to}


def _plant(rng, q: np.ndarray, t: np.ndarray):
    L = int(rng.integers(50, 2001))
    ident = float(rng.uniform(0.70, 0.95))
    L = min(L, q.size, t.size)
    qo = int(rng.integers(0, q.size - L + 1))
    seg = mutate(rng, q[qo: qo + L], ident, indel_rate=0.01)
    seg = seg[: t.size]
    to = int(rng.integers(0, t.size - seg.size + 1))
    t[to: to + seg.size] = seg
    return {"L": L, "ident": ident, "qo": qo, "to": to}


def uniform_pairs(n_pairs: int, seed: int = 1, length: int = 4096, gc: float = 0.41):
    """Returns (targets, queries) lists of bytes, each entry one 4096-bp record."""
    rng = np.random.default_rng(seed)
    T, Q = [], []
    for k in range(n_pairs):
        q = random_seq(rng, length, gc)
        t = random_seq(rng, length, gc)
        if k % 2 == 0:
            _plant(rng, q, t)
        T.append(t.tobytes())
        Q.append(q.tobytes())
    return T, Q


def uneven_pairs(n_pairs: int, seed: int = 2, lo: int = 256, hi: int = 16384):
    rng = np.random.default_rng(seed)
    T, Q = [], []
    for k in range(n_pairs):
        lq = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        lt = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        gc = float(rng.uniform(0.35, 0.55))
        q = random_seq(rng, lq, gc)
        t = random_seq(rng, lt, gc)
        if k % 2 == 0:
            _plant(rng, q, t)
        if rng.random() < 0.10:
            period = int(rng.integers(1, 7))
            rl = int(rng.integers(100, 1001))
            unit = random_seq(rng, period, gc)
            rep = np.resize(unit, rl)
            tgt = q if rng.random() < 0.5 else t
            rl = min(rl, tgt.size)
            off = int(rng.integers(0, tgt.size - rl + 1))
            tgt[off: off + rl] = rep[:rl]
        T.append(t.tobytes())
        Q.append(q.tobytes())
    return T, Q


def fft_len(q_len: int, t_len: int) -> int:
    n = 1
    while n < q_len + t_len:
        n *= 2
    return n


def write_fasta(path: str, records, names=None, width: int = 60) -> None:
    with open(path, "wb") as f:
        for i, s in enumerate(records):
            name = names[i] if names else f"r{i}"
            f.write(b">" + name.encode() + b"\n")
            for k in range(0, len(s), width):
                f.write(s[k: k + width] + b"\n")


def mutate_fast(rng: np.random.Generator, seg: np.ndarray, ident: float, indel_rate: float) -> np.ndarray:
    """mutate() for long segments: substitutions vectorised, indels (50/50 insertion or
    deletion, geometric length p = 0.5) applied event by event."""
    seg = seg.copy()
    sub = rng.random(seg.size) < (1.0 - ident)
    if sub.any():
        cur = np.searchsorted(ALPHA, seg[sub])
        seg[sub] = ALPHA[(cur + rng.integers(1, 4, size=cur.size)) % 4]
    ev = np.nonzero(rng.random(seg.size) < indel_rate)[0]
    if ev.size == 0:
        return seg
    lens = rng.geometric(0.5, size=ev.size)
    ins = rng.random(ev.size) < 0.5
    pieces, prev = [], 0
    for pos, ln, is_ins in zip(ev.tolist(), lens.tolist(), ins.tolist()):
        if pos < prev:
            continue
        pieces.append(seg[prev:pos + 1])
        if is_ins:
            pieces.append(ALPHA[rng.integers(0, 4, size=ln)])
            prev = pos + 1
        else:
            prev = pos + 1 + ln
    pieces.append(seg[prev:])
    return np.concatenate(pieces)


def revcomp(a: np.ndarray) -> np.ndarray:
    return ALPHA[3 - np.searchsorted(ALPHA, a[::-1])]


def random_seq_fast(rng: np.random.Generator, n: int, gc: float) -> np.ndarray:
    """random_seq by inverse CDF on uniform bytes (for genome-sized n)."""
    cum = np.array([(1 - gc) / 2, 0.5, (1 + gc) / 2]) * 256.0
    u = rng.integers(0, 256, size=n, dtype=np.uint8)
    return ALPHA[np.searchsorted(cum, u, side="right")]


def syntenic_genomes(length: int = 50_000_000, n_blocks: int = 2000, seed: int = 4, gc: float = 0.41):
    """(target bytes, query bytes, blocks) with blocks = list of (t_start, q_start, len, inverted)."""
    rng = np.random.default_rng(seed)
    t = random_seq_fast(rng, length, gc)
    q = random_seq_fast(rng, length, gc)
    lens = np.exp(rng.uniform(np.log(5_000), np.log(200_000), n_blocks)).astype(np.int64)
    blocks = []
    for k in range(n_blocks):
        ln = int(lens[k])
        ts = int(rng.integers(0, length - ln))
        qs = int(min(max(ts + rng.normal(0, 2e6), 0), length - ln))
        seg = mutate_fast(rng, t[ts:ts + ln], float(rng.uniform(0.75, 0.95)), 0.01)[:ln]
        inv = bool(rng.random() < 0.1)
        if inv:
            seg = revcomp(seg)
        q[qs:qs + seg.size] = seg
        blocks.append((ts, qs, int(seg.size), inv))
    return t.tobytes(), q.tobytes(), blocks


def large_genomes(length: int = 200_000_000, n_blocks: int = 5000, seed: int = 5, gc_window: int = 100_000,
                  repeat_frac: float = 0.10, n_families: int = 5, copies: int = 1000):
    """Two genomes for the large served run: GC uniform in [0.38, 0.45] per 100-kbp window;
    syntenic blocks as in syntenic_genomes (lengths log-uniform 5-200 kbp, identity 75-95%,
    1% indels, 10% inverted); then, in each genome independently, tandem repeats (period
    1-6, length 100-1000 bp) until they cover repeat_frac of it, and n_families
    transposon-like families (300-6000 bp) of `copies` copies each at 80-95% identity.
    Returns (target bytes, query bytes, blocks)."""
    rng = np.random.default_rng(seed)

    def base():
        g = np.empty(length, np.uint8)
        for w in range(0, length, gc_window):
            n = min(gc_window, length - w)
            g[w:w + n] = random_seq_fast(rng, n, float(rng.uniform(0.38, 0.45)))
        return g

    t, q = base(), base()
    lens = np.exp(rng.uniform(np.log(5_000), np.log(200_000), n_blocks)).astype(np.int64)
    blocks = []
    for k in range(n_blocks):
        ln = int(lens[k])
        ts = int(rng.integers(0, length - ln))
        qs = int(min(max(ts + rng.normal(0, 4e6), 0), length - ln))
        seg = mutate_fast(rng, t[ts:ts + ln], float(rng.uniform(0.75, 0.95)), 0.01)[:ln]
        inv = bool(rng.random() < 0.1)
        if inv:
            seg = revcomp(seg)
        q[qs:qs + seg.size] = seg
        blocks.append((ts, qs, int(seg.size), inv))
    fams = [random_seq_fast(rng, int(rng.integers(300, 6001)), 0.42) for _ in range(n_families)]
    for g in (t, q):
        covered = 0
        while covered < repeat_frac * length:
            period, ln = int(rng.integers(1, 7)), int(rng.integers(100, 1001))
            at = int(rng.integers(0, length - ln))
            g[at:at + ln] = np.resize(ALPHA[rng.integers(0, 4, period)], ln)
            covered += ln
        for fam in fams:
            for _ in range(copies):
                cp = mutate_fast(rng, fam, float(rng.uniform(0.80, 0.95)), 0.01)
                if rng.random() < 0.5:
                    cp = revcomp(cp)
                at = int(rng.integers(0, length - cp.size))
                g[at:at + cp.size] = cp
    return t.tobytes(), q.tobytes(), blocks
