"""FASTA pairs for the KMatch tests (test infrastructure): a target of random records and a
query built from its segments, some reverse-complemented, with point changes, soft-masked
(lowercase) and N stretches, segments repeated within each genome (k-mers above max_freq)
and line layouts that include a last line without a newline."""
from __future__ import annotations

import numpy as np

COMP = bytes.maketrans(b"ACGTacgtN", b"TGCAtgcaN")


def _rnd(rng, n):
    return np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, n)].tobytes()


def _mutate(rng, s: bytes, rate: float) -> bytes:
    a = np.frombuffer(s, np.uint8).copy()
    hit = rng.random(len(a)) < rate
    a[hit] = np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, int(hit.sum()))]
    return a.tobytes()


def _soft(rng, s: bytes) -> bytes:
    a = bytearray(s)
    for _ in range(len(s) // 20000 + 1):
        i = int(rng.integers(0, max(1, len(a) - 600)))
        a[i:i + 500] = a[i:i + 500].lower()
        j = int(rng.integers(0, max(1, len(a) - 200)))
        a[j:j + int(rng.integers(1, 120))] = b"N" * len(a[j:j + int(rng.integers(1, 120))])
    return bytes(a)


def genome_pair(seed: int, n_records: int = 3, size: int = 60_000):
    rng = np.random.default_rng(seed)
    target = []
    for _ in range(n_records):
        s = _rnd(rng, int(rng.integers(size // 2, size)))
        rep = s[1000:1800]
        s = s[:5000] + rep + s[5000:9000] + rep + s[9000:]  # a repeat: its k-mers occur twice
        target.append(_soft(rng, s))
    query = []
    for _ in range(n_records):
        parts = []
        for _ in range(int(rng.integers(3, 7))):
            t = target[int(rng.integers(len(target)))]
            i = int(rng.integers(0, len(t) - 3000))
            seg = t[i:i + int(rng.integers(800, 3000))].upper()
            seg = _mutate(rng, seg, float(rng.choice([0.0, 0.01, 0.05])))
            if rng.random() < 0.4:
                seg = seg.translate(COMP)[::-1]
            parts.append(seg)
            parts.append(_rnd(rng, int(rng.integers(100, 2000))))
        query.append(_soft(rng, b"".join(parts)))
    return target, query


def fasta_text(records, width: int = 80, final_newline: bool = True, prefix: str = "r") -> bytes:
    out = []
    for i, s in enumerate(records):
        out.append(f">{prefix}{i} some description\n".encode())
        w = width if width > 0 else len(s)
        for j in range(0, len(s), w):
            out.append(s[j:j + w] + b"\n")
    text = b"".join(out)
    return text if final_newline else text[:-1]
