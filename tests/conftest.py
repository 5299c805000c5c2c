import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI library)")


@pytest.fixture(scope="session")
def golden():
    def load(name):
        return np.load(os.path.join(GOLDEN, name), allow_pickle=False)

    return load


CODE = np.frombuffer(b"ACGT", np.uint8)


def unpack2(packed: np.ndarray, n: int) -> bytes:
    """Inverse of tests/golden/make_golden.py:pack2."""
    p = np.asarray(packed, np.uint8)
    c = np.stack([(p >> s) & 3 for s in (0, 2, 4, 6)], 1).reshape(-1)[:n]
    return CODE[c].tobytes()


def lcg_noise(n: int, seed: int) -> np.ndarray:
    out = np.empty(n, np.float32)
    x = seed & 0xFFFFFFFFFFFFFFFF
    for i in range(n):
        x = (x * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        out[i] = np.float32((x >> 40) / float(1 << 24) - 0.5)
    return out


def chunk_records(records, size, overlap):
    """ChunkManager::ChunkItSelect (analysis/SeqChunk.cc:72-165) on uppercased records:
    returns list of (record_id, start, bytes) with empty / all-N chunks cleared."""
    out = []
    for rid, s in enumerate(records):
        L = len(s)
        if L < 6:
            continue
        step = size - overlap
        for j in range(1 + L // step):
            start = j * step
            end = min((j + 1) * step + overlap, L)
            seg = s[start:end]
            nn = seg.count(b"N") + seg.count(b"X")
            if nn >= len(seg):
                seg = b""
            out.append((rid, start, seg))
    return out


def dense_run_pair(a=40, b=12, seed=3, length=4096):
    """A chunk repeating (a random bases, one fixed b-base block): on the diagonals at
    multiples of a + b the window crosses the threshold about once per period, so one
    16-lag epoch of the scan holds more runs (1,081) than the run queue (1,024 entries)."""
    rng = np.random.default_rng(seed)
    blk = rng.choice(list(b"ACGT"), b).astype(np.uint8)
    units = []
    for _ in range(length // (a + b) + 1):
        units.append(rng.choice(list(b"ACGT"), a).astype(np.uint8))
        units.append(blk)
    return np.concatenate(units)[:length].tobytes()


def n_pairs(g):
    """Chunk pairs of A/C/G/T/N from the golden uniform pairs (the homology planted in pair 0
    and 2 included): one N, N runs on one or both sides, N at the chunk ends, N inside the
    homologous region (runs that contain an N), an N every 50 bases; the last pair stays pure."""
    out = []
    for r in range(7):
        t = bytearray(unpack2(g["t"][r], 4096))
        q = bytearray(unpack2(g["q"][r], 4096))
        if r == 0:
            q[1234] = ord("N")
        elif r == 1:
            t[1000:1100] = b"N" * 100
        elif r == 2:
            t[2000:2300] = b"N" * 300
            q[2100:2150] = b"N" * 50
            for i in range(0, 4096, 97):
                t[i] = ord("N")
        elif r == 3:
            t[0:7] = b"N" * 7
            q[4090:4096] = b"N" * 6
        elif r == 4:
            for i in range(0, 4096, 50):
                q[i] = ord("N")
        elif r == 5:
            for i in range(0, 4096, 61):
                t[i] = ord("N")
                q[(i + 17) % 4096] = ord("N")
        out.append((bytes(t), bytes(q)))
    return out


def threshold_size(t, q, st, sq, length, min_prob=0.99):
    """The largest target size T (a double) at which the oracle (glibc erf/exp) still keeps
    the run: prob(T) >= min_prob > prob(nextafter(T))."""
    import oracle

    lo, hi = 1.0, 1.0
    while oracle.match_probability(t, q, st, sq, length, hi) >= min_prob:
        lo, hi = hi, hi * 2.0
        if hi > 1e300:  # kept at every size (prob = 1)
            return None
    if oracle.match_probability(t, q, st, sq, length, lo) < min_prob:
        return None
    for _ in range(200):  # bisection on doubles
        mid = 0.5 * (lo + hi)
        if mid in (lo, hi):
            break
        if oracle.match_probability(t, q, st, sq, length, mid) >= min_prob:
            lo = mid
        else:
            hi = mid
    return lo
