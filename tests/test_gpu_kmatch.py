"""KMatch on the GPU (include/sxc_kmatch.h, satsuma2_amd/lib/KMatch) against the reference
program compiled from kmatch/KMatch.cc (oracle/_ref/KMatch): the output files, t_result
records in the reference's order, are compared byte for byte.  The cases have soft-masked
and N stretches, repeats whose k-mers exceed max_freq, reverse-complemented segments, point
changes and a FASTA whose last line has no newline (the reference does not read it); they
run with SatsumaSynteny2's arguments (min_length K, jump K - 1, max_freq 1) at K = 11 .. 31
and with other lengths, jumps and max_freq 2 and 3.  tests/golden/kmatch_cases.npz holds a
fixed set of cases and the reference's outputs for machines without the program."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from kmatch_cases import fasta_text, genome_pair

pytestmark = pytest.mark.gpu

MINE = os.path.join(ROOT, "satsuma2_amd", "lib", "KMatch")
REF = os.path.join(ROOT, "oracle", "_ref", "KMatch")
GOLDEN = os.path.join(ROOT, "tests", "golden", "kmatch_cases.npz")


@pytest.fixture(scope="module")
def ctx():
    from satsuma2_amd.api import Context

    c = Context(0)
    yield c
    c.close()


def run(prog, q, t, K, out, min_len, jump, freq):
    subprocess.run([prog, str(q), str(t), str(K), str(out), str(min_len), str(jump), str(freq)], check=True,
                   stdout=subprocess.DEVNULL, timeout=300)
    with open(out, "rb") as f:
        return f.read()


FIELDS = ["query_id", "target_id", "query_size", "qstart", "tstart", "len", "reverse", "prob", "ident"]


def records(b: bytes) -> np.ndarray:
    """The t_result fields; the 7 padding bytes after `reverse` are left out (the reference
    writes whatever its stack held there)."""
    from satsuma2_amd._lib import RESULT_DTYPE
    assert len(b) % RESULT_DTYPE.itemsize == 0
    return np.frombuffer(b, RESULT_DTYPE)[FIELDS].copy()


def same_records(a: bytes, b: bytes, freq: int = 1) -> bool:
    """Equal records in equal order.  With max_freq > 1 a query position can join several
    target positions; the reference leaves their order to std::sort (equal keys), so there
    the records must be the same set and the sequence of query starts the same, while
    blocks sharing a query start may come in another order."""
    ra, rb = records(a), records(b)
    if len(ra) != len(rb):
        return False
    if freq <= 1:
        return all(np.array_equal(ra[f], rb[f]) for f in FIELDS)
    key = ["query_id", "qstart"]
    return (all(np.array_equal(ra[f], rb[f]) for f in key)
            and np.array_equal(np.sort(ra, order=FIELDS), np.sort(rb, order=FIELDS)))


def files(tmp_path, seed, final_newline=True):
    t, q = genome_pair(seed)
    tf, qf = tmp_path / f"t{seed}.fa", tmp_path / f"q{seed}.fa"
    tf.write_bytes(fasta_text(t, 60, final_newline, "t"))
    qf.write_bytes(fasta_text(q, 80, True, "q"))
    return qf, tf


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref/KMatch not built")
@pytest.mark.parametrize("K", [11, 15, 21, 31])
def test_kmatch_satsuma_args_match_reference(tmp_path, K):
    qf, tf = files(tmp_path, 1)
    print(f"[kmatch] K={K} seed=1", flush=True)
    mine = run(MINE, qf, tf, K, tmp_path / "mine.bin", K, K - 1, 1)
    ref = run(REF, qf, tf, K, tmp_path / "ref.bin", K, K - 1, 1)
    assert len(ref) > 72 * 10
    assert same_records(mine, ref)


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref/KMatch not built")
@pytest.mark.parametrize("K,min_len,jump,freq", [(13, 40, 5, 1), (17, 17, 0, 1), (11, 11, 10, 2), (15, 30, 14, 3)])
def test_kmatch_other_args_match_reference(tmp_path, K, min_len, jump, freq):
    qf, tf = files(tmp_path, 2, final_newline=False)
    print(f"[kmatch] K={K} min={min_len} jump={jump} freq={freq} seed=2", flush=True)
    mine = run(MINE, qf, tf, K, tmp_path / "mine.bin", min_len, jump, freq)
    ref = run(REF, qf, tf, K, tmp_path / "ref.bin", min_len, jump, freq)
    assert len(ref) > 72 * 10
    assert same_records(mine, ref, freq)


@pytest.mark.skipif(not os.path.exists(GOLDEN), reason="fixture not generated")
def test_kmatch_fixtures(tmp_path):
    g = np.load(GOLDEN)
    qf, tf = tmp_path / "q.fa", tmp_path / "t.fa"
    qf.write_bytes(g["query_fasta"].tobytes())
    tf.write_bytes(g["target_fasta"].tobytes())
    for i, (K, min_len, jump, freq) in enumerate(g["args"].tolist()):
        print(f"[kmatch] fixture K={K}", flush=True)
        mine = run(MINE, qf, tf, K, tmp_path / "mine.bin", min_len, jump, freq)
        assert same_records(mine, g[f"out{i}"].tobytes(), freq), (K, min_len, jump, freq)


def test_kmatch_records_api_equals_files(tmp_path, ctx):
    """The in-memory entry point gives the file path's records for the same sequences."""
    from satsuma2_amd.kmatch import kmatch
    t, q = genome_pair(3)
    qf, tf = tmp_path / "q.fa", tmp_path / "t.fa"
    qf.write_bytes(fasta_text([s.upper() for s in q], 80, True, "q"))
    tf.write_bytes(fasta_text([s.upper() for s in t], 80, True, "t"))
    a, sa = kmatch(ctx, str(qf), str(tf), 21, 21, 20, 1)
    b, sb = kmatch(ctx, [s.upper() for s in q], [s.upper() for s in t], 21, 21, 20, 1)
    assert len(a) > 10 and a.tobytes() == b.tobytes()
    assert sa["joined"] == sb["joined"] > 0
