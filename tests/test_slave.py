"""The slave executable (satsuma2_amd/csrc/host/slave_main.cpp): flags, FASTA loading,
chunking and the master exchange, on the CPU.

The program is linked against the workgroup emulator's copy of the C ABI (tests/hostbuild.py),
so everything it does — down to the kernels' own code — runs here without a GPU; the GPU
build of the same program is exercised in test_gpu_slave.py.  Expected records come from
the reference slave (tests/golden/samples_slice.npz, made by the reference's own
align_target over the same two blocks).
"""
import os
import subprocess

import numpy as np
import pytest

import hostbuild
import oracle
from conftest import ROOT, chunk_records
from master_stub import MasterStub
from satsuma2_amd import synth

REF_DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")


def write_inputs(golden, tmp_path):
    g = golden("samples_slice.npz")
    qf, tf = str(tmp_path / "q.fa"), str(tmp_path / "t.fa")
    synth.write_fasta(qf, [g["q"].tobytes()], ["dogX"])
    synth.write_fasta(tf, [g["t"].tobytes()], ["humanX"])
    return g, qf, tf


def run_slave(exe, qf, tf, blocks, expect, per_exchange=2, threads=1, extra=(), timeout=600, batch=False):
    """batch=True: the emulated library refuses the serving loop, so the slave takes its
    batch fallback."""
    m = MasterStub(blocks, expect=expect, per_exchange=per_exchange, deadline=timeout).start()
    env = dict(os.environ, SXC_EMU_NO_SERVER="1" if batch else "0")
    p = subprocess.run([exe, "-master", "127.0.0.1", "-port", str(m.port), "-sid", "3", "-p", str(threads),
                        "-q", qf, "-t", tf, *extra], capture_output=True, text=True, timeout=timeout, env=env)
    m.done.wait(10)
    assert m.error is None, m.error
    assert p.returncode == 0, p.stdout + p.stderr
    assert set(m.slave_ids) == {3}
    return m.results(), p


@pytest.mark.parametrize("per_exchange,batch", [(2, False), (1, False), (2, True), (1, True)])
def test_slave_samples_matches_reference_slave(golden, tmp_path, per_exchange, batch):
    g, qf, tf = write_inputs(golden, tmp_path)
    exe = hostbuild.emu_program("HomologyByXCorrSlave")
    got, p = run_slave(exe, qf, tf, [tuple(int(x) for x in r) for r in g["pairs"]], len(g["results"]),
                       per_exchange=per_exchange, batch=batch)
    assert ("aligning queued blocks in batches" in p.stdout) == batch
    exp = g["results"]
    assert len(got) == len(exp)
    assert got.tobytes() == exp.tobytes()


def test_slave_block_split_and_bad_block(golden, tmp_path):
    """The two blocks split into four, one per exchange, plus a block naming chunks that do
    not exist (skipped with a message, as no reference chunk exists for it)."""
    g, qf, tf = write_inputs(golden, tmp_path)
    exe = hostbuild.emu_program("HomologyByXCorrSlave")
    blocks = [(0, 6, 0, 1, 0), (0, 6, 2, 4, 0), (7, 13, 5, 7, 1), (7, 13, 8, 9, 1), (50, 60, 0, 1, 0)]
    for batch in (False, True):
        got, p = run_slave(exe, qf, tf, blocks, len(g["results"]), per_exchange=1, batch=batch)
        assert got.tobytes() == g["results"].tobytes()
        assert "outside the 14 target / 10 query chunks" in p.stdout


def test_slave_asan(golden, tmp_path):
    """AddressSanitizer build of the slave and the emulated kernels on a few blocks (the
    last chunks of each record, 64 and 3136 bytes long), against the oracle."""
    g, qf, tf = write_inputs(golden, tmp_path)
    t, q = g["t"].tobytes().upper(), g["q"].tobytes().upper()
    tch, qch = chunk_records([t], 4096, 1024), chunk_records([q], 4096, 0)
    blocks = [(12, 13, 8, 9, 0), (11, 13, 9, 9, 1)]
    pairs = [dict(tseq=tch[ti][2], qseq=qch[qi][2], t_chunk_start=tch[ti][1], q_chunk_start=qch[qi][1],
                  q_record_len=len(q), fast=fast)
             for tf_, tt, qf_, qt, fast in blocks for qi in range(qf_, qt + 1) for ti in range(tf_, tt + 1)]
    exp = np.concatenate(oracle.align_batch(pairs, oracle.make_params(target_total=float(len(t)))))
    got, _ = run_slave(hostbuild.emu_program("HomologyByXCorrSlave", sanitize=True), qf, tf, blocks, len(exp),
                       per_exchange=1)
    assert got.tobytes() == exp.tobytes()


def test_slave_min_len_flag(golden, tmp_path):
    g, qf, tf = write_inputs(golden, tmp_path)
    exe = hostbuild.emu_program("HomologyByXCorrSlave")
    exp = g["results"][g["results"]["len"] >= 200]
    got, _ = run_slave(exe, qf, tf, [tuple(int(x) for x in r) for r in g["pairs"]], len(exp), extra=("-l", "200"))
    assert got.tobytes() == exp.tobytes()


def test_slave_flags_like_reference():
    exe = hostbuild.emu_program("HomologyByXCorrSlave")
    p = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0 and "Available arguments" in p.stdout and "-cutoff_fast<double>" in p.stdout
    p = subprocess.run([exe, "-h"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0 and "-min_prob<double>" in p.stdout
    p = subprocess.run([exe, "-bogus", "1"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0 and "Invalid command-line arg: -bogus" in p.stdout
    p = subprocess.run([exe, "-master", "x", "-q", "a", "-t", "b"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 255 and "need to specify -sid" in p.stdout
    p = subprocess.run([exe, "-master", "x", "-sid", "1", "-q", "/nonexistent.fa", "-t", "b"], capture_output=True,
                       text=True, timeout=60)
    assert p.returncode == 1 and "cannot open /nonexistent.fa" in p.stdout


TRICKY_A = (b"ACGTNNxx\n>rec one\tdesc\nacgtACGTNNNN  ignored tokens\nGGCC\n\n>short\nACG\n>alln\n"
            b"NNNNNNNNNNXXXX\n>crlf\r\nACGTAC\r\nGG\n>empty\n>last\nTTTTTTTTTTAAAAAAAAAACCCCCCCCCCGG")
TRICKY_B = b">second file\nacgtacgtacgt\n"


def _lines(out):
    return [line for line in out.split("\n") if line[:2] in ("R ", "C ", "T ")]


def _stdout(cmd):
    """stdout as text with '\r' kept (text=True would translate it)."""
    return subprocess.run(cmd, capture_output=True, check=True).stdout.decode("latin-1")


def test_fasta_reader_and_chunker_known_answers(tmp_path):
    """Reference behaviour spelled out: bytes before the first header lead the first record,
    only a line's first token counts, names join tokens with '_', lowercase is uppercased,
    records under 6 bytes get no chunk, all-N/X chunks are cleared, files are appended."""
    a, b = tmp_path / "a.fa", tmp_path / "b.fa"
    a.write_bytes(TRICKY_A)
    b.write_bytes(TRICKY_B)
    out = _stdout([hostbuild.seqio_check(), f"{a},{b}", "8", "2"])
    lines = _lines(out)
    recs = [line.split(" ", 3) for line in lines if line.startswith("R ")]
    assert [(r[2], r[3]) for r in recs] == [("24", "rec_one_desc"), ("3", "short"), ("14", "alln"), ("9", "crlf\r"),
                                           ("0", "empty"), ("32", "last"), ("12", "second_file")]
    chunks = [line.split(" ") for line in lines if line.startswith("C ")]
    first = bytes.fromhex(chunks[0][4])
    assert first == b"ACGTNNXX"
    # record 1 ("short", 3 bytes) has no chunk; record 2 (all N/X) has three cleared chunks
    assert [c[1] for c in chunks].count("1") == 0
    assert [(c[1], c[3]) for c in chunks if c[1] == "2"] == [("2", "0")] * 3
    # step 6 = size 8 - overlap 2; chunk count 1 + len // step; end clipped
    assert [(int(c[2]), int(c[3])) for c in chunks if c[1] == "5"] == [(0, 8), (6, 8), (12, 8), (18, 8), (24, 8),
                                                                      (30, 2)]
    assert lines[-1] == "T 94.0"


@pytest.mark.skipif(not os.path.exists(REF_DRIVER), reason="reference driver not built (needs the reference sources)")
def test_fasta_reader_and_chunker_match_reference(tmp_path, golden):
    a, b = tmp_path / "a.fa", tmp_path / "b.fa"
    a.write_bytes(TRICKY_A)
    b.write_bytes(TRICKY_B)
    g, qf, tf = write_inputs(golden, tmp_path)
    for files, size, ov in ((f"{a},{b}", 8, 2), (f"{a}", 7, 0), (qf, 4096, 0), (tf, 4096, 1024)):
        ours = _stdout([hostbuild.seqio_check(), files, str(size), str(ov)])
        ref = _stdout([REF_DRIVER, "chunks", files, str(size), str(ov)])
        assert _lines(ours) == _lines(ref), (files, size, ov)


@pytest.mark.parametrize("cut", ["0.9999", "0.9"])
def test_slave_prob_table_matches_reference(golden, tmp_path, cut):
    """-prob_table (analysis/HomologyByXCorrSlave.cc:184-187, 413-421): records equal the
    reference slave's with the same flag (digest; 171,360 records at -min_prob 0.9999, none at
    0.9, which the slave's fixed keep threshold of 0.99 rejects)."""
    import hashlib

    g, qf, tf = write_inputs(golden, tmp_path)
    p = golden("prob_table.npz")
    name = "hi" if cut == "0.9999" else "lo"
    n = int(p[f"results_{name}_n"][0])
    got, _ = run_slave(hostbuild.emu_program("HomologyByXCorrSlave"), qf, tf,
                       [tuple(int(x) for x in r) for r in g["pairs"]], n,
                       extra=("-prob_table", "true", "-min_prob", cut))
    assert len(got) == n
    assert hashlib.sha256(np.ascontiguousarray(got).tobytes()).hexdigest() == str(p[f"results_{name}_sha"])
