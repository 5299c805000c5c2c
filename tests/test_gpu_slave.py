"""The slave executable built against libsxc_hip.so, on the GPU: the master exchange with
the stub master, records compared byte for byte with the reference slave's
(tests/golden/samples_slice.npz) and with the library's own sxc_align_batch on a larger
synthetic genome pair."""
import os
import subprocess

import numpy as np
import pytest

import hostbuild
from master_stub import MasterStub
from satsuma2_amd import synth

pytestmark = pytest.mark.gpu


def _run(exe, qf, tf, blocks, expect, per_exchange=2, threads=1, timeout=300):
    m = MasterStub(blocks, expect=expect, per_exchange=per_exchange, deadline=timeout).start()
    p = subprocess.run([exe, "-master", "127.0.0.1", "-port", str(m.port), "-sid", "1", "-p", str(threads), "-q", qf,
                        "-t", tf], capture_output=True, text=True, timeout=timeout)
    m.done.wait(10)
    assert m.error is None, m.error
    assert p.returncode == 0, p.stdout + p.stderr
    return m.results()


def test_gpu_slave_samples_matches_reference_slave(golden, tmp_path):
    g = golden("samples_slice.npz")
    qf, tf = str(tmp_path / "q.fa"), str(tmp_path / "t.fa")
    synth.write_fasta(qf, [g["q"].tobytes()], ["dogX"])
    synth.write_fasta(tf, [g["t"].tobytes()], ["humanX"])
    exe = hostbuild.product_program("HomologyByXCorrSlave")
    got = _run(exe, qf, tf, [tuple(int(x) for x in r) for r in g["pairs"]], len(g["results"]))
    exp = g["results"]
    assert len(got) == len(exp)
    # Coordinates, lengths, strand, identity: exact.  prob: FP64 erf/exp on the device,
    # relative 1e-11 (1e-14 where prob >= 0.5, which decides what is kept).
    for f in ("query_id", "target_id", "query_size", "qstart", "tstart", "len", "reverse", "ident"):
        assert np.array_equal(got[f], exp[f]), f
    rel = np.abs(got["prob"] - exp["prob"]) / np.maximum(np.abs(exp["prob"]), 1e-300)
    assert rel.max() <= 1e-11


def test_gpu_slave_many_blocks_vs_batch(tmp_path):
    """Two synthetic 400 kbp records per side; 12 blocks handed out 3 per exchange; the
    records must equal one sxc_align_batch over the same chunk pairs, block by block."""
    from satsuma2_amd.api import ChunkPair, Context, HomologyByXCorr, SlaveParams
    rng = np.random.default_rng(11)
    base_a = [synth.random_seq(rng, 400_000, 0.41) for _ in range(2)]
    base = [a.tobytes() for a in base_a]
    qs = [synth.mutate(rng, a, 0.85).tobytes() for a in base_a]
    qf, tf = str(tmp_path / "q.fa"), str(tmp_path / "t.fa")
    synth.write_fasta(qf, qs, ["q0", "q1"])
    synth.write_fasta(tf, base, ["t0", "t1"])
    # chunk tables as the slave builds them (query: 4096 step 4096; target: 4096 step 3072)
    qch = [(r, j * 4096, s[j * 4096:j * 4096 + 4096]) for r, s in enumerate(qs) for j in range(1 + len(s) // 4096)]
    tch = [(r, j * 3072, s[j * 3072:min(j * 3072 + 4096, len(s))]) for r, s in enumerate(base)
           for j in range(1 + len(s) // 3072)]
    blocks = []
    for k in range(12):
        q0 = k * 15
        t0 = (k * 19) % (len(tch) - 20)
        blocks.append((t0, t0 + 19, q0, q0 + 14, k % 3 == 2))
    ctx = Context(0)
    h = HomologyByXCorr(ctx, SlaveParams(target_total=float(sum(len(s) for s in base))))
    exp = []
    for (tf_, tt, qf_, qt, fast) in blocks:
        pairs = [ChunkPair(tch[ti][2], qch[qi][2], tch[ti][1], qch[qi][1], tch[ti][0], qch[qi][0], len(qs[qch[qi][0]]),
                           fast=int(fast)) for qi in range(qf_, qt + 1) for ti in range(tf_, tt + 1)]
        exp.append(h.align_batch(pairs))
    exp = np.concatenate(exp)
    exe = hostbuild.product_program("HomologyByXCorrSlave")
    got = _run(exe, qf, tf, blocks, len(exp), per_exchange=3, threads=2)
    assert len(got) == len(exp) > 0
    assert got.tobytes() == exp.tobytes()


def test_gpu_slave_prob_table(golden, tmp_path):
    """-prob_table: the records equal the reference slave's with the flag (171,360 on the
    samples slice; the table lookup has no floating-point tail, so they are bit-exact)."""
    import hashlib

    g = golden("samples_slice.npz")
    p = golden("prob_table.npz")
    qf, tf = str(tmp_path / "q.fa"), str(tmp_path / "t.fa")
    synth.write_fasta(qf, [g["q"].tobytes()], ["dogX"])
    synth.write_fasta(tf, [g["t"].tobytes()], ["humanX"])
    exe = hostbuild.product_program("HomologyByXCorrSlave")
    n = int(p["results_hi_n"][0])
    m = MasterStub([tuple(int(x) for x in r) for r in g["pairs"]], expect=n, per_exchange=2, deadline=300).start()
    r = subprocess.run([exe, "-master", "127.0.0.1", "-port", str(m.port), "-sid", "1", "-q", qf, "-t", tf,
                        "-prob_table", "true", "-min_prob", "0.9999"], capture_output=True, text=True, timeout=300)
    m.done.wait(10)
    assert m.error is None and r.returncode == 0, r.stdout + r.stderr
    got = m.results()
    assert len(got) == n
    assert hashlib.sha256(np.ascontiguousarray(got).tobytes()).hexdigest() == str(p["results_hi_sha"])
