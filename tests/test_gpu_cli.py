"""The standalone aligner built against libsxc_hip.so, on the GPU: its match files must
equal the reference CLI's (tools/analysis/HomologyByXCorr.cc) byte for byte, including
the reused-CCSignal tails of short chunks, which the library encodes on the device."""
import subprocess

import numpy as np
import pytest

import hostbuild
from satsuma2_amd import synth

pytestmark = pytest.mark.gpu

CLI_RUNS = ["default", "small_chunks", "same_only", "blocks"]


def _run(qf, tf, out, flags):
    exe = hostbuild.product_program("HomologyByXCorr")
    p = subprocess.run([exe, "-q", qf, "-t", tf, "-o", out, *flags], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout + p.stderr
    return open(out, "rb").read()


def _mm_matches(b: bytes) -> np.ndarray:
    """Match records of a MultiMatches v3 file (analysis/SequenceMatch.cc:320-360)."""
    o = 4
    for _ in range(2):
        n = int(np.frombuffer(b, "<i4", 1, o)[0])
        o += 4
        for _ in range(n):
            o += 8 + int(np.frombuffer(b, "<i8", 1, o)[0])
    n = int(np.frombuffer(b, "<i4", 1, o)[0])
    dt = np.dtype([("ints", "<i4", (7,)), ("matches", "<f8"), ("prob", "<f8"), ("ident", "<f8")])
    return np.frombuffer(b, dt, n, o + 4)


def _compare(got: bytes, exp: bytes):
    if got == exp:
        return
    # Only prob may differ (device erf/exp, relative 1e-11); every other byte must match.
    a, b = _mm_matches(got), _mm_matches(exp)
    assert len(got) == len(exp) and len(a) == len(b)
    assert np.array_equal(a["ints"], b["ints"]) and np.array_equal(a["ident"], b["ident"])
    assert np.array_equal(a["matches"], b["matches"])
    rel = np.abs(a["prob"] - b["prob"]) / np.maximum(np.abs(b["prob"]), 1e-300)
    assert rel.max() <= 1e-11


@pytest.mark.parametrize("run", CLI_RUNS)
def test_gpu_cli_records_match_reference(golden, tmp_path, run):
    g = golden("cli_records.npz")
    qf, tf = tmp_path / "q.fa", tmp_path / "t.fa"
    qf.write_bytes(g["q_fasta"].tobytes())
    tf.write_bytes(g["t_fasta"].tobytes())
    got = _run(str(qf), str(tf), str(tmp_path / "x.out"), str(g[f"flags_{run}"]).split())
    _compare(got, g[f"out_{run}"].tobytes())


def test_gpu_cli_samples_slice(golden, tmp_path):
    g = golden("samples_slice.npz")
    qf, tf = str(tmp_path / "q.fa"), str(tmp_path / "t.fa")
    synth.write_fasta(qf, [g["q"].tobytes()], ["dogX"])
    synth.write_fasta(tf, [g["t"].tobytes()], ["humanX"])
    got = _run(qf, tf, str(tmp_path / "x.out"), ["-min_prob", "0.99999"])
    _compare(got, g["match_file"].tobytes())
