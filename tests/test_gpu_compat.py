"""The reference-compatible C++ classes (include/satsuma_compat/) on the GPU: a program
written against the reference's own interface (CCSignal, CrossCorrelation, SeqAnalyzer,
GetMatchProbability, PrintMatch -- satsuma2_amd/csrc/host/compat_align.cpp) reproduces the
reference's runs and identities bit for bit on the single-pair fixture, and its
probabilities within the stated tolerance."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

PROB_RTOL = 1e-11
PROB_RTOL_KEEP = 1e-14
EXE = os.path.join(ROOT, "satsuma2_amd", "lib", "compat_align")


def test_compat_classes_match_reference(golden, tmp_path):
    g = golden("single_pair.npz")
    tf, qf = tmp_path / "t.txt", tmp_path / "q.txt"
    tf.write_bytes(g["t"].tobytes())
    qf.write_bytes(g["q"].tobytes())
    p = subprocess.run([EXE, str(tf), str(qf), "8192", "1.8", "4096"], capture_output=True, timeout=120)
    assert p.returncode == 0, p.stdout.decode() + p.stderr.decode()
    rows = [line.split() for line in p.stdout.decode().splitlines()]
    for s in (0, 1):
        got = [r for r in rows if int(r[0]) == s]
        gm = g[f"m{s}"]
        assert len(got) == len(gm) > 0, s
        runs = np.array([[int(r[1]), int(r[2]), int(r[3])] for r in got])
        assert np.array_equal(runs, np.stack([gm["st"], gm["sq"], gm["len"]], 1)), s
        ident = np.array([float.fromhex(r[5]) for r in got])
        assert np.array_equal(ident.view(np.uint64), np.asarray(gm["ident"], np.float64).view(np.uint64)), s
        prob = np.array([float.fromhex(r[4]) for r in got])
        ref = np.asarray(gm["prob"], np.float64)
        rel = np.abs(prob - ref) / np.maximum(np.abs(ref), 1e-300)
        assert rel.max() <= PROB_RTOL
        assert rel[ref >= 0.5].max(initial=0.0) <= PROB_RTOL_KEEP
