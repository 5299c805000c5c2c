"""The standalone aligner (satsuma2_amd/csrc/host/xcorr_main.cpp), on the CPU through the
workgroup emulator's copy of the C ABI: its MultiMatches v3 match files must equal the
reference CLI's byte for byte (tests/golden/cli_records.npz and samples_slice.npz, written
by tools/analysis/HomologyByXCorr.cc).  The record sets include chunks shorter than the
ones before them, whose signals carry the reused-CCSignal tails."""
import subprocess

import pytest

import hostbuild
from satsuma2_amd import synth

CLI_RUNS = ["default", "small_chunks", "same_only", "blocks"]


def write_cli_inputs(golden, tmp_path):
    g = golden("cli_records.npz")
    qf, tf = tmp_path / "q.fa", tmp_path / "t.fa"
    qf.write_bytes(g["q_fasta"].tobytes())
    tf.write_bytes(g["t_fasta"].tobytes())
    return g, str(qf), str(tf)


def run_cli(exe, qf, tf, out, flags):
    p = subprocess.run([exe, "-q", qf, "-t", tf, "-o", out, *flags], capture_output=True, text=True, timeout=900)
    assert p.returncode == 0, p.stdout + p.stderr
    return open(out, "rb").read()


@pytest.mark.parametrize("run", CLI_RUNS)
def test_cli_records_match_reference(golden, tmp_path, run):
    g, qf, tf = write_cli_inputs(golden, tmp_path)
    flags = str(g[f"flags_{run}"]).split()
    got = run_cli(hostbuild.emu_program("HomologyByXCorr"), qf, tf, str(tmp_path / "x.out"), flags)
    assert got == g[f"out_{run}"].tobytes()


def test_cli_samples_slice_match_file(golden, tmp_path):
    g = golden("samples_slice.npz")
    qf, tf = str(tmp_path / "q.fa"), str(tmp_path / "t.fa")
    synth.write_fasta(qf, [g["q"].tobytes()], ["dogX"])
    synth.write_fasta(tf, [g["t"].tobytes()], ["humanX"])
    exe = hostbuild.emu_program("HomologyByXCorr")
    got = run_cli(exe, qf, tf, str(tmp_path / "x.out"), ["-min_prob", "0.99999"])
    assert got == g["match_file"].tobytes()


def test_cli_asan_blocks(golden, tmp_path):
    """AddressSanitizer build on the block-restricted run (tails over skipped chunks)."""
    g, qf, tf = write_cli_inputs(golden, tmp_path)
    got = run_cli(hostbuild.emu_program("HomologyByXCorr", sanitize=True), qf, tf, str(tmp_path / "x.out"),
                  str(g["flags_blocks"]).split())
    assert got == g["out_blocks"].tobytes()


def test_cli_flags():
    exe = hostbuild.emu_program("HomologyByXCorr")
    p = subprocess.run([exe, "-q", "a", "-t", "b"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 255 and "need to specify -o" in p.stdout
    for extra in (["-chain", "true"], ["-guide", "g.out"], ["-proteins", "1"], ["-line", "0.5"], ["-pairs", "0:1"]):
        p = subprocess.run([exe, "-q", "a", "-t", "b", "-o", "x", *extra], capture_output=True, text=True, timeout=60)
        assert p.returncode == 1 and "is not supported by this build" in p.stdout, extra
