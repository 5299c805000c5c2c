"""The HIP kernel bodies run on the CPU workgroup emulator (tests/emu/sxc_emu.cpp) against
the oracle.  The emulator executes the exact code of satsuma2_amd/csrc/sxc_*.h with a
sequential executor; the AddressSanitizer build bounds-checks every LDS and global
index the kernels compute.  This is how kernel indexing is checked before a GPU runs it.
"""
import ctypes as C
import os
import struct
import subprocess

import numpy as np
import pytest

import hostbuild
import oracle
from conftest import dense_run_pair, lcg_noise, n_pairs, threshold_size, unpack2

COLS = ["query_id", "target_id", "query_size", "qstart", "tstart", "len", "reverse", "prob", "ident"]


@pytest.fixture(scope="module")
def emu_lib():
    return C.CDLL(hostbuild.emu_library())


@pytest.fixture(scope="module")
def emu_check():
    return hostbuild.emu_checker()


class Pair(C.Structure):
    _fields_ = [("tseq", C.c_char_p), ("tlen", C.c_int32), ("qseq", C.c_char_p), ("qlen", C.c_int32),
                ("t_chunk_start", C.c_int32), ("q_chunk_start", C.c_int32), ("t_record", C.c_int32),
                ("q_record", C.c_int32), ("q_record_len", C.c_int32), ("fast", C.c_int32)]


def emu_batch(lib, pairs, N=8192, strands=2, target_total=4096.0 * 64, want_xc=False):
    arr = (Pair * len(pairs))(*[Pair(p["tseq"], len(p["tseq"]), p["qseq"], len(p["qseq"]), 0, 0, p.get("t_record", 0),
                                     p.get("q_record", 0), len(p["qseq"]), p.get("fast", 0)) for p in pairs])
    prm = oracle.make_params(N=N, strands=strands, target_total=target_total)
    out = np.zeros(1 << 18, oracle.RESULT_DTYPE)
    n = C.c_int64(0)
    total = len(pairs) * strands
    xc = np.zeros((total, N), np.float32) if want_xc else None
    st = np.zeros((total, 2), np.int32)
    rc = lib.emu_align_batch(arr, len(pairs), C.byref(prm), out.ctypes.data_as(C.c_void_p), C.c_int64(len(out)),
                             C.byref(n), xc.ctypes.data_as(C.c_void_p) if want_xc else None,
                             st.ctypes.data_as(C.c_void_p))
    assert rc == 0
    return out[: n.value].copy(), xc, st


@pytest.mark.parametrize("nb", range(9, 16))
def test_emulated_fft(emu_lib, nb):
    N = 1 << nb
    x = lcg_noise(N, 3 * N + 7)
    out = np.zeros(N, np.float32)
    assert emu_lib.emu_fft_forward(x.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), N) == 0
    assert np.array_equal(out.view(np.uint32), oracle.fft_forward(x).view(np.uint32))


def test_emulated_batch_uniform(emu_lib, golden):
    g = golden("uniform_pairs.npz")
    P = 6
    pairs = [dict(tseq=unpack2(g["t"][r], 4096), qseq=unpack2(g["q"][r], 4096), t_record=r, q_record=r,
                  fast=int(g["fast"][r])) for r in range(P)]
    got, xc, st = emu_batch(emu_lib, pairs, want_xc=True)
    ref = np.concatenate(oracle.align_batch(
        [dict(p, q_record_len=4096) for p in pairs], oracle.make_params(target_total=4096.0 * 64)))
    for c in COLS:
        assert np.array_equal(got[c], ref[c]), c
    for ps in range(2 * P):
        p = pairs[ps // 2]
        qs = p["qseq"] if ps % 2 == 0 else oracle.reverse_complement(p["qseq"])
        oxc = oracle.xcorr(oracle.encode(p["tseq"], 8192), oracle.encode(qs, 8192))
        assert np.array_equal(xc[ps].view(np.uint32), oxc.view(np.uint32))
        assert st[ps, 0] == len(oracle.findtop(oxc, 2.9 if p["fast"] else 1.8))


def test_emulated_n_pairs(emu_lib, golden):
    """Pairs of A/C/G/T/N chunks take the bitmap scan with N masks (a position with an N
    scores 25): records and run counts equal the oracle's."""
    pairs = [dict(tseq=t, qseq=q, t_record=r, q_record=r) for r, (t, q) in enumerate(n_pairs(golden("uniform_pairs.npz")))]
    got, _, st = emu_batch(emu_lib, pairs)
    ref = np.concatenate(oracle.align_batch([dict(p, q_record_len=4096) for p in pairs],
                                            oracle.make_params(target_total=4096.0 * 64)))
    assert len(ref) > 0
    for c in COLS:
        assert np.array_equal(got[c], ref[c]), c
    for ps in range(2 * len(pairs)):
        p = pairs[ps // 2]
        qs = p["qseq"] if ps % 2 == 0 else oracle.reverse_complement(p["qseq"])
        oxc = oracle.xcorr(oracle.encode(p["tseq"], 8192), oracle.encode(qs, 8192))
        assert st[ps, 1] == len(oracle.matchup(qs, p["tseq"], oxc, 1.8)), ps


def test_emulated_dense_runs(emu_lib):
    """More runs in one epoch than the queue holds: the pure kernel extracts them in rounds
    that each fit the queue.  Records and run counts equal the oracle's."""
    t = dense_run_pair()
    pairs = [dict(tseq=t, qseq=t, t_record=0, q_record=0)]
    got, _, st = emu_batch(emu_lib, pairs)
    ref = oracle.align_batch([dict(pairs[0], q_record_len=4096)], oracle.make_params(target_total=4096.0 * 64))[0]
    assert len(ref) > 0
    for c in COLS:
        assert np.array_equal(got[c], ref[c]), c
    for strand in range(2):
        qs = t if strand == 0 else oracle.reverse_complement(t)
        oxc = oracle.xcorr(oracle.encode(t, 8192), oracle.encode(qs, 8192))
        assert st[strand, 1] == len(oracle.matchup(qs, t, oxc, 1.8))
    assert st[0, 1] > 1024


def test_emulated_resident_query_spectra(emu_lib, golden, monkeypatch):
    """The pair kernel reading both strands' query spectra precomputed by the spectra kernel
    (the serving loop's resident query spectra) gives the oracle's records and xc bits."""
    g = golden("uniform_pairs.npz")
    pairs = [dict(tseq=unpack2(g["t"][r], 4096), qseq=unpack2(g["q"][r], 4096), t_record=r, q_record=r,
                  fast=int(g["fast"][r])) for r in range(4)]
    monkeypatch.setenv("SXC_EMU_QSPEC", "1")
    got, xc, _ = emu_batch(emu_lib, pairs, want_xc=True)
    ref = np.concatenate(oracle.align_batch(
        [dict(p, q_record_len=4096) for p in pairs], oracle.make_params(target_total=4096.0 * 64)))
    for c in COLS:
        assert np.array_equal(got[c], ref[c]), c
    for ps in range(8):
        p = pairs[ps // 2]
        qs = p["qseq"] if ps % 2 == 0 else oracle.reverse_complement(p["qseq"])
        oxc = oracle.xcorr(oracle.encode(p["tseq"], 8192), oracle.encode(qs, 8192))
        assert np.array_equal(xc[ps].view(np.uint32), oxc.view(np.uint32))


@pytest.mark.parametrize("split", [2, 3])
def test_emulated_split_scan(emu_lib, golden, monkeypatch, split):
    """A (pair, strand)'s lags shared by several scan workgroups (the serving loop's small
    launches): the same records, candidate and run counts as one workgroup."""
    g = golden("uniform_pairs.npz")
    pairs = [dict(tseq=unpack2(g["t"][r], 4096), qseq=unpack2(g["q"][r], 4096), t_record=r, q_record=r,
                  fast=int(g["fast"][r])) for r in range(4)]
    one, _, st1 = emu_batch(emu_lib, pairs)
    monkeypatch.setenv("SXC_EMU_SPLIT", str(split))
    got, _, st = emu_batch(emu_lib, pairs)
    assert len(one) > 0 and got.tobytes() == one.tobytes()
    assert np.array_equal(st, st1)


# The alternative FFT geometries the library compiles for its autotuner (stage-A width,
# levels per LDS round), by log2 length; each must reproduce the oracle bit for bit.
# (log2 N, stage-A width, levels per LDS round, lean LDS plan); the register-budget variants
# run the same code as their plain counterparts on the CPU.
GEOMETRIES = ([(nb, 16, 2, 0) for nb in range(9, 15)] + [(15, 32, 2, 0)] + [(nb, 32, 3, 0) for nb in range(11, 15)]
              + [(12, 32, 3, 1), (12, 16, 2, 1), (13, 32, 3, 1), (14, 32, 3, 1), (11, 32, 3, 1), (11, 16, 2, 1),
                 (13, 32, 2, 1)])


@pytest.mark.parametrize("nb,sa,km,lean", GEOMETRIES)
def test_emulated_geometries(emu_lib, nb, sa, km, lean, monkeypatch):
    N = 1 << nb
    rng = np.random.default_rng(nb * 100 + sa + km + 7 * lean)
    pairs = []
    for r in range(2):
        t = rng.choice(np.frombuffer(b"ACGT", np.uint8), N // 2)
        q = rng.choice(np.frombuffer(b"ACGT", np.uint8), N // 2)
        ln = N // 8
        q[N // 5: N // 5 + ln] = t[N // 7: N // 7 + ln]  # a homology
        pairs.append(dict(tseq=t.tobytes(), qseq=q.tobytes(), t_record=r, q_record=r))
    monkeypatch.setenv("SXC_EMU_GEO", f"{sa},{km},{lean}")
    got, xc, _ = emu_batch(emu_lib, pairs, N=N, target_total=1e6, want_xc=True)
    ref = np.concatenate(oracle.align_batch([dict(p, q_record_len=N // 2) for p in pairs],
                                            oracle.make_params(N=N, target_total=1e6)))
    assert len(ref) > 0
    for c in COLS:
        assert np.array_equal(got[c], ref[c]), c
    for ps in range(4):
        p = pairs[ps // 2]
        qs = p["qseq"] if ps % 2 == 0 else oracle.reverse_complement(p["qseq"])
        oxc = oracle.xcorr(oracle.encode(p["tseq"], N), oracle.encode(qs, N))
        assert np.array_equal(xc[ps].view(np.uint32), oxc.view(np.uint32)), ps


def _write_input(path, pairs, N, target_total=1e6):
    with open(path, "wb") as f:
        f.write(struct.pack("<3i", N, 2, len(pairs)))
        f.write(struct.pack("<4d", 1.8, 2.9, 0.99, target_total))
        for t, q, fast in pairs:
            f.write(struct.pack("<8i", len(t), len(q), 0, 0, 0, 0, len(q), fast))
            f.write(t)
            f.write(q)


def _run_check(exe, tmp_path, pairs, N):
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_input(inp, pairs, N)
    r = subprocess.run([exe, str(inp), str(outp)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    b = outp.read_bytes()
    n = int(np.frombuffer(b, np.int64, 1)[0])
    got = np.frombuffer(b, oracle.RESULT_DTYPE, n, 8)
    ref = np.concatenate(oracle.align_batch([dict(tseq=t, qseq=q, fast=f) for t, q, f in pairs],
                                            oracle.make_params(N=N, target_total=1e6)))
    assert len(got) == len(ref)
    for c in COLS:
        assert np.array_equal(got[c], ref[c]), c


def test_asan_edge_pairs(emu_check, golden, tmp_path):
    g = golden("edge_pairs.npz")
    pairs = [(g[f"{n}_t"].tobytes(), g[f"{n}_q"].tobytes(), fast) for n in g["names"] for fast in (0, 1)]
    _run_check(emu_check, tmp_path, pairs, 8192)


def test_asan_uneven_pairs(emu_check, golden, tmp_path):
    g = golden("uneven_pairs.npz")
    for k in range(0, int(g["n"][0]), 3):
        N = int(g[f"N{k}"][0])
        lt, lq = g[f"len{k}"]
        _run_check(emu_check, tmp_path, [(unpack2(g[f"t{k}"], lt), unpack2(g[f"q{k}"], lq), 0)], N)


def test_library_prob_table_golden(emu_lib, golden):
    """The library's host-built ProbTable thresholds equal the reference's."""
    import hashlib

    g = golden("prob_table.npz")
    t = np.zeros((512, 2048), np.float64)
    emu_lib.emu_prob_table(C.c_double(float(g["target_size"][0])), t.ctypes.data_as(C.c_void_p))
    assert hashlib.sha256(t.tobytes()).hexdigest() == str(g["table_sha"])


def test_tail_float_prefilter_is_one_sided(emu_lib):
    """The drain drops a run on a float estimate of the tail's z only where the FP64 z is
    beyond the keep bound too: with the bound just above the exact z (by 1e-12 .. 1e-4,
    relative) the filter never fires on 10^6 random sums, and it does fire with the bound
    below z (so it is not vacuous)."""
    emu_lib.emu_check_z_filter.restype = C.c_int64
    fired = C.c_int64(0)
    bad = emu_lib.emu_check_z_filter(C.c_int64(1_000_000), C.c_uint64(7), C.byref(fired))
    assert bad == 0
    assert fired.value > 900_000


def test_window_flags_bytewise(emu_lib):
    """The pure A/C/G/T scan tests the 46-wide window four positions per byte: over 4000
    diagonals of 256 positions with match densities from 0 to 1 every flag and every carried
    count equals the window counted position by position, and flags are neither all set nor
    all clear."""
    emu_lib.emu_check_window_flags.restype = C.c_int64
    flagged = C.c_int64(0)
    bad = emu_lib.emu_check_window_flags(C.c_int64(4000), C.c_int(8), C.c_uint64(11), C.byref(flagged))
    assert bad == 0
    assert 0.2 * 4000 * 256 < flagged.value < 0.8 * 4000 * 256


def test_window_flags_mk_bytewise(emu_lib):
    """The A/C/G/T/N scan tests the window (100 per identical A/C/G/T base, 25 per position
    with an N, above 1889) four positions per byte: over 4000 diagonals of 256 positions with
    match and N densities from 0 to 1 every flag and every carried value equals the window
    scored position by position, and flags are neither all set nor all clear."""
    emu_lib.emu_check_window_flags_mk.restype = C.c_int64
    flagged = C.c_int64(0)
    bad = emu_lib.emu_check_window_flags_mk(C.c_int64(4000), C.c_int(8), C.c_uint64(13), C.byref(flagged))
    assert bad == 0
    assert 0.1 * 4000 * 256 < flagged.value < 0.9 * 4000 * 256


def test_emulated_bdhv_pairs(emu_lib, golden):
    """Chunks dense in B/D/H/V: the kernels read the window entropies the host computed with
    its log (sxc::window_entropy_row); signals (through the xc bits) and records equal the
    oracle's at every N of the fixture up to 8192."""
    g = golden("bdhv_pairs.npz")
    for name in g["names"]:
        t, q, N = g[f"{name}_t"].tobytes(), g[f"{name}_q"].tobytes(), int(g[f"{name}_N"][0])
        if N > 8192:
            continue
        pairs = [dict(tseq=t, qseq=q, t_record=0, q_record=0)]
        got, xc, _ = emu_batch(emu_lib, pairs, N=N, want_xc=True)
        ref = oracle.align_batch([dict(pairs[0], q_record_len=len(q))],
                                 oracle.make_params(N=N, target_total=4096.0 * 64))[0]
        assert len(ref) > 0
        for c in COLS:
            assert np.array_equal(got[c], ref[c]), (name, c)
        for s, qs in ((0, q), (1, oracle.reverse_complement(q))):
            oxc = oracle.xcorr(oracle.encode(t, N), oracle.encode(qs, N))
            assert np.array_equal(xc[s].view(np.uint32), oxc.view(np.uint32)), (name, s)


def test_window_entropy_rows(emu_lib, golden):
    """The host's window entropies of a B/D/H/V chunk, both strands, turned into the signal by
    the reference's SeqToPCM formula, give the oracle's encoded signal bits."""
    g = golden("bdhv_pairs.npz")
    name = str(g["names"][0])
    seq, N = g[f"{name}_t"].tobytes(), int(g[f"{name}_N"][0])
    row = np.zeros(1024, np.float32)
    emu_lib.emu_window_entropy(seq, len(seq), N, row.ctypes.data_as(C.c_void_p))
    win = N // 512
    codec = {b"A": (1, 0, 0, 0), b"C": (0, 1, 0, 0), b"G": (0, 0, 1, 0), b"T": (0, 0, 0, 1),
             b"K": (0, 0, .5, .5), b"M": (.5, .5, 0, 0), b"R": (.5, 0, .5, 0), b"Y": (0, .5, 0, .5),
             b"S": (0, .5, .5, 0), b"W": (.5, 0, 0, .5), b"B": (0, 1 / 3, 1 / 3, 1 / 3), b"V": (1 / 3, 1 / 3, 1 / 3, 0),
             b"H": (1 / 3, 1 / 3, 0, 1 / 3), b"D": (1 / 3, 0, 1 / 3, 1 / 3), b"N": (.25,) * 4, b"X": (.25,) * 4}
    for strand, s in ((0, seq), (1, oracle.reverse_complement(seq))):
        vals = np.array([codec.get(bytes([c]), (0, 0, 0, 0)) for c in s], np.float64)
        ent = np.repeat(row[strand * 512: strand * 512 + (len(s) + win - 1) // win], win)[: len(s)]
        sig = np.zeros((4, N), np.float32)
        for ch in range(4):
            off = 0.0
            for v in vals[:, ch]:  # the reference's sequential sum
                off += v
            off /= len(s)
            sig[ch, : len(s)] = (ent.astype(np.float64) * (vals[:, ch] - off)).astype(np.float32)
        assert np.array_equal(sig.view(np.uint32), oracle.encode(s, N).view(np.uint32)), strand


def test_emulated_keep_band(emu_lib, golden):
    """Target sizes that put one run's reference probability on either side of min_prob by
    one double: the scan flags the run, the host decides it with its libm, and the records
    equal the oracle's, probabilities bit for bit."""
    emu_lib.emu_last_near.restype = C.c_int64
    g = golden("single_pair.npz")
    t, q = g["t"].tobytes(), g["q"].tobytes()
    m = g["m0"]
    r = next(r for r in m[m["len"] >= 40]
             if (lambda T: T is not None and 1e3 < T < 1e8)(threshold_size(t, q, int(r["st"]), int(r["sq"]),
                                                                           int(r["len"]))))
    T = threshold_size(t, q, int(r["st"]), int(r["sq"]), int(r["len"]))
    for size in (T, np.nextafter(T, np.inf)):
        got, _, _ = emu_batch(emu_lib, [dict(tseq=t, qseq=q)], strands=1, target_total=size)
        assert emu_lib.emu_last_near() >= 1
        ref = oracle.align_batch([dict(tseq=t, qseq=q)], oracle.make_params(strands=1, target_total=size))[0]
        for c in COLS:
            assert np.array_equal(got[c], ref[c]), (size, c)
        sel = (ref["tstart"] == r["st"]) & (ref["qstart"] == r["sq"]) & (ref["len"] == r["len"])
        assert sel.sum() == (1 if size == T else 0)
        assert not got["pad_"].any()
    lo, hi = C.c_double(0), C.c_double(0)
    emu_lib.emu_keep_band(C.c_double(0.99), C.c_double(16777216.0), C.byref(lo), C.byref(hi))
    # (16 T + 64) 2^-52 either side: 6e-8 at the 4096-pair workload's size
    assert lo.value < 0.99 < hi.value and (hi.value - lo.value) / 0.99 < 1.3e-7
