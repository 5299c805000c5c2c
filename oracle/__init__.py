"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the GPU path.  Imported by tests/,
``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg, never by the
``satsuma2_amd`` package.  The library is a restatement of the reference's
chunk-pair path (see oracle/sxc_oracle.c for per-function citations).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
REF_DIR = os.path.join(_HERE, "_ref")


class SeqMatch(C.Structure):
    _fields_ = [("start_target", C.c_int32), ("start_query", C.c_int32), ("len", C.c_int32)]


class Result(C.Structure):
    """Field layout of the reference's t_result (analysis/WorkQueue.h:23-33), 72 bytes."""

    _fields_ = [
        ("query_id", C.c_uint64),
        ("target_id", C.c_uint64),
        ("query_size", C.c_uint64),
        ("qstart", C.c_uint64),
        ("tstart", C.c_uint64),
        ("len", C.c_uint64),
        ("reverse", C.c_uint8),
        ("pad_", C.c_uint8 * 7),
        ("prob", C.c_double),
        ("ident", C.c_double),
    ]


class Pair(C.Structure):
    _fields_ = [
        ("tseq", C.c_char_p),
        ("tlen", C.c_int32),
        ("qseq", C.c_char_p),
        ("qlen", C.c_int32),
        ("t_chunk_start", C.c_int32),
        ("q_chunk_start", C.c_int32),
        ("t_record", C.c_int32),
        ("q_record", C.c_int32),
        ("q_record_len", C.c_int32),
        ("fast", C.c_int32),
    ]


class Params(C.Structure):
    _fields_ = [
        ("N", C.c_int32),
        ("q_chunk", C.c_int32),
        ("min_len", C.c_int32),
        ("strands", C.c_int32),
        ("cutoff", C.c_double),
        ("cutoff_fast", C.c_double),
        ("min_prob", C.c_double),
        ("target_total", C.c_double),
        ("ptable", C.c_void_p),
        ("table_prob", C.c_double),
    ]


RESULT_DTYPE = np.dtype(
    [
        ("query_id", "<u8"),
        ("target_id", "<u8"),
        ("query_size", "<u8"),
        ("qstart", "<u8"),
        ("tstart", "<u8"),
        ("len", "<u8"),
        ("reverse", "u1"),
        ("pad_", "u1", (7,)),
        ("prob", "<f8"),
        ("ident", "<f8"),
    ]
)
assert RESULT_DTYPE.itemsize == 72

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        fp = C.POINTER(C.c_float)
        L.sxo_encode.argtypes = [C.c_char_p, C.c_int32, C.c_int32, fp]
        L.sxo_fft_forward.argtypes = [C.c_int32, fp, fp]
        L.sxo_fft_inverse.argtypes = [C.c_int32, fp, fp]
        L.sxo_fft_rescale.argtypes = [C.c_int32, fp]
        L.sxo_fft_rescale.restype = None
        L.sxo_xcorr.argtypes = [fp, fp, C.c_int32, fp]
        L.sxo_findtop.argtypes = [fp, C.c_int32, C.c_double, C.POINTER(C.c_int32), C.c_int32]
        L.sxo_findtop.restype = C.c_int32
        L.sxo_matchup.argtypes = [C.c_char_p, C.c_int32, C.c_char_p, C.c_int32, fp, C.c_int32, C.c_double,
                                  C.POINTER(SeqMatch), C.c_int32]
        L.sxo_matchup.restype = C.c_int32
        L.sxo_match_probability.argtypes = [C.c_char_p, C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_double]
        L.sxo_match_probability.restype = C.c_double
        L.sxo_print_match_ident.argtypes = [C.c_char_p, C.c_char_p, C.c_int32, C.c_int32, C.c_int32]
        L.sxo_print_match_ident.restype = C.c_double
        L.sxo_dna_equal.argtypes = [C.c_int, C.c_int]
        L.sxo_dna_equal.restype = C.c_double
        L.sxo_dna_equal_amb.argtypes = [C.c_int, C.c_int]
        L.sxo_dna_equal_amb.restype = C.c_double
        L.sxo_reverse_complement.argtypes = [C.c_char_p, C.c_int32, C.c_char_p]
        L.sxo_reverse_complement.restype = None
        L.sxo_align_pair.argtypes = [C.POINTER(Pair), C.POINTER(Params), C.POINTER(Result), C.c_int32]
        L.sxo_align_pair.restype = C.c_int32
        L.sxo_align_batch.argtypes = [C.POINTER(Pair), C.c_int32, C.POINTER(Params), C.c_int32,
                                      C.POINTER(Result), C.c_int32, C.POINTER(C.c_int32)]
        L.sxo_align_batch.restype = C.c_int64
        L.sxo_align_batch_tails.argtypes = [C.POINTER(Pair), C.c_int32, C.POINTER(Params), C.c_void_p, C.c_void_p,
                                            C.c_int32, C.POINTER(Result), C.c_int32, C.POINTER(C.c_int32)]
        L.sxo_align_batch_tails.restype = C.c_int64
        L.sxo_prob_table.argtypes = [C.c_double, C.c_int32, C.POINTER(C.c_double)]
        L.sxo_prob_table.restype = None
        _lib = L
    return _lib


def _fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


def encode(seq: bytes, N: int) -> np.ndarray:
    out = np.zeros((4, N), np.float32)
    rc = lib().sxo_encode(seq, len(seq), N, _fptr(out))
    if rc:
        raise ValueError(f"sxo_encode failed ({rc})")
    return out


def fft_forward(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    f = np.zeros_like(x)
    if lib().sxo_fft_forward(x.size, _fptr(x), _fptr(f)):
        raise ValueError("bad FFT length")
    return f


def fft_inverse(f: np.ndarray) -> np.ndarray:
    f = np.ascontiguousarray(f, np.float32)
    x = np.zeros_like(f)
    if lib().sxo_fft_inverse(f.size, _fptr(f), _fptr(x)):
        raise ValueError("bad FFT length")
    return x


def fft_rescale(x: np.ndarray) -> np.ndarray:
    x = np.array(x, np.float32, copy=True)
    lib().sxo_fft_rescale(x.size, _fptr(x))
    return x


def xcorr(tsig: np.ndarray, qsig: np.ndarray) -> np.ndarray:
    tsig = np.ascontiguousarray(tsig, np.float32)
    qsig = np.ascontiguousarray(qsig, np.float32)
    N = tsig.shape[1]
    xc = np.zeros(N, np.float32)
    lib().sxo_xcorr(_fptr(tsig), _fptr(qsig), N, _fptr(xc))
    return xc


def findtop(xc: np.ndarray, cutoff: float) -> np.ndarray:
    xc = np.ascontiguousarray(xc, np.float32)
    top = np.zeros(xc.size, np.int32)
    n = lib().sxo_findtop(_fptr(xc), xc.size, cutoff, top.ctypes.data_as(C.POINTER(C.c_int32)), xc.size)
    return top[:n].copy()


def matchup(q: bytes, t: bytes, xc: np.ndarray, cutoff: float) -> np.ndarray:
    """Returns an (n, 3) int32 array of (start_target, start_query, len)."""
    xc = np.ascontiguousarray(xc, np.float32)
    cap = 1 << 15
    while True:
        buf = (SeqMatch * cap)()
        n = lib().sxo_matchup(q, len(q), t, len(t), _fptr(xc), xc.size, cutoff, buf, cap)
        if n <= cap:
            break
        cap = n
    arr = np.frombuffer(buf, dtype=np.int32, count=3 * n).reshape(n, 3)
    return arr.copy()


def match_probability(t: bytes, q: bytes, st: int, sq: int, length: int, target_size: float) -> float:
    return lib().sxo_match_probability(t, q, st, sq, length, target_size)


def print_match_ident(q: bytes, t: bytes, st: int, sq: int, length: int) -> float:
    return lib().sxo_print_match_ident(q, t, st, sq, length)


def reverse_complement(seq: bytes) -> bytes:
    out = C.create_string_buffer(len(seq) + 1)
    lib().sxo_reverse_complement(seq, len(seq), out)
    return out.raw[: len(seq)]


def make_params(N=8192, q_chunk=4096, min_len=0, strands=2, cutoff=1.8, cutoff_fast=2.9, min_prob=0.99,
                target_total=4096.0, ptable=None, table_prob=0.0) -> Params:
    """ptable: a float64 array from prob_table() (kept alive by the caller) for -prob_table."""
    return Params(N, q_chunk, min_len, strands, cutoff, cutoff_fast, min_prob, target_total,
                  ptable.ctypes.data if ptable is not None else None, table_prob)


def prob_table(target_size: float, threads: int = 8) -> np.ndarray:
    """ProbTable::Setup (analysis/ProbTable.cc:15-52) as a (512, 2048) float64 array."""
    t = np.zeros((512, 2048), np.float64)
    lib().sxo_prob_table(C.c_double(target_size), threads, t.ctypes.data_as(C.POINTER(C.c_double)))
    return t


class TailSeg(C.Structure):
    _fields_ = [("end", C.c_int32), ("len", C.c_int32), ("seq", C.c_char_p), ("strand", C.c_int32),
                ("pad_", C.c_int32)]


class Tail(C.Structure):
    _fields_ = [("segs", C.POINTER(TailSeg)), ("nseg", C.c_int32), ("pad_", C.c_int32)]


def _tails(lst, keep):
    """lst: per pair a list of (end, seq bytes, strand) segments (sxo_tail_seg)."""
    arr = (Tail * len(lst))()
    for k, segs in enumerate(lst):
        if not segs:
            continue
        sa = (TailSeg * len(segs))(*[TailSeg(e, len(q), q, st, 0) for e, q, st in segs])
        keep += [sa] + [q for _, q, _ in segs]
        arr[k] = Tail(C.cast(sa, C.POINTER(TailSeg)), len(segs), 0)
    return arr


def align_batch(pairs, params: Params, threads: int = 1, cap_per_pair: int = 1 << 14, t_tails=None, q_tails=None):
    """pairs: list of dicts with keys tseq, qseq, t_chunk_start, q_chunk_start, t_record,
    q_record, q_record_len, fast.  t_tails / q_tails (optional): per pair the signal tail
    segments (end, sequence, strand) of sxo_align_batch_tails.  Returns a list of numpy
    structured arrays (RESULT_DTYPE)."""
    n = len(pairs)
    arr = (Pair * n)()
    keep = []
    for k, p in enumerate(pairs):
        t, q = p["tseq"], p["qseq"]
        keep += [t, q]
        arr[k] = Pair(t, len(t), q, len(q), p.get("t_chunk_start", 0), p.get("q_chunk_start", 0),
                      p.get("t_record", 0), p.get("q_record", 0), p.get("q_record_len", len(q)), p.get("fast", 0))
    out = np.zeros(n * cap_per_pair, RESULT_DTYPE)
    counts = np.zeros(n, np.int32)
    tt = _tails(t_tails, keep) if t_tails is not None else None
    qt = _tails(q_tails, keep) if q_tails is not None else None
    lib().sxo_align_batch_tails(arr, n, C.byref(params), C.cast(tt, C.c_void_p) if tt is not None else None,
                                C.cast(qt, C.c_void_p) if qt is not None else None, threads,
                                out.ctypes.data_as(C.POINTER(Result)), cap_per_pair,
                                counts.ctypes.data_as(C.POINTER(C.c_int32)))
    return [out[k * cap_per_pair: k * cap_per_pair + counts[k]].copy() for k in range(n)]
