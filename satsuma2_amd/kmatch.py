"""KMatch seeding (kmatch/KMatch.cc) with its k-mer tables on the GPU (include/sxc_kmatch.h).

`kmatch` returns KMatch's t_result records (RESULT_DTYPE) for two sets of records, or two
FASTA files, and the run's statistics.  `satsuma_seeds` runs it as SatsumaSynteny2 does,
for K = 11, 13, ..., 31 with min_length K, jump K - 1 and the given max_freq
(analysis/SatsumaSynteny2.cc:417-463), and concatenates the records in that order.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence, Tuple

import numpy as np

from . import _lib


class Params(C.Structure):
    _fields_ = [("K", C.c_int32), ("min_length", C.c_int32), ("max_jump", C.c_int32), ("max_freq", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("query_kmers", C.c_int64), ("target_kmers", C.c_int64), ("query_kept", C.c_int64),
                ("target_kept", C.c_int64), ("joined", C.c_int64), ("blocks", C.c_int64), ("read_ms", C.c_double),
                ("device_ms", C.c_double), ("sweep_ms", C.c_double)]


_declared = False


def _L():
    global _declared
    L = _lib.load()
    if not _declared:
        vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        cpp, pi64 = C.POINTER(C.c_char_p), C.POINTER(i64)
        L.sxc_kmatch_files.argtypes = [vp, C.c_char_p, C.c_char_p, C.POINTER(Params), C.POINTER(vp)]
        L.sxc_kmatch_records.argtypes = [vp, i32, cpp, pi64, i32, cpp, pi64, C.POINTER(Params), C.POINTER(vp)]
        L.sxc_kmatch_results.argtypes = [vp, vp, i64, pi64]
        L.sxc_kmatch_get_stats.argtypes = [vp, C.POINTER(Stats)]
        L.sxc_kmatch_destroy.argtypes = [vp]
        L.sxc_kmatch_destroy.restype = None
        _declared = True
    return L


def _collect(h) -> Tuple[np.ndarray, dict]:
    L = _L()
    try:
        n = C.c_int64(0)
        rc = L.sxc_kmatch_results(h, None, 0, C.byref(n))
        out = np.zeros(n.value, _lib.RESULT_DTYPE)
        if n.value:
            _lib.check(L.sxc_kmatch_results(h, out.ctypes.data, n.value, C.byref(n)))
        elif rc not in (_lib.SXC_OK, _lib.SXC_ERR_CAPACITY):
            _lib.check(rc)
        st = Stats()
        _lib.check(L.sxc_kmatch_get_stats(h, C.byref(st)))
        return out, {f: getattr(st, f) for f, _ in Stats._fields_}
    finally:
        L.sxc_kmatch_destroy(h)


def kmatch(ctx, query, target, K: int, min_length: int, max_jump: int, max_freq: int = 1) -> Tuple[np.ndarray, dict]:
    """query/target: FASTA paths, or sequences of records (bytes, record i has index i + 1)."""
    L = _L()
    p = Params(K, min_length, max_jump, max_freq)
    h = C.c_void_p()
    if isinstance(query, str):
        _lib.check(L.sxc_kmatch_files(ctx.handle, query.encode(), target.encode(), C.byref(p), C.byref(h)))
    else:
        def arrays(recs: Sequence[bytes]):
            seqs = (C.c_char_p * max(1, len(recs)))(*[bytes(r) for r in recs])
            lens = np.array([len(r) for r in recs] or [0], np.int64)
            return seqs, lens
        qs, ql = arrays(query)
        ts, tl = arrays(target)
        _lib.check(L.sxc_kmatch_records(ctx.handle, len(query), qs, ql.ctypes.data_as(C.POINTER(C.c_int64)),
                                        len(target), ts, tl.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(p),
                                        C.byref(h)))
    return _collect(h)


def satsuma_seeds(ctx, query, target, max_freq: int = 1) -> Tuple[np.ndarray, dict]:
    """The eleven KMatch runs SatsumaSynteny2 starts, K = 11 .. 31 (odd), records concatenated."""
    outs, st = [], {"device_ms": 0.0, "sweep_ms": 0.0, "blocks": 0}
    for k in range(11, 32, 2):
        r, s = kmatch(ctx, query, target, k, k, k - 1, max_freq)
        outs.append(r)
        for f in st:
            st[f] += s[f]
    return (np.concatenate(outs) if outs else np.zeros(0, _lib.RESULT_DTYPE)), st
