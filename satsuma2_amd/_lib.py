"""ctypes binding of libsxc_hip.so (the C ABI in include/sxc.h).

Loading fails loudly when the library is missing or no HIP device is present: there is
no host fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libsxc_hip.so")

SXC_OK = 0
SXC_ERR_ARG = -1
SXC_ERR_HIP = -2
SXC_ERR_CAPACITY = -3
SXC_ERR_NOMEM = -4
SXC_PENDING = 1


class SxcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sxc error {code}: {msg}")
        self.code = code


class SeqMatch(C.Structure):
    _fields_ = [("start_target", C.c_int32), ("start_query", C.c_int32), ("len", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("query_id", C.c_uint64), ("target_id", C.c_uint64), ("query_size", C.c_uint64),
                ("qstart", C.c_uint64), ("tstart", C.c_uint64), ("len", C.c_uint64), ("reverse", C.c_uint8),
                ("pad_", C.c_uint8 * 7), ("prob", C.c_double), ("ident", C.c_double)]


class Params(C.Structure):
    _fields_ = [("N", C.c_int32), ("q_chunk", C.c_int32), ("min_len", C.c_int32), ("strands", C.c_int32),
                ("cutoff", C.c_double), ("cutoff_fast", C.c_double), ("min_prob", C.c_double),
                ("target_total", C.c_double), ("prob_table", C.c_int32), ("pad_", C.c_int32),
                ("table_prob", C.c_double)]


class Pair(C.Structure):
    _fields_ = [("tseq", C.c_char_p), ("tlen", C.c_int32), ("qseq", C.c_char_p), ("qlen", C.c_int32),
                ("t_chunk_start", C.c_int32), ("q_chunk_start", C.c_int32), ("t_record", C.c_int32),
                ("q_record", C.c_int32), ("q_record_len", C.c_int32), ("fast", C.c_int32)]


class Timing(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("upload_ms", C.c_double), ("spectra_ms", C.c_double),
                ("pair_ms", C.c_double), ("finalize_ms", C.c_double), ("download_ms", C.c_double),
                ("candidates", C.c_int64), ("runs", C.c_int64), ("correlate_ms", C.c_double),
                ("scan_ms", C.c_double), ("device_ms", C.c_double), ("graph", C.c_int32), ("pad_", C.c_int32),
                ("near_threshold", C.c_int64)]


class TailSeg(C.Structure):
    _fields_ = [("end", C.c_int32), ("len", C.c_int32), ("seq", C.c_char_p), ("strand", C.c_int32),
                ("pad_", C.c_int32)]


class SignalTail(C.Structure):
    _fields_ = [("segs", C.POINTER(TailSeg)), ("nseg", C.c_int32), ("pad_", C.c_int32)]


class ChunkDesc(C.Structure):
    _fields_ = [("seq", C.c_char_p), ("len", C.c_int32), ("start", C.c_int32), ("record", C.c_int32),
                ("record_len", C.c_int32)]


class Block(C.Structure):
    _fields_ = [("target_from", C.c_int32), ("target_to", C.c_int32), ("query_from", C.c_int32),
                ("query_to", C.c_int32), ("fast", C.c_int32)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_int64), ("max_pairs_per_launch", C.c_int32), ("max_slots", C.c_int32),
                ("no_query_spectra", C.c_int32), ("pad_", C.c_int32)]


class ServerInfo(C.Structure):
    _fields_ = [("requests_done", C.c_int64), ("pairs_done", C.c_int64), ("launches", C.c_int64),
                ("results", C.c_int64), ("resident_bytes", C.c_int64), ("slot_bytes", C.c_int64),
                ("slots", C.c_int32), ("max_pairs_per_launch", C.c_int32), ("device_ms", C.c_double),
                ("query_spectra_bytes", C.c_int64), ("near_threshold", C.c_int64)]


class Geometry(C.Structure):
    _fields_ = [("stage_a", C.c_int32), ("round_levels", C.c_int32), ("three_per_cu", C.c_int32),
                ("lean_lds", C.c_int32)]


class TuneEntry(C.Structure):
    _fields_ = [("N", C.c_int32), ("pairs", C.c_int32), ("geo", Geometry), ("correlate_ms", C.c_double),
                ("spectra_ms", C.c_double), ("chosen_correlate", C.c_int32), ("chosen_spectra", C.c_int32)]


SXC_KERNEL_PAIRS = 0
SXC_KERNEL_SPECTRA = 1

RESULT_DTYPE = np.dtype([("query_id", "<u8"), ("target_id", "<u8"), ("query_size", "<u8"), ("qstart", "<u8"),
                         ("tstart", "<u8"), ("len", "<u8"), ("reverse", "u1"), ("pad_", "u1", (7,)),
                         ("prob", "<f8"), ("ident", "<f8")])
assert RESULT_DTYPE.itemsize == C.sizeof(Result) == 72

EXPORTS = ["sxc_create", "sxc_destroy", "sxc_last_error", "sxc_version", "sxc_encode_signal", "sxc_fft_forward",
           "sxc_xcorr", "sxc_matchup", "sxc_match_probability", "sxc_align_pair", "sxc_align_pair_static",
           "sxc_align_batch",
           "sxc_batch_create", "sxc_batch_create_tails", "sxc_batch_run", "sxc_batch_set_graph", "sxc_batch_results",
           "sxc_batch_stats", "sxc_batch_lags", "sxc_findtop",
           "sxc_batch_destroy", "sxc_last_timing", "sxc_server_create", "sxc_server_submit", "sxc_server_wait",
           "sxc_server_results", "sxc_server_stats", "sxc_server_destroy", "sxc_autotune", "sxc_set_geometry",
           "sxc_get_geometry"]

_lib = None
_lock = threading.Lock()


def load() -> C.CDLL:
    """Loads the library and declares signatures (no device call)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m satsuma2_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double
        fp = C.POINTER(C.c_float)
        dp = C.POINTER(C.c_double)
        L.sxc_create.argtypes = [C.POINTER(vp), C.c_int]
        L.sxc_destroy.argtypes = [vp]
        L.sxc_destroy.restype = None
        L.sxc_last_error.restype = C.c_char_p
        L.sxc_version.restype = C.c_char_p
        L.sxc_encode_signal.argtypes = [vp, C.c_char_p, i32, i32, fp]
        L.sxc_fft_forward.argtypes = [vp, fp, fp, i32, i32]
        L.sxc_xcorr.argtypes = [vp, fp, fp, i32, fp]
        L.sxc_matchup.argtypes = [vp, C.c_char_p, i32, C.c_char_p, i32, fp, i32, dbl, C.POINTER(SeqMatch), i32,
                                  C.POINTER(i32)]
        L.sxc_match_probability.argtypes = [vp, C.c_char_p, i32, C.c_char_p, i32, C.POINTER(SeqMatch), i32, dbl, dp,
                                            dp]
        L.sxc_align_pair.argtypes = [vp, C.POINTER(Pair), C.POINTER(Params), C.POINTER(Result), i32, C.POINTER(i32)]
        L.sxc_align_pair_static.argtypes = [vp, C.POINTER(Pair), C.POINTER(Params), C.POINTER(Result), i32,
                                            C.POINTER(i32)]
        L.sxc_align_batch.argtypes = [vp, C.POINTER(Pair), i32, C.POINTER(Params), C.POINTER(Result), i64,
                                      C.POINTER(i64)]
        L.sxc_batch_create.argtypes = [vp, C.POINTER(Pair), i32, C.POINTER(Params), C.POINTER(vp)]
        L.sxc_batch_run.argtypes = [vp]
        L.sxc_batch_set_graph.argtypes = [vp, i32]
        L.sxc_autotune.argtypes = [vp, i32, i32, C.POINTER(TuneEntry), i32, C.POINTER(i32)]
        L.sxc_set_geometry.argtypes = [vp, i32, i32, C.POINTER(Geometry)]
        L.sxc_get_geometry.argtypes = [vp, i32, i32, C.POINTER(Geometry)]
        L.sxc_batch_results.argtypes = [vp, C.POINTER(Result), i64, C.POINTER(i64)]
        L.sxc_batch_stats.argtypes = [vp, C.POINTER(i64), C.POINTER(i64)]
        L.sxc_batch_lags.argtypes = [vp, i32, i32, C.POINTER(i32), i32, C.POINTER(i32)]
        L.sxc_findtop.argtypes = [vp, fp, i32, dbl, C.POINTER(i32), i32, C.POINTER(i32)]
        L.sxc_batch_destroy.argtypes = [vp]
        L.sxc_batch_destroy.restype = None
        L.sxc_last_timing.argtypes = [vp, C.POINTER(Timing)]
        L.sxc_batch_create_tails.argtypes = [vp, C.POINTER(Pair), i32, C.POINTER(Params), C.POINTER(SignalTail),
                                             C.POINTER(SignalTail), C.POINTER(vp)]
        L.sxc_server_create.argtypes = [vp, C.POINTER(Params), C.POINTER(ChunkDesc), i32, C.POINTER(ChunkDesc), i32,
                                        C.POINTER(ServerConfig), C.POINTER(vp)]
        L.sxc_server_submit.argtypes = [vp, C.POINTER(Block), i32, C.POINTER(C.c_uint64)]
        L.sxc_server_wait.argtypes = [vp, C.c_uint64, i32, C.POINTER(i64), dp]
        L.sxc_server_results.argtypes = [vp, C.c_uint64, C.POINTER(Result), i64, C.POINTER(i64)]
        L.sxc_server_stats.argtypes = [vp, C.POINTER(ServerInfo)]
        L.sxc_server_destroy.argtypes = [vp]
        L.sxc_server_destroy.restype = None
        _lib = L
        return L


def check(rc: int) -> int:
    if rc != SXC_OK:
        raise SxcError(rc, load().sxc_last_error().decode(errors="replace"))
    return rc
