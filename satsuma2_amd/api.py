"""Python mirror of the reference's chunk-pair interfaces, computed on the GPU.

Names and argument meaning follow the reference classes so code and tests read like
the reference's own:

  CCSignal.SetSequence(seq, size)          analysis/CrossCorr.h:16, CrossCorr.cc:179-206
  CrossCorrelation.CrossCorrelate(t, q)    analysis/CrossCorr.h:151, CrossCorr.cc:386-403
  SeqAnalyzer.MatchUp(query, target, xc)   analysis/CrossCorr.h:261, CrossCorr.cc:583-605
  GetMatchProbability(t, q, sT, sQ, len, targetSize)   analysis/AlignProbability.h:10-15
  PrintMatch(q, t, m) -> identity          analysis/CrossCorr.h:293, CrossCorr.cc:967-1037
  HomologyByXCorr.Align / align_target     analysis/HomologyByXCorrSlave.cc:221-300

Every call goes through libsxc_hip.so; a missing library or device raises.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, NamedTuple, Sequence

import numpy as np

from . import _lib
from ._lib import RESULT_DTYPE, Pair, Params, Result, SeqMatch, Timing, check


def _fp(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _b(s) -> bytes:
    return s if isinstance(s, bytes) else bytes(s)


class Geometry(NamedTuple):
    """A compiled FFT kernel geometry (sxc_geometry)."""
    stage_a: int = 16
    round_levels: int = 3
    three_per_cu: int = 0
    lean_lds: int = 0

    @staticmethod
    def of(g) -> "Geometry":
        return Geometry(g.stage_a, g.round_levels, g.three_per_cu, g.lean_lds)


class Context:
    """One HIP device, its stream and constant tables (sxc_create)."""

    def __init__(self, device: int = 0):
        self.L = _lib.load()
        self._h = C.c_void_p()
        check(self.L.sxc_create(C.byref(self._h), device))

    def close(self):
        if self._h:
            self.L.sxc_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # ---- FFT kernel geometry (autotuner)
    def autotune(self, N_lo: int = 8192, N_hi: int = 0) -> list:
        """Times every compiled FFT geometry of the pair and target-spectra kernels for each
        length in [N_lo, N_hi] (sxc_autotune) and keeps the fastest; returns one dict per
        (length, geometry): N, pairs, geo (a Geometry tuple), correlate_ms, spectra_ms,
        chosen_correlate, chosen_spectra."""
        N_hi = N_hi or N_lo
        cap = 128
        out = (_lib.TuneEntry * cap)()
        n = C.c_int32(0)
        check(self.L.sxc_autotune(self._h, N_lo, N_hi, out, cap, C.byref(n)))
        res = []
        for i in range(n.value):
            e = out[i]
            res.append(dict(N=e.N, pairs=e.pairs, geo=Geometry.of(e.geo), correlate_ms=e.correlate_ms,
                            spectra_ms=e.spectra_ms, chosen_correlate=e.chosen_correlate,
                            chosen_spectra=e.chosen_spectra))
        return res

    def set_geometry(self, N: int, kernel: str, geo: "Geometry"):
        """kernel: "pairs" or "spectra"."""
        k = _lib.SXC_KERNEL_PAIRS if kernel == "pairs" else _lib.SXC_KERNEL_SPECTRA
        g = _lib.Geometry(*geo)
        check(self.L.sxc_set_geometry(self._h, N, k, C.byref(g)))

    def geometry(self, N: int, kernel: str = "pairs") -> "Geometry":
        k = _lib.SXC_KERNEL_PAIRS if kernel == "pairs" else _lib.SXC_KERNEL_SPECTRA
        g = _lib.Geometry()
        check(self.L.sxc_get_geometry(self._h, N, k, C.byref(g)))
        return Geometry.of(g)

    # ---- plain entry points (one reference call each)
    def encode(self, seq: bytes, N: int) -> np.ndarray:
        out = np.zeros((4, N), np.float32)
        check(self.L.sxc_encode_signal(self._h, _b(seq), len(seq), N, _fp(out)))
        return out

    def fft_forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(np.atleast_2d(x), np.float32)
        out = np.zeros_like(x)
        check(self.L.sxc_fft_forward(self._h, _fp(x), _fp(out), x.shape[1], x.shape[0]))
        return out

    def xcorr(self, tsig: np.ndarray, qsig: np.ndarray) -> np.ndarray:
        tsig = np.ascontiguousarray(tsig, np.float32)
        qsig = np.ascontiguousarray(qsig, np.float32)
        N = tsig.shape[1]
        xc = np.zeros(N, np.float32)
        check(self.L.sxc_xcorr(self._h, _fp(tsig), _fp(qsig), N, _fp(xc)))
        return xc

    def matchup(self, q: bytes, t: bytes, xc: np.ndarray, cutoff: float) -> np.ndarray:
        xc = np.ascontiguousarray(xc, np.float32)
        cap = 1 << 14
        while True:
            buf = (SeqMatch * cap)()
            n = C.c_int32(0)
            rc = self.L.sxc_matchup(self._h, _b(q), len(q), _b(t), len(t), _fp(xc), xc.size, cutoff, buf, cap,
                                    C.byref(n))
            if rc == _lib.SXC_ERR_CAPACITY and n.value > cap:
                cap = n.value
                continue
            check(rc)
            break
        return np.frombuffer(buf, np.int32, 3 * n.value).reshape(-1, 3).copy()

    def findtop(self, xc: np.ndarray, cutoff: float) -> np.ndarray:
        """SeqAnalyzer::FindTop (analysis/CrossCorr.cc:878-944): accepted lag indices."""
        xc = np.ascontiguousarray(xc, np.float32)
        out = np.zeros(xc.size, np.int32)
        n = C.c_int32(0)
        check(self.L.sxc_findtop(self._h, _fp(xc), xc.size, cutoff, out.ctypes.data_as(C.POINTER(C.c_int32)),
                                 out.size, C.byref(n)))
        return out[: n.value].copy()

    def match_probability(self, t: bytes, q: bytes, runs: np.ndarray, target_size: float):
        runs = np.ascontiguousarray(runs, np.int32).reshape(-1, 3)
        n = runs.shape[0]
        prob = np.zeros(n, np.float64)
        ident = np.zeros(n, np.float64)
        check(self.L.sxc_match_probability(self._h, _b(t), len(t), _b(q), len(q),
                                           runs.ctypes.data_as(C.POINTER(SeqMatch)), n, target_size,
                                           prob.ctypes.data_as(C.POINTER(C.c_double)),
                                           ident.ctypes.data_as(C.POINTER(C.c_double))))
        return prob, ident

    def timing(self) -> Timing:
        t = Timing()
        check(self.L.sxc_last_timing(self._h, C.byref(t)))
        return t


@dataclass
class SlaveParams:
    """HomologyByXCorrSlave flags (analysis/HomologyByXCorrSlave.cc:334-384)."""

    t_chunk: int = 4096
    q_chunk: int = 4096
    min_len: int = 0
    cutoff: float = 1.8
    cutoff_fast: float = 2.9
    min_prob: float = 0.99  # m_minProb is fixed at 0.99 on the slave path (:76)
    strands: int = 2
    target_total: float = 0.0
    prob_table: bool = False  # -prob_table: ProbTable lookup (analysis/ProbTable.cc:105-141)
    table_prob: float = 0.9999  # -min_prob: what a table hit reports (the table's cutoff)

    def to_c(self) -> Params:
        return Params(2 * self.t_chunk, self.q_chunk, self.min_len, self.strands, self.cutoff, self.cutoff_fast,
                      self.min_prob, self.target_total, 1 if self.prob_table else 0, 0, self.table_prob)


@dataclass
class ChunkPair:
    tseq: bytes
    qseq: bytes
    t_chunk_start: int = 0
    q_chunk_start: int = 0
    t_record: int = 0
    q_record: int = 0
    q_record_len: int = -1
    fast: int = 0


def _pairs_array(pairs: Sequence[ChunkPair]):
    arr = (Pair * max(1, len(pairs)))()
    keep = []
    for k, p in enumerate(pairs):
        keep += [p.tseq, p.qseq]
        qrl = p.q_record_len if p.q_record_len >= 0 else len(p.qseq)
        arr[k] = Pair(p.tseq, len(p.tseq), p.qseq, len(p.qseq), p.t_chunk_start, p.q_chunk_start, p.t_record,
                      p.q_record, qrl, p.fast)
    return arr, keep


class HomologyByXCorr:
    """The slave's per-pair work (Align + FilterMatches) as plain, batched and resident calls."""

    def __init__(self, ctx: Context, params: SlaveParams):
        self.ctx = ctx
        self.params = params

    def align_pair(self, pair: ChunkPair) -> np.ndarray:
        arr, keep = _pairs_array([pair])
        prm = self.params.to_c()
        cap = 1024
        while True:
            out = np.zeros(cap, RESULT_DTYPE)
            n = C.c_int32(0)
            rc = self.ctx.L.sxc_align_pair(self.ctx.handle, arr, C.byref(prm), out.ctypes.data_as(C.POINTER(Result)),
                                           cap, C.byref(n))
            if rc == _lib.SXC_ERR_CAPACITY and n.value > cap:
                cap = n.value
                continue
            check(rc)
            return out[: n.value].copy()

    def align_batch(self, pairs: Sequence[ChunkPair]) -> np.ndarray:
        arr, keep = _pairs_array(pairs)
        prm = self.params.to_c()
        cap = max(1024, 4 * len(pairs))
        while True:
            out = np.zeros(cap, RESULT_DTYPE)
            n = C.c_int64(0)
            rc = self.ctx.L.sxc_align_batch(self.ctx.handle, arr, len(pairs), C.byref(prm),
                                            out.ctypes.data_as(C.POINTER(Result)), cap, C.byref(n))
            if rc == _lib.SXC_ERR_CAPACITY and n.value > cap:
                cap = n.value
                continue
            check(rc)
            return out[: n.value].copy()


class ResidentBatch:
    """Chunk pairs uploaded once to HBM (sxc_batch_create); run() repeats the device work."""

    def __init__(self, ctx: Context, pairs: Sequence[ChunkPair], params: SlaveParams):
        self.ctx = ctx
        self._arr, self._keep = _pairs_array(pairs)
        self._prm = params.to_c()
        self._h = C.c_void_p()
        check(ctx.L.sxc_batch_create(ctx.handle, self._arr, len(pairs), C.byref(self._prm), C.byref(self._h)))

    def run(self):
        check(self.ctx.L.sxc_batch_run(self._h))

    def set_graph(self, on: bool):
        """Replay runs from a captured HIP graph (the default from the second run on) or
        launch kernel by kernel (phase timings available)."""
        check(self.ctx.L.sxc_batch_set_graph(self._h, 1 if on else 0))

    def results(self) -> np.ndarray:
        cap = 1 << 12
        while True:
            out = np.zeros(cap, RESULT_DTYPE)
            n = C.c_int64(0)
            rc = self.ctx.L.sxc_batch_results(self._h, out.ctypes.data_as(C.POINTER(Result)), cap, C.byref(n))
            if rc == _lib.SXC_ERR_CAPACITY and n.value > cap:
                cap = n.value
                continue
            check(rc)
            return out[: n.value].copy()

    def stats(self):
        c, r = C.c_int64(0), C.c_int64(0)
        check(self.ctx.L.sxc_batch_stats(self._h, C.byref(c), C.byref(r)))
        return c.value, r.value

    def lags(self, pair: int, strand: int) -> np.ndarray:
        """The lags the pair kernel accepted (FindTop) for one pair and strand in the last run."""
        cap = 1 << 15
        out = np.zeros(cap, np.int32)
        n = C.c_int32(0)
        check(self.ctx.L.sxc_batch_lags(self._h, pair, strand, out.ctypes.data_as(C.POINTER(C.c_int32)), cap,
                                        C.byref(n)))
        return out[: n.value].copy()

    def close(self):
        if self._h:
            self.ctx.L.sxc_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Chunk:
    """One chunk of a resident genome (ChunkManager, analysis/SeqChunk.cc:72-165)."""

    seq: bytes
    start: int = 0
    record: int = 0
    record_len: int = -1


class Server:
    """The serving loop (sxc_server_*): genomes resident in HBM, requests of t_pair blocks
    (target_from, target_to, query_from, query_to, fast) taken as they arrive."""

    def __init__(self, ctx: Context, params: SlaveParams, targets: Sequence[Chunk], queries: Sequence[Chunk],
                 hbm_budget: int = 0, max_pairs_per_launch: int = 0, max_slots: int = 0,
                 query_spectra: bool = True):
        self.ctx = ctx
        self._keep = [c.seq for c in targets] + [c.seq for c in queries]

        def descs(chunks):
            arr = (_lib.ChunkDesc * max(1, len(chunks)))()
            for k, c in enumerate(chunks):
                arr[k] = _lib.ChunkDesc(c.seq, len(c.seq), c.start, c.record,
                                        c.record_len if c.record_len >= 0 else len(c.seq))
            return arr

        self._t, self._q = descs(targets), descs(queries)
        self._prm = params.to_c()
        cfg = _lib.ServerConfig(hbm_budget, max_pairs_per_launch, max_slots, 0 if query_spectra else 1, 0)
        self._h = C.c_void_p()
        check(ctx.L.sxc_server_create(ctx.handle, C.byref(self._prm), self._t, len(targets), self._q, len(queries),
                                      C.byref(cfg), C.byref(self._h)))

    def submit(self, blocks) -> int:
        arr = (_lib.Block * max(1, len(blocks)))(*[_lib.Block(*[int(x) for x in b]) for b in blocks])
        t = C.c_uint64(0)
        check(self.ctx.L.sxc_server_submit(self._h, arr, len(blocks), C.byref(t)))
        return t.value

    def wait(self, ticket: int, timeout_ms: int = -1):
        """(number of records, latency in ms), or None if not done within timeout_ms."""
        n, lat = C.c_int64(0), C.c_double(0)
        rc = self.ctx.L.sxc_server_wait(self._h, ticket, timeout_ms, C.byref(n), C.byref(lat))
        if rc == _lib.SXC_PENDING:
            return None
        check(rc)
        return n.value, lat.value

    def results(self, ticket: int) -> np.ndarray:
        n, _ = self.wait(ticket)
        out = np.zeros(max(1, n), RESULT_DTYPE)
        got = C.c_int64(0)
        check(self.ctx.L.sxc_server_results(self._h, ticket, out.ctypes.data_as(C.POINTER(Result)), len(out),
                                            C.byref(got)))
        return out[: got.value].copy()

    def stats(self) -> dict:
        i = _lib.ServerInfo()
        check(self.ctx.L.sxc_server_stats(self._h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in _lib.ServerInfo._fields_}

    def close(self):
        if self._h:
            self.ctx.L.sxc_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- reference-named mirror

_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class CCSignal:
    """CCSignal (analysis/CrossCorr.h:10-71): four float channels A, C, G, T."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self._sig = np.zeros((4, 0), np.float32)

    def SetSequence(self, seq: bytes, size: int) -> None:
        self._sig = self.ctx.encode(seq, size)

    def GetFullSize(self) -> int:
        return self._sig.shape[1]

    def GetCount(self) -> int:
        return 4

    def Get(self, i: int) -> np.ndarray:
        return self._sig[i]

    @property
    def channels(self) -> np.ndarray:
        return self._sig


class CrossCorrelation:
    """CrossCorrelation (analysis/CrossCorr.h:146-161)."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx or default_context()

    def CrossCorrelate(self, one: CCSignal, two: CCSignal) -> np.ndarray:
        return self.ctx.xcorr(one.channels, two.channels)


class SeqAnalyzer:
    """SeqAnalyzer (analysis/CrossCorr.h:257-290)."""

    def __init__(self, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.m_topCutoff = 1.8

    def SetTopCutoff(self, c: float) -> None:
        self.m_topCutoff = c

    def MatchUp(self, query: bytes, target: bytes, xc: np.ndarray) -> np.ndarray:
        return self.ctx.matchup(query, target, xc, self.m_topCutoff)


def GetMatchProbability(target: bytes, query: bytes, startTarget: int, startQuery: int, length: int,
                        targetSize: float, ctx: Context | None = None) -> float:
    ctx = ctx or default_context()
    prob, _ = ctx.match_probability(target, query, np.array([[startTarget, startQuery, length]], np.int32), targetSize)
    return float(prob[0])


def PrintMatch(query: bytes, target: bytes, m, ctx: Context | None = None) -> float:
    """Identity returned by PrintMatch(query, target, m, true); m = (startTarget, startQuery, len)."""
    ctx = ctx or default_context()
    _, ident = ctx.match_probability(target, query, np.array([list(m)], np.int32), 0.0)
    return float(ident[0])
