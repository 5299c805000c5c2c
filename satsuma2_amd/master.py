"""The master's side of Satsuma2's search over libsxc_hip.so (include/sxc_master.h).

`MatchSet` is MultiMatches (analysis/SequenceMatch.h:176-257) with Sort, Collapse,
LengthFilter, the chaining dynamic programme (RunMatchDynProg / RunMatchDynProgMult,
analysis/MatchDynProg.cc) and the synteny interpolation (analysis/SatsumaSynteny2.cc:18-146);
`Coverage` is TargetCoverageTracker; `Grid` is GridSearch (analysis/GridSearch.cc), whose
collected targets are the blocks of chunk pairs the chunk-pair server aligns next;
`SearchIteration` is one turn of the master's loop (analysis/SatsumaSynteny2.cc:504-541).
These are host code: they run without a GPU.
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib

MATCH_DTYPE = np.dtype([("target_id", "<i4"), ("query_id", "<i4"), ("query_len", "<i4"), ("start_target", "<i4"),
                        ("start_query", "<i4"), ("length", "<i4"), ("rc", "<i4"), ("reserved", "<i4"),
                        ("matches", "<f8"), ("prob", "<f8"), ("ident", "<f8")])
assert MATCH_DTYPE.itemsize == 56
GRID_TARGET_DTYPE = np.dtype([("target_first", "<i4"), ("target_last", "<i4"), ("query_first", "<i4"),
                              ("query_last", "<i4"), ("x", "<i4"), ("y", "<i4"), ("y2", "<i4"), ("count", "<i4"),
                              ("fast", "<i4")])

_declared = False


def _L():
    global _declared
    L = _lib.load()
    if not _declared:
        vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
        pi32, pi64, pu64 = C.POINTER(i32), C.POINTER(i64), C.POINTER(C.c_uint64)
        L.sxc_matchset_create.argtypes = [i32, pi32, i32, pi32]
        L.sxc_matchset_create.restype = vp
        L.sxc_matchset_destroy.argtypes = [vp]
        L.sxc_matchset_destroy.restype = None
        L.sxc_matchset_add.argtypes = [vp, vp, i64]
        L.sxc_matchset_collect.argtypes = [vp, vp, i64, C.c_int]
        L.sxc_matchset_count.argtypes = [vp]
        L.sxc_matchset_count.restype = i64
        L.sxc_matchset_get.argtypes = [vp, vp, i64]
        L.sxc_matchset_get.restype = i64
        for f in ("sxc_matchset_sort", "sxc_matchset_collapse", "sxc_matchset_interpolate"):
            getattr(L, f).argtypes = [vp]
        L.sxc_matchset_length_filter.argtypes = [vp, i64]
        L.sxc_matchset_copy.argtypes = [vp, vp]
        L.sxc_matchset_chain.argtypes = [vp, C.c_int, vp]
        L.sxc_coverage_create.argtypes = []
        L.sxc_coverage_create.restype = vp
        L.sxc_coverage_destroy.argtypes = [vp]
        L.sxc_coverage_destroy.restype = None
        L.sxc_coverage_update.argtypes = [vp, vp, pu64, pu64]
        cpp = C.POINTER(C.c_char_p)
        L.sxc_grid_create.argtypes = [i32, i32, i32, cpp, cpp, pi64, i32, cpp, cpp, pi64]
        L.sxc_grid_create.restype = vp
        L.sxc_grid_create_fasta.argtypes = [i32, i32, C.c_char_p, C.c_char_p]
        L.sxc_grid_create_fasta.restype = vp
        L.sxc_grid_destroy.argtypes = [vp]
        L.sxc_grid_destroy.restype = None
        L.sxc_grid_chunks.argtypes = [vp, pi32, pi32]
        L.sxc_grid_chunk.argtypes = [vp, C.c_int, i32, pi32, pi32]
        L.sxc_grid_update.argtypes = [vp, vp]
        L.sxc_grid_collect.argtypes = [vp, i32, vp, i32, pi32, pi32]
        L.sxc_grid_last.argtypes = [vp, vp, i32, pi32]
        _declared = True
    return L


def _raise_null(p, what: str):
    if not p:
        raise _lib.SxcError(_lib.SXC_ERR_ARG, _lib.load().sxc_last_error().decode(errors="replace") or what)
    return p


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


class MatchSet:
    """MultiMatches: record sizes plus a list of matches (MATCH_DTYPE rows)."""

    def __init__(self, target_sizes: Sequence[int], query_sizes: Sequence[int]):
        L = _L()
        self.target_sizes, self.query_sizes = _i32(target_sizes), _i32(query_sizes)
        self._h = _raise_null(L.sxc_matchset_create(
            len(self.target_sizes), self.target_sizes.ctypes.data_as(C.POINTER(C.c_int32)), len(self.query_sizes),
            self.query_sizes.ctypes.data_as(C.POINTER(C.c_int32))), "sxc_matchset_create failed")

    def close(self):
        if getattr(self, "_h", None):
            _L().sxc_matchset_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def add(self, matches: np.ndarray) -> "MatchSet":
        a = np.ascontiguousarray(matches, dtype=MATCH_DTYPE)
        _lib.check(_L().sxc_matchset_add(self._h, a.ctypes.data, len(a)))
        return self

    def collect(self, results: np.ndarray, from_file: bool = False) -> "MatchSet":
        """Slave records (RESULT_DTYPE) as WorkQueue::collect_new_matches adds them."""
        r = np.ascontiguousarray(results, dtype=_lib.RESULT_DTYPE)
        _lib.check(_L().sxc_matchset_collect(self._h, r.ctypes.data, len(r), 1 if from_file else 0))
        return self

    def __len__(self) -> int:
        return int(_L().sxc_matchset_count(self._h))

    def matches(self) -> np.ndarray:
        n = len(self)
        out = np.zeros(n, MATCH_DTYPE)
        got = _L().sxc_matchset_get(self._h, out.ctypes.data, n)
        if got < 0:
            _lib.check(int(got))
        return out

    def sort(self) -> "MatchSet":
        _lib.check(_L().sxc_matchset_sort(self._h))
        return self

    def collapse(self) -> "MatchSet":
        _lib.check(_L().sxc_matchset_collapse(self._h))
        return self

    def length_filter(self, min_length: int) -> "MatchSet":
        _lib.check(_L().sxc_matchset_length_filter(self._h, int(min_length)))
        return self

    def interpolate(self) -> "MatchSet":
        _lib.check(_L().sxc_matchset_interpolate(self._h))
        return self

    def copy(self) -> "MatchSet":
        out = MatchSet(self.target_sizes, self.query_sizes)
        _lib.check(_L().sxc_matchset_copy(self._h, out._h))
        return out

    def chain(self, dups: bool = False) -> "MatchSet":
        """RunMatchDynProg (or RunMatchDynProgMult with dups) into a new set."""
        out = MatchSet(self.target_sizes, self.query_sizes)
        _lib.check(_L().sxc_matchset_chain(self._h, 1 if dups else 0, out._h))
        return out


class Coverage:
    """TargetCoverageTracker: bases added and removed against the previous set."""

    def __init__(self):
        self._h = _raise_null(_L().sxc_coverage_create(), "sxc_coverage_create failed")

    def close(self):
        if getattr(self, "_h", None):
            _L().sxc_coverage_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def update(self, s: MatchSet) -> Tuple[int, int]:
        a, r = C.c_uint64(0), C.c_uint64(0)
        _lib.check(_L().sxc_coverage_update(self._h, s._h, C.byref(a), C.byref(r)))
        return int(a.value), int(r.value)


class Grid:
    """GridSearch over target and query records: chunk `chunk`, `blocks` x `blocks` cells."""

    def __init__(self, handle, keep=None):
        self._h = _raise_null(handle, "sxc_grid_create failed")
        self._keep = keep

    @classmethod
    def from_records(cls, chunk: int, blocks: int, targets: Sequence[Tuple[str, bytes]],
                     queries: Sequence[Tuple[str, bytes]]) -> "Grid":
        def arrays(recs):
            names = (C.c_char_p * max(1, len(recs)))(*[n.encode() for n, _ in recs])
            seqs = (C.c_char_p * max(1, len(recs)))(*[bytes(s) for _, s in recs])
            lens = np.array([len(s) for _, s in recs] or [0], np.int64)
            return names, seqs, lens
        tn, ts, tl = arrays(targets)
        qn, qs, ql = arrays(queries)
        h = _L().sxc_grid_create(chunk, blocks, len(targets), tn, ts, tl.ctypes.data_as(C.POINTER(C.c_int64)),
                                 len(queries), qn, qs, ql.ctypes.data_as(C.POINTER(C.c_int64)))
        return cls(h, keep=(tn, ts, tl, qn, qs, ql))

    @classmethod
    def from_fasta(cls, chunk: int, blocks: int, target_fasta: str, query_fasta: str) -> "Grid":
        return cls(_L().sxc_grid_create_fasta(chunk, blocks, target_fasta.encode(), query_fasta.encode()))

    def close(self):
        if getattr(self, "_h", None):
            _L().sxc_grid_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def chunks(self) -> Tuple[int, int]:
        t, q = C.c_int32(0), C.c_int32(0)
        _lib.check(_L().sxc_grid_chunks(self._h, C.byref(t), C.byref(q)))
        return int(t.value), int(q.value)

    def chunk(self, query: bool, index: int) -> Tuple[int, int]:
        r, s = C.c_int32(0), C.c_int32(0)
        _lib.check(_L().sxc_grid_chunk(self._h, 1 if query else 0, index, C.byref(r), C.byref(s)))
        return int(r.value), int(s.value)

    def update(self, chained: MatchSet):
        _lib.check(_L().sxc_grid_update(self._h, chained._h))

    def collect(self, n: int) -> Tuple[np.ndarray, int]:
        """CollectTargets(n): the targets (GRID_TARGET_DTYPE) and the candidate cell count."""
        cap = max(16, 2 * n)
        out = np.zeros(cap, GRID_TARGET_DTYPE)
        count, cand = C.c_int32(0), C.c_int32(0)
        rc = _L().sxc_grid_collect(self._h, n, out.ctypes.data, cap, C.byref(count), C.byref(cand))
        if rc == _lib.SXC_ERR_CAPACITY:
            out = np.zeros(count.value, GRID_TARGET_DTYPE)
            rc = _L().sxc_grid_last(self._h, out.ctypes.data, len(out), C.byref(count))
        _lib.check(rc)
        return out[: count.value].copy(), int(cand.value)


class SearchIteration:
    """The master's loop body (analysis/SatsumaSynteny2.cc:504-541): the collected records
    join the match set, which is sorted, collapsed, coverage-tracked and chained; the chain
    replaces the set, its interpolation weights the grid, and the grid yields the next
    targets.  `step` returns (targets, candidate cells, coverage (added, removed))."""

    def __init__(self, grid: Grid, target_sizes: Sequence[int], query_sizes: Sequence[int], dups: bool = False):
        self.grid, self.dups = grid, dups
        self.matches = MatchSet(target_sizes, query_sizes)
        self.coverage = Coverage()
        self.chained: Optional[MatchSet] = None
        self.first = True

    def seed(self, results: np.ndarray, min_seed_length: int = 24):
        """Seeds from KMatch result files (WorkQueue::results_from_file), then Sort, Collapse
        and LengthFilter (analysis/SatsumaSynteny2.cc:466-477)."""
        self.matches.collect(results, from_file=True)
        if len(self.matches):
            self.matches.sort().collapse().length_filter(min_seed_length)

    def step(self, results: np.ndarray, n_targets: int) -> Tuple[np.ndarray, int, Tuple[int, int]]:
        if len(results):
            self.matches.collect(results)
        cov = (0, 0)
        if self.first or len(results):
            self.first = False
            self.matches.sort().collapse()
            cov = self.coverage.update(self.matches)
            chained = self.matches.chain(self.dups)
            self.matches.close()
            self.matches = chained.copy()
            chained.interpolate()
            self.grid.update(chained)
            if self.chained is not None:
                self.chained.close()
            self.chained = chained
        targets, cand = self.grid.collect(n_targets)
        return targets, cand, cov
