"""The master's side of the search (include/sxc_master.h) against the reference's own code.

oracle/_ref/ref_master runs the reference's MultiMatches, RunMatchDynProg(Mult),
SyntenyInterpolator, TargetCoverageTracker, WorkQueue collection and GridSearch, compiled
from its sources (oracle/Makefile).  Where it is built, randomized cases compare every field
of every output bit for bit; the committed fixtures (tests/golden/master_cases.npz, written
by tests/golden/make_golden.py with the same driver) hold a fixed set of cases for machines
without the reference.  The generators aim at what decides outputs: neighbours within the
collapse distance, ties in target start (std::sort's order), runs of >= 40 collinear matches
(interpolation), repeats (the chain's repeat weights), both strands, matches of one target
to several queries (RunMatchDynProgMult), records of N only, records too short for chunks
and records sharing a name (the grid).
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from satsuma2_amd import _lib
from satsuma2_amd.master import GRID_TARGET_DTYPE, MATCH_DTYPE, Coverage, Grid, MatchSet, SearchIteration

REF_MASTER = os.path.join(ROOT, "oracle", "_ref", "ref_master")
GOLDEN = os.path.join(ROOT, "tests", "golden", "master_cases.npz")
need_ref = pytest.mark.skipif(not os.path.exists(REF_MASTER), reason="oracle/_ref/ref_master not built")


# ---- the driver's file formats (oracle/ref_master.cc)

def set_bytes(tsizes, qsizes, m) -> bytes:
    return (np.array([len(tsizes), len(qsizes)], np.int32).tobytes() + np.asarray(tsizes, np.int32).tobytes()
            + np.asarray(qsizes, np.int32).tobytes() + np.array([len(m)], np.int64).tobytes()
            + np.ascontiguousarray(m, MATCH_DTYPE).tobytes())


def parse_set(b: bytes, off: int = 0):
    nt, nq = np.frombuffer(b, np.int32, 2, off)
    off += 8
    ts = np.frombuffer(b, np.int32, nt, off)
    off += 4 * int(nt)
    qs = np.frombuffer(b, np.int32, nq, off)
    off += 4 * int(nq)
    n = int(np.frombuffer(b, np.int64, 1, off)[0])
    off += 8
    m = np.frombuffer(b, MATCH_DTYPE, n, off).copy()
    return ts, qs, m, off + n * MATCH_DTYPE.itemsize


def parse_targets(b: bytes, off: int):
    ptar, n = np.frombuffer(b, np.int32, 2, off)
    off += 8
    t = np.frombuffer(b, GRID_TARGET_DTYPE, n, off).copy()
    return int(ptar), t, off + int(n) * GRID_TARGET_DTYPE.itemsize


def run_ref(tmp_path, *args) -> bytes:
    out = str(tmp_path / "ref.out")
    subprocess.run([REF_MASTER, *[str(a) for a in args], out] if args[0] not in ("coverage",)
                   else [REF_MASTER, args[0], out, *[str(a) for a in args[1:]]],
                   check=True, stdout=subprocess.DEVNULL)
    with open(out, "rb") as f:
        return f.read()


def same(a: np.ndarray, b: np.ndarray) -> bool:
    return a.dtype == b.dtype and a.shape == b.shape and a.tobytes() == b.tobytes()


# ---- generators

def random_set(seed: int, nt: int = 3, nq: int = 4, runs: int = 12, noise: int = 300):
    """Matches along collinear runs (some long enough to interpolate), their near copies
    (collapse), repeats and noise, on both strands."""
    rng = np.random.default_rng(seed)
    tsizes = rng.integers(400_000, 2_000_000, nt).astype(np.int32)
    qsizes = rng.integers(400_000, 2_000_000, nq).astype(np.int32)
    rows = []
    for _ in range(runs):
        t, q = int(rng.integers(nt)), int(rng.integers(nq))
        rc = int(rng.integers(2))
        k = int(rng.choice([8, 30, 45, 70]))
        step = int(rng.integers(300, 4000))
        st = int(rng.integers(0, max(1, tsizes[t] - k * step - 1000)))
        sq = int(rng.integers(0, max(1, qsizes[q] - k * step - 1000)))
        for i in range(k):
            ln = int(rng.integers(50, 260))
            jitter = int(rng.integers(-30, 31))
            rows.append((t, q, qsizes[q], st + i * step, max(0, sq + i * step + jitter), ln, rc,
                         float(rng.uniform(0.99, 1.0)), float(rng.uniform(0.6, 1.0))))
            if rng.random() < 0.15:  # a near copy: collapses with it
                rows.append((t, q, qsizes[q], st + i * step + int(rng.integers(-3, 4)),
                             max(0, sq + i * step + jitter + int(rng.integers(-3, 4))), ln + int(rng.integers(0, 40)),
                             rc, float(rng.uniform(0.99, 1.0)), float(rng.uniform(0.6, 1.0))))
    for _ in range(noise):
        t, q = int(rng.integers(nt)), int(rng.integers(nq))
        ln = int(rng.integers(40, 400))
        st = int(rng.integers(0, tsizes[t] - ln - 1))
        if rng.random() < 0.2 and rows:  # the same target start as another match: a sort tie
            st = min(rows[int(rng.integers(len(rows)))][3], int(tsizes[t]) - ln - 1)
        rows.append((t, q, qsizes[q], st, int(rng.integers(0, qsizes[q] - ln - 1)), ln, int(rng.integers(2)),
                     float(rng.uniform(0.99, 1.0)), float(rng.uniform(0.5, 1.0))))
    # a repeat: many matches of one target region to different places
    t = int(rng.integers(nt))
    st = int(rng.integers(0, tsizes[t] - 1000))
    for _ in range(30):
        q = int(rng.integers(nq))
        rows.append((t, q, qsizes[q], st + int(rng.integers(0, 50)), int(rng.integers(0, qsizes[q] - 500)), 200,
                     int(rng.integers(2)), 0.999, 0.8))
    m = np.zeros(len(rows), MATCH_DTYPE)
    for i, r in enumerate(rows):
        m[i]["target_id"], m[i]["query_id"], m[i]["query_len"], m[i]["start_target"], m[i]["start_query"] = r[:5]
        m[i]["length"], m[i]["rc"], m[i]["prob"], m[i]["ident"] = r[5:]
        m[i]["matches"] = r[8] * r[5]
    rng.shuffle(m)
    return tsizes, qsizes, m


def prepared(seed: int):
    """Sorted and collapsed as the master hands sets to the chaining."""
    ts, qs, m = random_set(seed)
    s = MatchSet(ts, qs).add(m).sort().collapse()
    return ts, qs, s.matches()


def mine(op: str, ts, qs, m, *arg) -> np.ndarray:
    s = MatchSet(ts, qs).add(m)
    if op == "sort":
        s.sort()
    elif op == "collapse":
        s.collapse()
    elif op == "lengthfilter":
        s.length_filter(arg[0])
    elif op == "interpolate":
        s.interpolate()
    elif op in ("chain", "chainmult"):
        s = s.chain(dups=op == "chainmult")
    return s.matches()


def random_results(rng, tsizes, qsizes, n: int) -> np.ndarray:
    r = np.zeros(n, _lib.RESULT_DTYPE)
    r["target_id"] = rng.integers(0, len(tsizes), n)
    r["query_id"] = rng.integers(0, len(qsizes), n)
    r["query_size"] = np.asarray(qsizes, np.int64)[r["query_id"].astype(np.int64)]
    r["len"] = rng.integers(40, 400, n)
    r["tstart"] = [int(rng.integers(0, int(tsizes[t]) - 500)) for t in r["target_id"]]
    r["qstart"] = [int(rng.integers(0, int(qsizes[q]) - 500)) for q in r["query_id"]]
    r["reverse"] = rng.integers(0, 2, n)
    r["prob"] = rng.uniform(0.99, 1.0, n)
    r["ident"] = rng.uniform(0.5, 1.0, n)
    return r


def write_fasta(path, recs):
    with open(path, "w") as f:
        for name, seq in recs:
            f.write(f">{name}\n")
            for i in range(0, len(seq), 80):
                f.write(seq[i:i + 80].decode() + "\n")


def genomes(seed: int):
    """Target and query records: random bases, a stretch of N covering whole chunks, a record
    of 5 bases (no chunks) and two records sharing a name."""
    rng = np.random.default_rng(seed)
    acgt = np.frombuffer(b"ACGT", np.uint8)

    def rnd(n):
        return acgt[rng.integers(0, 4, n)].tobytes()
    t = [("chrA", rnd(300_000)), ("chrB", rnd(150_000)), ("tiny", rnd(5)), ("chrB", rnd(90_000))]
    nblock = b"N" * 40_000
    t[0] = ("chrA", t[0][1][:100_000] + nblock + t[0][1][140_000:])
    q = [("q1", rnd(260_000)), ("q2", rnd(120_000)), ("q3", rnd(70_000))]
    return t, q


# ---- tests against the reference driver

@need_ref
@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("op", ["sort", "collapse", "interpolate"])
def test_set_ops_match_reference(tmp_path, op, seed):
    ts, qs, m = random_set(seed)
    if op != "sort":
        m = MatchSet(ts, qs).add(m).sort().matches()  # as the master calls them
    p = tmp_path / "in.bin"
    p.write_bytes(set_bytes(ts, qs, m))
    _, _, want, _ = parse_set(run_ref(tmp_path, op, p))
    assert same(mine(op, ts, qs, m), want)


@need_ref
def test_collapse_unsorted_and_tiny_sets(tmp_path):
    """Collapse on the set's own order, and its end case: the last group is never emitted
    (a set of one match collapses to none)."""
    ts, qs, m = random_set(9)
    for arr in (m, m[:2], m[:1]):
        p = tmp_path / "in.bin"
        p.write_bytes(set_bytes(ts, qs, arr))
        _, _, want, _ = parse_set(run_ref(tmp_path, "collapse", p))
        assert same(mine("collapse", ts, qs, arr), want)
    assert len(mine("collapse", ts, qs, m[:1])) == 0


@need_ref
@pytest.mark.parametrize("min_len", [0, 100, 220])
def test_length_filter_matches_reference(tmp_path, min_len):
    ts, qs, m = random_set(4)
    p = tmp_path / "in.bin"
    p.write_bytes(set_bytes(ts, qs, m))
    _, _, want, _ = parse_set(run_ref(tmp_path, "lengthfilter", p, min_len))
    assert same(mine("lengthfilter", ts, qs, m, min_len), want)


@need_ref
@pytest.mark.parametrize("seed", [5, 6, 7, 8])
@pytest.mark.parametrize("op", ["chain", "chainmult"])
def test_chaining_matches_reference(tmp_path, op, seed):
    ts, qs, m = prepared(seed)
    p = tmp_path / "in.bin"
    p.write_bytes(set_bytes(ts, qs, m))
    wts, wqs, want, _ = parse_set(run_ref(tmp_path, op, p))
    got = mine(op, ts, qs, m)
    assert len(want) > 0 and same(got, want)
    assert np.array_equal(wts, ts) and np.array_equal(wqs, qs)


@need_ref
@pytest.mark.parametrize("op", ["chain", "chainmult"])
def test_chaining_many_targets_matches_reference(tmp_path, op):
    """Targets are chained on several host threads; the result keeps the reference's target
    order."""
    ts, qs, m = random_set(12, nt=24, nq=5, runs=60, noise=600)
    m = MatchSet(ts, qs).add(m).sort().collapse().matches()
    assert len(np.unique(m["target_id"])) == 24
    p = tmp_path / "in.bin"
    p.write_bytes(set_bytes(ts, qs, m))
    _, _, want, _ = parse_set(run_ref(tmp_path, op, p))
    assert len(want) > 0 and same(mine(op, ts, qs, m), want)


def limits_set(seed: int):
    """A set that reaches the chaining's look-ahead limits: more than 2,000 matches within
    250 kbp (the count limit), a repeat-dense stretch whose matches weigh almost nothing (skip
    sums stay tiny, so far sources keep competing), and sparse matches more than 1 Mbp apart
    (the distance break after 11 sources, and TransPenalty's square-root term)."""
    rng = np.random.default_rng(seed)
    tsizes = np.array([6_000_000], np.int32)
    qsizes = np.array([6_000_000, 3_000_000], np.int32)
    rows = []
    pos = 100
    for i in range(9000):  # dense collinear run with jitter, both queries, both strands
        q = int(rng.integers(2)) if rng.random() < 0.1 else 0
        rc = 1 if rng.random() < 0.05 else 0
        ln = int(rng.integers(40, 200))
        rows.append((0, q, qsizes[q], pos, min(int(qsizes[q]) - 300, max(0, pos + int(rng.integers(-20, 21)))), ln,
                     rc, float(rng.uniform(0.98, 1.0)), 0.9))
        pos += int(rng.integers(20, 160))
    base = pos + 5000
    for i in range(3000):  # a repeat: many copies of one region, piled on one spot
        st = base + int(rng.integers(0, 400))
        rows.append((0, 1, qsizes[1], st, int(rng.integers(0, qsizes[1] - 400)), 300, int(rng.integers(2)),
                     0.999, 0.8))
    pos = base + 600_000
    while pos < tsizes[0] - 500:  # sparse: gaps of 0.3 - 1.6 Mbp
        rows.append((0, 0, qsizes[0], pos, min(int(qsizes[0]) - 300, pos // 2), 150, 0, 0.995, 0.85))
        pos += int(rng.integers(300_000, 1_600_000))
    m = np.zeros(len(rows), MATCH_DTYPE)
    for i, r in enumerate(rows):
        m[i]["target_id"], m[i]["query_id"], m[i]["query_len"], m[i]["start_target"], m[i]["start_query"] = r[:5]
        m[i]["length"], m[i]["rc"], m[i]["prob"], m[i]["ident"] = r[5:]
        m[i]["matches"] = r[8] * r[5]
    rng.shuffle(m)
    return tsizes, qsizes, MatchSet(tsizes, qsizes).add(m).sort().collapse().matches()


@need_ref
@pytest.mark.parametrize("op", ["chain", "chainmult"])
def test_chaining_at_look_ahead_limits_matches_reference(tmp_path, op):
    ts, qs, m = limits_set(31)
    st = np.sort(m["start_target"])
    assert len(m) > 12_000 and np.max(np.searchsorted(st, st + 250_000) - np.arange(len(st))) > 2000
    assert np.max(np.diff(st)) > 1_000_000
    p = tmp_path / "in.bin"
    p.write_bytes(set_bytes(ts, qs, m))
    _, _, want, _ = parse_set(run_ref(tmp_path, op, p))
    got = mine(op, ts, qs, m)
    assert len(want) > 1000 and same(got, want)


def tied_set(seed: int):
    """A collinear run in which every 7th match has a twin: same target, query, strand, target
    start and length, query start one base on.  Twins tie under SingleMatch::operator<, so a
    chain holding both is put in order by std::sort's own tie order."""
    rng = np.random.default_rng(seed)
    tsizes = np.array([3_000_000], np.int32)
    qsizes = np.array([3_000_000, 2_000_000], np.int32)
    rows = []
    pos = 1000
    for i in range(3000):
        q = 1 if rng.random() < 0.05 else 0
        rc = 1 if rng.random() < 0.1 else 0
        ln = int(rng.integers(60, 240))
        sq = min(int(qsizes[q]) - 300, pos + int(rng.integers(-4, 5)))
        rows.append((0, q, qsizes[q], pos, sq, ln, rc, float(rng.uniform(0.99, 1.0)), 0.9))
        if i % 7 == 0:
            rows.append((0, q, qsizes[q], pos, sq + 1, ln, rc, float(rng.uniform(0.99, 1.0)), 0.9))
        pos += int(rng.integers(200, 800))
    m = np.zeros(len(rows), MATCH_DTYPE)
    for i, r in enumerate(rows):
        m[i]["target_id"], m[i]["query_id"], m[i]["query_len"], m[i]["start_target"], m[i]["start_query"] = r[:5]
        m[i]["length"], m[i]["rc"], m[i]["prob"], m[i]["ident"] = r[5:]
        m[i]["matches"] = r[8] * r[5]
    rng.shuffle(m)
    return tsizes, qsizes, m


@need_ref
@pytest.mark.parametrize("op", ["chain", "chainmult"])
def test_chaining_with_ties_matches_reference(tmp_path, op):
    """The chain is put in SingleMatch::operator< order by grouping when that order is strict,
    and by std::sort otherwise: a chain holding tied matches, against the reference."""
    ts, qs, m = tied_set(41)
    m = MatchSet(ts, qs).add(m).sort().matches()  # sorted, not collapsed: the twins stay
    p = tmp_path / "tied.bin"
    p.write_bytes(set_bytes(ts, qs, m))
    _, _, want, _ = parse_set(run_ref(tmp_path, op, p))
    key = ["target_id", "query_id", "rc", "start_target", "length"]
    ties = np.all([want[k][1:] == want[k][:-1] for k in key], axis=0)
    assert len(want) > 1000 and ties.sum() > 10  # the reference's own chain holds ties
    assert same(mine(op, ts, qs, m), want)


@need_ref
def test_interpolation_adds_guesses(tmp_path):
    """A chained set with a run of >= 40 collinear matches gets two guesses on each side."""
    ts, qs, m = prepared(11)
    chained = MatchSet(ts, qs).add(m).chain().matches()
    p = tmp_path / "in.bin"
    p.write_bytes(set_bytes(ts, qs, chained))
    _, _, want, _ = parse_set(run_ref(tmp_path, "interpolate", p))
    got = mine("interpolate", ts, qs, chained)
    assert same(got, want)
    assert len(got) > len(chained) and np.all(got[len(chained):]["prob"] == 0.999)


@need_ref
def test_coverage_matches_reference(tmp_path):
    files, sets = [], []
    for k, seed in enumerate([21, 22, 22, 23]):
        ts, qs, m = prepared(seed)
        p = tmp_path / f"c{k}.bin"
        p.write_bytes(set_bytes(ts, qs, m))
        files.append(p)
        sets.append((ts, qs, m))
    want = np.frombuffer(run_ref(tmp_path, "coverage", *files), np.uint64).reshape(-1, 2)
    cov = Coverage()
    got = np.array([cov.update(MatchSet(ts, qs).add(m)) for ts, qs, m in sets], np.uint64)
    assert np.array_equal(got, want)
    assert want[2].tolist() == [0, 0]  # the same set again adds and removes nothing


@need_ref
@pytest.mark.parametrize("from_file", [False, True])
def test_collect_matches_reference(tmp_path, from_file):
    """WorkQueue's collection: the running m_matches sum; seed files mirror reverse starts.
    The reference path names only seed files, so the slave path is checked by construction."""
    rng = np.random.default_rng(31)
    ts, qs = [500_000, 700_000], [400_000, 300_000, 200_000]
    r = random_results(rng, ts, qs, 500)
    sizes = tmp_path / "sizes.bin"
    sizes.write_bytes(set_bytes(ts, qs, np.zeros(0, MATCH_DTYPE)))
    res = tmp_path / "res.bin"
    res.write_bytes(r.tobytes())
    got = MatchSet(ts, qs).collect(r, from_file=True).matches()
    if from_file:
        _, _, want, _ = parse_set(run_ref(tmp_path, "collect", sizes, res))
        assert same(got, want)
    else:
        plain = MatchSet(ts, qs).collect(r).matches()
        rev = r["reverse"] != 0
        q = r["qstart"].astype(np.int64)
        assert np.array_equal(plain["start_query"], q)
        assert np.array_equal(got["start_query"][rev], (r["query_size"] - r["len"] - r["qstart"])[rev])
        assert np.array_equal(plain["matches"], np.cumsum(r["ident"] * r["len"].astype(np.float64)))


@need_ref
def test_grid_matches_reference(tmp_path):
    """Seed lines on an empty grid, then the heaviest cells of chained sets, over three
    rounds (cells once submitted stay out; used target columns stay used)."""
    t, q = genomes(41)
    tf, qf = tmp_path / "t.fa", tmp_path / "q.fa"
    write_fasta(tf, t)
    write_fasta(qf, q)
    chunk, blocks = 4096, 4
    tsz = [len(s) for _, s in t]
    qsz = [len(s) for _, s in q]
    rng = np.random.default_rng(42)
    rounds = []
    for k, n in enumerate([40, 25, 60]):
        m = np.zeros(0, MATCH_DTYPE)
        if k:
            r = random_results(rng, [tsz[0], tsz[1]], qsz, 120)
            m = MatchSet(tsz, qsz).collect(r).sort().collapse().chain().matches()
        rounds.append((n, m))
    script = np.array([len(rounds)], np.int32).tobytes()
    for n, m in rounds:
        script += np.array([n], np.int32).tobytes() + set_bytes(tsz, qsz, m)
    sp = tmp_path / "script.bin"
    sp.write_bytes(script)
    ref = run_ref(tmp_path, "grid", tf, qf, chunk, blocks, sp)
    g = Grid.from_fasta(chunk, blocks, str(tf), str(qf))
    g2 = Grid.from_records(chunk, blocks, [(n, s) for n, s in t], [(n, s) for n, s in q])
    off = 0
    for n, m in rounds:
        ptar, want, off = parse_targets(ref, off)
        for grid in (g, g2):
            grid.update(MatchSet(tsz, qsz).add(m))
            got, cand = grid.collect(n)
            assert cand == ptar and same(got, want)
    assert off == len(ref)


@need_ref
@pytest.mark.parametrize("dups", [0, 1])
def test_search_iteration_matches_reference(tmp_path, dups):
    """The master's loop body over four rounds (seeds, then slave results): the kept match
    set, the interpolated chain, the coverage pair and the collected targets."""
    t, q = genomes(51)
    tf, qf = tmp_path / "t.fa", tmp_path / "q.fa"
    write_fasta(tf, t)
    write_fasta(qf, q)
    tsz = [len(s) for _, s in t]
    qsz = [len(s) for _, s in q]
    rng = np.random.default_rng(52)
    rounds = [(30, 1, random_results(rng, [tsz[0], tsz[1]], qsz, 300))]
    for n in (20, 50, 10):
        rounds.append((n, 0, random_results(rng, [tsz[0], tsz[1]], qsz, 200)))
    script = np.array([len(rounds)], np.int32).tobytes()
    for n, kind, r in rounds:
        script += np.array([n, kind], np.int32).tobytes() + np.array([len(r)], np.int64).tobytes() + r.tobytes()
    sp = tmp_path / "script.bin"
    sp.write_bytes(script)
    ref = run_ref(tmp_path, "iterate", tf, qf, 4096, 4, dups, sp)
    it = SearchIteration(Grid.from_fasta(4096, 4, str(tf), str(qf)), tsz, qsz, dups=bool(dups))
    off = 0
    for k, (n, kind, r) in enumerate(rounds):
        if kind == 1:
            it.matches.collect(r, from_file=True)
            targets, cand, cov = it.step(np.zeros(0, _lib.RESULT_DTYPE), n)
        else:
            targets, cand, cov = it.step(r, n)
        ptar, want_t, off = parse_targets(ref, off)
        want_cov = np.frombuffer(ref, np.uint64, 2, off)
        off += 16
        _, _, want_m, off = parse_set(ref, off)
        _, _, want_c, off = parse_set(ref, off)
        assert cand == ptar and same(targets, want_t), k
        assert tuple(int(v) for v in want_cov) == cov, k
        assert same(it.matches.matches(), want_m), k
        assert same(it.chained.matches(), want_c), k
    assert off == len(ref)


# ---- committed fixtures (no reference needed)

@pytest.mark.skipif(not os.path.exists(GOLDEN), reason="fixture not generated")
def test_master_fixtures():
    g = np.load(GOLDEN)
    for name in g["cases"]:
        name = str(name)
        op = str(g[f"{name}_op"])
        ts, qs, m = g[f"{name}_ts"], g[f"{name}_qs"], g[f"{name}_in"]
        want = g[f"{name}_out"]
        assert same(mine(op, ts, qs, m), want), name


def test_errors_are_reported():
    s = MatchSet([1000], [1000])
    bad = np.zeros(1, MATCH_DTYPE)
    bad["target_id"], bad["query_id"] = 3, 0
    s.add(bad)
    with pytest.raises(_lib.SxcError):
        s.chain()
    with pytest.raises(_lib.SxcError):
        Grid.from_fasta(4096, 4, "/nonexistent.fa", "/nonexistent.fa")
