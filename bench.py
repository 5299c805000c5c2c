"""Chunk-pair throughput of the MI355X cross-correlation path on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (uniform_pairs): 4096 chunk pairs of 4096 bp, i.i.d. bases at GC 0.41, half of
them carrying one planted homology (50-2000 bp, 70-95% identity, 1% indels), FFT length
8192, both strands, slave parameters (cutoff 1.8, keep threshold 0.99).  Inputs are
uploaded to HBM before the timed region (sxc_batch_create); one step = one
sxc_batch_run over all 4096 pairs (target spectra, fused pair kernel, result ordering),
replayed from the HIP graph the library captures on the second run.  The per-kernel times
(phase_ms, roofline) come from the same number of steps launched kernel by kernel.

Serving (config of the serving loop): two synthetic 50-Mbp genomes sharing 2,000 syntenic
blocks, chunked as the slave chunks them (16,277 target chunks of 4096 bp every 3072 bp,
12,208 query chunks), target spectra resident in HBM.  The requests are SatsumaSynteny2's
own: KMatch seeds on the GPU (K = 11 .. 31), then the master's search loop (the restated
grid search, match-set steps and chaining of satsuma2_amd.master) with the server as its
slaves, in synchronous rounds for a time budget (`search`: rounds, chunk pairs per second
through the whole loop, master and alignment seconds).  The requests it made (24 x 24-chunk
cells and seed lines) are then replayed as streams offering 25%, 50%, 75%, 100%, 150% and
200% of the batched throughput in chunk pairs per second (exponential gaps scaled by each
request's pairs); each rate reports request latency p50/p99 counted from the request's
scheduled arrival, the server's own service latency, and chunk pairs per second; the highest
completed rate is the saturation point.  The headline p50/p99 is at 50%.

Uneven (config of the uneven batch): 20,000 pairs with target and query lengths log-uniform
in [256, 16384] bp, each at its own FFT length (1024 ... 32768), in one resident batch.

Large served run (config of the large served run; --no-large skips it): two 200-Mbp
genomes with GC drawn per 100-kbp window, tandem repeats over 10% of each, five
transposon-like families of 1,000 copies and 5,000 syntenic blocks: the same seeds and search
loop, its requests served at 25%, 50% and 100% of the batched rate.

KMatch: the GPU KMatch program beside the reference's (compiled from kmatch/KMatch.cc) on
two 10-Mbp genomes at K = 21, same files in and out.

Printed: one JSON line with the batched throughput (`value`), the plain one-pair-per-call
throughput and the one-shot (host buffers in and out, PCIe included) batch rate beside it,
the serving latencies, the dominant kernel's roofline figures from HIP events against the
8 TB/s spec and against a device copy measured here, and the reference CPU path timed on
this host's cores (all that are available, and one GPU's share of the host) on a bounded
sample.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "chunk-pairs/s at 4096bp (batched vs plain, 1 GPU); serve p50/p99 ms"
N_PAIRS = 4096
HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per chunk pair (both strands, N = 8192, 4 channels, FP32 spectra
# materialised, target spectra resident, peak search fused): forward
# 2*4*8192*4 + 1024 = 263,168 B; inverse 262,144 + 131,072 + 65,536 = 458,752 B.
B_PAIR = 263_168 + 458_752
# HBM bytes the pair kernel moves per chunk pair, from rocprofv3 FETCH_SIZE and WRITE_SIZE
# counter runs on the 4096-pair launch, each scaled by what the counter reads for a plain
# device copy of known size: (271,704 KiB x 2.0 + 4,916 KiB) / 4096 pairs
# (profiles/uniform_pairs_4096_kernels.md, the build named there; kernels changed since need
# a new counter run before this figure holds for them).
TRAFFIC_PER_PAIR = 137_081
TRAFFIC_SOURCE = "profiles/uniform_pairs_4096_kernels.md (FETCH_SIZE x 2.0 + WRITE_SIZE)"
# offered chunk pairs/s as fractions of the batched rate (0.5: the headline); the serving loop
# keeps query spectra resident, so its saturation lies above the batched rate
SERVE_LOADS = (0.25, 0.5, 0.75, 1.0, 1.5, 2.0)
LARGE_LOADS = (0.25, 0.5, 1.0)   # offered loads of the large served run


def kernel_name(ctx, N: int = 8192) -> str:
    """The pair kernel's instantiation name for the context's geometry at length N (a plain
    log2 length for the default geometry, else the packed geometry code)."""
    sa, km, occ, lean = ctx.geometry(N, "pairs")
    nb = N.bit_length() - 1
    default = (32 if nb >= 15 else 16, 3, 0, 0)
    code = nb if (sa, km, occ, lean) == default else (
        nb | (32 if sa == 32 else 0) | (km << 6) | (256 if occ else 0) | (512 if lean else 0))
    return f"k_correlate<{code}>"


def make_workload(n_pairs: int, seed: int = 1):
    from satsuma2_amd import synth
    from satsuma2_amd.api import ChunkPair

    T, Q = synth.uniform_pairs(n_pairs, seed=seed)
    pairs = [ChunkPair(T[r], Q[r], t_record=r, q_record=r, q_record_len=4096) for r in range(n_pairs)]
    return T, Q, pairs


def host_cpus():
    """(CPUs of the host, CPUs this process may run on)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return os.cpu_count() or 1, avail


_COPY_PROBE = """
import json, sys, torch
gib, reps = float(sys.argv[1]), int(sys.argv[2])
n = int(gib * (1 << 30))
a = torch.empty(n, dtype=torch.uint8, device="cuda")
b = torch.empty_like(a)
a.fill_(1)
b.copy_(a)
torch.cuda.synchronize()
e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
e0.record()
for _ in range(reps):
    b.copy_(a)
e1.record()
torch.cuda.synchronize()
print(json.dumps({"bytes": 2 * n, "ms": e0.elapsed_time(e1) / reps}))
"""


def copy_bandwidth(gib: float = 1.0, reps: int = 20):
    """Device-to-device copy of `gib` GiB on this GPU (read + write bytes per second), the
    measured ceiling the roofline fraction is also quoted against.  Measured in a child
    process (torch's HIP runtime is not mixed with the library's in one process)."""
    try:
        r = subprocess.run([sys.executable, "-c", _COPY_PROBE, str(gib), str(reps)], capture_output=True, text=True,
                           timeout=180)
        d = json.loads(r.stdout.strip().splitlines()[-1])
    except Exception:  # no torch, no GPU visible to it: no figure
        return None
    return {"gb_per_s": round(d["bytes"] / (d["ms"] * 1e-3) / 1e9, 1), "bytes": d["bytes"], "ms": round(d["ms"], 4),
            "how": f"torch copy_ of {gib:g} GiB, device to device, mean of {reps}, in a child process"}


def cpu_baseline(T, Q, sample: int, threads: int):
    """The reference slave path (compiled from source into oracle/_ref) on `sample` pairs,
    `threads` worker threads; the oracle restatement when that binary is absent."""
    drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    from satsuma2_amd import synth

    if os.path.exists(drv):
        with tempfile.TemporaryDirectory() as tmp:
            qf, tf, pf, of = (os.path.join(tmp, x) for x in ("q.fa", "t.fa", "p.txt", "r.bin"))
            synth.write_fasta(qf, Q[:sample])
            synth.write_fasta(tf, T[:sample])
            with open(pf, "w") as f:
                for r in range(sample):
                    f.write(f"{2 * r} {2 * r} {2 * r} {2 * r} 0\n")
            out = subprocess.run([drv, "pairs", qf, tf, pf, of, str(threads)], check=True, capture_output=True,
                                 text=True).stdout.strip().splitlines()[-1]
            kv = dict(x.split("=") for x in out.split())
            secs = float(kv["seconds"])
            return {"value": sample / secs, "unit": "chunk-pairs/s", "cores": threads, "kind": "reference",
                    "sample": f"{sample} pairs of the same workload, both strands, HomologyByXCorr::align_target"}
    import oracle

    prm = oracle.make_params(target_total=float(4096 * N_PAIRS))
    pairs = [dict(tseq=T[r], qseq=Q[r], t_record=r, q_record=r, q_record_len=4096) for r in range(sample)]
    t0 = time.perf_counter()
    oracle.align_batch(pairs, prm, threads=threads)
    secs = time.perf_counter() - t0
    return {"value": sample / secs, "unit": "chunk-pairs/s", "cores": threads, "kind": "port",
            "sample": f"{sample} pairs of the same workload, both strands, oracle restatement"}


def serve_bench(ctx, genome_bp: int, duration: float, capacity: float, loads, seed: int = 4, large: bool = False,
                search_s: float = 15.0):
    """The served path under SatsumaSynteny2's own request stream: KMatch seeds (GPU,
    K = 11 .. 31), then the master's search loop (restated grid search and chaining) with the
    server as its slaves, in synchronous rounds under a time budget; the requests it made are
    then replayed as a Poisson stream at each offered load for latency percentiles."""
    from satsuma2_amd import synth
    from satsuma2_amd.api import Server, SlaveParams
    from satsuma2_amd.kmatch import satsuma_seeds
    from satsuma2_amd.serve_load import Replay, genome_chunks, run_search, run_stream

    print(f"[bench] serving: generating two {genome_bp / 1e6:.0f}-Mbp genomes", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    if large:
        T, Q, _ = synth.large_genomes(length=genome_bp, n_blocks=max(1, genome_bp * 5000 // 200_000_000), seed=seed)
    else:
        T, Q, _ = synth.syntenic_genomes(length=genome_bp, n_blocks=max(1, genome_bp * 2000 // 50_000_000),
                                         seed=seed)
    gen_s = time.perf_counter() - t0
    tch = genome_chunks(T, 4096, 3072)
    qch = genome_chunks(Q, 4096, 4096)
    print("[bench] serving: KMatch seeds, K = 11 .. 31", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    seeds, kst = satsuma_seeds(ctx, [Q], [T])
    seeds_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    srv = Server(ctx, SlaveParams(target_total=float(len(T))), tch, qch)
    create_s = time.perf_counter() - t0
    print(f"[bench] serving: search loop ({search_s:.0f} s budget)", file=sys.stderr, flush=True)
    reqs, search = run_search(srv, T, Q, seeds, budget_s=search_s)
    runs = []
    for k, load in enumerate(loads):
        print(f"[bench] serving: {load * capacity:.0f} chunk pairs/s offered ({load:.0%} of the batched rate) "
              f"for {duration:.0f} s", file=sys.stderr, flush=True)
        r = run_stream(srv, Replay(reqs), 0.0, duration, seed=seed + k, pairs_rate=load * capacity)
        r["offered_load"] = load
        runs.append(r)
    info = srv.stats()
    srv.close()
    head = next((r for r in runs if r["offered_load"] == 0.5), runs[0])
    sat = max(runs, key=lambda r: r["pairs_per_s"])
    kind = ("GC 0.38-0.45 per 100-kbp window, 10% tandem repeats, 5 transposon families x 1000 copies, "
            "5,000 syntenic blocks" if large else "2,000 syntenic blocks")
    npairs = [sum((b[1] - b[0] + 1) * (b[3] - b[2] + 1) for b in rq) for rq in reqs]
    return {"workload": f"two {genome_bp / 1e6:.0f}-Mbp synthetic genomes ({kind}), {len(tch)} target x "
                        f"{len(qch)} query chunks; requests: the targets SatsumaSynteny2's search loop collects "
                        "(24x24-chunk cells, seed lines) from KMatch seeds, replayed as a Poisson stream",
            "generate_s": round(gen_s, 1),
            "p50_ms": head["p50_ms"], "p99_ms": head["p99_ms"], "at_load": head["offered_load"],
            "latency": "from each request's scheduled (Poisson) arrival to its completion; service_*: from "
                       "submission, as the server measures it",
            "saturation_pairs_per_s": sat["pairs_per_s"], "saturation_at_load": sat["offered_load"],
            "saturation_p99_ms": sat["p99_ms"],
            "requests": {"recorded": len(reqs), "pairs_p50": int(np.percentile(npairs, 50)) if npairs else 0,
                         "pairs_max": int(max(npairs)) if npairs else 0,
                         "seed_lines": int(sum(b[4] for rq in reqs for b in rq))},
            "kmatch": {"seconds": round(seeds_s, 2), "records": int(len(seeds)),
                       "device_ms": round(kst["device_ms"], 1), "sweep_ms": round(kst["sweep_ms"], 1)},
            "search": search,
            "rates": runs, "create_s": round(create_s, 2), "resident_gb": round(info["resident_bytes"] / 1e9, 3),
            "slots": info["slots"], "max_pairs_per_launch": info["max_pairs_per_launch"],
            "launches": info["launches"]}


def kmatch_bench(ctx, genome_bp: int = 10_000_000, K: int = 21, seed: int = 6):
    """KMatch on one genome pair at SatsumaSynteny2's arguments for one K: the GPU program
    (satsuma2_amd/lib/KMatch) beside the reference program compiled from kmatch/KMatch.cc
    (oracle/_ref/KMatch, its two loader threads), same FASTA files, same output."""
    import subprocess
    import tempfile

    from satsuma2_amd import synth

    ref = os.path.join(ROOT, "oracle", "_ref", "KMatch")
    mine = os.path.join(ROOT, "satsuma2_amd", "lib", "KMatch")
    T, Q, _ = synth.syntenic_genomes(length=genome_bp, n_blocks=max(1, genome_bp // 25_000), seed=seed)
    out = {"workload": f"two {genome_bp / 1e6:.0f}-Mbp synthetic genomes, K = {K}, min_length {K}, jump {K - 1}, "
                       "max_freq 1 (FASTA files in, t_result file out)"}
    with tempfile.TemporaryDirectory() as d:
        tf, qf = os.path.join(d, "t.fa"), os.path.join(d, "q.fa")
        for path, seq, name in ((tf, T, b">t\n"), (qf, Q, b">q\n")):
            with open(path, "wb") as f:
                f.write(name)
                for i in range(0, len(seq), 80):
                    f.write(seq[i:i + 80] + b"\n")
        for key, prog in (("gpu_s", mine), ("reference_cpu_s", ref)):
            if not os.path.exists(prog):
                continue
            o = os.path.join(d, key + ".bin")
            t0 = time.perf_counter()
            subprocess.run([prog, qf, tf, str(K), o, str(K), str(K - 1), "1"], check=True, stdout=subprocess.DEVNULL)
            out[key] = round(time.perf_counter() - t0, 3)
            out[key.replace("_s", "_records")] = os.path.getsize(o) // 72
    if "gpu_s" in out and "reference_cpu_s" in out:
        out["speedup"] = round(out["reference_cpu_s"] / out["gpu_s"], 1)
        out["reference_cores"] = 2
    return out


def uneven_bench(ctx, n_pairs: int, steps: int, seed: int = 3):
    """Uneven pairs in one batch: lengths log-uniform in [256, 16384] bp, each pair at its own
    FFT length (next power of two >= q + t), one sub-batch per length class, the classes
    running side by side on their own streams."""
    import collections

    from satsuma2_amd import synth
    from satsuma2_amd.api import ChunkPair, ResidentBatch, SlaveParams

    print(f"[bench] uneven: generating {n_pairs} pairs", file=sys.stderr, flush=True)
    T, Q = synth.uneven_pairs(n_pairs, seed=seed)
    pairs = [ChunkPair(T[r], Q[r], t_record=r, q_record=r, q_record_len=len(Q[r])) for r in range(n_pairs)]
    classes = collections.Counter(synth.fft_len(len(q), len(t)) for t, q in zip(T, Q))
    rb = ResidentBatch(ctx, pairs, SlaveParams(t_chunk=0, target_total=float(sum(len(t) for t in T))))
    rb.run()
    print(f"[bench] uneven: timing {steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        rb.run()
    wall = (time.perf_counter() - t0) / steps
    n_res = len(rb.results())
    rb.close()
    # each length class alone, launched kernel by kernel (device ms per phase; the batch runs
    # the classes side by side)
    per_class = {}
    tot = float(sum(len(t) for t in T))
    for N in sorted(classes):
        idx = [r for r in range(n_pairs) if synth.fft_len(len(Q[r]), len(T[r])) == N]
        cb = ResidentBatch(ctx, [pairs[r] for r in idx], SlaveParams(t_chunk=0, target_total=tot))
        cb.set_graph(False)
        cb.run()
        t1 = time.perf_counter()
        cb.run()
        tm = ctx.timing()
        per_class[str(N)] = {"pairs": len(idx), "ms": round(1e3 * (time.perf_counter() - t1), 3),
                             "spectra_ms": round(tm.spectra_ms, 3), "correlate_ms": round(tm.correlate_ms, 3),
                             "scan_ms": round(tm.scan_ms, 3), "pair_kernel": kernel_name(ctx, N)}
        cb.close()
    return {"workload": f"uneven_pairs: {n_pairs} pairs, q and t log-uniform in [256, 16384] bp, GC 0.35-0.55, "
                        "FFT length per pair, both strands",
            "pairs_per_s": round(n_pairs / wall, 1), "ms_per_step": round(1e3 * wall, 3), "steps": steps,
            "fft_len_classes": {str(k): v for k, v in sorted(classes.items())},
            "bp": int(sum(len(t) + len(q) for t, q in zip(T, Q))), "results": n_res, "classes_alone": per_class}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--pairs", type=int, default=N_PAIRS)
    ap.add_argument("--plain-pairs", type=int, default=256)
    ap.add_argument("--cpu-sample", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serve-bp", type=int, default=50_000_000)
    ap.add_argument("--serve-seconds", type=float, default=4.0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the large served run (two 200-Mbp genomes with repeats; about a minute)")
    ap.add_argument("--large-bp", type=int, default=200_000_000)
    ap.add_argument("--no-autotune", action="store_true", help="default FFT kernel geometries (for profiling)")
    ap.add_argument("--uneven-pairs", type=int, default=20_000)
    ap.add_argument("--uneven-steps", type=int, default=5)
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="only the device copy of this many GiB (the HBM counters' calibration), then exit")
    args = ap.parse_args()
    if args.copy_probe > 0:
        sys.argv = [sys.argv[0], str(args.copy_probe), "2"]
        exec(_COPY_PROBE, {})
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (chunk pairs shard by target chunk "
              f"across GPUs with no collective: each GPU holds its targets' spectra and takes their pairs)",
              file=sys.stderr)
        sys.exit(2)

    from satsuma2_amd.api import Context, HomologyByXCorr, ResidentBatch, SlaveParams

    print(f"[bench] generating {args.pairs} pairs", file=sys.stderr, flush=True)
    T, Q, pairs = make_workload(args.pairs)
    params = SlaveParams(target_total=float(4096 * args.pairs))
    ctx = Context(0)
    tune, tune_s = [], 0.0
    if not args.no_autotune:
        print("[bench] autotuning the FFT kernel geometries", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        tune = ctx.autotune(512, 32768)
        tune_s = time.perf_counter() - t0
    rb = ResidentBatch(ctx, pairs, params)

    print("[bench] warmup", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        rb.run()
    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    graphed = 0
    for _ in range(args.steps):
        rb.run()  # returns after the step's last copy has completed (a captured-graph replay)
        graphed += ctx.timing().graph
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    value = args.pairs / (wall / args.steps)
    n_res = len(rb.results())
    cands, runs = rb.stats()
    # the same steps launched kernel by kernel, for the per-kernel times (HIP events)
    rb.set_graph(False)
    spec_ms, fin_ms, cor_ms, scan_ms = [], [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rb.run()
        tm = ctx.timing()
        spec_ms.append(tm.spectra_ms)
        fin_ms.append(tm.finalize_ms)
        cor_ms.append(tm.correlate_ms)
        scan_ms.append(tm.scan_ms)
    eager_ms = 1e3 * (time.perf_counter() - t0) / args.steps
    assert len(rb.results()) == n_res
    rb.close()

    # one-shot batch: host buffers in and out, PCIe transfers included (sxc_align_batch)
    h = HomologyByXCorr(ctx, params)
    h.align_batch(pairs[:64])
    t0 = time.perf_counter()
    one = h.align_batch(pairs)
    one_s = time.perf_counter() - t0
    tm1 = ctx.timing()
    oneshot = {"pairs_per_s": round(args.pairs / one_s, 1), "ms": round(1e3 * one_s, 3),
               "upload_ms": round(tm1.upload_ms, 3), "download_ms": round(tm1.download_ms, 3),
               "results": int(len(one)),
               "what": "sxc_align_batch on the same pairs: pack on the host, upload, run, download"}

    # plain drop-in: one chunk pair per call, host buffers in and out (as the slave calls Align)
    plain = float("nan")
    if args.plain_pairs > 0:
        h = HomologyByXCorr(ctx, params)
        for p in pairs[:4]:
            h.align_pair(p)
        t0 = time.perf_counter()
        for p in pairs[: args.plain_pairs]:
            h.align_pair(p)
        plain = args.plain_pairs / (time.perf_counter() - t0)

    k_ms = float(np.mean(cor_ms))
    achieved = B_PAIR * args.pairs / (k_ms * 1e-3) / 1e9
    copy = copy_bandwidth()
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "chunk-pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "uniform_pairs: 4096-bp target x 4096-bp query chunk pairs, N=8192, both strands",
                   "pairs": args.pairs, "chunk_bp": 4096, "fft_len": 8192, "strands": 2, "cutoff": 1.8,
                   "keep_prob": 0.99},
        "graph_steps": graphed,
        "eager_ms_per_step": round(eager_ms, 3),
        "plain_pairs_per_s": round(plain, 1) if plain == plain else None,
        "batched_over_plain": round(value / plain, 2) if plain == plain else None,
        "oneshot": oneshot,
        "phase_ms": {"target_spectra": round(float(np.mean(spec_ms)), 3),
                     "correlate": round(float(np.mean(cor_ms)), 3), "scan": round(float(np.mean(scan_ms)), 3),
                     "finalize": round(float(np.mean(fin_ms)), 3)},
        "results": n_res,
        "candidate_lags": cands,
        "diagonal_runs": runs,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": int(TRAFFIC_PER_PAIR * args.pairs) if TRAFFIC_PER_PAIR else None,
                     "traffic_source": TRAFFIC_SOURCE if TRAFFIC_PER_PAIR else None,
                     "kernel": kernel_name(ctx), "kernel_ms": round(k_ms, 4), "bytes_per_pair": B_PAIR,
                     "copy_gbs": copy["gb_per_s"] if copy else None,
                     "frac_vs_copy": round(achieved / copy["gb_per_s"], 4) if copy else None},
    }
    if tune:
        out["autotune"] = {"seconds": round(tune_s, 2), "classes": {
            str(N): {"pairs_kernel": list(ctx.geometry(N, "pairs")), "spectra_kernel": list(ctx.geometry(N, "spectra")),
                     "candidates": [list(e["geo"]) + [round(e["correlate_ms"], 4), round(e["spectra_ms"], 4)]
                                    for e in tune if e["N"] == N]}
            for N in sorted({e["N"] for e in tune})},
            "columns": "stage_a, round_levels, three_per_cu, lean_lds, pair kernel ms, spectra kernel ms "
                       "(synthetic batch per length)"}
    if args.uneven_pairs > 0:
        out["uneven"] = uneven_bench(ctx, args.uneven_pairs, args.uneven_steps)
    if not args.no_serve:
        sv = serve_bench(ctx, args.serve_bp, args.serve_seconds, value, SERVE_LOADS)
        out["serve_p50_ms"] = sv["p50_ms"]
        out["serve_p99_ms"] = sv["p99_ms"]
        out["serve"] = sv
    if not args.no_large:
        out["serve_large"] = serve_bench(ctx, args.large_bp, args.serve_seconds, value, LARGE_LOADS, seed=5, large=True,
                                         search_s=12.0)
    if not args.no_serve:
        print("[bench] kmatch: GPU program beside the reference", file=sys.stderr, flush=True)
        out["kmatch"] = kmatch_bench(ctx)
    if copy:
        out["hbm_copy"] = copy
    if not args.no_cpu:
        ncpu, avail = host_cpus()
        threads = max(1, avail)
        print(f"[bench] cpu baseline on {threads} threads (host: {ncpu} CPUs)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(T, Q, min(args.cpu_sample, args.pairs), threads)
        out["cpu_baseline"]["host_cpus"] = ncpu
        share = max(1, min(avail, ncpu // 8))  # one of the node's eight GPUs' share of its host cores
        if share != threads:
            print(f"[bench] cpu baseline on {share} threads (one GPU's share)", file=sys.stderr, flush=True)
            out["cpu_baseline_gpu_share"] = cpu_baseline(T, Q, min(args.cpu_sample // 2, args.pairs), share)
        else:
            out["cpu_baseline_gpu_share"] = dict(out["cpu_baseline"])
        out["cpu_baseline_gpu_share"]["share"] = f"min(available CPUs, host CPUs / 8) = min({avail}, {ncpu} // 8)"
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
