"""overlapInCore throughput on one MI355X: DP cells/s and overlaps/s, learned-binned vs plain.

Workload (default c3): 50 Mbp genome, 25 % of it interspersed repeats (300 families of 3-8 kbp,
copies 1-5 % diverged) and 20 tandem arrays of a 171 bp unit (50-200 kbp, 0.5 % unit divergence);
30x lognormal reads (12 kbp mean, 5 kbp sd, <= 60 kbp), 1.5/1.5/1.0 % sub/ins/del; k=22,
--maxerate 0.045, --minlength 500; one hash block over all reads (--hashbits 26 --hashload 0.8
--hashdatalen 2e9).  Reads are packed into HBM and the k-mer table is built before timing; a step
is one overlapInCore search of a batch of ref reads (both orientations) against the whole table,
records copied back to host memory.

DP cells are the (e, d) updates of the forward / reverse DP calls whose results Process_Matches
uses (the reference's count without the reverse DPs it runs after a forward DP that did not reach
the end); the engine's count equals the oracle's exactly.  The learned schedule's cost model is
trained first, on the GPU, on traces of a fresh c3-generator instance (seed 4).

Prints one JSON line.  --gpus must be 1: this build runs on one GPU.
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_INT_TOPS = 256 * 64 * 2.4e9 / 1e12    # 256 CUs x 64 int32 lanes/clk x 2.4 GHz
PEAK_HBM_GBS = 8000.0

WORKLOADS = {
    "c1": dict(preset="c1", hashbits=23, hashload=0.8, hashdatalen=200_000_000,
               desc="2 Mbp uniform genome, 15x lognormal 8k/3k (min 1k), 0.5/0.5/0.5% sub/ins/del"),
    "c2": dict(preset="c2", hashbits=26, hashload=0.8, hashdatalen=2_000_000_000,
               desc="50 Mbp uniform genome, 30x lognormal 12k/5k (max 60k), 1.5/1.5/1.0% sub/ins/del"),
    "c3": dict(preset="c3", hashbits=26, hashload=0.8, hashdatalen=2_000_000_000,
               desc="c2 with 25% repeats: 300 families 3-8 kbp (1-5% diverged), 20 satellite arrays of a 171 bp unit"),
    "c5": dict(preset="c5", hashbits=27, hashload=0.8, hashdatalen=4_000_000_000, maxerate=0.01, minlength=1000,
               desc="100 Mbp genome with the c3 repeat mix, 30x lognormal 18k/4k, 0.2% errors mostly homopolymer "
                    "indels, homopolymer-compressed reads"),
}
# The CPU baseline's sample: ref reads of the same instance.  The oracle (like the reference) gives
# each thread whole ref reads, so one read from a satellite array or a large repeat family is one
# thread's serial job of up to minutes (read 22 of c3: 3.2e9 DP cells, about 280 s on one core).  On
# c3 the sample is refs 1-1200 without the 55 reads whose search needs more than 5e8 DP cells (as
# the engine counts them; the generator is seeded, so the list is fixed): 7.8e9 DP cells in the
# 1145 others, 1.4e11 in those 55.
C3_HEAVY = (22, 43, 137, 172, 173, 180, 184, 186, 230, 233, 237, 250, 269, 301, 336, 339, 381, 386, 393, 408,
            410, 426, 427, 452, 459, 476, 494, 508, 510, 541, 557, 575, 588, 609, 662, 685, 692, 733, 741, 805,
            851, 859, 887, 916, 967, 1021, 1030, 1044, 1070, 1074, 1117, 1147, 1175, 1182, 1193)
CPU_SAMPLE = {
    "c1": ((1, 0), "c1 is searched whole (all reads against all)."),
    "c2": ((1, 24), ""),
    "c3": ([r for r in range(1, 1201) if r not in C3_HEAVY],
           "Refs 1-1200 without the 55 whose search needs more than 5e8 DP cells each (satellite and "
           "repeat-family reads: 1.4e11 DP cells together, against 7.8e9 in the 1145 others)."),
    "c5": ((1, 16), ""),
}
for _k, _w in WORKLOADS.items():
    _w.setdefault("maxerate", 0.045)
    _w.setdefault("minlength", 500)
    _w["cpu_refs"], _w["cpu_note"] = CPU_SAMPLE[_k]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_from_profiles(kernel):
    """HBM bytes per dispatch of kernel from the committed PMC summary, or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"\s*(\S+)\s+hbm_bytes_per_dispatch=(\d+)", line)
        if m and m.group(1) == kernel.replace(" ", ""):
            return int(m.group(2))
    return None


def host_cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(W, R, refs, threads):
    """The oracle ('port') on the same instance: ref reads refs[0]..refs[1] (both orientations)
    against an index over every read, timed on host cores.  The oracle indexes only the k-mers
    those ref reads hold (its restricted-index mode), which searches exactly like the full
    index; the index build is not timed."""
    from oracle import oracle
    kw = dict(kmer=22, maxErate=W["maxerate"], minOlap=W["minlength"], hashBits=28, hashLoad=0.8,
              hashDataLen=8_000_000_000, nThreads=threads, ompThreads=threads, restrictIndex=1)
    if isinstance(refs, list):
        out = oracle.run(R, refs=refs, **kw)
        what = f"{len(refs)} ref reads ({refs[0]}-{refs[-1]})"
    else:
        b, e = refs
        out = oracle.run(R, bgnRef=b, endRef=e, **kw)
        what = f"ref reads {b}-{e}"
    return dict(value=out["cells_used"] / out["t_search"], unit="DP cells/s", cores=threads, kind="port",
                cpu=host_cpu_model(),
                sample=f"oracle (C restatement, OpenMP over ref reads) searching {what} of the same "
                       f"{W['preset']} instance against all {R.n} reads: {out['cells_used']:.3g} DP cells in "
                       f"{out['t_search']:.1f} s on {threads} threads ({out['pairs']} candidate pairs, "
                       f"{len(out['records'])} overlaps); index build {out['t_index']:.1f} s, not counted. " + W["cpu_note"],
                overlaps_per_s=len(out["records"]) / out["t_search"])


def run_steps(eng, n, batch, steps, warmup, first):
    per = []
    c0 = eng.counters()
    recs = 0
    for i in range(warmup + steps):
        bgn = first + i * batch
        end = min(n, bgn + batch - 1)
        cs = eng.counters()
        t0 = time.perf_counter()
        res = eng.search(bgn, end)
        dt = time.perf_counter() - t0
        if i >= warmup:
            per.append(dt)
            recs += len(res.records)
        if i == warmup - 1:
            c0 = eng.counters()
        ce = eng.counters()
        stages = " ".join(f"{k[3:]} {ce[k] - cs[k]:.0f}" for k in ("ms_gather", "ms_lists", "ms_extend", "ms_alloc"))
        log(f"  step {i - warmup:+d}: ref {bgn}-{end} {dt*1e3:.1f} ms ({stages} ms; {ce['allocs'] - cs['allocs']} allocations), "
            f"{len(res.records)} overlaps")
    c1 = eng.counters()
    d = {k: c1[k] - c0[k] for k in c1}
    return per, recs, d


def timed_sample(eng, bgn, end):
    """One search over a fixed ref range: (seconds, records, counter deltas)."""
    c0 = eng.counters()
    t0 = time.perf_counter()
    res = eng.search(bgn, end)
    dt = time.perf_counter() - t0
    c1 = eng.counters()
    return dt, res.records, {k: c1[k] - c0[k] for k in c1}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--batch", type=int, default=2500, help="ref reads per step")
    ap.add_argument("--schedule", default="learned", choices=["plain", "batched", "learned"])
    ap.add_argument("--no-plain", action="store_true", help="skip the plain one-pair-per-wavefront comparison")
    ap.add_argument("--plain-refs", type=int, default=200, help="ref reads of the plain-vs-scheduled comparison sample")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--fwd-lanes", type=int, default=0, help="fix the packed forward kernel's lanes per DP (skips autotuning)")
    ap.add_argument("--fwd-tile", type=int, default=0, help="fix its band tile (with --fwd-lanes)")
    ap.add_argument("--fwd-loop", type=int, default=-1,
                    help="fix its loop form (0 row-stepped, 1 per task, 2 flat rows; with --fwd-lanes; default: "
                         "flat rows where the shape allows them, else row-stepped)")
    ap.add_argument("--fwd-counters", action="store_true", help="count rows, lane slots and slide iterations of the packed DPs")
    ap.add_argument("--tune-refs", type=int, default=200, help="ref reads the autotuner times each launch shape on")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--train-genome", type=int, default=12_000_000, help="genome of the c3 instance the model learns from")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only copy BYTES between two device buffers (the reference kernel for HBM byte counters)")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch
        n = args.copy_calibration
        src = torch.randint(0, 255, (n,), dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n, "launches": 3, "ok": bool(torch.equal(src, dst))}), flush=True)
        return

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(overlap jobs shard by hash block x ref range with no exchange between GPUs)", file=sys.stderr)
        sys.exit(2)

    from canu_amd import oic, readsim
    W = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    R = readsim.simulate(readsim.PRESETS[W["preset"]])
    log(f"simulated {R.n} reads, {int(R.length.sum())/1e9:.3f} Gbp in {time.perf_counter()-t0:.1f} s")

    sched = dict(plain=oic.SCHED_PLAIN, batched=oic.SCHED_BATCHED, learned=oic.SCHED_LEARNED)
    kw = dict(kmer=22, maxerate=W["maxerate"], minlength=W["minlength"], hashbits=W["hashbits"], hashload=W["hashload"],
              hashdatalen=W["hashdatalen"], threads=1)
    eng = oic.Engine(schedule=sched[args.schedule], **kw)
    t0 = time.perf_counter()
    eng.load_reads(R)
    t_load = time.perf_counter() - t0
    t0 = time.perf_counter()
    last = eng.build_hash_index(1, R.n)
    t_index = time.perf_counter() - t0
    log(f"load {t_load:.2f} s, index over reads 1-{last} in {t_index:.2f} s")
    info = None
    if args.schedule == "learned":
        from canu_amd import learned
        t0 = time.perf_counter()
        # the workload's own generator with a fresh seed (c3: the configured cost model's training set)
        glen = min(args.train_genome, readsim.PRESETS[W["preset"]].genome_len)
        info = learned.train_for(eng, engine_kw=kw, genome_len=glen, preset=W["preset"])
        info["trained_on"] = f"{W['preset']} generator, seed 4, {glen / 1e6:g} Mbp genome"
        log(f"cost model: {info} ({time.perf_counter() - t0:.1f} s)")

    first = 1
    timed_end = first + (args.warmup + args.steps) * args.batch - 1
    if args.fwd_lanes:
        shape = dict(fwd_lanes=args.fwd_lanes, fwd_tile=args.fwd_tile or 128)
        flat_ok = shape["fwd_lanes"] in (4, 8) and shape["fwd_tile"] in (32, 64, 128)
        shape["fwd_loop"] = args.fwd_loop if args.fwd_loop >= 0 else (2 if flat_ok else 0)
        eng.set_tuning(**shape)
    if args.fwd_counters:
        eng.set_tuning(fwd_counters=1)
    tuning = eng.tuning()
    if not args.no_autotune and not args.fwd_lanes and args.schedule != "plain":
        tb = min(R.n - 1, timed_end + 1)
        t0 = time.perf_counter()
        tuning, trials = eng.autotune(tb, min(R.n, tb + args.tune_refs - 1))
        log(f"autotune over refs {tb}-{tb + args.tune_refs - 1}: {tuning} ({time.perf_counter() - t0:.1f} s)")
        for t in trials:
            log(f"    {t}")

    per, recs, d = run_steps(eng, R.n, args.batch, args.steps, args.warmup, first)
    tot = sum(per)
    cells_s = d["dp_cells"] / tot
    olap_s = recs / tot
    steps = args.steps
    # the dominant DP kernel by measured device time: the packed forward kernel or Process_Matches
    lanes, tile = tuning["fwd_lanes"], tuning["fwd_tile"]
    stats_flag = "true" if tuning.get("fwd_counters") else "false"
    if lanes == 1:
        fwd_name = f"k_fwd_lane<{tile}, {stats_flag}>"
    else:
        form = {0: "k_fwd_sm", 1: "k_fwd_tasks", 2: "k_fwd_flat"}[tuning.get("fwd_loop", 0)]
        fwd_name = f"{form}<{lanes}, {tile}, {stats_flag}>"
    kernels = [dict(kernel="k_process_pairs<512>", ms=d["ms_pp_kernel"], cells=d["pp_cells"], bases=d["pp_bases"],
                    launches=d["pp_launches"],
                    timing="HIP events around each launch on its stream")]
    if args.schedule != "plain":
        kernels.append(dict(kernel=fwd_name, ms=d["ms_fwd_narrow"], cells=d["fwd_cells"], bases=d["fwd_bases"],
                            launches=d["fwd_launches"],
                            timing="HIP events around each forward-pass hipGraph on its stream; inside it, each packed "
                                   "launch's span from its first wavefront's start to its last one's end (device wall clock)"))
    dom = max(kernels, key=lambda k: k["ms"])
    kname, kms, nl = dom["kernel"], dom["ms"], dom["launches"]
    ops = 4 * dom["cells"] + 2 * dom["bases"] / 16.0
    achieved = ops / (kms / 1e3) / 1e12 if kms > 0 else 0.0
    traffic = traffic_from_profiles(kname)
    hbm = None
    if traffic and nl:
        gbs = traffic / (kms / nl / 1e3) / 1e9
        hbm = {"achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": gbs / PEAK_HBM_GBS,
               "bytes_per_launch": traffic, "source": "profiles/pmc_traffic.txt (rocprofv3 FETCH_SIZE, WRITE_SIZE)"}
    line = {
        "metric": "overlaps/s + DP cells/s on 1 GPU, learned-binned vs plain, bit-exact",
        "value": cells_s, "unit": "DP cells/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * tot / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32", "data": f"synthetic ({W['desc']})",
        "config": {"workload": f"{args.workload}: {W['desc']}; k=22 --maxerate {W['maxerate']} --minlength {W['minlength']}",
                   "ref_reads_per_step": args.batch, "reads": R.n, "schedule": args.schedule},
        "overlaps_per_s": olap_s,
        "dp_cells_per_step": d["dp_cells"] / steps,
        "index_build_s": t_index,
        "tuning": tuning,
        "cost_model": info,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_INT_TOPS, "unit": "TOPS",
                     "frac": achieved / PEAK_INT_TOPS, "traffic": traffic, "hbm": hbm,
                     "kernel": kname, "kernel_ms_per_step": kms / steps, "int_ops_per_step": ops / steps,
                     "launches_per_step": (nl / steps) if nl else None, "timing": dom["timing"],
                     "graph_ms_per_step": d.get("ms_fwd_graph", 0.0) / steps,
                     "avg_launch_ms": (kms / nl) if nl else None,
                     "int_ops_per_launch": (ops / nl) if nl else None,
                     "kernels_ms_per_step": {k["kernel"]: k["ms"] / steps for k in kernels}},
        "counters_per_step": {k: (v / steps) for k, v in d.items()},
    }
    if not args.no_plain and args.schedule != "plain":
        # bounded comparison: the same ref reads through the scheduled engine and the plain one
        pb, pe = first, min(R.n, first + args.plain_refs - 1)
        ts, rs, ds = timed_sample(eng, pb, pe)
        eng.close()
        eng = oic.Engine(schedule=oic.SCHED_PLAIN, **kw)
        eng.load_reads(R)
        eng.build_hash_index(1, R.n)
        log(f"plain one-pair-per-wavefront path on refs {pb}-{pe}:")
        tp, rp, dp = timed_sample(eng, pb, pe)
        assert ds["dp_cells"] == dp["dp_cells"], "plain and scheduled paths disagree on DP cells"
        a, b = oic.sort_records(rs), oic.sort_records(rp)
        assert len(a) == len(b) and all((a[f] == b[f]).all() for f in ("a_iid", "b_iid", "w0", "w1")), \
            "plain and scheduled paths disagree"
        line["plain"] = {"sample": f"ref reads {pb}-{pe} against the whole index", "dp_cells": dp["dp_cells"],
                         "plain_s": tp, "scheduled_s": ts, "plain_dp_cells_per_s": dp["dp_cells"] / tp,
                         "scheduled_dp_cells_per_s": ds["dp_cells"] / ts}
        line["learned_vs_plain" if args.schedule == "learned" else "batched_vs_plain"] = tp / ts
        log(f"  plain {tp:.2f} s, scheduled {ts:.2f} s: {tp / ts:.2f}x")
    eng.close()
    if not args.no_cpu:
        threads = min(16, len(os.sched_getaffinity(0)))
        refs = W["cpu_refs"] if (isinstance(W["cpu_refs"], list) or W["cpu_refs"][1]) else (1, R.n)
        log(f"cpu baseline: oracle on {len(refs) if isinstance(refs, list) else refs} refs, {threads} threads")
        line["cpu_baseline"] = cpu_baseline(W, R, refs, threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
