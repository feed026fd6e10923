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
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_from_profiles(kernel):
    """HBM bytes per dispatch of kernel from the committed PMC summary, or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        m = re.match(r"\s*(\S+)\s+hbm_bytes_per_dispatch=(\d+)", line)
        if m and m.group(1) == kernel:
            return int(m.group(2))
    return None


def cpu_baseline(preset, genome_len, threads):
    """The oracle ('port') on the same generator at a smaller genome, timed on host cores."""
    from canu_amd import readsim
    from oracle import oracle
    cfg = readsim.PRESETS[preset].scaled(genome_len, seed=readsim.PRESETS[preset].seed)
    R = readsim.simulate(cfg)
    out = oracle.run(R, kmer=22, maxErate=0.045, minOlap=500, nThreads=threads, ompThreads=threads)
    return dict(value=out["cells_used"] / out["t_search"], unit="DP cells/s", cores=threads, kind="port",
                sample=f"oracle on the {preset} generator at a {genome_len/1e6:g} Mbp genome ({R.n} reads, all-vs-all): "
                       f"{out['cells_used']:.3g} used cells ({out['cells']:.3g} incl. unused reverse DPs) in "
                       f"{out['t_search']:.1f} s of search (+{out['t_index']:.1f} s single-thread index build, not counted); "
                       f"{len(out['records'])} overlaps",
                overlaps_per_s=len(out["records"]) / out["t_search"])


def run_steps(eng, n, batch, steps, warmup, first):
    per = []
    c0 = eng.counters()
    recs = 0
    for i in range(warmup + steps):
        bgn = first + i * batch
        end = min(n, bgn + batch - 1)
        t0 = time.perf_counter()
        res = eng.search(bgn, end)
        dt = time.perf_counter() - t0
        if i == warmup - 1 or (warmup == 0 and i == 0 and not per):
            pass
        if i >= warmup:
            per.append(dt)
            recs += len(res.records)
        if i == warmup - 1:
            c0 = eng.counters()
        log(f"  step {i - warmup:+d}: ref {bgn}-{end} {dt*1e3:.1f} ms, {len(res.records)} overlaps")
    c1 = eng.counters()
    d = {k: c1[k] - c0[k] for k in c1}
    return per, recs, d


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--batch", type=int, default=2500, help="ref reads per step")
    ap.add_argument("--schedule", default="learned", choices=["plain", "batched", "learned"])
    ap.add_argument("--no-plain", action="store_true", help="skip the plain one-pair-per-wavefront comparison")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-genome", type=int, default=500_000)
    ap.add_argument("--train-genome", type=int, default=12_000_000, help="genome of the c3 instance the model learns from")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the overlap jobs shard by hash block x ref range with no exchange, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from canu_amd import oic, readsim
    W = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    R = readsim.simulate(readsim.PRESETS[W["preset"]])
    log(f"simulated {R.n} reads, {int(R.length.sum())/1e9:.3f} Gbp in {time.perf_counter()-t0:.1f} s")

    sched = dict(plain=oic.SCHED_PLAIN, batched=oic.SCHED_BATCHED, learned=oic.SCHED_LEARNED)
    kw = dict(kmer=22, maxerate=0.045, minlength=500, hashbits=W["hashbits"], hashload=W["hashload"],
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
        info = learned.train_for(eng, engine_kw=kw, genome_len=args.train_genome)
        log(f"cost model: {info} ({time.perf_counter() - t0:.1f} s)")

    first = 1
    per, recs, d = run_steps(eng, R.n, args.batch, args.steps, args.warmup, first)
    tot = sum(per)
    cells_s = d["dp_cells"] / tot
    olap_s = recs / tot
    kern_s = d["ms_dp_kernel"] / 1e3
    ops = 4 * d["dp_cells"] + 2 * d["dp_bases"] / 16.0
    achieved = ops / kern_s / 1e12 if kern_s > 0 else 0.0
    launches = max(1, args.steps)
    line = {
        "metric": "overlaps/s + DP cells/s on 1 GPU, learned-binned vs plain, bit-exact",
        "value": cells_s, "unit": "DP cells/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * tot / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32", "data": f"synthetic ({W['desc']})",
        "config": {"workload": f"{args.workload}: {W['desc']}; k=22 --maxerate 0.045 --minlength 500",
                   "ref_reads_per_step": args.batch, "reads": R.n, "schedule": args.schedule},
        "overlaps_per_s": olap_s,
        "dp_cells_per_step": d["dp_cells"] / args.steps,
        "index_build_s": t_index,
        "cost_model": info if args.schedule == "learned" else None,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_INT_TOPS, "unit": "TOPS",
                     "frac": achieved / PEAK_INT_TOPS, "traffic": traffic_from_profiles("k_process_pairs"),
                     "kernel": "k_process_pairs", "kernel_ms_per_step": d["ms_dp_kernel"] / launches,
                     "ops_per_step": ops / launches},
    }
    if not args.no_plain and args.schedule != "plain":
        eng.close()
        eng = oic.Engine(schedule=oic.SCHED_PLAIN, **kw)
        eng.load_reads(R)
        eng.build_hash_index(1, R.n)
        log("plain one-pair-per-wavefront path:")
        per_p, recs_p, dp = run_steps(eng, R.n, args.batch, args.steps, args.warmup, first)
        tp = sum(per_p)
        line["plain"] = {"dp_cells_per_s": dp["dp_cells"] / tp, "overlaps_per_s": recs_p / tp, "ms_per_step": 1e3 * tp / args.steps}
        line["learned_vs_plain"] = cells_s / (dp["dp_cells"] / tp) if dp["dp_cells"] else None
        assert recs_p == recs, "plain and scheduled paths disagree"
    eng.close()
    if not args.no_cpu:
        threads = min(16, len(os.sched_getaffinity(0)))
        log(f"cpu baseline: oracle, {threads} threads")
        line["cpu_baseline"] = cpu_baseline(W["preset"], args.cpu_genome, threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
