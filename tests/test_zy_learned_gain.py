"""What the trained cost model gains over the method without it, measured on the GPU.

The model is trained on traces of one C3-generator instance (seed 4) and used on others it has not
seen.  Against the static length-based estimate (OIC_SCHED_BATCHED) it must:
  * finish the DP kernels of a held-out C3-generator search sooner (device time, best of three runs
    each, same launch shapes), records and DP cells unchanged;
  * run no more speculative DP cells, in fewer speculative tasks, where most pairs are true overlaps
    (C1-generator reads): after a dovetail probe the static rule speculates on every pair with 16 or
    more match nodes, the model only where it expects more than twice the probe's cells; records and
    DP cells unchanged.  The gate decides per pair, and C1's speculative waste is per node (a pair
    has about 62 match nodes, of which Process_Matches extends about 13), so the cut is small: 1-7 %
    of the speculative cells with models trained on 1-2 Mbp instances.
"""
import numpy as np
import pytest

from canu_amd import learned, oic, readsim

pytestmark = pytest.mark.gpu


def same(a, b):
    a, b = oic.sort_records(a.records), oic.sort_records(b.records)
    return len(a) == len(b) and all(np.array_equal(a[f], b[f]) for f in ("a_iid", "b_iid", "w0", "w1"))


def best_of(eng, runs=3):
    best, res = None, None
    for _ in range(runs):
        r = eng.run()
        ms = r.counters["ms_dp_kernel"]
        if best is None or ms < best:
            best, res = ms, r
    return best, res


def test_learned_order_beats_static_in_dp_time():
    R = readsim.simulate(readsim.PRESETS["c3"].scaled(6_000_000, seed=9))
    kw = dict(kmer=22, maxerate=0.045, minlength=500, hashbits=26, hashload=0.8, hashdatalen=2_000_000_000, threads=1)
    static = oic.Engine(schedule=oic.SCHED_BATCHED, **kw)
    model = oic.Engine(schedule=oic.SCHED_LEARNED, **kw)
    info = learned.train_for(model, engine_kw=kw, genome_len=4_000_000, seed=4, epochs=10, max_traces=600_000)
    print(f"\ncost model: {info}")
    out = {}
    for name, eng in (("static", static), ("learned", model)):
        eng.set_tuning(fwd_lanes=4, fwd_tile=64, fwd_loop=2, fwd_graph=1)
        eng.load_reads(R)
        out[name] = best_of(eng)
        print(f"{name}: DP kernels {out[name][0]:.2f} ms, cells {out[name][1].counters['dp_cells']}")
    assert out["learned"][1].counters["learned_pairs"] == out["learned"][1].counters["pairs"] > 10_000
    assert same(out["static"][1], out["learned"][1])
    assert out["static"][1].counters["dp_cells"] == out["learned"][1].counters["dp_cells"]
    assert out["learned"][0] < out["static"][0]


def test_learned_speculation_gate_runs_fewer_tasks_on_true_overlaps():
    R = readsim.simulate(readsim.PRESETS["c1"].scaled(1_000_000, seed=19))
    kw = dict(kmer=22, maxerate=0.045, minlength=500, threads=1)
    static = oic.Engine(schedule=oic.SCHED_BATCHED, **kw)
    model = oic.Engine(schedule=oic.SCHED_LEARNED, **kw)
    info = learned.train_for(model, engine_kw=kw, genome_len=2_000_000, seed=4, epochs=10, max_traces=400_000, preset="c1")
    print(f"\ncost model: {info}")
    res = {}
    for name, eng in (("static", static), ("learned", model)):
        eng.load_reads(R)
        res[name] = eng.run()
        c = res[name].counters
        print(f"{name}: speculative cells {c['spec_cells']}, used cells {c['dp_cells']}, records {len(res[name].records)}")
    assert len(res["static"].records) > 1000
    assert same(res["static"], res["learned"])
    assert res["static"].counters["dp_cells"] == res["learned"].counters["dp_cells"]
    assert res["learned"].counters["spec_cells"] <= res["static"].counters["spec_cells"]
    assert res["learned"].counters["spec_tasks"] < res["static"].counters["spec_tasks"]
