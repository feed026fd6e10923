"""The learned cost model: traces from the exact aligner, GPU training, and what it buys.

Traces come from the engine itself (Engine.set_tracing): for every candidate pair, 12 seed
features and the DP cells its Process_Matches consumed plus the exit diagonal of its first
extension.  The model (12 -> 64 -> 64 -> 2 MLP, FP32) is trained on the GPU by the library's own
forward / backward / Adam kernels.  It only orders work: OIC_SCHED_LEARNED sorts candidate pairs
and the packed DP tasks longest-predicted-first; OIC_SCHED_BATCHED uses the static length-based
estimate static_cost() in its place, which is also the yardstick for the model's prediction error.
"""
import time

import numpy as np

from canu_amd import oic, readsim


def static_cost(feat):
    """Length-based estimate (the fallback): one extension per match node over the shorter read."""
    nnodes = np.expm1(feat[:, 3])
    minlen = np.expm1(feat[:, 10])
    consistent = feat[:, 8] > 0.5
    return np.where(consistent, 1.0, nnodes) * minlen


def collect_traces(preset="c3", genome_len=25_000_000, seed=4, engine_kw=None, batch=4000, max_traces=None):
    """Run the exact aligner over a fresh instance of a workload generator and return its traces."""
    cfg = readsim.PRESETS[preset].scaled(genome_len, seed=seed)
    R = readsim.simulate(cfg)
    kw = dict(kmer=22, maxerate=0.045, minlength=500, hashbits=26, hashload=0.8, hashdatalen=2_000_000_000)
    if engine_kw:
        kw.update(engine_kw)
    kw["schedule"] = oic.SCHED_BATCHED
    eng = oic.Engine(**kw)
    eng.load_reads(R)
    eng.build_hash_index(1, R.n)
    eng.set_tracing(True)
    for bgn in range(1, R.n + 1, batch):
        eng.search(bgn, min(R.n, bgn + batch - 1))
        if max_traces and eng.traces()[0].shape[0] >= max_traces:
            break
    f, t = eng.traces()
    eng.close()
    return f, t


def makespan(costs, order, workers):
    """List scheduling: workers take jobs in order as they free up; returns the finish time."""
    import heapq
    h = [0.0] * workers
    heapq.heapify(h)
    for j in order:
        t = heapq.heappop(h)
        heapq.heappush(h, t + costs[j])
    return max(h)


def ordering_gain(feat, targ, pred, workers=4096, chunk=80_000, rng=0):
    """Makespans of discovery, static-estimate and learned longest-first orders on true costs.

    Jobs are taken in chunks of the size of one search batch, as the engine schedules them."""
    r = np.random.default_rng(rng)
    true = np.expm1(targ[:, 0].astype(np.float64))
    st = static_cost(feat)
    out = {"plain": 0.0, "static": 0.0, "learned": 0.0, "ideal": 0.0}
    n = len(true)
    for c0 in range(0, n, chunk):
        idx = np.arange(c0, min(n, c0 + chunk))
        c = true[idx]
        out["plain"] += makespan(c, np.arange(len(idx)), workers)
        out["static"] += makespan(c, np.argsort(-st[idx], kind="stable"), workers)
        out["learned"] += makespan(c, np.argsort(-pred[idx], kind="stable"), workers)
        out["ideal"] += makespan(c, np.argsort(-c, kind="stable"), workers)
    return out


def prediction_error(targ, pred, feat):
    """Rank correlation of predicted vs true cells, for the model and the static estimate."""
    from scipy.stats import spearmanr
    return {"learned_spearman": float(spearmanr(pred, targ[:, 0]).correlation),
            "static_spearman": float(spearmanr(static_cost(feat), targ[:, 0]).correlation),
            "learned_rmse_log": float(np.sqrt(np.mean((pred - targ[:, 0]) ** 2)))}


def train_for(eng, R=None, engine_kw=None, genome_len=25_000_000, seed=4, epochs=12, max_traces=2_000_000, preset="c3"):
    """Collect traces on a fresh instance (another seed) of a workload generator, train eng's model on
    the GPU; report quality."""
    t0 = time.perf_counter()
    f, t = collect_traces(preset, genome_len, seed, {k: v for k, v in (engine_kw or {}).items() if k != "schedule"},
                          max_traces=max_traces)
    t_tr = time.perf_counter() - t0
    n = len(f)
    perm = np.random.default_rng(seed).permutation(n)
    cut = int(0.9 * n)
    tr, te = perm[:cut], perm[cut:]
    t0 = time.perf_counter()
    loss = eng.train_model(f[tr], t[tr], epochs=epochs)
    t_fit = time.perf_counter() - t0
    pred = eng.predict(f[te])[:, 0]
    info = {"traces": n, "trace_s": round(t_tr, 1), "train_s": round(t_fit, 2), "train_loss": round(float(loss), 4)}
    info.update({k: round(v, 4) for k, v in prediction_error(t[te], pred, f[te]).items()})
    return info
