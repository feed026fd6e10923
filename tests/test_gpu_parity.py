"""GPU engine (liboic_hip.so through the C ABI) against the oracle, bit for bit.

DP units: prefixEditDistance::forward / ::reverse / Extend_Alignment on random and mutated
windows (with 'n', exact matches, zero error limits, long reads whose bands need the wide tile).
Whole runs: sorted 24-byte records and the .stats counters on small versions of the configured
workloads and on every option the engine supports.
"""
import numpy as np
import pytest

from canu_amd import oic, readsim
from oracle import oracle

pytestmark = pytest.mark.gpu


def mutate(rng, s, rate):
    out = []
    for ch in s:
        u = rng.random()
        if u < rate / 3:
            out.append(rng.choice([c for c in "ACGT" if c != ch]))
        elif u < 2 * rate / 3:
            continue
        elif u < rate:
            out.append(rng.choice(list("ACGT")))
            out.append(ch)
        else:
            out.append(ch)
    return "".join(out)


def dp_cases(seed, n, lmin, lmax):
    rng = np.random.default_rng(seed)
    seqs = []
    for _ in range(n):
        L = int(rng.integers(lmin, lmax))
        a = "".join(rng.choice(list("ACGT"), L))
        b = mutate(rng, a, float(rng.choice([0.0, 0.01, 0.03, 0.06, 0.1, 0.25])))
        if rng.random() < 0.3:
            b += "".join(rng.choice(list("ACGT"), int(rng.integers(0, 300))))
        if rng.random() < 0.1 and L > 4:
            k = int(rng.integers(0, L))
            a = a[:k] + "N" + a[k + 1:]
        seqs += [a, b]
    return seqs


def check_dp(seqs, rate, limit_slack):
    R = readsim.from_strings(seqs)
    eng = oic.Engine(kmer=22, maxerate=rate)
    eng.load_reads(R)
    od = oracle.OracleDP(rate)
    tasks = []
    for i in range(0, len(seqs), 2):
        ia, ib = i + 1, i + 2
        if len(seqs[i]) > len(seqs[i + 1]):
            ia, ib = ib, ia
        m, n = len(seqs[ia - 1]), len(seqs[ib - 1])
        lim = min(int(np.ceil(m * od.tables.maxErate)) + limit_slack, od.tables.MAX_ERRORS - 1)
        tasks.append((ia, ib, 0, 0, 0, 0, m, n, lim, 0))
        tasks.append((ia, ib, 0, 0, m - 1, n - 1, m, n, lim, 1))
        tasks.append((ia, ib, 0, 0, 0, 0, m, n, 0, 0))
    t = np.array(tasks, oic.DP_TASK)
    res, delta = eng.prefix_dp(t, delta_stride=8192)
    od.reset_counters()
    for k, tk in enumerate(t):
        A = seqs[tk["a_id"] - 1].lower().encode()
        B = seqs[tk["b_id"] - 1].lower().encode()
        r = res[k]
        assert r["status"] == 0
        if tk["dir"] == 0:
            e, ae, te, mte, d = od.forward(A, B, int(tk["error_limit"]))
        else:
            e, ae, te, lo, mte, d = od.reverse(A, B, int(tk["error_limit"]))
            assert r["leftover"] == lo, k
        assert (r["errors"], r["a_end"], r["t_end"], bool(r["match_to_end"])) == (e, ae, te, mte), (k, tk)
        assert r["delta_len"] == len(d) and np.array_equal(delta[k][:len(d)], d), k
    assert int(res["cells"].sum()) == od.counters()["cells"]


def test_dp_units_short():
    check_dp(dp_cases(1, 300, 1, 600), 0.045, 3)


def test_dp_units_long_wide_band():
    # 8-30 kbp windows at up to 25 % divergence: bands grow past the small tile
    check_dp(dp_cases(2, 24, 8000, 30000), 0.045, 3)


def test_dp_units_other_rates():
    check_dp(dp_cases(3, 120, 50, 3000), 0.01, 2)
    check_dp(dp_cases(4, 120, 50, 3000), 0.12, 2)


def test_extend_alignment_units():
    rng = np.random.default_rng(5)
    R = readsim.simulate(readsim.PRESETS["c2"].scaled(150_000, seed=31))
    eng = oic.Engine(kmer=22, maxerate=0.045, minlength=500)
    eng.load_reads(R)
    od = oracle.OracleDP(0.045)
    seqs = [None] + [R.read(i) for i in range(1, R.n + 1)]
    comp = bytes.maketrans(b"ACGT", b"TGCA")
    tasks = []
    for _ in range(400):
        s, t = (int(x) for x in rng.integers(1, R.n + 1, 2))
        if s == t:
            continue
        sd = int(rng.integers(0, 2))
        S = seqs[s] if sd == 0 else seqs[s].translate(comp)[::-1]
        T = seqs[t]
        ln = int(rng.integers(1, 40))
        st = int(rng.integers(0, len(S) - ln))
        of = int(rng.integers(0, len(T) - ln))
        tasks.append((s, t, sd, st, of, ln, S, T))
    arr = np.array([x[:6] for x in tasks], oic.EXT_TASK)
    res, delta = eng.extend_alignment(arr, delta_stride=8192)
    for k, (s, t, sd, st, of, ln, S, T) in enumerate(tasks):
        kind, slo, shi, tlo, thi, err, d = od.extend(S.lower(), T.lower(), st, of, ln)
        r = res[k]
        assert r["status"] == 0
        assert (r["kind"], r["s_lo"], r["s_hi"], r["t_lo"], r["t_hi"], r["errors"]) == (kind, slo, shi, tlo, thi, err), k
        assert np.array_equal(delta[k][:r["delta_len"]], d), k


def run_both(R, threads=4, schedule=oic.SCHED_PLAIN, **kw):
    ekw = dict(kmer=22, maxerate=0.045, minlength=500, threads=threads, schedule=schedule)
    okw = dict(kmer=22, maxErate=0.045, minOlap=500, nThreads=threads, ompThreads=16)
    m = dict(partial=("partial", "partial"), unique=("unique", "unique"), hopeless=("hopeless", "hopeless"),
             minkmers=("minkmers", "minkmers"), hashbits=("hashbits", "hashBits"), hashload=("hashload", "hashLoad"),
             maxerate=("maxerate", "maxErate"), minlength=("minlength", "minOlap"))
    ranges = {}
    tuning = kw.pop("tuning", None)
    for k, v in kw.items():
        if k in ("bgn_ref", "end_ref", "bgn_hash", "end_hash"):
            ranges[k] = v
            continue
        ek, ok = m[k]
        ekw[ek] = v
        okw[ok] = int(v) if isinstance(v, bool) else v
    eng = oic.Engine(**ekw)
    if tuning:
        eng.set_tuning(**tuning)
    eng.load_reads(R)
    res = eng.run(ranges.get("bgn_hash", 1), ranges.get("end_hash", 0xFFFFFFFF), ranges.get("bgn_ref", 1),
                  ranges.get("end_ref", 0xFFFFFFFF))
    ref = oracle.run(R, bgnHash=ranges.get("bgn_hash", 1), endHash=ranges.get("end_hash", 0xFFFFFFFF),
                     bgnRef=ranges.get("bgn_ref", 1), endRef=ranges.get("end_ref", 0xFFFFFFFF), **okw)
    return res, ref


def assert_same(res, ref):
    a = oic.sort_records(res.records)
    b = oracle.sort_records(ref["records"])
    assert len(a) == len(b), (len(a), len(b))
    for f in ("a_iid", "b_iid", "w0", "w1"):
        assert np.array_equal(a[f], b[f]), f
    assert list(res.stats) == list(ref["stats"])


@pytest.fixture(scope="module")
def c1s():
    return readsim.simulate(readsim.PRESETS["c1"].scaled(200_000, seed=41))


@pytest.fixture(scope="module")
def c2s():
    return readsim.simulate(readsim.PRESETS["c2"].scaled(120_000, seed=42))


@pytest.fixture(scope="module")
def c3s():
    return readsim.simulate(readsim.PRESETS["c3"].scaled(300_000, seed=43))


def test_run_c1_small(c1s):
    res, ref = run_both(c1s)
    assert_same(res, ref)
    # the DP work Process_Matches consumes, cell for cell
    assert res.counters["dp_cells"] == ref["cells_used"]
    assert res.counters["extensions"] == ref["extensions"]


@pytest.mark.parametrize("schedule", [oic.SCHED_PLAIN, oic.SCHED_BATCHED])
def test_run_c2_small(c2s, schedule):
    res, ref = run_both(c2s, schedule=schedule)
    assert_same(res, ref)
    assert res.counters["dp_cells"] == ref["cells_used"]
    assert res.counters["extensions"] == ref["extensions"]


@pytest.mark.parametrize("schedule", [oic.SCHED_PLAIN, oic.SCHED_BATCHED])
def test_run_c3_small_repeats(c3s, schedule):
    res, ref = run_both(c3s, schedule=schedule)
    assert_same(res, ref)
    assert res.counters["dp_cells"] == ref["cells_used"]


def test_run_c5_small_hifi_hpc():
    """C5 generator: homopolymer-compressed HiFi-like reads, --maxerate 0.01 --minlength 1000."""
    R = readsim.simulate(readsim.PRESETS["c5"].scaled(400_000, seed=46))
    res, ref = run_both(R, schedule=oic.SCHED_BATCHED, maxerate=0.01, minlength=1000)
    assert len(ref["records"]) > 100
    assert_same(res, ref)
    assert res.counters["dp_cells"] == ref["cells_used"]


@pytest.mark.parametrize("opt", [dict(unique=False), dict(partial=True), dict(hopeless=False), dict(minkmers=True),
                                 dict(maxerate=0.01, minlength=1000), dict(maxerate=0.12)])
def test_run_options(c1s, opt):
    assert_same(*run_both(c1s, **opt))


def test_run_multiple_hash_blocks(c1s):
    res, ref = run_both(c1s, hashbits=12, hashload=0.3)
    assert len(ref["hash_end"]) > 2
    assert_same(res, ref)


def test_run_ranges_and_thread_chunking(c1s):
    n = c1s.n
    for t in (1, 3, 7, 8):
        assert_same(*run_both(c1s, threads=t, bgn_ref=5, end_ref=n - 7, bgn_hash=3, end_hash=n - 2))


def test_run_schedules_identical(c2s):
    base = run_both(c2s)[0]
    for sch in (oic.SCHED_BATCHED, oic.SCHED_LEARNED):
        eng = oic.Engine(kmer=22, maxerate=0.045, minlength=500, threads=4, schedule=sch)
        eng.load_reads(c2s)
        r = eng.run()
        a, b = oic.sort_records(r.records), oic.sort_records(base.records)
        assert len(a) == len(b) and all(np.array_equal(a[f], b[f]) for f in ("a_iid", "b_iid", "w0", "w1"))
        assert r.stats == base.stats


@pytest.mark.parametrize("fixture", ["c1s", "c2s", "c3s"])
@pytest.mark.parametrize("min_hits", [0, 1 << 31])
def test_list_builders(request, fixture, min_hits):
    """Match lists built by one wavefront per pair (min_hits 0) or one thread per pair (2^31)."""
    R = request.getfixturevalue(fixture)
    res, ref = run_both(R, schedule=oic.SCHED_BATCHED, tuning=dict(list_wave_min_hits=min_hits))
    assert_same(res, ref)
    assert res.counters["dp_cells"] == ref["cells_used"]
    if min_hits == 0:
        assert res.counters["list_wave_pairs"] == res.counters["pairs"]
    else:
        assert res.counters["list_wave_pairs"] == 0


def test_forward_launch_shapes_identical(c3s):
    """Every lanes-per-DP x band-tile shape of the packed forward kernel, in both loop forms, gives
    the oracle's output."""
    ref = oracle.run(c3s, kmer=22, maxErate=0.045, minOlap=500, nThreads=4, ompThreads=16)
    eng = oic.Engine(kmer=22, maxerate=0.045, minlength=500, threads=4, schedule=oic.SCHED_BATCHED)
    eng.load_reads(c3s)
    for tile in (16, 32, 64):
        # one DP per lane; the 16-cell ring sends many bands to the retry levels
        eng.set_tuning(fwd_lanes=1, fwd_tile=tile, fwd_loop=0)
        res = eng.run()
        assert_same(res, ref)
        assert res.counters["dp_cells"] == ref["cells_used"], (1, tile)
        if tile == 16:
            assert res.counters["overflow_retries"] > 0
    for loop in (0, 1):
        for lanes in (4, 8, 16, 32):
            for tile in (32, 64, 128, 256, 512):
                if (lanes == 4 and tile == 512) or (tile <= 64 and lanes not in (8, 16)):
                    continue
                eng.set_tuning(fwd_lanes=lanes, fwd_tile=tile, fwd_loop=loop)
                res = eng.run()
                assert_same(res, ref)
                assert res.counters["dp_cells"] == ref["cells_used"], (loop, lanes, tile)
    # flat rows: 4 or 8 lanes per DP for its own state, every row's cells over the whole wavefront
    for lanes in (4, 8):
        for tile in (32, 64, 128):
            eng.set_tuning(fwd_lanes=lanes, fwd_tile=tile, fwd_loop=2)
            res = eng.run()
            assert_same(res, ref)
            assert res.counters["dp_cells"] == ref["cells_used"], (2, lanes, tile)
            if tile == 32:
                assert res.counters["overflow_retries"] > 0
    eng.set_tuning(fwd_lanes=16, fwd_tile=128, fwd_order=1, fwd_loop=0)
    res = eng.run()
    assert_same(res, ref)
    assert res.counters["dp_cells"] == ref["cells_used"]
    # the forward passes as one hipGraph per batch (default) and as direct launches
    for graph in (1, 0):
        eng.set_tuning(fwd_order=0, fwd_graph=graph)
        res = eng.run()
        assert_same(res, ref)
        assert res.counters["dp_cells"] == ref["cells_used"]
        assert (res.counters["graphs_launched"] > 0) == bool(graph)
    with pytest.raises(oic.OicError):
        eng.set_tuning(fwd_lanes=64)
    with pytest.raises(oic.OicError):
        eng.set_tuning(fwd_lanes=1, fwd_tile=128)
    with pytest.raises(oic.OicError):
        eng.set_tuning(fwd_loop=2)       # flat rows with 16 lanes per DP
    with pytest.raises(oic.OicError):
        eng.set_tuning(fwd_loop=3)


def test_forward_loop_forms_do_the_same_work(c3s):
    """The row-stepped loop runs exactly the per-task loop's DPs: the same cells, rows, lane slots
    and bases per speculative pass, in fewer wavefront rows (its slides compare 64-base windows after
    the first 32 bases, so window counts differ); the flat-row form the same cells and rows in fewer
    band chunk iterations."""
    eng = oic.Engine(kmer=22, maxerate=0.045, minlength=500, threads=4, schedule=oic.SCHED_BATCHED)
    eng.load_reads(c3s)
    got = {}
    for loop in (1, 0):
        eng.set_tuning(fwd_lanes=8, fwd_tile=64, fwd_loop=loop, fwd_counters=1)
        got[loop] = eng.run().counters
    for k in ("spec_cells", "spec_rows", "spec_lane_slots", "spec_bases", "dp_cells"):
        assert got[0][k] == got[1][k], k
    assert got[0]["spec_rows"] > 0
    assert got[0]["spec_wave_rows"] < got[1]["spec_wave_rows"]
    # flat rows run the same DPs with their cells spread over the wavefront: fewer chunk iterations
    eng.set_tuning(fwd_lanes=8, fwd_tile=64, fwd_loop=2, fwd_counters=1)
    flat = eng.run().counters
    for k in ("spec_cells", "spec_rows", "dp_cells"):
        assert flat[k] == got[0][k], k
    # a row that reaches an end runs all its cells at once, so the cells past the reaching diagonal
    # compare their bases too
    assert got[0]["spec_bases"] <= flat["spec_bases"] <= 1.001 * got[0]["spec_bases"]
    assert 0 < flat["spec_wave_chunks"] < got[0]["spec_wave_chunks"]


def test_autotune_keeps_results(c3s):
    ref = oracle.run(c3s, kmer=22, maxErate=0.045, minOlap=500, nThreads=1, ompThreads=16)
    eng = oic.Engine(kmer=22, maxerate=0.045, minlength=500, threads=1, schedule=oic.SCHED_BATCHED)
    eng.load_reads(c3s)
    eng.build_hash_index(1, c3s.n)
    best, trials = eng.autotune(1, min(c3s.n, 60))
    # warm-up; the incumbent (flat rows, 4 lanes x 64 cells) and the 18 other forward shapes; the
    # loop form: the incumbent, the other of row-stepped and per-task (a winning shape of more than
    # one lane per DP that is not flat) and flat rows with 4 and 8 lanes per DP unless incumbent; the
    # task order; the incumbent list threshold and 4 others
    assert len(trials) in (1 + 19 + 2 + 2 + 5, 1 + 19 + 3 + 2 + 5, 1 + 19 + 4 + 2 + 5)
    shapes = {(t["fwd_lanes"], t["fwd_tile"]) for t in trials if t["stage"] == "fwd"}
    assert len(shapes) == 19 and (best["fwd_lanes"], best["fwd_tile"]) in shapes
    flat = {(t["fwd_lanes"], t["fwd_tile"]) for t in trials if t["stage"] == "fwd" and t["fwd_loop"] == 2}
    assert flat == {(4, 64), (8, 64)}
    assert eng.tuning() == best
    assert_same(eng.run(), ref)


def test_list_builder_hands_back_crowded_offsets():
    """An S k-mer with more T occurrences than the wavefront builder holds (550 copies of a 20 bp
    unit) is handed back to the one-thread builder; results stay exact."""
    rng = np.random.default_rng(12)
    unit = "".join(rng.choice(list("ACGT"), 20))
    other = "".join(rng.choice(list("ACGT"), 4000))
    t = other[:2000] + unit * 550 + other[2000:]
    s = unit * 30 + other[2000:2600]
    R = readsim.from_strings([s, t])
    res, ref = run_both(R, threads=1, schedule=oic.SCHED_BATCHED, tuning=dict(list_wave_min_hits=0))
    assert_same(res, ref)
    assert res.counters["list_handed_back"] >= 1
    assert res.counters["dp_cells"] == ref["cells_used"]


def test_list_builder_second_level():
    """An offset with 300 T occurrences (between the first wavefront level's 128 and the second's
    512) is built by the second level; results stay exact and nothing is handed back."""
    rng = np.random.default_rng(13)
    unit = "".join(rng.choice(list("ACGT"), 20))
    other = "".join(rng.choice(list("ACGT"), 4000))
    t = other[:2000] + unit * 300 + other[2000:]
    s = unit * 30 + other[2000:2600]
    R = readsim.from_strings([s, t])
    res, ref = run_both(R, threads=1, schedule=oic.SCHED_BATCHED, tuning=dict(list_wave_min_hits=0))
    assert_same(res, ref)
    assert res.counters["list_second_level"] >= 1 and res.counters["list_handed_back"] == 0
    assert res.counters["dp_cells"] == ref["cells_used"]


def top_kmers(R, n, seed=0):
    """The n most frequent k-mers of a read set (ties by text), half written reverse-complemented."""
    comp = str.maketrans("ACGTacgt", "TGCAtgca")
    counts = {}
    for i in range(1, R.n + 1):
        s = R.read(i).decode().upper()
        for j in range(len(s) - 22 + 1):
            w = s[j:j + 22]
            counts[w] = counts.get(w, 0) + 1
    top = sorted(counts, key=lambda w: (-counts[w], w))[:n]
    return [w if j % 2 else w.translate(comp)[::-1] for j, w in enumerate(top)]


@pytest.mark.parametrize("fixture", ["c1s", "c3s"])
@pytest.mark.parametrize("hopeless", [True, False])
def test_skip_kmer_file(request, fixture, hopeless):
    """-k <file> (Mark_Skip_Kmers): skipped k-mers give no hits and screen read ends for the
    hopeless check; the file is parsed by the engine, the oracle gets the same k-mers."""
    R = request.getfixturevalue(fixture)
    km = top_kmers(R, 400)
    text = "".join(f">{i}\n{w} 99\n" if i % 3 == 0 else f"{w.lower()}\t{i}\n" for i, w in enumerate(km))
    assert oracle.parse_skip_file(text, 22) == [w.lower() for w in km]
    eng = oic.Engine(kmer=22, maxerate=0.045, minlength=500, threads=4, hopeless=hopeless, schedule=oic.SCHED_BATCHED)
    eng.set_skip_kmers(text)
    eng.load_reads(R)
    res = eng.run()
    ref = oracle.run(R, kmer=22, maxErate=0.045, minOlap=500, nThreads=4, ompThreads=16, hopeless=int(hopeless), skip=km)
    base = oracle.run(R, kmer=22, maxErate=0.045, minOlap=500, nThreads=4, ompThreads=16, hopeless=int(hopeless))
    assert ref["hits"] < base["hits"]
    assert_same(res, ref)
    assert res.counters["dp_cells"] == ref["cells_used"]
    with pytest.raises(oic.OicError):
        eng.set_skip_kmers("ACGTACGT\n")


@pytest.mark.parametrize("schedule,tuning", [
    (oic.SCHED_PLAIN, None),
    (oic.SCHED_BATCHED, dict(fwd_lanes=4, fwd_tile=64, fwd_loop=2)),
    (oic.SCHED_BATCHED, dict(fwd_lanes=8, fwd_tile=64, fwd_loop=0)),
    (oic.SCHED_BATCHED, dict(fwd_lanes=1, fwd_tile=32, fwd_loop=0)),
], ids=["plain", "flat", "row-stepped", "lane"])
def test_run_reads_with_n_and_short_reads(schedule, tuning):
    """Reads with N (the packed kernels' N-mask paths) and reads too short to hash or search."""
    rng = np.random.default_rng(9)
    R = readsim.simulate(readsim.PRESETS["c1"].scaled(60_000, seed=44))
    seqs = [R.read(i).decode() for i in range(1, R.n + 1)]
    for i in range(0, len(seqs), 3):
        s = list(seqs[i])
        for p in rng.integers(0, len(s), 5):
            s[p] = "N"
        seqs[i] = "".join(s)
    seqs[1] = seqs[1][:300]     # shorter than --minlength: neither hashed nor searched
    seqs.append("ACGT" * 3)     # shorter than k
    R2 = readsim.from_strings(seqs)
    res, ref = run_both(R2, schedule=schedule, tuning=tuning)
    assert_same(res, ref)
    assert res.counters["dp_cells"] == ref["cells_used"]


def test_run_single_and_two_reads():
    R = readsim.simulate(readsim.PRESETS["c1"].scaled(60_000, seed=45))
    one = R.subset(1)
    res, ref = run_both(one)
    assert len(res.records) == 0 and len(ref["records"]) == 0
    two = R.subset(2)
    assert_same(*run_both(two))


def test_run_c1_full_bit_exact():
    R = readsim.simulate(readsim.PRESETS["c1"])
    res, ref = run_both(R, threads=8)
    assert len(res.records) > 50000
    assert_same(res, ref)


def test_short_sink_keeps_records_for_fetch(c1s):
    """A sink too small for a search: the call runs to its end and returns OIC_ECAPACITY with
    records_needed; oic_fetch_pending hands over the held records without a second search, and
    the counters are those of one search."""
    import ctypes
    eng = oic.Engine(kmer=22, maxerate=0.045, minlength=500, threads=4, schedule=oic.SCHED_BATCHED)
    eng.load_reads(c1s)
    full = eng.run()
    n = len(full.records)
    assert n > 100
    eng.build_hash_index(1, c1s.n)
    c0 = eng.counters()
    small = np.zeros(7, oic.RECORD_DTYPE)
    s = oic.Sink(small.ctypes.data, 7, 0, 0)
    rc = oic.lib().oic_search(eng._ctx, 1, c1s.n, ctypes.byref(s))
    assert rc == oic.OIC_ECAPACITY
    assert s.count == 7 and s.records_needed == n
    assert list(s.stats) == list(full.stats)
    c1 = eng.counters()
    # still too small: refused, nothing lost
    rc = oic.lib().oic_fetch_pending(eng._ctx, ctypes.byref(s))
    assert rc == oic.OIC_ECAPACITY and s.records_needed == n and s.count == 7
    big = np.zeros(n, oic.RECORD_DTYPE)
    big[:7] = small
    s.records, s.capacity = big.ctypes.data, n
    assert oic.lib().oic_fetch_pending(eng._ctx, ctypes.byref(s)) == oic.OIC_OK
    assert s.count == n
    assert eng.counters()["dp_cells"] == c1["dp_cells"]
    assert c1["dp_cells"] - c0["dp_cells"] == full.counters["dp_cells"]
    a, b = oic.sort_records(big), oic.sort_records(full.records)
    assert all(np.array_equal(a[f], b[f]) for f in ("a_iid", "b_iid", "w0", "w1"))
    # a NULL sink only learns the size
    s2 = oic.Sink(None, 0, 0, 0)
    assert oic.lib().oic_search(eng._ctx, 1, c1s.n, ctypes.byref(s2)) == oic.OIC_ECAPACITY
    assert s2.count == 0 and s2.records_needed == n
    # the Python wrapper takes the held records the same way
    r = eng.search(1, c1s.n, cap=5)
    a = oic.sort_records(r.records)
    assert len(a) == n and all(np.array_equal(a[f], b[f]) for f in ("a_iid", "b_iid", "w0", "w1"))


def test_learned_schedule_with_nan_model_falls_back(c2s):
    """A model whose predictions are NaN cannot order pairs: the learned schedule falls back to the
    static estimate for the batch and the output stays the oracle's."""
    eng = oic.Engine(kmer=22, maxerate=0.045, minlength=500, threads=4, schedule=oic.SCHED_LEARNED)
    w = np.full(oic.NPARAM, np.nan, np.float32)
    eng.set_model(w, np.zeros(oic.NFEAT, np.float32), np.ones(oic.NFEAT, np.float32))
    assert np.isnan(eng.predict(np.ones((3, oic.NFEAT), np.float32))).all()
    eng.load_reads(c2s)
    res = eng.run()
    ref = oracle.run(c2s, kmer=22, maxErate=0.045, minOlap=500, nThreads=4, ompThreads=16)
    assert_same(res, ref)
    assert res.counters["dp_cells"] == ref["cells_used"]


def test_learned_schedule_with_trained_model_matches_oracle(c3s):
    """The learned path end to end: the cost model is trained on the GPU on traces of a fresh
    C3-generator instance (seed 4), then orders the pairs and packed DP tasks of a search of another
    C3-generator instance.  Records, .stats and DP cells equal the oracle's; every pair was ordered by
    the model's prediction."""
    from canu_amd import learned
    eng = oic.Engine(kmer=22, maxerate=0.045, minlength=500, threads=4, schedule=oic.SCHED_LEARNED)
    info = learned.train_for(eng, genome_len=2_000_000, seed=4, epochs=8, max_traces=400_000)
    print(f"\ncost model: {info}")
    assert info["traces"] > 20_000
    assert info["learned_spearman"] > info["static_spearman"]
    w, mu, sd = eng.model()
    assert np.isfinite(w).all()
    eng.load_reads(c3s)
    res = eng.run()
    assert res.counters["pairs"] > 1000
    assert res.counters["learned_pairs"] == res.counters["pairs"]
    ref = oracle.run(c3s, kmer=22, maxErate=0.045, minOlap=500, nThreads=4, ompThreads=16)
    assert_same(res, ref)
    assert res.counters["dp_cells"] == ref["cells_used"]
    assert res.counters["extensions"] == ref["extensions"]
