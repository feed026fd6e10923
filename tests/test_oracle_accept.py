"""Which candidate pairs the oracle extends, and which extensions it keeps, against statements of the
rules that take their inputs from brute force rather than from the oracle:

* -minkmers (Process_String_Overlaps.C:22-33, 615-618; overlapInCore.C:410-411): a pair is skipped,
  before Process_Matches sees it, when its hit count is below max(F, floor(exp(-k e) (span - k +
  1))), span = last minus first S offset of its hits (0 below k), F = floor(exp(-k e) (minOlap - k
  + 1)); without -minkmers none is;
* the hopeless check (:383-411): a pair whose hits form one exact match (one maximal diagonal run)
  is dropped when, before the match or after it, the shorter of the two reads' overhangs exceeds 90
  bases (no read end is screened here: there are no skip k-mers); -h off, or a second run, keeps it;
* every other pair is extended at least once, and a consistent pair exactly once, from its longest
  match (:413-424, 479-480);
* an extension is kept when it is a dovetail, both aligned spans are at least minOlap long, and
  its errors are at most Error_Bound[1 + min(S span, T span)] (:446-468): for consistent pairs
  with one longest match the kept count is 0 or 1 by this rule, the extension computed by the
  oracle's Extend_Alignment (pinned by its own tests) on the brute-force longest run.

Hits and runs come from exact k-mer matching of the reads, as in tests/test_oracle_search.py.
"""
import math

import numpy as np
import pytest

from canu_amd import readsim
from oracle import oracle

K = 22
ERATE = 0.045
MINOLAP = 500
HOPELESS = 90
COMP = bytes.maketrans(b"ACGTacgt", b"TGCAtgca")


def rc(s):
    return s.translate(COMP)[::-1]


@pytest.fixture(scope="module")
def reads():
    cfg = readsim.PRESETS["c1"].scaled(30000, seed=23)
    cfg.ln_mu, cfg.ln_sigma, cfg.min_len = 7.6, 0.35, 700
    return readsim.simulate(cfg)


@pytest.fixture(scope="module")
def planted():
    """Random reads sharing short exact segments (22-60 bases) planted at random places, some on the
    reverse strand: pairs whose hits form one exact match, near a read end or deep inside both."""
    rng = np.random.default_rng(5)
    acgt = np.frombuffer(b"ACGT", np.uint8)
    n = 60
    seqs = [bytearray(acgt[rng.integers(0, 4, size=int(rng.integers(500, 1500)))].tobytes()) for _ in range(n)]
    for _ in range(150):
        a, b = (int(x) for x in rng.choice(n, size=2, replace=False))
        ln = int(rng.integers(K, 61))
        seg = acgt[rng.integers(0, 4, size=ln)].tobytes()
        mode = int(rng.integers(0, 3))     # 0: start of a, end of b; 1: anywhere in both; 2: start of a, anywhere in b
        for r in (a, b):
            L = len(seqs[r])
            near_start, near_end, anywhere = int(rng.integers(0, 80)), L - ln - int(rng.integers(0, 80)), int(rng.integers(0, L - ln))
            where = (near_start if r == a else near_end) if mode == 0 else anywhere if mode == 1 else \
                (near_start if r == a else anywhere)
            piece = seg if (r == a or rng.random() < 0.5) else rc(seg)
            seqs[r][where:where + ln] = piece
    return readsim.from_strings([bytes(x) for x in seqs])


def brute_hits(R):
    seqs = [None] + [R.read(i).upper() for i in range(1, R.n + 1)]
    idx = []
    for i in range(1, R.n + 1):
        d = {}
        s = seqs[i]
        for p in range(len(s) - K + 1):
            w = s[p:p + K]
            if all(c in b"ACGT" for c in w):
                d.setdefault(w, []).append(p)
        idx.append(d)
    res = {}
    for s in range(1, R.n + 1):
        for dr, S in ((0, seqs[s]), (1, rc(seqs[s]))):
            for t in range(s + 1, R.n + 1):
                tk = idx[t - 1]
                hits = [(o, toff) for o in range(len(S) - K + 1) for toff in tk.get(S[o:o + K], ())]
                if hits:
                    res[(s, t, dr)] = hits
    return seqs, res


@pytest.fixture(scope="module")
def brute(reads):
    return brute_hits(reads)


@pytest.fixture(scope="module")
def brute_planted(planted):
    return brute_hits(planted)


def runs(hits):
    """Maximal diagonal runs: (S start, T start, hits)."""
    hs = set(hits)
    out = []
    for o, t in hits:
        if (o - 1, t - 1) in hs:
            continue
        n = 1
        while (o + n, t + n) in hs:
            n += 1
        out.append((o, t, n))
    return out


def hopeless(run, s_len, t_len):
    o, t, n = run
    ln = K + n - 1
    head = min(o, t)                             # the shorter overhang before the match
    tail = min(s_len - o - ln + 1, t_len - t - ln + 1)
    return head > HOPELESS or tail > HOPELESS


def min_kmers(hits):
    f = math.floor(math.exp(-K * ERATE) * (MINOLAP - K + 1))
    span = max(o for o, _ in hits) - min(o for o, _ in hits)
    exp_ = 0 if span < K else math.floor(math.exp(-K * ERATE) * (span - K + 1))
    return max(f, exp_)


def trace(reads, **kw):
    out = oracle.run(reads, collect_trace=True, kmer=K, maxErate=ERATE, minOlap=MINOLAP, nThreads=1, **kw)
    return {(int(x["s_id"]), int(x["t_id"]), int(x["dir"])): x for x in out["trace"]}


@pytest.mark.parametrize("which", ["simulated", "planted"])
def test_hopeless_check_and_at_least_one_extension(which, reads, brute, planted, brute_planted):
    R, (seqs, bp) = (reads, brute) if which == "simulated" else (planted, brute_planted)
    tr = trace(R)
    assert set(tr) == set(bp)
    n_hope = n_single = 0
    for key, hits in bp.items():
        x = tr[key]
        rr = runs(hits)
        s_len, t_len = len(seqs[key[0]]), len(seqs[key[1]])
        if len(rr) == 1 and hopeless(rr[0], s_len, t_len):
            assert x["n_ext"] == 0, key
            n_hope += 1
        else:
            assert x["n_ext"] >= 1, key
            n_single += len(rr) == 1
        if x["consistent"]:
            assert x["n_ext"] <= 1, key
    if which == "planted":
        print(f"hopeless {n_hope}, extended single-run pairs {n_single}")
        assert n_hope > 50 and n_single > 10   # both outcomes for single-run pairs


def test_no_hopeless_check_extends_every_pair(reads, brute):
    _, bp = brute
    tr = trace(reads, hopeless=0)
    assert set(tr) == set(bp)
    assert all(tr[k]["n_ext"] >= 1 for k in bp)


def test_minkmers_skip_rule(reads, brute):
    seqs, bp = brute
    tr = trace(reads, minkmers=1)
    n_skip = 0
    for key, hits in bp.items():
        skip = min_kmers(hits) > len(hits)
        assert (key not in tr) == skip, key     # a skipped pair never reaches Process_Matches
        n_skip += skip
        if not skip:
            rr = runs(hits)
            hope = len(rr) == 1 and hopeless(rr[0], len(seqs[key[0]]), len(seqs[key[1]]))
            assert (tr[key]["n_ext"] == 0) == hope, key
    assert set(tr) <= set(bp)
    print(f"skipped {n_skip}, searched {len(tr)}")
    assert n_skip > 20 and len(tr) > 20


def test_kept_extension_follows_accept_rule(reads, brute):
    seqs, bp = brute
    tr = trace(reads)
    dp = oracle.OracleDP(ERATE)
    bound = dp.tables.error_bound
    n_keep = n_drop = 0
    for key, hits in bp.items():
        x = tr[key]
        if not x["consistent"] or x["n_ext"] != 1:
            continue
        rr = runs(hits)
        lens = sorted((n for _, _, n in rr), reverse=True)
        if len(lens) > 1 and lens[0] == lens[1]:
            continue                             # the longest is not unique: the list order decides
        o, t, n = max(rr, key=lambda r: r[2])
        s, tt, dr = key
        S = seqs[s] if dr == 0 else rc(seqs[s])
        kind, slo, shi, tlo, thi, err, _ = dp.extend(S.lower(), seqs[tt].lower(), o, t, K + n - 1, s, tt)
        ln = 1 + min(shi - slo, thi - tlo)
        keep = kind == 3 and 1 + shi - slo >= MINOLAP and 1 + thi - tlo >= MINOLAP and err <= bound[ln]
        assert x["n_olaps"] == int(keep), key
        n_keep += keep
        n_drop += not keep
    print(f"kept {n_keep}, dropped {n_drop}")
    assert n_keep > 20 and n_drop > 5


@pytest.mark.parametrize("head,tail", [(90, 40), (91, 40), (40, 90), (40, 91), (90, 90), (91, 91)])
def test_hopeless_threshold_by_hand(head, tail):
    """One 30-base exact match between two random reads, HEAD bases into the first read (the second
    read's overhang before it is longer) and TAIL bases before the second read's end: extended
    exactly when both shorter overhangs are at most 90."""
    rng = np.random.default_rng(head * 1000 + tail)
    acgt = np.frombuffer(b"ACGT", np.uint8)
    a = bytearray(acgt[rng.integers(0, 4, size=800)].tobytes())
    b = bytearray(acgt[rng.integers(0, 4, size=900)].tobytes())
    seg = acgt[rng.integers(0, 4, size=30)].tobytes()
    a[head:head + 30] = seg
    t = 900 - 30 - tail + 1
    b[t:t + 30] = seg
    # flanks that differ, so the exact match is exactly the 30 planted bases
    comp = {65: 67, 67: 65, 71: 84, 84: 71}
    b[t - 1] = comp[a[head - 1]]
    b[t + 30] = comp[a[head + 30]]
    R = readsim.from_strings([bytes(a), bytes(b)])
    tr = trace(R)
    assert list(tr) == [(1, 2, 0)]
    extended = tr[(1, 2, 0)]["n_ext"] > 0
    assert extended == (head <= HOPELESS and tail <= HOPELESS)


@pytest.mark.gpu
@pytest.mark.parametrize("opt", [dict(), dict(hopeless=False), dict(minkmers=True)])
def test_gpu_planted_pairs_match_oracle(planted, opt):
    """The engine on the planted-segment reads (mostly single-match pairs, hopeless or not): the same
    records, statistics and extension count as the oracle."""
    from canu_amd import oic

    eng = oic.Engine(kmer=K, maxerate=ERATE, minlength=MINOLAP, threads=1, schedule=oic.SCHED_BATCHED, **opt)
    eng.load_reads(planted)
    res = eng.run()
    ref = oracle.run(planted, kmer=K, maxErate=ERATE, minOlap=MINOLAP, nThreads=1, **{k: int(v) for k, v in opt.items()})
    assert list(res.stats) == list(ref["stats"])
    assert res.counters["extensions"] == ref["extensions"]
    a, b = oic.sort_records(res.records), oracle.sort_records(ref["records"])
    assert len(a) == len(b)
    for f in ("a_iid", "b_iid", "w0", "w1"):
        assert np.array_equal(a[f], b[f]), f
