"""The oracle's seed index, Find_Overlaps match lists, hash blocks and output records.

Pinned against brute force on small read sets:
  * seed hits of each (S, orientation, T) pair = windows of k ACGT bases shared by S and T with
    S_ID < T_ID (Find_Overlaps.C:251-332 with Hash_Find's exact k-mer compare);
  * match nodes = maximal runs of hits on one diagonal at consecutive S offsets, Len = k + run - 1
    (Add_Match, Find_Overlaps.C:24-96);
  * hash blocks stop before the read at which the distinct k-mer count reaches
    hashload * 2^hashbits * 21 or the loaded length reaches hashdatalen + 2^21 - 1
    (Build_Hash_Index.C:445,504-506), and the last hash read is never indexed when a block ends
    one short of it (overlapInCore.C:204);
  * records: the ovOverlap layout invariants and the simulator's truth.
"""
import numpy as np
import pytest

from canu_amd import readsim
from oracle import oracle

K = 22
COMP = bytes.maketrans(b"ACGTacgt", b"TGCAtgca")


def rc(s):
    return s.translate(COMP)[::-1]


@pytest.fixture(scope="module")
def small():
    cfg = readsim.PRESETS["c1"].scaled(30000, seed=21)
    cfg.ln_mu, cfg.ln_sigma, cfg.min_len = 7.6, 0.3, 800
    R = readsim.simulate(cfg)
    out = oracle.run(R, collect_trace=True, kmer=K, maxErate=0.045, minOlap=500, nThreads=1)
    return R, out


def kmer_positions(s):
    d = {}
    for i in range(len(s) - K + 1):
        w = s[i:i + K]
        if all(c in b"ACGT" for c in w):
            d.setdefault(w, []).append(i)
    return d


def brute_pairs(R):
    seqs = [None] + [R.read(i).upper() for i in range(1, R.n + 1)]
    idx = [None] + [kmer_positions(seqs[i]) for i in range(1, R.n + 1)]
    res = {}
    for s in range(1, R.n + 1):
        for dr, S in ((0, seqs[s]), (1, rc(seqs[s]))):
            for t in range(s + 1, R.n + 1):
                hits = []
                tk = idx[t]
                for o in range(len(S) - K + 1):
                    for toff in tk.get(S[o:o + K], ()):
                        hits.append((o, toff))
                if hits:
                    res[(s, t, dr)] = hits
    return res


def runs(hits):
    hs = set(hits)
    starts = [h for h in hits if (h[0] - 1, h[1] - 1) not in hs]
    lens = []
    for o, t in starts:
        n = 1
        while (o + n, t + n) in hs:
            n += 1
        lens.append(n)
    return lens


def test_hits_and_nodes_match_brute_force(small):
    R, out = small
    tr = out["trace"]
    bp = brute_pairs(R)
    got = {(int(x["s_id"]), int(x["t_id"]), int(x["dir"])): x for x in tr}
    assert set(got) == set(bp)
    for key, hits in bp.items():
        x = got[key]
        assert x["n_hits"] == len(hits), key
        lens = runs(hits)
        assert x["n_nodes"] == len(lens), key
        assert x["longest"] == K + max(lens) - 1, key
        assert x["diag_min"] == min(t - o for o, t in hits)
        assert x["diag_max"] == max(t - o for o, t in hits)
    assert out["hits"] == sum(len(h) for h in bp.values())


def test_stats_consistent(small):
    R, out = small
    st = out["stats"]
    tr = out["trace"]
    assert st[0] + st[1] == len(tr)           # every processed pair counted once
    assert st[3] == len(out["records"]) == st[4] + st[5]
    assert st[6] == st[7] == 0


def decode(recs):
    w0, w1 = recs["w0"].astype(np.uint64), recs["w1"].astype(np.uint64)
    F = np.uint64(0x1FFFFF)
    return dict(ahg5=w0 & F, ahg3=(w0 >> np.uint64(21)) & F, evalue=(w0 >> np.uint64(42)) & np.uint64(0xFFFF),
                flipped=(w0 >> np.uint64(58)) & np.uint64(1), obt=(w0 >> np.uint64(59)) & np.uint64(1),
                dup=(w0 >> np.uint64(60)) & np.uint64(1), utg=(w0 >> np.uint64(61)) & np.uint64(1),
                bhg5=w1 & F, bhg3=(w1 >> np.uint64(21)) & F, span=(w1 >> np.uint64(42)) & F,
                extra=(w0 >> np.uint64(62)) | (w1 >> np.uint64(63)))


def test_record_layout_invariants(small):
    R, out = small
    recs = out["records"]
    f = decode(recs)
    L = R.length.astype(np.int64)
    assert np.all(f["utg"] == 1) and np.all(f["obt"] == 0) and np.all(f["dup"] == 0) and np.all(f["extra"] == 0)
    assert np.all(recs["a_iid"] != recs["b_iid"])
    # at most one of ahg5/bhg5 and of ahg3/bhg3 is positive: dovetail or contained (Output.C:66-68)
    assert np.all((f["ahg5"] == 0) | (f["bhg5"] == 0))
    assert np.all((f["ahg3"] == 0) | (f["bhg3"] == 0))
    # errors <= Error_Bound[len] = ceil(0.045 len) with len >= 500: evalue < 1e5 * (0.045 + 1/500)
    assert np.all(f["evalue"] <= 4700)
    # aligned spans fit the reads
    alen = L[recs["a_iid"]] - f["ahg5"].astype(np.int64) - f["ahg3"].astype(np.int64)
    blen = L[recs["b_iid"]] - f["bhg5"].astype(np.int64) - f["bhg3"].astype(np.int64)
    assert np.all(alen >= 500) and np.all(blen >= 500)
    assert np.all(np.abs(f["span"].astype(np.int64) - (alen + blen) // 2) <= 0.05 * (alen + blen))


def test_overlaps_match_truth(small):
    """Every reported pair overlaps on the genome, and long true overlaps are all found."""
    R, out = small
    recs = out["records"]
    pos = R.pos.astype(np.int64)
    tl = np.array([0] + [int(R.length[i]) for i in range(1, R.n + 1)], np.int64)
    found = {(min(a, b), max(a, b)) for a, b in zip(recs["a_iid"].tolist(), recs["b_iid"].tolist())}
    for a, b in found:
        ov = min(pos[a] + tl[a], pos[b] + tl[b]) - max(pos[a], pos[b])
        assert ov > 300, (a, b, ov)
    true_long = []
    for a in range(1, R.n + 1):
        for b in range(a + 1, R.n + 1):
            ov = min(pos[a] + tl[a], pos[b] + tl[b]) - max(pos[a], pos[b])
            if ov >= 1500:
                true_long.append((a, b))
    hit = sum(p in found for p in true_long)
    assert hit >= 0.97 * len(true_long), (hit, len(true_long))


def test_hash_blocks_follow_distinct_kmer_limit():
    R = readsim.simulate(readsim.PRESETS["c1"].scaled(20000, seed=3))
    # distinct k-mers of reads in order, brute force
    seen, cum = set(), [0]
    for i in range(1, R.n + 1):
        s = R.read(i).upper()
        if len(s) >= 500:
            for j in range(len(s) - K + 1):
                w = s[j:j + K]
                if all(c in b"ACGT" for c in w):
                    seen.add(w)
        cum.append(len(seen))
    hb, load = 10, 0.2
    limit = int(load * (1 << hb) * 21)
    # block ends: first read r (after the block start) with distinct-before-r >= limit
    ends, bgn = [], 1
    while bgn < R.n:
        base = cum[bgn - 1]
        end = R.n
        for r in range(bgn, R.n + 1):
            if cum[r - 1] - base >= limit:
                end = r - 1
                break
        # distinct counts restart per block: recompute from the block's own reads
        seen_b, end_b = set(), R.n
        for r in range(bgn, R.n + 1):
            if len(seen_b) >= limit:
                end_b = r - 1
                break
            s = R.read(r).upper()
            if len(s) >= 500:
                for j in range(len(s) - K + 1):
                    w = s[j:j + K]
                    if all(c in b"ACGT" for c in w):
                        seen_b.add(w)
        ends.append(end_b)
        bgn = end_b + 1
    out = oracle.run(R, kmer=K, maxErate=0.045, minOlap=500, hashBits=hb, hashLoad=load)
    assert out["hash_end"] == ends
    assert len(ends) > 2


def test_thread_chunking_skips_chunk_at_end():
    """-r 1-E with -t T: E is never searched when (E - 1) % perThread == 0 (perThread = 1 + (E-1)/T/8)."""
    R = readsim.simulate(readsim.PRESETS["c1"].scaled(40000, seed=5))
    E = R.n - 3
    t_skip = next(t for t in range(1, 64) if (E - 1) % (1 + (E - 1) // t // 8) == 0)
    t_keep = next(t for t in range(1, 64) if (E - 1) % (1 + (E - 1) // t // 8) != 0)
    a = oracle.run(R, collect_trace=True, kmer=K, maxErate=0.045, minOlap=500, endRef=E, nThreads=t_skip)
    b = oracle.run(R, collect_trace=True, kmer=K, maxErate=0.045, minOlap=500, endRef=E, nThreads=t_keep)
    assert E not in set(a["trace"]["s_id"].tolist())
    sb = set(b["trace"]["s_id"].tolist())
    assert max(sb) <= E and (E in sb or not any((R.pos[E] < R.pos[t] + R.length[t]) and (R.pos[t] < R.pos[E] + R.length[E]) for t in range(E + 1, R.n + 1)))


def test_skip_file_parsing():
    text = "ACGTACGTACGTACGTACGTAC 17\n>header\nttttttttttttttttttttta\tx\nGGGGGGGGGGGGGGGGGGGGGG\n"
    assert oracle.parse_skip_file(text, K) == ["acgtacgtacgtacgtacgtac", "ttttttttttttttttttttta",
                                                "gggggggggggggggggggggg"]
    with pytest.raises(ValueError):
        oracle.parse_skip_file("ACGT\n", K)


def test_skip_kmers_remove_hits_both_strands(small):
    """-k <file>: a skipped k-mer (or its reverse complement) yields no seed hits
    (Mark_Skip_Kmers, Build_Hash_Index.C:185-257; Hash_Find's Empty entries, Find_Overlaps.C:177-222)."""
    R, _ = small
    # skip the most frequent k-mers of the read set, some given as reverse complements
    counts = {}
    for i in range(1, R.n + 1):
        for w in kmer_positions(R.read(i).upper()):
            counts[w] = counts.get(w, 0) + 1
    top = sorted(counts, key=lambda w: (-counts[w], w))[:300]
    skip = [w.decode() if j % 2 else rc(w).decode() for j, w in enumerate(top)]
    out = oracle.run(R, collect_trace=True, kmer=K, maxErate=0.045, minOlap=500, nThreads=1, skip=skip)
    banned = set(top) | {rc(w) for w in top}
    bp = brute_pairs(R)
    want = {}
    for key, hits in bp.items():
        s, t, dr = key
        S = R.read(s).upper()
        S = rc(S) if dr else S
        kept = [(o, toff) for o, toff in hits if S[o:o + K] not in banned]
        if kept:
            want[key] = kept
    got = {(int(x["s_id"]), int(x["t_id"]), int(x["dir"])): x for x in out["trace"]}
    assert set(got) == set(want)
    for key, hits in want.items():
        assert got[key]["n_hits"] == len(hits), key
        assert got[key]["n_nodes"] == len(runs(hits)), key
    assert out["hits"] < small[1]["hits"]


def _same_run(a, b):
    ra, rb = oracle.sort_records(a["records"]), oracle.sort_records(b["records"])
    assert len(ra) == len(rb)
    for f in ("a_iid", "b_iid", "w0", "w1"):
        assert np.array_equal(ra[f], rb[f]), f
    assert a["stats"] == b["stats"]
    for k in ("cells", "cells_used", "bases", "dp_calls", "extensions", "hits", "pairs"):
        assert a[k] == b[k], k


@pytest.mark.parametrize("case", ["c3_mid_refs", "c3_low_refs_skip", "c5_partial", "c2_no_hopeless_skip"])
def test_restricted_index_searches_like_the_full_index(case):
    """The oracle's restricted index (only the k-mers of the searched ref reads and the skip
    k-mers) gives the full index's records, .stats and work counters."""
    preset, glen, seed, kw = {
        "c3_mid_refs": ("c3", 400_000, 41, dict(bgnRef=300, endRef=340)),
        "c3_low_refs_skip": ("c3", 300_000, 42, dict(bgnRef=1, endRef=25)),
        "c5_partial": ("c5", 300_000, 43, dict(bgnRef=100, endRef=130, partial=1)),
        "c2_no_hopeless_skip": ("c2", 300_000, 44, dict(bgnRef=50, endRef=90, hopeless=0)),
    }[case]
    R = readsim.simulate(readsim.PRESETS[preset].scaled(glen, seed=seed))
    rate, minolap = (0.01, 1000) if preset == "c5" else (0.045, 500)
    base = dict(kmer=K, maxErate=rate, minOlap=minolap, hashBits=26, hashLoad=0.8, hashDataLen=int(2e9),
                nThreads=4, ompThreads=4, collect_trace=False, **kw)
    if "skip" in case:
        # frequent k-mers of a few reads, plus one absent and one with an N
        s = R.read(kw["endRef"]).upper().decode()
        base["skip"] = [s[i:i + K] for i in (0, 40, 500, 1000)] + ["A" * K, "ACGTN" + "A" * (K - 5)]
    full = oracle.run(R, **base)
    part = oracle.run(R, restrictIndex=1, **base)
    assert full["hash_end"] == part["hash_end"] == [R.n]
    assert full["pairs"] > 0 and full["cells"] > 0
    assert len(full["records"]) > 0 or preset != "c5"
    _same_run(full, part)


def test_restricted_index_refuses_more_than_one_block():
    R = readsim.simulate(readsim.PRESETS["c1"].scaled(100_000, seed=45))
    with pytest.raises(RuntimeError):
        oracle.run(R, kmer=K, maxErate=0.045, minOlap=500, hashBits=10, hashLoad=0.6, bgnRef=1, endRef=5, restrictIndex=1)


def test_record_orientation_and_hangs_match_truth():
    """Output_Overlap's A/B choice, N/I/O orientation and hangs (overlapInCore-Output.C:71-150), read
    back as ovOverlap's a_hang()/b_hang()/flipped (ovOverlap.H:162-166), describe where the two
    reads lie on the simulated genome: B is reverse-complemented against A exactly when their
    strands differ, and in A's frame B starts a_hang and ends b_hang bases after A does (read
    coordinates drift from genome coordinates by the reads' own indels)."""
    cfg = readsim.PRESETS["c1"].scaled(60_000, seed=23)
    cfg.ln_mu, cfg.ln_sigma, cfg.min_len = 7.8, 0.35, 800
    R = readsim.simulate(cfg)
    out = oracle.run(R, kmer=K, maxErate=0.045, minOlap=500, nThreads=4, ompThreads=4)
    recs = out["records"]
    f = decode(recs)
    ah = f["ahg5"].astype(np.int64) - f["bhg5"].astype(np.int64)
    bh = f["bhg3"].astype(np.int64) - f["ahg3"].astype(np.int64)
    pos, tl, st = R.pos.astype(np.int64), R.tlen.astype(np.int64), R.strand.astype(np.int64)
    kinds = set()
    for k in range(len(recs)):
        a, b = int(recs["a_iid"][k]), int(recs["b_iid"][k])
        assert int(f["flipped"][k]) == int(st[a] != st[b]), k
        if st[a] == 0:
            ea, eb = pos[b] - pos[a], (pos[b] + tl[b]) - (pos[a] + tl[a])
        else:
            ea, eb = (pos[a] + tl[a]) - (pos[b] + tl[b]), pos[a] - pos[b]
        assert abs(ah[k] - ea) <= 0.03 * max(abs(ea), 1) + 60, (k, a, b, ah[k], ea)
        assert abs(bh[k] - eb) <= 0.03 * max(abs(eb), 1) + 60, (k, a, b, bh[k], eb)
        kinds.add((int(f["flipped"][k]), int(np.sign(ah[k])), int(np.sign(bh[k])), int(a < b)))
    # every orientation, both hang signs (dovetails and containments) and both ID orders occur
    assert len(recs) > 300
    assert {x[0] for x in kinds} == {0, 1} and {x[3] for x in kinds} == {0, 1}
    assert {(x[1], x[2]) for x in kinds} >= {(1, 1), (-1, -1), (1, -1)}


def test_ref_list_searches_exactly_the_listed_reads():
    """The oracle's explicit ref list (restricted index) gives the union of single-read searches
    (a range [r, r + 1] under -t 1 searches read r only: the chunk rule skips r + 1)."""
    R = readsim.simulate(readsim.PRESETS["c3"].scaled(300_000, seed=47))
    refs = [5, 9, 17, 40]
    base = dict(kmer=K, maxErate=0.045, minOlap=500, hashBits=26, hashLoad=0.8, hashDataLen=int(2e9), nThreads=1,
                ompThreads=4)
    got = oracle.run(R, refs=refs, restrictIndex=1, **base)
    parts = [oracle.run(R, bgnRef=r, endRef=r + 1, **base) for r in refs]
    recs = np.concatenate([p["records"] for p in parts])
    _same_run(got, dict(records=recs, stats=list(np.sum([p["stats"] for p in parts], axis=0)),
                        **{k: sum(p[k] for p in parts) for k in ("cells", "cells_used", "bases", "dp_calls", "extensions",
                                                                  "hits", "pairs")}))
    assert got["pairs"] > 0
