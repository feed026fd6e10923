"""The full-size workloads the bench runs (C3: 50 Mbp repeat-rich genome, 125k reads; C5: 100 Mbp,
homopolymer-compressed HiFi-like reads, 167k reads), checked against the oracle: chosen ref reads
searched against an index of every read, through the HIP engine (one hash block over all reads)
and through the oracle (its restricted index: the k-mers of those ref reads, each with its complete
occurrence list, which searches exactly like the full index).

The ref reads were chosen, from single-read searches of refs 1-1200, to cover what the small
instances do not: on C3, reads from repeat families (up to 214k match nodes against one read), a
40 kbp read, reads with overlaps and plain reads; on C5, reads with pairs in which an offset has more
than 128 seed hits (satellite arrays: the 512-hit list level) and reads with many overlaps.  Each
read's oracle search stays under about 1e8 DP cells, so the whole check runs in seconds on the host.
Read IDs are those of the fixed-seed generators; a change of generator shows as a failed coverage
assertion, not as a silent pass.
"""
import numpy as np
import pytest

from canu_amd import oic, readsim
from oracle import oracle

pytestmark = pytest.mark.gpu

C3_REFS = [1, 2, 3, 46, 85, 376, 388, 489, 903]
C5_REFS = [1, 2, 22, 89, 90]


def search_refs(R, refs, rate, minlen, schedule, tuning):
    eng = oic.Engine(kmer=22, maxerate=rate, minlength=minlen, hashbits=27, hashload=0.8, hashdatalen=8_000_000_000,
                     threads=1, schedule=schedule)
    eng.load_reads(R)
    assert eng.build_hash_index(1, R.n) == R.n
    eng.set_tuning(**tuning)
    recs, stats = [], np.zeros(8, np.int64)
    c0 = eng.counters()
    for r in refs:
        print(f"  search ref {r} ({tuning})", flush=True)
        res = eng.search(r, r + 1)   # [r, r + 1] with -t 1 searches read r only (OverlapDriver's chunk rule)
        recs.append(res.records)
        stats += np.array(res.stats, np.int64)
    c1 = eng.counters()
    eng.close()
    return np.concatenate(recs), stats, {k: c1[k] - c0[k] for k in c1}


def check(R, refs, rate, minlen, runs):
    print(f"oracle on refs {refs}", flush=True)
    ref = oracle.run(R, refs=refs, restrictIndex=1, kmer=22, maxErate=rate, minOlap=minlen, hashBits=28, hashLoad=0.8,
                     hashDataLen=8_000_000_000, nThreads=1, ompThreads=16)
    b = oracle.sort_records(ref["records"])
    out = []
    for schedule, tuning in runs:
        recs, stats, ctr = search_refs(R, refs, rate, minlen, schedule, tuning)
        a = oic.sort_records(recs)
        assert len(a) == len(b), (schedule, tuning, len(a), len(b))
        for f in ("a_iid", "b_iid", "w0", "w1"):
            assert np.array_equal(a[f], b[f]), (schedule, tuning, f)
        assert list(stats) == list(ref["stats"]), (schedule, tuning)
        assert ctr["dp_cells"] == ref["cells_used"], (schedule, tuning)
        assert ctr["extensions"] == ref["extensions"], (schedule, tuning)
        assert ctr["hits"] == ref["hits"] and ctr["pairs"] == ref["pairs"], (schedule, tuning)
        print(f"  hits {ctr['hits']} / oracle {ref['hits']}, pairs {ctr['pairs']} / oracle {ref['pairs']}, "
              f"overflow retries {ctr['overflow_retries']}, second list level {ctr['list_second_level']}", flush=True)
        out.append(ctr)
    return ref, out


def test_c3_full_size_refs_match_oracle():
    R = readsim.simulate(readsim.PRESETS["c3"])
    assert R.n == 125023
    ref, ctr = check(R, C3_REFS, 0.045, 500, [
        (oic.SCHED_BATCHED, dict(fwd_lanes=4, fwd_tile=64, fwd_loop=2, list_wave_min_hits=256)),
        (oic.SCHED_BATCHED, dict(fwd_lanes=8, fwd_tile=64, fwd_loop=0, list_wave_min_hits=2048)),
        # a 16-cell ring sends most bands through the 512-cell retry level
        (oic.SCHED_BATCHED, dict(fwd_lanes=1, fwd_tile=16, fwd_loop=0, list_wave_min_hits=2048)),
        (oic.SCHED_PLAIN, dict(list_wave_min_hits=1 << 31)),
    ])
    assert ref["pairs"] > 2500 and ref["extensions"] > 100_000 and len(ref["records"]) > 0
    assert ctr[0]["list_wave_pairs"] > 1000
    assert ctr[2]["overflow_retries"] > 1000


def test_c5_full_size_refs_match_oracle():
    R = readsim.simulate(readsim.PRESETS["c5"])
    assert R.n == 166527
    ref, ctr = check(R, C5_REFS, 0.01, 1000, [
        (oic.SCHED_BATCHED, dict(fwd_lanes=4, fwd_tile=64, fwd_loop=2, list_wave_min_hits=256)),
        (oic.SCHED_BATCHED, dict(fwd_lanes=16, fwd_tile=128, fwd_loop=0, list_wave_min_hits=256)),
        (oic.SCHED_BATCHED, dict(fwd_lanes=1, fwd_tile=16, fwd_loop=0, list_wave_min_hits=256)),
    ])
    assert len(ref["records"]) > 100
    # pairs with an offset of more than 128 seed hits ran at the 512-hit list level
    assert ctr[0]["list_second_level"] > 0 and ctr[2]["list_second_level"] > 0
    assert ctr[2]["overflow_retries"] > 0
