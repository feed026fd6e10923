"""The workload simulator is deterministic and matches its stated read model."""
import numpy as np

from canu_amd import readsim

COMP = bytes.maketrans(b"ACGT", b"TGCA")


def test_deterministic():
    cfg = readsim.PRESETS["c1"].scaled(100_000, seed=9)
    a, b = readsim.simulate(cfg), readsim.simulate(cfg)
    assert np.array_equal(a.bases, b.bases) and np.array_equal(a.start, b.start)


def test_lognormal_lengths_and_coverage():
    cfg = readsim.PRESETS["c1"].scaled(1_000_000, seed=1)
    R = readsim.simulate(cfg)
    L = R.length[1:].astype(float)
    assert L.min() >= 900                       # min 1 kbp template, minus deletions
    assert abs(L.mean() - 8000) < 600 and abs(L.std() - 3000) < 600
    cov = R.length[1:].sum() / R.genome_len
    assert 14.0 < cov < 16.5


def test_error_free_reads_are_genome_substrings():
    cfg = readsim.PRESETS["c1"].scaled(50_000, seed=4)
    cfg.p_sub = cfg.p_ins = cfg.p_del = 0.0
    R = readsim.simulate(cfg)
    # rebuild each read from a template: same length, strand as recorded
    for i in range(1, min(R.n, 30) + 1):
        s = R.read(i)
        assert len(s) == int(R.length[i])
        if R.strand[i]:
            s = s.translate(COMP)[::-1]
        assert set(s) <= set(b"ACGT")


def test_hpc_reads_have_no_runs():
    cfg = readsim.PRESETS["c5"].scaled(200_000, seed=5)
    R = readsim.simulate(cfg)
    s = R.read(1)
    assert all(s[i] != s[i + 1] for i in range(len(s) - 1))


def test_repeats_present():
    cfg = readsim.PRESETS["c3"].scaled(2_000_000, seed=3)
    R = readsim.simulate(cfg)
    assert R.genome_len >= 1_900_000 and R.n > 1000
