"""The step of Process_Matches that turns extension results into output overlaps: Add_Overlap
(overlapInCore-Process_String_Overlaps.C:175-244) folds each result into at most three distinct
overlaps, then the pair's selection (:501-513) keeps them all (-partial without -u), picks one
(Choose_Best_Partial :289-311 under -partial -u, Combine_Into_One_Olap :42-96 under -u) or merges
intersecting ones (Merge_Intersecting_Olaps :106-162).

The oracle is checked against statements of these rules written as geometry and orderings rather
than as the reference's loops:

* a result joins an entry when both lie on the same side of the main diagonal (diagonal > 0, or
  <= 0) and its diagonal runs through the entry's bounds for at least MIN_INTERSECTION + 1 = 11
  positions, counted from the bounds' left S edge to their top T edge (diagonal > 0) or from their
  bottom T edge to their right S edge (diagonal <= 0); the first such entry takes it, its bounds
  grow to cover the result, and its box and quality become the result's when the result is
  strictly better; a result that joins nothing opens a new entry if fewer than three exist;
* Combine keeps the first entry of least (quality, -length), length = 1 + min(S span, T span),
  with the union of every entry's bounds and diagonal range;
* Choose measures matching bases as (1 - quality) * (2 + S span + T span) and compares every
  entry with the FIRST entry's count, which is set once (:296) and never raised: it keeps the last
  entry that has more matching bases than the first, or as many with a lower quality than the pick
  so far, else the first.  (This is not the entry with the most matching bases when a middle entry
  has more than a later one that still beats the first; parity keeps the reference's pick.)
* Merge keeps, in any group of intersecting entries, the best quality, later entries winning ties,
  and the kept entries' bounds cover every entry's bounds.

Boxes and qualities are drawn from small sets so that ties, threshold cases and the three-entry
cap occur often.  The GPU test runs the same groups through oic_distinct_olaps_batch.
"""
import numpy as np
import pytest

from oracle import oracle

MIN_INTER = 10
CAP = 3
QUALS = (0.0, 0.01, 0.02, 0.02, 0.025, 0.03)


def crossing(d, b):
    """Positions of diagonal d inside bounds b = (s_left, s_right, t_left, t_right) in the half the
    rule reads, counted by enumeration."""
    s = np.arange(-5000, 5000)
    t = s + d
    if d > 0:
        return int(np.count_nonzero((s >= b[0]) & (t <= b[3])))
    return int(np.count_nonzero((t >= b[2]) & (s <= b[1])))


def add_model(results, partial=False):
    """Entries after the results by the statement above: dicts of box, quality, bounds, diagonals."""
    ents = []
    for (slo, shi, tlo, thi), q in results:
        d = tlo - slo
        hit = None
        if not partial:
            for e in ents:
                od = e["box"][2] - e["box"][0]
                if (d > 0) == (od > 0) and crossing(d, e["bounds"]) >= MIN_INTER + 1:
                    hit = e
                    break
        if hit is not None:
            b = hit["bounds"]
            hit["bounds"] = (min(b[0], slo), max(b[1], shi), min(b[2], tlo), max(b[3], thi))
            hit["diag"] = (min(hit["diag"][0], d), max(hit["diag"][1], d))
            if q < hit["q"]:
                hit["box"], hit["q"] = (slo, shi, tlo, thi), q
            continue
        if len(ents) < CAP:
            ents.append(dict(box=(slo, shi, tlo, thi), q=q, bounds=(slo, shi, tlo, thi), diag=(d, d)))
    return ents


def length(box):
    return 1 + min(box[1] - box[0], box[3] - box[2])


def rows(ents):
    return [list(e["box"]) + list(e["diag"]) + list(e["bounds"]) for e in ents]


def random_group(rng):
    """One pair's extension results: boxes near a few diagonals (some on each side of the main
    diagonal, some on it), spans from a few bases to a read."""
    diags = [int(x) for x in rng.choice([-700, -300, -40, -12, -1, 0, 1, 9, 12, 60, 250, 800], size=int(rng.integers(1, 4)))]
    out = []
    for _ in range(int(rng.integers(1, 7))):
        d = diags[int(rng.integers(0, len(diags)))] + int(rng.choice([0, 0, 0, -1, 1, -3, 3, -15, 15]))
        slo = int(rng.integers(0, 1500))
        span = int(rng.choice([5, 12, 20, 60, 400, 1500]))
        shi = slo + span
        tlo = slo + d
        thi = shi + d + int(rng.integers(-4, 5))
        out.append(((slo, shi, tlo, thi), float(rng.choice(QUALS))))
    return out


def run_oracle(group, partial=False, unique=False):
    r, q = oracle.distinct_olaps([b for b, _ in group], [x for _, x in group], partial=partial, unique=unique)
    return r, q


def test_add_overlap_matches_statement():
    rng = np.random.default_rng(7)
    joins = drops = 0
    for case in range(3000):
        g = random_group(rng)
        r, q = run_oracle(g, partial=False, unique=False)
        # Merge may delete and widen entries: compare before the selection by keeping -partial off
        # and reading the entries the merge left alone only when nothing merged
        ents = add_model(g)
        assert len(r) == len(ents), case
        if not r[:, 10].any():
            assert r[:, :10].tolist() == rows(ents), case
            assert q.tolist() == [e["q"] for e in ents], case
        joins += len(g) > len(ents)
        drops += len(ents) == CAP and len(g) > CAP
        # -partial: no joins, the first three results
        rp, qp = run_oracle(g, partial=True, unique=False)
        assert rp[:, :4].tolist() == [list(b) for b, _ in g[:CAP]] and not rp[:, 10].any(), case
        assert qp.tolist() == [x for _, x in g[:CAP]], case
    assert joins > 500 and drops > 50


def test_combine_keeps_first_least_quality_then_longest():
    rng = np.random.default_rng(11)
    ties = 0
    for case in range(3000):
        g = random_group(rng)
        r, q = run_oracle(g, unique=True)
        ents = add_model(g)
        key = sorted(range(len(ents)), key=lambda i: (ents[i]["q"], -length(ents[i]["box"]), i))
        best = key[0]
        assert r[:, 10].tolist() == [int(i != best) for i in range(len(ents))], case
        allb = np.array([e["bounds"] for e in ents])
        alld = np.array([e["diag"] for e in ents])
        assert r[best, 4] == alld[:, 0].min() and r[best, 5] == alld[:, 1].max(), case
        assert r[best, 6:10].tolist() == [allb[:, 0].min(), allb[:, 1].max(), allb[:, 2].min(), allb[:, 3].max()], case
        assert r[best, :4].tolist() == list(ents[best]["box"]) and q[best] == ents[best]["q"], case
        ties += len(ents) > 1 and ents[key[0]]["q"] == ents[key[1]]["q"]
    assert ties > 100


def test_choose_best_partial_keeps_most_matching_bases():
    rng = np.random.default_rng(12)
    not_max = 0
    for case in range(3000):
        g = random_group(rng)
        r, q = run_oracle(g, partial=True, unique=True)
        boxes = [b for b, _ in g[:CAP]]
        quals = [x for _, x in g[:CAP]]
        mb = [(1.0 - x) * (2 + b[1] - b[0] + b[3] - b[2]) for b, x in zip(boxes, quals)]
        best = 0
        for i in range(1, len(boxes)):
            if mb[i] > mb[0] or (mb[i] == mb[0] and quals[i] < quals[best]):
                best = i
        assert r[:, 10].tolist() == [int(i != best) for i in range(len(boxes))], case
        top = max(range(len(boxes)), key=lambda i: (mb[i], -i))
        not_max += best != top
    assert not_max > 20         # the streams reach the case where the pick is not the maximum


def test_merge_keeps_best_and_covers_bounds():
    rng = np.random.default_rng(13)
    merged = 0
    for case in range(3000):
        g = random_group(rng)
        r, q = run_oracle(g, unique=False)
        ents = add_model(g)
        kept = r[:, 10] == 0
        assert kept.any(), case
        merged += (~kept).any()
        for side in (True, False):          # entries with min diagonal <= 0, then > 0
            cls = np.array([(e["diag"][0] <= 0) == side for e in ents], bool)
            if not cls.any():
                continue
            assert (kept & cls).any(), case
            allb = np.array([e["bounds"] for e in ents])[cls]
            kb = r[kept & cls, 6:10]
            assert kb[:, 0].min() == allb[:, 0].min() and kb[:, 1].max() == allb[:, 1].max(), case
            assert kb[:, 2].min() == allb[:, 2].min() and kb[:, 3].max() == allb[:, 3].max(), case
            qs = np.array([e["q"] for e in ents])
            m = qs[cls].min()
            if np.count_nonzero(qs[cls] == m) == 1:            # a unique best survives
                assert kept[np.flatnonzero(cls & (qs == m))[0]], case
    assert merged > 20


def test_hand_worked_cases():
    # one diagonal, two results: one entry, the better box, bounds over both
    r, q = run_oracle([((0, 999, 100, 1099), 0.02), ((5, 990, 105, 1090), 0.01)])
    assert r.tolist() == [[5, 990, 105, 1090, 100, 100, 0, 999, 100, 1099, 0]] and q.tolist() == [0.01]
    # diagonal 600 runs through bounds s 0..999, t 100..1099 for 1099 - 600 - 0 + 1 = 500 positions
    r, _ = run_oracle([((0, 999, 100, 1099), 0.02), ((0, 999, 600, 1599), 0.03)])
    assert r.tolist() == [[0, 999, 100, 1099, 100, 600, 0, 999, 100, 1599, 0]]
    # the threshold: 11 positions join, 10 do not (bounds s 0..100, t 50..150; diagonal 140 / 141)
    r, _ = run_oracle([((0, 100, 50, 150), 0.02), ((0, 10, 140, 150), 0.03)])
    assert len(r) == 1
    r, _ = run_oracle([((0, 100, 50, 150), 0.02), ((0, 9, 141, 150), 0.03)])
    assert len(r) == 2
    # diagonal 0 belongs with the negative side: a diagonal-5 entry never takes it
    r, _ = run_oracle([((0, 500, 5, 505), 0.02), ((0, 500, 0, 500), 0.01)])
    assert len(r) == 2
    # the rule reads one corner of the bounds only: diagonal -200 never meets the box s 0..100,
    # t -600..-500, yet runs for 501 positions with t >= -600 and s <= 100, so it joins
    r, _ = run_oracle([((0, 100, -600, -500), 0.02), ((0, 100, -200, -100), 0.03)])
    assert len(r) == 1 and r[0, 4:10].tolist() == [-600, -200, 0, 100, -600, -100]
    # in the other order it does not: diagonal -600 has 100 - 600 + 200 + 1 < 11 positions with
    # t >= -200 and s <= 100; a fourth distinct result is dropped
    g = [((0, 100, d, 100 + d), 0.02) for d in (-200, -600, 300, 800)]
    r, _ = run_oracle(g, partial=False, unique=False)
    assert len(r) == 3 and (r[:, 2] - r[:, 0]).tolist() == [-200, -600, 300]
    # Combine: equal quality, the longer wins; equal quality and length, the first wins
    g = [((0, 100, -300, -200), 0.02), ((0, 300, 400, 700), 0.02)]
    r, _ = run_oracle(g, unique=True)
    assert r[:, 10].tolist() == [1, 0] and r[1, 6:10].tolist() == [0, 300, -300, 700]
    g = [((0, 100, -300, -200), 0.02), ((0, 100, 400, 500), 0.02)]
    r, _ = run_oracle(g, unique=True)
    assert r[:, 10].tolist() == [0, 1]
    # Choose: 2 x 101 bases at 0.02 beat 2 x 91 at 0.0; equal matching bases, the lower quality wins
    g = [((0, 90, 300, 390), 0.0), ((0, 100, -300, -200), 0.02)]
    r, _ = run_oracle(g, partial=True, unique=True)
    assert r[:, 10].tolist() == [1, 0]      # 0.98 * 202 = 197.96 > 182
    g = [((0, 99, 300, 399), 0.5), ((0, 49, -300, -250), 0.0)]
    r, _ = run_oracle(g, partial=True, unique=True)
    assert r[:, 10].tolist() == [1, 0]      # 0.5 * 200 = 1.0 * 100, the second has the lower quality
    # matching bases 9.7, 2923.05, 21.56: the third beats the first, so it is kept, not the second
    g = [((1443, 1448, 1500, 1503), 0.03), ((1042, 2542, 1117, 2613), 0.025), ((283, 295, -17, -9), 0.02)]
    r, _ = run_oracle(g, partial=True, unique=True)
    assert r[:, 10].tolist() == [1, 1, 0]
    # Merge: entries that became intersecting after a join; equal quality keeps the later one.
    # Entry 0 (diag 100, s 0..50) and entry 1 (diag 160, s 400..600) do not meet; a result on
    # diagonal 100 at s 300..700 widens entry 0's bounds to s 0..700, t 100..800, which diagonal
    # 160 crosses for 800 - 160 - 0 + 1 = 641 positions
    g = [((0, 50, 100, 150), 0.02), ((400, 600, 560, 760), 0.02), ((300, 700, 400, 800), 0.03)]
    r, _ = run_oracle(g)
    assert len(r) == 2 and r[:, 10].tolist() == [1, 0]
    assert r[1, 4:10].tolist() == [100, 160, 0, 700, 100, 800]
    # the same with the first entry strictly better: it survives and takes the other's bounds
    g = [((0, 50, 100, 150), 0.01), ((400, 600, 560, 760), 0.02), ((300, 700, 400, 800), 0.03)]
    r, _ = run_oracle(g)
    assert r[:, 10].tolist() == [0, 1] and r[0, 4:10].tolist() == [100, 160, 0, 700, 100, 800]


def test_planted_mistake_is_caught():
    """A Combine that lets the later entry win quality ties fails the statement."""
    rng = np.random.default_rng(11)
    caught = 0
    for _ in range(300):
        ents = add_model(random_group(rng))
        right = sorted(range(len(ents)), key=lambda i: (ents[i]["q"], -length(ents[i]["box"]), i))[0]
        wrong = sorted(range(len(ents)), key=lambda i: (ents[i]["q"], -length(ents[i]["box"]), -i))[0]
        caught += right != wrong
    assert caught > 0


@pytest.mark.gpu
@pytest.mark.parametrize("partial,unique", [(False, True), (False, False), (True, True), (True, False)])
def test_gpu_distinct_olaps_match_oracle(partial, unique):
    from canu_amd import oic

    rng = np.random.default_rng(21 + 2 * partial + unique)
    groups = [random_group(rng) for _ in range(5000)] + [[]]
    eng = oic.Engine(partial=partial, unique=unique)
    got = eng.distinct_olaps([[(*b, q) for b, q in g] for g in groups])
    for i, g in enumerate(groups):
        if not g:
            assert len(got[i]) == 0
            continue
        r, q = run_oracle(g, partial=partial, unique=unique)
        x = got[i]
        cols = ["s_lo", "s_hi", "t_lo", "t_hi", "min_diag", "max_diag", "s_left", "s_right", "t_left", "t_right", "deleted"]
        assert np.stack([x[c] for c in cols], 1).tolist() == r.tolist(), i
        assert x["quality"].tolist() == q.tolist(), i
