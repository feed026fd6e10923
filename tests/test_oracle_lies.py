"""Lies_On_Alignment (overlapInCore-Process_String_Overlaps.C:260-281): after a dovetail extension,
Process_Matches drops every match node that lies on the alignment (:476-490).  SHIFT_SLACK = 1.

The oracle's function is checked against:
* the empty delta (no indels): a node lies on the alignment exactly when its diagonal is within 1
  of the alignment's;
* alignments written out base by base from a delta (the convention test_oracle_dp.py pins: |d| - 1
  diagonal steps, then a T base inserted for d < 0 or an S base deleted for d > 0) whose diagonal
  runs are all longer than twice the number of indels: a node on the path lies on it, and a node
  whose diagonal is at least 3 from every run's diagonal does not;
* hand cases at the run boundaries.  The reference's S cursor (s_lo) advances |d| per entry plus
  one more after a deletion, one base beyond the path per entry, so near a boundary a node is
  judged against the run before: with delta [5, 5] from (0, 0) the path is on diagonal -2 from S
  position 10 on, yet a node at S 10 is judged against diagonal -1; with delta [-5] the path is on
  diagonal 1 from S 4, yet a node at S 4 is judged against diagonal 0.
"""
import numpy as np

from oracle import oracle

SLACK = 1


def path(s_lo, t_lo, delta, tail):
    """The alignment's (S, T) positions, base by base, then tail diagonal steps."""
    i, j = s_lo, t_lo
    pts = []
    for d in delta:
        for _ in range(abs(d) - 1):
            pts.append((i, j))
            i += 1
            j += 1
        if d < 0:
            j += 1
        else:
            i += 1
    for _ in range(tail):
        pts.append((i, j))
        i += 1
        j += 1
    return pts


def test_no_indels_closed_form():
    rng = np.random.default_rng(1)
    for _ in range(5000):
        s_lo, t_lo = (int(x) for x in rng.integers(0, 5000, 2))
        start = int(rng.integers(s_lo, s_lo + 3000))
        offset = start + (t_lo - s_lo) + int(rng.integers(-4, 5))
        want = abs((offset - start) - (t_lo - s_lo)) <= SLACK
        assert oracle.lies_on_alignment(start, offset, s_lo, t_lo, []) == want


def test_nodes_on_and_off_long_runs():
    rng = np.random.default_rng(2)
    on = off = 0
    for case in range(400):
        n = int(rng.integers(1, 8))
        run = 2 * n + 2
        delta = [int(rng.integers(run + 1, run + 40)) * int(rng.choice([-1, 1])) for _ in range(n)]
        s_lo, t_lo = (int(x) for x in rng.integers(0, 3000, 2))
        pts = path(s_lo, t_lo, delta, int(rng.integers(run, run + 40)))
        diags = sorted({t - s for s, t in pts})
        for _ in range(30):
            s, t = pts[int(rng.integers(0, len(pts)))]
            assert oracle.lies_on_alignment(s, t, s_lo, t_lo, delta), (case, s, t)
            on += 1
            g = int(rng.integers(diags[0] - 20, diags[-1] + 21))
            if min(abs(g - d) for d in diags) >= 3:
                assert not oracle.lies_on_alignment(s, s + g, s_lo, t_lo, delta), (case, s, g)
                off += 1
    assert on > 10000 and off > 2000


def test_boundary_cases_by_hand():
    # [5, 5] from (0, 0): S 0-3 on diagonal 0, S 4 deleted, S 5-8 on -1, S 9 deleted, S 10 on -2.
    # The reference judges S positions below 5 against 0, 5-10 against -1 and from 11 on against -2.
    assert oracle.lies_on_alignment(3, 3, 0, 0, [5, 5])
    assert oracle.lies_on_alignment(10, 8, 0, 0, [5, 5])          # on the path: -2 is within 1 of -1
    assert not oracle.lies_on_alignment(10, 7, 0, 0, [5, 5])      # -3: 1 from the path, 2 from -1
    assert oracle.lies_on_alignment(11, 8, 0, 0, [5, 5])          # -3 from S 11 on: within 1 of -2
    assert oracle.lies_on_alignment(4, 5, 0, 0, [5, 5])           # diagonal 1 at S 4, judged against 0
    assert not oracle.lies_on_alignment(4, 6, 0, 0, [5, 5])       # diagonal 2 there: 2 from 0
    # [-5] from (0, 0): S 0-3 on diagonal 0, a T base inserted, S 4 on diagonal 1.
    assert oracle.lies_on_alignment(3, 4, 0, 0, [-5])
    assert not oracle.lies_on_alignment(4, 6, 0, 0, [-5])          # 2: 1 from the path, 2 from 0
    assert oracle.lies_on_alignment(5, 7, 0, 0, [-5])              # from S 5 on: against 1
    # the slack is symmetric and exactly 1
    for g in (-2, -1, 0, 1, 2):
        assert oracle.lies_on_alignment(20, 20 + g, 0, 0, []) == (abs(g) <= 1)
