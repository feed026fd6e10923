"""The oracle's Add_Match (overlapInCore-Find_Overlaps.C:26-96) against a statement of its effect,
derived from that code but written as operations on blocks of nodes rather than its pointer walk:

* a node is a maximal run of seed hits at consecutive S offsets on one diagonal (T - S offset):
  Offset, Start = the run's first hit, Len = k + hits - 1, expected start es = last offset + 1;
* the list stays ordered by es, descending: when offset o is processed, the nodes with es = o (last
  touched at o - 1, "waiting") follow the nodes already touched at o (es = o + 1, the "block");
* a hit on the diagonal of a waiting node extends it; the node joins the block's back if it was the
  first waiting node (it stays where it is), else the block's front (moved to the front);
* any other hit creates a node at the block's front; the pair stops being consistent if the list
  was not empty and either the block or the waiting nodes are not empty (nodes scanned before the
  walk stops), or the first older node (es < o; the last node if there is none) lies more than 3
  diagonals away or o < its es + k - 2.

Hits arrive by S offset ascending, T offset descending, as the engine's hit sort leaves them.
Streams mix long runs on a few diagonals (true overlap and repeat copies), nearby diagonals
(indels), crowded offsets (satellites) and isolated hits.
"""
import numpy as np
import pytest

from oracle import oracle

K = 22


class Node:
    __slots__ = ("off", "start", "len", "es", "diag")

    def __init__(self, off, start, length, es, diag):
        self.off, self.start, self.len, self.es, self.diag = off, start, length, es, diag


def block_model(hits, k=K, always_front=False):
    """The list and consistent flag after the hits, by the block statement above (always_front:
    a planted mistake, every extended node moved to the front)."""
    lst, consistent = [], True
    i = 0
    while i < len(hits):
        o = hits[i][0]
        j = i
        while j < len(hits) and hits[j][0] == o:
            j += 1
        block = []
        waiting = [n for n in lst if n.es == o]
        rest = [n for n in lst if n.es < o]
        assert len(waiting) + len(rest) == len(lst)
        for _, t in hits[i:j]:
            d = t - o
            x = next((n for n in waiting if n.diag == d), None)
            if x is not None:
                first = waiting[0] is x
                waiting.remove(x)
                x.len += 1
                x.es = o + 1
                if first and not always_front:
                    block.append(x)
                else:
                    block.insert(0, x)
                continue
            cur = block + waiting + rest
            if cur:
                stop = rest[0] if rest else cur[-1]
                if len(block) + len(waiting) > 0 or abs(stop.diag - d) > 3 or o < stop.es + k - 2:
                    consistent = False
            block.insert(0, Node(t, o, k, o + 1, d))
        lst = block + waiting + rest
        i = j
    return np.array([(n.off, n.len, n.start) for n in lst], np.int32).reshape(-1, 3), consistent


def random_stream(rng):
    """One pair's seed hits: (S offset, T offset), sorted as the engine emits them."""
    hits = set()
    span = int(rng.integers(200, 3000))
    for _ in range(int(rng.integers(1, 6))):                  # runs on a few diagonals
        d = int(rng.integers(-500, 500))
        o = int(rng.integers(0, span))
        for _ in range(int(rng.integers(1, 8))):
            run = int(rng.integers(1, 60))
            for q in range(run):
                if o + q + d >= 0:
                    hits.add((o + q, o + q + d))
            o += run + int(rng.integers(1, 40))
            d += int(rng.choice([0, 0, 0, -1, 1, -2, 2, 5]))  # indels shift the diagonal
    for _ in range(int(rng.integers(0, 4))):                  # a crowded stretch (satellite copies)
        o0 = int(rng.integers(0, span))
        period = int(rng.integers(2, 12))
        for q in range(int(rng.integers(5, 40))):
            for c in range(int(rng.integers(2, 9))):
                t = o0 + q + c * period + int(rng.integers(0, 3))
                hits.add((o0 + q, t))
    for _ in range(int(rng.integers(0, 30))):                 # isolated hits
        o = int(rng.integers(0, span))
        hits.add((o, int(rng.integers(0, span + 600))))
    return sorted(hits, key=lambda h: (h[0], -h[1]))


def oracle_list(hits):
    o = [h[0] for h in hits]
    t = [h[1] for h in hits]
    return oracle.match_list(K, o, t)


def test_small_cases_by_hand():
    # one run of three hits: one node, Len k + 2
    lst, cons = oracle_list([(0, 100), (1, 101), (2, 102)])
    assert lst.tolist() == [[100, K + 2, 0]] and cons
    # a second run on the same diagonal after a gap shorter than k - 2 past the first node's end:
    # a new node at the front, no longer consistent
    lst, cons = oracle_list([(0, 100), (1, 101), (2, 102), (10, 110), (11, 111)])
    assert lst.tolist() == [[110, K + 1, 10], [100, K + 2, 0]] and not cons
    # the same after a gap of at least k - 2: still consistent
    lst, cons = oracle_list([(0, 100), (1, 101), (2, 102), (30, 130)])
    assert lst.tolist() == [[130, K, 30], [100, K + 2, 0]] and cons
    # two hits at one offset on diagonals 100 and 50 (T offset descending), both extended at the
    # next offset: at offset 1 the node of diagonal 100 (the list's second) is found after the
    # waiting node of diagonal 50, so it moves to the front
    lst, cons = oracle_list([(0, 100), (0, 50), (1, 101), (1, 51)])
    assert lst.tolist() == [[100, K + 1, 0], [50, K + 1, 0]] and not cons


def test_block_model_matches_oracle_on_random_streams():
    rng = np.random.default_rng(2026)
    n_moves = n_incons = 0
    for case in range(3000):
        hits = random_stream(rng)
        a, ca = oracle_list(hits)
        b, cb = block_model(hits)
        assert np.array_equal(a, b), case
        assert ca == cb, case
        es = a[:, 2] + a[:, 1] - K + 1
        assert np.all(np.diff(es) <= 0), case                 # ordered by expected start, descending
        n_incons += not ca
        c, _ = block_model(hits, always_front=True)
        n_moves += not np.array_equal(a, c)
    # the streams exercise both orders of an extended node and both consistent outcomes
    assert n_moves > 300 and 100 < n_incons < 3000
