"""The two formulations of Add_Match the engine's list builders compute (canu_amd/csrc/oic_search.h),
modelled on the CPU and checked against the oracle's Add_Match (overlapInCore-Find_Overlaps.C:26-96)
on the random hit streams of test_oracle_lists.py.

The wavefront builder (k_build_lists_wave) works per S offset, with the live nodes (those touched
at the previous offset) and their list positions:
* each hit finds the live node on its diagonal, if any; tag each list position with the hit that
  touches it (none: the node dies);
* a touched node stays where it is exactly when no earlier list position is dead or touched by a
  later hit (the exclusive prefix maximum of the tags is below its own tag), else it moves to the
  front; a hit with no live node creates a node at the front;
* the new list: the front nodes in reverse hit order, then the nodes kept in place in their old
  order; dead nodes leave for the list's tail in list order; an offset that does not follow the
  previous one kills every live node first.
The model keeps the kernel's encoding: list positions counted from the front for live nodes, and
a dead-queue index for the others, turned into final positions at the end.

The one-thread builder (k_build_lists) works per hit in O(1): the live nodes are kept sorted by
diagonal (the previous offset's hits in order), so a hit finds its node by a merge; a pointer to
the first untouched live list position says whether an extended node moves to the front; the
nodes not touched at an offset form a dead block, and the final list is the live nodes followed by
the dead blocks, newest first.
"""
import numpy as np

from oracle import oracle
from test_oracle_lists import K, random_stream

INF = 1 << 30


def formulation(hits, k=K):
    live = []          # hit order of the last offset: [offset, len, start, list position, diagonal]
    dq = []            # dead nodes: [offset, len, start, w]
    consistent, nnode, prev_o, head = True, 0, -2, None
    i = 0
    while i < len(hits):
        o = hits[i][0]
        j = i
        while j < len(hits) and hits[j][0] == o:
            j += 1
        group = hits[i:j]
        i = j
        nA = len(live)
        if o != prev_o + 1 and nA > 0:               # a gap: every live node dies, in list order
            order = sorted(live, key=lambda r: r[3])
            ndq = len(dq)
            for kk, r in enumerate(order):
                dq.append([r[0], r[1], r[2], ndq + nA - kk])
            head = order[0]
            live, nA = [], 0
        c = len(group)
        g = [t - o for _, t in group]
        by_diag = {r[4]: q for q, r in enumerate(live)}
        qs = [by_diag.get(gj, -1) for gj in g]
        tag = [INF] * nA
        for jj, q in enumerate(qs):
            if q >= 0:
                tag[live[q][3]] = jj
        pm, run = [], -1                               # exclusive prefix maximum of the tags
        for pos in range(nA):
            pm.append(run)
            run = max(run, tag[pos])
        cf = [q < 0 or pm[live[q][3]] > jj for jj, q in enumerate(qs)]
        rank, nfr = [], 0
        for f in cf:
            rank.append(nfr)
            nfr += f
        ncre = sum(q < 0 for q in qs)
        bad = any(q < 0 and (jj > 0 or nA > 0) for jj, q in enumerate(qs))
        if nA == 0 and nnode > 0:
            es, dg = head[2] + head[1] - k + 1, head[0] - head[2]
            bad = bad or abs(dg - g[0]) > 3 or o < es + k - 2
        consistent = consistent and not bad
        sz = nA - (c - ncre)
        ndq = len(dq)
        inpl_rank, ni, nd = {}, 0, 0
        pos_to_q = {r[3]: q for q, r in enumerate(live)}
        dead_recs = [None] * sz
        for pos in range(nA):
            dead = tag[pos] == INF
            if not dead and pm[pos] < tag[pos]:
                inpl_rank[pos] = ni
                ni += 1
            if dead:
                r = live[pos_to_q[pos]]
                dead_recs[nd] = [r[0], r[1], r[2], ndq + sz - nd]
                nd += 1
        dq.extend(dead_recs)
        new = []
        for jj, q in enumerate(qs):
            if q < 0:
                rec = [g[jj] + o, k, o]
                P = nfr - 1 - rank[jj]
            else:
                r = live[q]
                rec = [r[0], r[1] + 1, r[2]]
                P = nfr - 1 - rank[jj] if cf[jj] else nfr + inpl_rank[r[3]]
            new.append(rec + [P, g[jj]])
        nnode += ncre
        live = new
        prev_o = o
    nA, ndq = len(live), len(dq)
    out = [None] * (nA + ndq)
    for r in live:
        out[r[3]] = (r[0], r[1], r[2])
    for r in dq:
        out[nA + ndq - r[3]] = (r[0], r[1], r[2])
    return np.array(out, np.int32).reshape(-1, 3), consistent


def o1_builder(hits, k=K):
    nodes = []                 # [offset, len, start]
    s_cur, o_cur = [], []      # live node ids by diagonal (descending) and in list order
    pos = {}                   # live node id -> list position
    dq, db = [], []            # dead node ids, dead block starts
    consistent, prev_o, i = True, -2, 0
    while i < len(hits):
        o = hits[i][0]
        if o != prev_o + 1 and o_cur:
            db.append(len(dq))
            dq.extend(o_cur)
            s_cur, o_cur = [], []
        nA = len(o_cur)
        fl = [0] * nA
        f = q = nmoved = 0
        fr, s_new = [], []
        while i < len(hits) and hits[i][0] == o:
            g = hits[i][1] - o
            i += 1
            while q < nA and nodes[s_cur[q]][0] - nodes[s_cur[q]][2] > g:
                q += 1
            if q < nA and nodes[s_cur[q]][0] - nodes[s_cur[q]][2] == g:
                x = s_cur[q]
                ix = pos[x]
                nodes[x][1] += 1
                fl[ix] = 1
                if f < ix:
                    fl[ix] = 2
                    fr.append(x)
                    nmoved += 1
                while f < nA and fl[f]:
                    f += 1
                s_new.append(x)
                q += 1
            else:
                if nodes and consistent:
                    if len(fr) + nA - nmoved > 0:
                        consistent = False
                    else:
                        hd = nodes[dq[db[-1]]]
                        es, dg = hd[2] + hd[1] - k + 1, hd[0] - hd[2]
                        if abs(dg - g) > 3 or o < es + k - 2:
                            consistent = False
                nodes.append([g + o, k, o])
                fr.append(len(nodes) - 1)
                s_new.append(len(nodes) - 1)
        o_new = list(reversed(fr))
        dead_block = False
        for ix in range(nA):
            if fl[ix] == 1:
                o_new.append(o_cur[ix])
            elif fl[ix] == 0:
                if not dead_block:
                    db.append(len(dq))
                    dead_block = True
                dq.append(o_cur[ix])
        pos = {x: n for n, x in enumerate(o_new)}
        s_cur, o_cur = s_new, o_new
        prev_o = o
    out = list(o_cur)
    for b in range(len(db) - 1, -1, -1):
        e = db[b + 1] if b + 1 < len(db) else len(dq)
        out.extend(dq[db[b]:e])
    return np.array([nodes[x] for x in out], np.int32).reshape(-1, 3), consistent


def test_o1_builder_matches_oracle_on_random_streams():
    rng = np.random.default_rng(2027)
    for case in range(3000):
        hits = random_stream(rng)
        a, ca = oracle.match_list(K, [h[0] for h in hits], [h[1] for h in hits])
        b, cb = o1_builder(hits)
        assert np.array_equal(a, b), case
        assert ca == cb, case


def test_formulation_matches_oracle_on_random_streams():
    rng = np.random.default_rng(2026)
    for case in range(3000):
        hits = random_stream(rng)
        a, ca = oracle.match_list(K, [h[0] for h in hits], [h[1] for h in hits])
        b, cb = formulation(hits)
        assert np.array_equal(a, b), case
        assert ca == cb, case


def test_formulation_hand_cases():
    # one run: one node; two nodes at one offset both extended: the second moves to the front
    assert formulation([(0, 100), (1, 101), (2, 102)])[0].tolist() == [[100, K + 2, 0]]
    lst, cons = formulation([(0, 100), (0, 50), (1, 101), (1, 51)])
    assert lst.tolist() == [[100, K + 1, 0], [50, K + 1, 0]] and not cons
