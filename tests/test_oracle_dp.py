"""The oracle's prefixEditDistance::forward / ::reverse / Extend_Alignment.

Pinned against:
  * textbook edit distance (an O(mn) matrix with 'n' matching anything): with pruning off
    (Edit_Match_Limit = 0) and Error_Limit <= 10 no branch point can fire (forward.C:189
    needs e > 10), so forward returns the first e for which a diagonal d (lowest first) has
    max{i : D(i, i+d) <= e} reaching the end of A or T, plus the forced-mismatch shift
    (forward.C:215-221); reverse() is the same on the reversed strings;
  * the delta invariant: walking Right_Delta / Left_Delta with Display_Alignment's rules
    (|delta|-1 paired bases, then one base of T for delta < 0 or of A for delta > 0)
    reproduces an alignment with exactly e edits ending where the call says;
  * no state carried between calls: with every row poisoned first, no unwritten cell is read.
"""
import numpy as np
import pytest

from oracle import oracle

ZERO_LIMIT = np.zeros(200000, np.int32)


def mutate(rng, s, rate, alphabet="acgt"):
    out = []
    for ch in s:
        u = rng.random()
        if u < rate / 3:
            out.append(rng.choice([c for c in alphabet if c != ch]))
        elif u < 2 * rate / 3:
            continue
        elif u < rate:
            out.append(rng.choice(list(alphabet)))
            out.append(ch)
        else:
            out.append(ch)
    return "".join(out)


def edit_matrix(A, T):
    m, n = len(A), len(T)
    D = np.zeros((m + 1, n + 1), np.int64)
    D[:, 0] = np.arange(m + 1)
    D[0, :] = np.arange(n + 1)
    for i in range(1, m + 1):
        for j in range(1, n + 1):
            eq = A[i - 1] == T[j - 1] or A[i - 1] == "n" or T[j - 1] == "n"
            D[i, j] = min(D[i - 1, j] + 1, D[i, j - 1] + 1, D[i - 1, j - 1] + (0 if eq else 1))
    return D


def brute_prefix(A, T, limit, forward=True):
    """(e, d, row) of the first end-reaching cell, with forward()'s mismatch shift, or None."""
    m, n = len(A), len(T)
    D = edit_matrix(A, T)

    def fr(e, d):
        best = None
        for i in range(0, m + 1):
            j = i + d
            if 0 <= j <= n and D[i, j] <= e:
                best = i
        return best

    for e in range(0, limit + 1):
        for d in range(-e, e + 1):
            i = fr(e, d)
            if i is None:
                continue
            if i == m or (e > 0 and i + d == n):
                dd = d
                if forward and e > 0 and i == m and d < e:
                    p = fr(e - 1, d + 1) if abs(d + 1) <= e - 1 else None
                    if p is not None and p + 1 == i:
                        dd = d + 1
                return e, dd, i
    return None


def walk_delta(A, T, delta, i, j):
    edits = 0
    for dl in delta:
        for _ in range(abs(int(dl)) - 1):
            if i >= len(A) or j >= len(T):
                return None
            edits += not (A[i] == T[j] or A[i] == "n" or T[j] == "n")
            i += 1
            j += 1
        if dl < 0:
            j += 1
        else:
            i += 1
        edits += 1
    return i, j, edits


def diag_edits(A, T, i, j, k):
    return sum(not (A[i + t] == T[j + t] or A[i + t] == "n" or T[j + t] == "n") for t in range(k))


def random_pair(rng, alphabet="acgt"):
    A = "".join(rng.choice(list(alphabet), int(rng.integers(1, 40))))
    T = mutate(rng, A, float(rng.choice([0.0, 0.05, 0.15, 0.3])), alphabet) + "".join(
        rng.choice(list(alphabet), int(rng.integers(0, 8))))
    if len(T) == 0:
        T = "a"
    if len(A) > len(T):
        A, T = T, A
    return A, T


@pytest.mark.parametrize("seed", range(4))
def test_forward_matches_textbook(seed):
    rng = np.random.default_rng(seed)
    od = oracle.OracleDP(0.045, match_limit=ZERO_LIMIT)
    n_end = 0
    for _ in range(150):
        A, T = random_pair(rng, "acgt" if seed % 2 == 0 else "acgtn")
        limit = int(rng.integers(0, 11))
        e, ae, te, mte, delta = od.forward(A.encode(), T.encode(), limit)
        b = brute_prefix(A, T, limit, forward=True)
        if b is None:
            assert not mte
            continue
        n_end += 1
        be, bd, brow = b
        assert mte and e == be, (A, T, limit, e, be)
        assert ae == brow and te == brow + bd, (A, T, limit, ae, te, brow, bd)
    assert n_end > 50


@pytest.mark.parametrize("seed", range(4))
def test_reverse_matches_textbook(seed):
    rng = np.random.default_rng(100 + seed)
    od = oracle.OracleDP(0.045, match_limit=ZERO_LIMIT)
    for _ in range(150):
        A, T = random_pair(rng)
        limit = int(rng.integers(0, 11))
        e, ae, te, lo, mte, delta = od.reverse(A.encode(), T.encode(), limit)
        b = brute_prefix(A[::-1], T[::-1], limit, forward=False)
        if b is None:
            assert not mte
            continue
        be, bd, brow = b
        assert mte and e == be
        assert ae == -brow
        # Set_Left_Delta may trade a leading indel for a mismatch and move T_End by one
        assert te in (-brow - bd, -brow - bd - 1)


@pytest.mark.parametrize("rate", [0.045, 0.01])
def test_delta_describes_alignment(rate):
    rng = np.random.default_rng(7)
    od = oracle.OracleDP(rate)
    checked = 0
    for _ in range(300):
        L = int(rng.integers(20, 400))
        A = "".join(rng.choice(list("acgt"), L))
        T = mutate(rng, A, float(rng.choice([0.01, 0.03, 0.08]))) + "".join(rng.choice(list("acgt"), int(rng.integers(0, 30))))
        if len(A) > len(T):
            A, T = T, A
        limit = int(np.ceil(len(A) * 0.06)) + 2
        e, ae, te, mte, delta = od.forward(A.encode(), T.encode(), limit)
        if mte:
            r = walk_delta(A, T, delta, 0, 0)
            assert r is not None
            i, j, ed = r
            assert ae - i == te - j >= 0
            assert ed + diag_edits(A, T, i, j, ae - i) == e
            checked += 1
        e, ae, te, lo, mte, delta = od.reverse(A.encode(), T.encode(), limit)
        if mte:
            m, n = len(A), len(T)
            r = walk_delta(A, T, delta, m + ae, n + te)
            assert r is not None
            i, j, ed = r
            assert m - i == n - j == lo
            assert ed + diag_edits(A, T, i, j, lo) == e
            checked += 1
    assert checked > 200


def test_exact_match_is_zero():
    od = oracle.OracleDP(0.045)
    s = b"acgtacgtttgacca" * 20
    assert od.forward(s, s + b"acg", 10)[:4] == (0, len(s), len(s), True)
    e, ae, te, lo, mte, d = od.reverse(s, b"ttt" + s, 10)
    assert (e, ae, te, lo, mte) == (0, -len(s), -len(s), len(s), True)


def test_no_state_between_calls():
    rng = np.random.default_rng(3)
    od = oracle.OracleDP(0.045)
    od.poison(True)
    for _ in range(200):
        L = int(rng.integers(50, 2000))
        A = "".join(rng.choice(list("acgt"), L))
        T = mutate(rng, A, float(rng.choice([0.02, 0.08, 0.2])))
        if len(A) > len(T):
            A, T = T, A
        od.forward(A.encode(), T.encode(), int(len(A) * 0.045) + 1)
        od.reverse(A.encode(), T.encode(), int(len(A) * 0.045) + 1)
    assert od.poison_hits() == 0


def test_extend_dovetail_alignment():
    """A DOVETAIL Extend_Alignment's merged Left_Delta walks from (S_Lo, T_Lo) to (S_Hi, T_Hi)."""
    rng = np.random.default_rng(11)
    od = oracle.OracleDP(0.045)
    seen = 0
    for _ in range(200):
        G = "".join(rng.choice(list("acgt"), 3000))
        a0, b0 = int(rng.integers(0, 1000)), int(rng.integers(0, 1000))
        S = mutate(rng, G[a0:a0 + 1800], 0.01)
        T = mutate(rng, G[b0:b0 + 1800], 0.01)
        # seed: a shared 22-mer
        for start in range(0, len(S) - 22, 7):
            off = T.find(S[start:start + 22])
            if off >= 0:
                break
        else:
            continue
        kind, s_lo, s_hi, t_lo, t_hi, err, delta = od.extend(S.encode(), T.encode(), start, off, 22)
        if kind != 3:
            continue
        r = walk_delta(S, T, delta, s_lo, t_lo)
        assert r is not None
        i, j, ed = r
        assert s_hi + 1 - i == t_hi + 1 - j >= 0
        assert ed + diag_edits(S, T, i, j, s_hi + 1 - i) == err
        assert (s_lo == 0 or t_lo == 0) and (s_hi == len(S) - 1 or t_hi == len(T) - 1)
        seen += 1
    assert seen > 50


# ---------------------------------------------------------------------------------------------
# Branch points and pruning, pinned by a level-by-level search over the edit grid.
#
# grid_prefix() computes what prefixEditDistance::forward / ::reverse return without their
# furthest-reaching recurrence: at error level e it keeps the SET of grid cells (i on A, j on T)
# reachable with at most e edits (one edit per mismatch, insertion or deletion step, then any run
# of equal bases for free) from the cells of level e - 1 that lie on diagonals d = j - i of that
# level's band.  Its furthest cell per diagonal is what the DP's Edit_Array row must hold.  On
# those values it applies the rules as prefixEditDistance-forward.C:147-313 states them: the first
# cell (lowest d) at the end of A or T ends the call, unless the branch-point test fires
# (Score = Row * Branch_Match_Value - e, Tail_Len = Row - Max_Score_Len,
# slope = (Max_Score - Score) / Tail_Len, e > 10, Tail_Len >= 20 as uint32, slope >= min slope;
# -partial also aborts when Score < Max_Score); after a row, the band is pruned from both ends by
# Edit_Match_Limit (prefixEditDistance.H:249-257, forward.C:245-280), the first maximal cell
# (lowest d) of the pruned band updates Longest, and Score = Longest * Branch_Match_Value - e
# updates Max_Score.  Nothing here rolls the furthest-reaching max(1 + prev[d], prev[d - 1],
# 1 + prev[d + 1]) or slides bases one by one.
# ---------------------------------------------------------------------------------------------

def _eq(a, b):
    return a == b or a == "n" or b == "n"


def grid_prefix(A, T, limit, ml, bmv, slope_min, forward=True, partial=False):
    """(e, a_end, t_end, mte, why) of forward() (or reverse() on reversed strings, forward=False)."""
    m, n = len(A), len(T)
    E = limit + 2
    ND = 2 * E + 5
    off = E + 2
    ii = np.arange(m + 1)
    Aa = np.frombuffer(A.encode(), np.uint8)
    Ta = np.frombuffer(T.encode(), np.uint8)
    N = ord("n")
    # match[x, i]: bases A[i] and T[i + d] (d = x - off) are equal, so (i, i+d) -> (i+1, i+d+1) is free
    match = np.zeros((ND, m + 1), bool)
    valid = np.zeros((ND, m + 1), bool)
    for x in range(ND):
        d = x - off
        j = ii + d
        valid[x] = (j >= 0) & (j <= n)
        ok = (ii < m) & (j >= 0) & (j < n)
        if ok.any():
            ai = Aa[ii[ok]]
            tj = Ta[j[ok]]
            match[x, ok] = (ai == tj) | (ai == N) | (tj == N)

    def close(R):
        for i in range(m):
            R[:, i + 1] |= R[:, i] & match[:, i]
        R &= valid
        return R

    def far(R, x):
        w = np.nonzero(R[x])[0]
        return int(w[-1]) if len(w) else None

    R = np.zeros((ND, m + 1), bool)
    R[off, 0] = True
    R = close(R)
    row0 = far(R, off)
    if row0 == m:
        return 0, m, m, True, "exact"
    fr_prev = {0: row0}
    L, Rt = 0, 0
    Longest = Best_d = Best_e = 0
    MSL = MSd = MSe = 0
    MaxScore = 0.0
    for e in range(1, limit + 1):
        src = np.zeros_like(R)
        for d in range(L, Rt + 1):
            src[d + off] = R[d + off]
        new = src.copy()
        for d in range(L, Rt + 1):
            x = d + off
            s = src[x]
            j = ii + d
            step = s & (ii < m) & (j < n)            # mismatch: (i, j) -> (i + 1, j + 1)
            new[x, 1:] |= step[:-1]
            step = s & (ii < m)                      # A base skipped: (i, j) -> (i + 1, j), diagonal d - 1
            new[x - 1, 1:] |= step[:-1]
            step = s & (j < n)                       # T base skipped: (i, j) -> (i, j + 1), diagonal d + 1
            new[x + 1] |= step
        new = close(new)
        L, Rt = max(L - 1, -e), min(Rt + 1, e)
        fr = {d: far(new, d + off) for d in range(L, Rt + 1)}
        assert all(v is not None for v in fr.values())
        for d in range(L, Rt + 1):
            row = fr[d]
            if row == m or row + d == n:
                Score = row * bmv - e
                Tail = row - MSL
                with np.errstate(divide="ignore", invalid="ignore"):
                    slope = np.float64(MaxScore - Score) / np.float64(Tail)
                abort = (partial and Score < MaxScore) or (e > 10 and (Tail >= 20 or Tail < 0) and slope >= slope_min)
                if abort:
                    return MSe, MSL, MSL + MSd, False, "branch"
                dd = d
                if forward and row == m and d < Rt:
                    p = fr_prev.get(d + 1, -2) if L + 1 <= d + 1 <= Rt - 1 else -2
                    if 1 + p == row:
                        dd = d + 1
                return e, row, row + dd, True, "end"
        # prune both ends (forward.C:245-280)
        def pr(d, extra):
            return fr[d] + extra < ml[e]
        while L <= Rt and L < 0 and pr(L, 0):
            L += 1
        if L >= 0:
            while L <= Rt and pr(L, L):
                L += 1
        if L > Rt:
            return MSe, MSL, MSL + MSd, False, "emptied"
        while Rt > 0 and pr(Rt, Rt):
            Rt -= 1
        if Rt <= 0:
            while pr(Rt, 0):
                Rt -= 1
        for d in range(L, Rt + 1):
            if fr[d] > Longest:
                Best_d, Best_e, Longest = d, e, fr[d]
        Score = Longest * bmv - e
        if Score > MaxScore:
            MaxScore, MSL, MSd, MSe = Score, Longest, Best_d, Best_e
        fr_prev = {d: fr[d] for d in range(L, Rt + 1)}
        R = new
    return MSe, MSL, MSL + MSd, False, "limit"


def branchy_pair(rng, rate_lo, rate_hi):
    """A shared stretch, then tails that diverge (often unrelated): branch points and pruning."""
    G = "".join(rng.choice(list("acgt"), int(rng.integers(20, 320))))
    A = mutate(rng, G, float(rng.uniform(rate_lo, rate_hi)))
    T = mutate(rng, G, float(rng.uniform(rate_lo, rate_hi)))
    kind = rng.integers(0, 8)
    if kind == 6:      # a block of 5-40 bases inserted into one copy: the band drifts to |d| >> 0
        G = "".join(rng.choice(list("acgt"), int(rng.integers(150, 350))))
        p = int(rng.integers(0, 60))
        ins = "".join(rng.choice(list("acgt"), int(rng.integers(5, 40))))
        A, T = mutate(rng, G, 0.02), mutate(rng, G[:p] + ins + G[p:], 0.02)
        if rng.random() < 0.5:
            A, T = T, A
        T += "".join(rng.choice(list("acgt"), int(rng.integers(0, 80))))
    elif kind == 7:    # A ends inside T, its last bases noisy
        G = "".join(rng.choice(list("acgt"), int(rng.integers(60, 300))))
        k = int(rng.integers(20, len(G)))
        A = G[:k - 8] + mutate(rng, G[k - 8:k], 0.5)
        T = mutate(rng, G, 0.03)
    elif kind >= 3:    # short unrelated tails after a long clean stretch: ends reached through a branch
        G = "".join(rng.choice(list("acgt"), int(rng.integers(120, 320))))
        A = mutate(rng, G, 0.01) + "".join(rng.choice(list("acgt"), int(rng.integers(20, 45))))
        T = mutate(rng, G, 0.01) + "".join(rng.choice(list("acgt"), int(rng.integers(20, 60))))
    elif kind == 0:    # unrelated tails
        A += "".join(rng.choice(list("acgt"), int(rng.integers(0, 120))))
        T += "".join(rng.choice(list("acgt"), int(rng.integers(0, 160))))
    elif kind == 1:    # a noisy shared tail
        H = "".join(rng.choice(list("acgt"), int(rng.integers(10, 120))))
        A += mutate(rng, H, 0.3)
        T += mutate(rng, H, 0.3)
    if rng.random() < 0.15 and len(A) > 3:
        k = int(rng.integers(0, len(A)))
        A = A[:k] + "n" + A[k + 1:]
    if len(A) > len(T):
        A, T = T, A
    return A, T


@pytest.mark.parametrize("rate,prune,partial,seed", [
    (0.045, False, False, 1), (0.045, True, False, 2), (0.045, True, False, 3),
    (0.01, True, False, 4), (0.12, True, False, 5), (0.045, True, True, 6)])
def test_forward_reverse_match_grid_search(rate, prune, partial, seed):
    """forward()/reverse() at error limits up to 60, with the branch-point abort, the Edit_Match_Limit
    pruning (or none) and -partial, against grid_prefix()."""
    rng = np.random.default_rng(seed)
    od = oracle.OracleDP(rate, partial=partial, match_limit=None if prune else ZERO_LIMIT)
    tb = od.tables
    ml = tb.match_limit if prune else ZERO_LIMIT
    seen = {"branch": 0, "end": 0, "emptied": 0, "limit": 0, "exact": 0}
    narrowed = 0
    for k in range(90):
        A, T = branchy_pair(rng, 0.0, 0.12 if k % 2 else 0.03)
        limit = int(min(60, max(1, rng.integers(5, 61))))
        limit = min(limit, tb.MAX_ERRORS - 1)
        for fwd in (True, False):
            if fwd:
                e, ae, te, mte, _ = od.forward(A.encode(), T.encode(), limit)
                g = grid_prefix(A, T, limit, ml, tb.Branch_Match_Value, tb.MIN_BRANCH_TAIL_SLOPE, True, partial)
            else:
                e, ae, te, lo, mte, _ = od.reverse(A.encode(), T.encode(), limit)
                g = grid_prefix(A[::-1], T[::-1], limit, ml, tb.Branch_Match_Value, tb.MIN_BRANCH_TAIL_SLOPE, False, partial)
                ae, te = -ae, -te
            ge, ga, gt, gm, why = g
            seen[why] += 1
            assert (e, ae, mte) == (ge, ga, gm), (k, fwd, A, T, limit, (e, ae, te, mte), g)
            if fwd:
                assert te == gt, (k, A, T, limit, te, g)
            else:
                # Set_Left_Delta may trade a leading indel for a mismatch and move T_End by one
                assert te in (gt, gt + 1), (k, A, T, limit, te, g)
    assert seen["end"] > 2, seen
    if rate == 0.045:
        # (at 0.12 a branch needs a tail with more edits than bases: slope >= 1.0)
        assert seen["branch"] > 5, seen
    if prune:
        assert seen["emptied"] + seen["branch"] > 5, seen
