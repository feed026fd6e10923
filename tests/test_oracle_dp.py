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
