"""The flat-row kernel's in-place band update (flat_row in canu_amd/csrc/oic_flat.h), modelled on
the CPU against a row computed from two buffers.

One row e of the furthest-reaching recurrence (prefixEditDistance-forward.C:155-167): for each
diagonal d of the band [Left, Right],
    Row = max(1 + prev[d], prev[d - 1], 1 + prev[d + 1]), then the slide along d.
The kernel keeps ONE ring buffer of CAP cells per DP (cell d at d mod CAP, cells 0 and 1 mirrored at
CAP and CAP + 1, so the three reads are three consecutive words) and overwrites it in place.  The
band's cells are spread over nl lanes and C chunk iterations: lane j takes d = Left + j + nl k in
chunk k.  In a chunk every lane reads before any lane writes; the cell d - 1 a run's first lane
needs in chunk k > 0 was overwritten in chunk k - 1, so it takes the value the run's last lane read
there (p1, its prev[d]) by a shuffle.  The model runs that schedule and must equal the two-buffer
row for every band, lane count and prior row.
"""
import numpy as np

CAP, MASK = 64, 63


def slide(A, T, d, row):
    m, n = len(A), len(T)
    while row < m and row + d < n and A[row] == T[row + d]:
        row += 1
    return row


def two_buffer_row(prev, left, right, A, T):
    return {d: slide(A, T, d, max(1 + prev[d], prev[d - 1], 1 + prev[d + 1])) for d in range(left, right + 1)}


def in_place_row(ring, left, right, nl, A, T):
    """The kernel's schedule on the ring (a list of CAP + 2 ints, updated in place)."""
    w = right - left + 1
    C = -(-w // nl)
    p1 = [-2] * nl
    for k in range(C):
        carried = p1[nl - 1]                       # the run's last lane's p1 from chunk k - 1
        reads = {}
        for j in range(nl):                        # every lane reads ...
            d = left + j + nl * k
            if d > right:
                continue
            b = (d - 1) & MASK
            l0, l1, l2 = ring[b], ring[b + 1], ring[b + 2]
            x2 = carried if (j == 0 and k > 0) else l0
            p1[j] = l1
            reads[j] = (d, max(1 + l1, x2, 1 + l2))
        for j, (d, row) in reads.items():          # ... before any lane writes
            row = slide(A, T, d, row)
            pos = d & MASK
            ring[pos] = row
            if pos < 2:
                ring[pos + CAP] = row
    return {d: ring[d & MASK] for d in range(left, right + 1)}


def test_in_place_row_equals_two_buffer_row():
    rng = np.random.default_rng(4)
    for case in range(20000):
        e = int(rng.integers(1, 200))
        w = int(rng.integers(1, min(2 * e + 1, CAP - 3) + 1))
        left = int(rng.integers(-e, e - w + 2))
        right = left + w - 1
        base = int(rng.integers(0, 3000))
        prev = {d: base + int(rng.integers(-3, 40)) for d in range(left + 1, right)}
        for d in (left - 1, left, right, right + 1):
            prev[d] = -2                           # the widened band's ends and their neighbours
        L = base + 300
        A = rng.choice(list(b"acgt"), L)
        T = A.copy()
        T[rng.random(L) < float(rng.choice([0.02, 0.1, 0.3]))] = ord("a")
        T = np.concatenate([rng.choice(list(b"acgt"), max(0, right) + 5), T])
        A = A.tobytes()
        T = T.tobytes()
        ring = [int(rng.integers(-5, 5000))] * (CAP + 2)   # stale cells of older rows
        for d, v in prev.items():
            ring[d & MASK] = v
            if (d & MASK) < 2:
                ring[(d & MASK) + CAP] = v
        nl = int(rng.integers(1, w + 1))
        want = two_buffer_row(prev, left, right, A, T)
        got = in_place_row(ring, left, right, nl, A, T)
        assert got == want, (case, left, right, nl)


def test_dropping_the_carried_neighbour_is_caught():
    """Without the shuffle (a run's first lane reading its overwritten left neighbour from the ring)
    the rows differ as soon as a row needs two chunks."""
    rng = np.random.default_rng(5)
    A = rng.choice(list(b"acgt"), 400).tobytes()
    T = A
    prev = {d: 100 + (7 * d) % 13 for d in range(-9, 10)}
    prev[-11] = prev[-10] = prev[10] = prev[11] = -2
    ring = [0] * (CAP + 2)
    for d, v in prev.items():
        ring[d & MASK] = v
        if (d & MASK) < 2:
            ring[(d & MASK) + CAP] = v
    want = two_buffer_row(prev, -10, 10, A, T)
    assert in_place_row(list(ring), -10, 10, 4, A, T) == want
    # the same schedule reading the ring for the first lane too
    w, nl = 21, 4
    r2 = list(ring)
    got = {}
    for k in range(-(-w // nl)):
        reads = {}
        for j in range(nl):
            d = -10 + j + nl * k
            if d > 10:
                continue
            b = (d - 1) & MASK
            reads[j] = (d, max(1 + r2[b + 1], r2[b], 1 + r2[b + 2]))
        for j, (d, row) in reads.items():
            row = slide(A, T, d, row)
            r2[d & MASK] = row
            if (d & MASK) < 2:
                r2[(d & MASK) + CAP] = row
            got[d] = row
    assert got != want
