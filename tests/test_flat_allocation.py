"""The flat-row kernel's lane allocation (flat_row in canu_amd/csrc/oic_flat.h), checked on the CPU.

Every row, the wavefront's active DPs (A of them, band widths w, total W) share 64 lanes over C
chunk iterations, ceil(w / C) lanes each.  The kernel:
* computes ceil(w / C) from a float reciprocal (v_rcp_f32, within 1 ulp), truncation and one
  correction step; this is exact for w, C up to 2^12 because a quotient that is not an integer is
  at least 1/w away from one, far more than the float error, and an integer quotient that comes
  out one low is corrected;
* tries C = ceil(W / 64), then C + 1, then ceil(W / (64 - A)), which always fits: the lane sum
  is below W / C + A.
Both claims are checked here, with the reciprocal rounded to nearest and one ulp either way.
"""
import numpy as np

WAVE = 64


def ceil_by_rcp(w, C, ulp):
    """The kernel's ceil(w / C): w, C int arrays; ulp -1, 0, 1 shifts the float reciprocal."""
    rc = (np.float32(1.0) / C.astype(np.float32)).astype(np.float32)
    if ulp:
        rc = np.nextafter(rc, np.float32(np.inf) if ulp > 0 else np.float32(0)).astype(np.float32)
    q = (w.astype(np.float32) * rc).astype(np.float32).astype(np.int64)   # float32 product, truncated
    return q + (q * C < w)


def test_rcp_ceiling_is_exact():
    w = np.arange(0, 4097, dtype=np.int64)
    for C in range(1, 4097):
        Cs = np.full_like(w, C)
        want = -(-w // C)
        for ulp in (-1, 0, 1):
            assert np.array_equal(ceil_by_rcp(w, Cs, ulp), want), (C, ulp)


def allocate(widths):
    """The chunk count the kernel settles on and the lanes it hands out."""
    A, W = len(widths), sum(widths)
    C = -(-W // WAVE)
    csafe = -(-W // (WAVE - A))
    for it in range(3):
        nv = [-(-w // C) for w in widths]
        if sum(nv) <= WAVE:
            return C, nv
        C = C + 1 if it == 0 else csafe
    raise AssertionError("no fit")


def test_allocation_always_fits():
    rng = np.random.default_rng(3)
    minimal = total = 0
    for _ in range(200000):
        A = int(rng.integers(1, 17))
        cap = int(rng.choice([32, 64, 128]))
        widths = [int(x) for x in rng.integers(1, cap - 2, A)]
        C, nv = allocate(widths)
        assert sum(nv) <= WAVE
        assert all(n * C >= w for n, w in zip(nv, widths))          # every cell gets a lane slot
        best = next(c for c in range(1, C + 1) if sum(-(-w // c) for w in widths) <= WAVE)
        minimal += best == C
        total += 1
    # the fewest chunk iterations in most rows (about 95 % of these random ones, whose bands are far
    # wider than a typical row's); the rest take the always-fitting count, a few iterations more
    assert minimal > 0.9 * total
