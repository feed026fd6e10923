"""oic_prefix_dp_batch_noalloc (no device allocation, a library-held scratch array) against
oic_prefix_dp_batch, through the C ABI: two calls in a row, the second with fewer tasks, must give
identical results and deltas; sizes beyond the array's bounds are refused."""
import ctypes

import numpy as np
import pytest

from canu_amd import oic, readsim

pytestmark = pytest.mark.gpu

MAX_TASKS = 16384          # tasks per call without allocation
MAX_LIMIT = 1021           # error limit per task without allocation


def _mutate(rng, s, rate):
    out = []
    for ch in s:
        u = rng.random()
        if u < rate / 3:
            out.append("ACGT"[(("ACGT".index(ch)) + int(rng.integers(1, 4))) % 4])
        elif u < 2 * rate / 3:
            continue
        elif u < rate:
            out.append("ACGT"[int(rng.integers(0, 4))])
            out.append(ch)
        else:
            out.append(ch)
    return "".join(out)


def _reads(seed, npairs, lmin, lmax):
    rng = np.random.default_rng(seed)
    seqs = []
    for _ in range(npairs):
        a = "".join(rng.choice(list("ACGT"), int(rng.integers(lmin, lmax))))
        b = _mutate(rng, a, float(rng.choice([0.0, 0.01, 0.04, 0.08, 0.2])))
        if rng.random() < 0.3:
            b += "".join(rng.choice(list("ACGT"), int(rng.integers(0, 200))))
        seqs += [a, b]
    return seqs


def _tasks(seqs, eb):
    t = []
    for i in range(0, len(seqs), 2):
        ia, ib = i + 1, i + 2
        if len(seqs[i]) > len(seqs[i + 1]):
            ia, ib = ib, ia
        m, n = len(seqs[ia - 1]), len(seqs[ib - 1])
        lim = min(int(eb[m]) + 3, MAX_LIMIT)
        t.append((ia, ib, 0, 0, 0, 0, m, n, lim, 0))
        t.append((ia, ib, 0, 0, m - 1, n - 1, m, n, lim, 1))
    return np.array(t, oic.DP_TASK)


def _noalloc(eng, tasks, stride):
    L = oic.lib()
    f = L.oic_prefix_dp_batch_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]
    res = np.zeros(len(tasks), oic.DP_RESULT)
    delta = np.full((len(tasks), stride), -7, np.int32)
    rc = f(eng._ctx, tasks.ctypes.data, len(tasks), res.ctypes.data, delta.ctypes.data, stride)
    return rc, res, delta


def _same(a_res, a_delta, b_res, b_delta):
    assert np.array_equal(a_res, b_res)
    for k in range(len(a_res)):
        n = int(a_res["delta_len"][k])
        assert np.array_equal(a_delta[k][:n], b_delta[k][:n]), k


def test_noalloc_matches_prefix_dp_batch_on_two_calls():
    seqs = _reads(71, 300, 20, 2500)
    R = readsim.from_strings(seqs)
    eng = oic.Engine(kmer=22, maxerate=0.045)
    eng.load_reads(R)
    eb = oic.dp_tables(0.045)[0]
    tasks = _tasks(seqs, eb)
    stride = 1100
    print(f"noalloc: first call, {len(tasks)} tasks", flush=True)
    ref_res, ref_delta = eng.prefix_dp(tasks, delta_stride=stride)
    rc, res, delta = _noalloc(eng, tasks, stride)
    assert rc == oic.OIC_OK, oic.lib().oic_last_error(eng._ctx)
    assert (ref_res["status"] == 0).all()
    _same(ref_res, ref_delta, res, delta)
    for k in range(len(res)):
        assert (delta[k][int(res["delta_len"][k]):] == 0).all(), k   # cleared, not left over

    # second call: fewer tasks, lower error limits; nothing of the first call may show through
    sub = tasks[::7].copy()
    sub["error_limit"] = np.maximum(sub["error_limit"] - 2, 0)
    print(f"noalloc: second call, {len(sub)} tasks", flush=True)
    ref_res2, ref_delta2 = eng.prefix_dp(sub, delta_stride=stride)
    rc, res2, delta2 = _noalloc(eng, sub, stride)
    assert rc == oic.OIC_OK, oic.lib().oic_last_error(eng._ctx)
    _same(ref_res2, ref_delta2, res2, delta2)
    assert int(res2["cells"].sum()) == int(ref_res2["cells"].sum())


def test_noalloc_refuses_sizes_beyond_its_bounds():
    seqs = _reads(72, 4, 50, 400)
    R = readsim.from_strings(seqs)
    eng = oic.Engine(kmer=22, maxerate=0.045)
    eng.load_reads(R)
    eb = oic.dp_tables(0.045)[0]
    tasks = _tasks(seqs, eb)
    many = np.resize(tasks, MAX_TASKS + 1)
    rc, _, _ = _noalloc(eng, many, 1)
    assert rc == oic.OIC_EUNSUPPORTED
    rc, _, _ = _noalloc(eng, tasks, (1 << 20) // len(tasks) + 1)
    assert rc == oic.OIC_EUNSUPPORTED
    big = tasks.copy()
    big["error_limit"][0] = MAX_LIMIT + 1
    rc, _, _ = _noalloc(eng, big, 16)
    assert rc == oic.OIC_EUNSUPPORTED
    # a refused call leaves the entry point usable
    rc, res, _ = _noalloc(eng, tasks, 16)
    assert rc == oic.OIC_OK and (res["status"] == 0).all()
