"""DP tables (prefixEditDistance.C:23-88, Binomial_Bound.C:35-195), oracle and engine.

Pinned by: the strtof rounding of --maxerate (overlapInCore.C:380) seen through exact
multiples; scipy's binomial and Poisson tails for the definition of Edit_Match_Limit
(smallest n with P[>= e-1 errors in n trials] > 1e-4, the normal shortcut z <= 3.62 above
35 trials); the fitted line beyond e = 2000; and engine == oracle bit for bit.
"""
import math

import numpy as np
import pytest
from scipy import stats

from canu_amd import oic
from oracle import oracle


@pytest.fixture(scope="module", params=[0.045, 0.01, 0.06, 0.12])
def tables(request):
    return request.param, oracle.OracleTables(request.param)


def test_cli_rate_is_float(tables):
    e, T = tables
    assert T.maxErate == float(np.float32(e))
    assert T.MAX_ERRORS == 1 + math.ceil(T.maxErate * oracle.MAX_READLEN)


def test_error_bound_float_quirk():
    T = oracle.OracleTables(0.045)
    # 200 * 0.045 = 9 exactly in decimal, but the float 0.045 is slightly larger: ceil gives 10
    assert T.error_bound[200] == 10
    assert T.error_bound[1000] == 46
    i = np.arange(0, 1 << 21, 997, dtype=np.int64)
    exp = np.ceil(i * np.float64(np.float32(0.045))).astype(np.int32)
    assert np.array_equal(T.error_bound[i], exp)


def _criterion(n, e_arg, p):
    """True iff Binomial_Bound would return n: P[>= e_arg errors in n trials] > 1e-4.

    Above 35 trials the reference sums Poisson terms with (n p)^k accumulated in a double:
    once (n p)^(e_arg - 1) exceeds DBL_MAX the sum is inf or NaN and the test fails, so only
    the normal shortcut can accept n.  Returns None on the overflow border (not decidable here).
    """
    if n <= 35:
        return stats.binom.sf(e_arg - 1, n, p) > 1e-4
    z = (e_arg - 0.5 - n * p) / math.sqrt(n * p * (1 - p))
    if z <= 3.62:
        return True
    mag = (e_arg - 1) * math.log10(n * p) if e_arg > 1 else 0.0
    if mag > 308.6:
        return False
    if mag > 307.5:
        return None
    return stats.poisson.sf(e_arg - 1, n * p) > 1e-4


def test_match_limit_definition(tables):
    e, T = tables
    p = T.maxErate
    ml = T.match_limit
    assert ml[0] == 0 and ml[1] == 0
    prev_n = 1
    checked = 0
    for ee in range(2, min(2000, T.MAX_ERRORS)):
        n = int(ml[ee]) + 1
        e_arg = ee - 1
        assert n >= max(prev_n, e_arg)
        assert _criterion(n, e_arg, p) is not False, (ee, n)
        lo = max(prev_n, e_arg)
        if n - 1 >= lo:
            assert _criterion(n - 1, e_arg, p) is not True, (ee, n)
            checked += 1
        prev_n = n
    assert checked > 100
    assert np.all(np.diff(ml[:2000]) >= 0)


def test_match_limit_linear_tail():
    T = oracle.OracleTables(0.045)
    p = T.maxErate
    ml = T.match_limit.astype(np.float64)
    sl = 0.982064188397525 / p + 0.067835741959926
    e = np.arange(2000, T.MAX_ERRORS)
    closed = np.ceil(ml[1999] + (e - 1999) * sl)
    # the reference accumulates the slope step by step; rounding may move a ceil by one
    assert np.max(np.abs(ml[2000:] - closed)) <= 1
    assert np.all(np.diff(ml) >= 0)


def test_small_limits_by_hand():
    # e=3 -> Binomial_Bound(2, p, 1): n=2 gives P[2 errors] = p^2 > 1e-4 -> ml = 1
    # e=4 -> Binomial_Bound(3, p, 2): n=3: p^3 = 9.1e-5 < 1e-4, n=4: 4p^3q + p^4 > 1e-4 -> ml = 3
    T = oracle.OracleTables(0.045)
    assert list(T.match_limit[:5]) == [0, 0, 0, 1, 3]


@pytest.mark.parametrize("rate,noise", [(0.045, 1.0), (0.01, 1.0), (0.06, 1.0), (0.12, 1.0), (0.045, 0.8)])
def test_engine_tables_equal_oracle(rate, noise):
    eb, ml = oic.dp_tables(rate, noise)
    T = oracle.OracleTables(rate, noise)
    assert np.array_equal(eb, T.error_bound)
    assert np.array_equal(ml, T.match_limit[:len(ml)])
    assert len(ml) == T.MAX_ERRORS
