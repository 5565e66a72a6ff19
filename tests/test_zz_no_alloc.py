"""cf_query_noalloc (no device allocation, fixed per-device scratch) against cf_query through the C
ABI: identical residues on two calls in a row, the second with fewer queries and fewer primes
needed, then a call too large for the fixed scratch is refused before any device work."""
import random

import pytest

import cuckoo_amd as cf
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tabs():
    eng = cf.Engine()
    t = cf.Tables(eng, 3, 300, 900, cf.prime_range(0, 200))
    yield t
    t.close()
    eng.close()


def test_noalloc_equals_query_on_two_calls(tabs):
    rng = random.Random(11)
    np_ = len(tabs.primes)
    big = [(rng.randint(1, 300), 0) for _ in range(40)]
    big = [(n, rng.randint(n + 1, 900)) for n, _ in big] + [(300, 900)]
    need_big = [rng.randint(1, np_) for _ in big[:-1]] + [np_]
    small = [(120, 121), (7, 800), (300, 301), (64, 65), (1, 2)]
    need_small = [3, np_, 17, 1, 0]
    for qs, need in ((big, need_big), (small, need_small)):
        want = tabs.query(qs, need=need)
        got = tabs.query(qs, need=need, noalloc=True)
        assert got == want
    # one column against the oracle as well (the two entry points share their kernels)
    assert [want[b][2] for b in range(2)] == [oracle.nbad(3, p, [small[2]])[0] for p in tabs.primes[:2]]


def test_noalloc_refuses_calls_beyond_its_scratch(tabs):
    qs = [(300, 900)] * 3000  # ~14 KB of partial sums and results per query at 200 primes
    with pytest.raises(cf.CfError) as ex:
        tabs.query(qs, noalloc=True)
    assert ex.value.code == cf.CF_E_ARGS
    # and the next call still works
    assert tabs.query([(300, 900)], noalloc=True) == tabs.query([(300, 900)])
