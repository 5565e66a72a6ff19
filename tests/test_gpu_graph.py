"""Graph mode (cf_set_graphs): a call's upload, launches and download captured as one hipGraph and
replayed for later calls with the same launch plan. Every replay is checked against an engine
that launches one kernel at a time, with that engine's eager work and synchronisations between
the replays."""
import pytest

import cuckoo_amd as cf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engines():
    g, e = cf.Engine(), cf.Engine()
    g.set_graphs(True)
    yield g, e
    g.close()
    e.close()


def test_fused_replays_match_eager(engines):
    g, e = engines
    queries = [(200, 400), (64, 65), (1000, 3000)]
    for rep, first in enumerate([0, 8, 16, 0, 24]):
        primes = cf.prime_range(first, 8)
        print(f"fused replay {rep}: primes {first}..{first + 7}", flush=True)
        got = g.run(queries, primes, 3)
        assert g.stats().graph == (1 if rep == 0 else 2)
        assert got == e.run(queries, primes, 3)
        assert e.stats().graph == 0


def test_batch_replays_match_eager(engines):
    g, e = engines
    P = cf.prime_range(0, 40)

    def spec(k):
        # a queue launch (2^10 .. 2^12) and a 2^13 launch, on shifted prime sets
        return [(3, 0, 4000, 9000, P[k:k + 3]), (2, 1, 700, 900, P[k + 3:k + 6]), (4, 0, 90, 91, P[k + 6:k + 8]),
                (3, 0, 5000, 12000, P[k + 8:k + 10])]

    for rep, k in enumerate([0, 10, 20, 0]):
        print(f"batch replay {rep}: offset {k}", flush=True)
        got = g.run_batch(spec(k))
        st = g.stats()
        assert st.launches == 2 and st.graph == (1 if rep == 0 else 2)
        assert got == e.run_batch(spec(k))


def test_shape_change_recaptures(engines):
    g, e = engines
    a, b = [(300, 700)], [(300, 700), (31, 32)]
    P = cf.prime_range(0, 6)
    assert g.run(a, P, 2) == e.run(a, P, 2) and g.stats().graph == 1
    assert g.run(b, P, 2) == e.run(b, P, 2) and g.stats().graph == 1
    assert g.run(a, P[::-1], 2) == e.run(a, P[::-1], 2) and g.stats().graph == 1
    assert g.run(a, P, 2) == e.run(a, P, 2) and g.stats().graph == 2
    mask = [[1, 0], [0, 1], [1, 1], [0, 0], [1, 0], [0, 1]]
    assert g.run(b, P, 2, mask=mask) == e.run(b, P, 2, mask=mask)


def test_cooperative_launches_stay_eager(engines):
    g, e = engines
    P = cf.prime_range(3, 1)
    got = g.run([(16385, 32770)], P, 3)
    assert g.stats().graph == 0 and g.stats().log2_length == 16
    assert got == e.run([(16385, 32770)], P, 3)
