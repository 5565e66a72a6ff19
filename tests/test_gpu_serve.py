"""Resident table sets, batched queries and the serving loop, through the C ABI.

Pinned against the reference's own residues ((200, 400) and (64, 65), d = 3, over their whole
prime sets), the oracle (other d, the replacement model, 32-bit primes), and the fused
build-and-query call; a query's residues must be the same alone, in a batch, and inside a
served stream of other requests."""
import math
import random

import pytest

import cuckoo_amd as cf
import oracle
from golden_data import fast_cases, residue_lines

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = cf.Engine()
    yield e
    e.close()


def _golden_full(n, M):
    for rec in fast_cases():
        if rec["d"] == 3 and [list(q) for q in rec["queries"]] == [[n, M]] and rec["first"] == 0:
            lines = residue_lines(rec["stdout"])
            if len(lines) == min(rec["count"], oracle.need(n, M, 3)):  # the whole need() set
                return lines
    raise LookupError((n, M))


@pytest.fixture(scope="module")
def tabs(eng):
    # d = 3, rows up to 256, M up to 600, primes 0..139 (need(200, 400) = 97, need(256, 600) = 133)
    t = cf.Tables(eng, 3, 256, 600, cf.prime_range(0, 140))
    yield t
    t.close()


def test_tables_vs_reference_residues(tabs):
    for n, M in ((200, 400), (64, 65)):
        lines = _golden_full(n, M)
        need = tabs.need(n, M)
        assert need == len(lines)
        res = tabs.query([(n, M)], need=[need])
        for (i, p, _, _, r) in lines:
            assert tabs.primes[i] == p and res[i][0] == r, (n, M, i)
        assert all(res[b][0] == 0 for b in range(need, len(tabs.primes)))


def test_tables_equal_fused_path(eng, tabs):
    qs = [(256, 600), (255, 256), (100, 250), (17, 18), (3, 10), (2, 5), (129, 300)]
    primes = tabs.primes[:5]
    want = eng.run(qs, primes, 3)
    got = tabs.query(qs)
    for b in range(5):
        assert [got[b][j] for j in range(len(qs))] == want[b], b


def test_batch_alone_equals_in_mix(tabs):
    rng = random.Random(5)
    qs = [(rng.randint(1, 256), 0) for _ in range(40)]  # 40 > one pass of 16
    qs = [(n, rng.randint(n + 1, 600)) for n, _ in qs]
    np_ = len(tabs.primes)
    need = [rng.randint(0, np_) for _ in qs]
    mix = tabs.query(qs, need=need)
    for j in (0, 7, 16, 31, 39):
        alone = tabs.query([qs[j]], need=[need[j]])
        assert [mix[b][j] for b in range(np_)] == [alone[b][0] for b in range(np_)], j


@pytest.mark.parametrize("d,model", [(1, 0), (2, 0), (4, 0), (5, 0), (2, 1), (3, 1)])
def test_tables_vs_oracle(eng, d, model):
    primes = cf.prime_range(3, 2) + [998244353]
    with cf.Tables(eng, d, 70, 160, primes, model=model) as t:
        qs = [(70, 160), (70, 71), (33, 40), (9, 150), (d, d + 1), (1, 2)]
        got = t.query(qs)
        for b, p in enumerate(primes):
            want = oracle.nbad(d, p, qs, rep=bool(model))
            assert [got[b][j] for j in range(len(qs))] == want, (d, model, p)


def test_server_replay_matches_queries(tabs):
    rng = random.Random(9)
    k = 300
    n = [rng.randint(1, 256) for _ in range(k)]
    M = [rng.randint(x + 1, 600) for x in n]
    arrival = [i * 0.0002 for i in range(k)]  # 5000 requests/s: batches form
    with cf.Server(tabs, slots=2, max_batch=32, crt_threads=4) as s:
        lat, l2, res, st = s.replay(n, M, arrival, residues=True)
        assert st.batches >= 1 and st.mean_batch >= 1 and st.p99_ms >= st.p50_ms > 0
        assert all(x > 0 for x in lat)
        # the same request alone
        for i in (0, 1, 150, 299):
            lat1, l21, res1, st1 = s.replay([n[i]], [M[i]], [0.0], residues=True)
            assert res1[0] == res[i] and l21[0] == l2[i] and st1.batches == 1
    for i in range(0, k, 37):
        need = tabs.need(n[i], M[i])
        alone = tabs.query([(n[i], M[i])], need=[need])
        assert res[i] == [alone[b][0] for b in range(len(tabs.primes))]
        # log2 p_fail from the exact CRT value
        N, _ = oracle.crt(zip(tabs.primes[:need], res[i][:need]))
        if N == 0:
            assert l2[i] == -math.inf
        else:
            shift = max(N.bit_length() - 60, 0)
            want = math.log2(N >> shift) + shift - n[i] * oracle.log2_q(M[i], 3)
            assert abs(l2[i] - want) < 1e-9 * max(1.0, abs(want)), (i, l2[i], want)


def test_server_rejects(tabs):
    with cf.Server(tabs, slots=1, max_batch=4, crt_threads=1) as s:
        with pytest.raises(cf.CfError) as ei:
            s.replay([300], [600], [0.0])  # n beyond the set
        assert ei.value.code == cf.CF_E_ARGS
    with cf.Tables(tabs.engine, 3, 256, 600, tabs.primes[:100]) as small, cf.Server(small, 1, 4, 1) as s:
        with pytest.raises(cf.CfError) as ei:
            s.replay([256], [600], [0.0])  # needs 133 primes, the set holds 100
        assert ei.value.code == cf.CF_E_PRIMES


def test_tables_admission_and_args(eng):
    with pytest.raises(cf.CfError) as ei:
        cf.Tables(eng, 3, 16384, 19000, cf.prime_range(0, 4000))  # about 4.3 TB: not admitted
    assert ei.value.code == cf.CF_E_NOMEM
    with pytest.raises(cf.CfError) as ei:
        cf.Tables(eng, 3, 100, 100, cf.prime_range(0, 2))
    assert ei.value.code == cf.CF_E_ARGS
    with cf.Tables(eng, 3, 40, 80, cf.prime_range(0, 2)) as t:
        with pytest.raises(cf.CfError) as ei:
            t.query([(41, 80)])
        assert ei.value.code == cf.CF_E_ARGS
        assert t.bytes() > 0
