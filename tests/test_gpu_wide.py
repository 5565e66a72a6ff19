"""Tables of more than 16384 rows: transforms of 2^16 spread over 16 workgroups per prime.

Pinned against the reference's own residues (its host build, one prime per run) for n from 16385
to 65536, several queries per table, d = 2, 3, 4, and a batch of tables in one launch."""
import pytest

import cuckoo_amd as cf
from golden_data import load, residue_lines

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = cf.Engine()
    yield e
    e.close()


def test_wide_tables_vs_reference(eng):
    recs = load("fast_wide.json")
    # group the reference's single-prime runs by (d, queries) into one engine call each
    groups = {}
    for r in recs:
        key = (r["d"], tuple(tuple(q) for q in r["queries"]))
        groups.setdefault(key, []).extend(residue_lines(r["stdout"]))
    for (d, qs), lines in groups.items():
        primes = sorted({(i, p) for i, p, _, _, _ in lines})
        print(f"d={d} queries={qs} primes={[i for i, _ in primes]}", flush=True)
        got = eng.run(list(qs), [p for _, p in primes], d)
        st = eng.stats()
        assert st.log2_length == 16
        for i, p, n, M, r in lines:
            b = [x for x, _ in primes].index(i)
            assert got[b][list(qs).index((n, M))] == r, (d, i, n, M, "tasks split over XCDs", st.wide_split)


def test_wide_n65536_reference_prime0(eng):
    rec = load("fast_65536.json")
    (i, p, n, M, r), = residue_lines(rec["stdout"])
    got = eng.run([(n, M)], [p], 3)
    assert got[0][0] == r


def test_wide_mixed_with_short_tables(eng):
    # one call with tables of both kinds: each prime's table is as long as its longest query
    recs = load("fast_wide.json")
    rec = next(r for r in recs if r["d"] == 3 and [list(q) for q in r["queries"]] == [[16385, 16386]])
    (i, p, n, M, r), = residue_lines(rec["stdout"])
    got = eng.run([(n, M), (200, 400)], [p, 1125899865948161], 3, mask=[[1, 0], [0, 1]])
    assert got[0][0] == r and got[1][1] == 918297158731096


def test_wide_device_scope_barriers(eng, monkeypatch):
    # the device-scope group barrier, the form a task takes when its workgroups span XCDs, forced
    monkeypatch.setenv("CF_WIDE_DEVICE_BARRIERS", "1")
    lines = [x for r in load("fast_wide.json") if r["d"] == 3 and [list(q) for q in r["queries"]] == [[16385, 16386]]
             for x in residue_lines(r["stdout"])]
    primes = sorted({(i, p) for i, p, _, _, _ in lines})
    got = eng.run([(16385, 16386)], [p for _, p in primes], 3)
    want = {i: r for i, _, _, _, r in lines}
    assert [g[0] for g in got] == [want[i] for i, _ in primes]


def test_wide_table_with_more_than_eight_queries(eng):
    # 13 queries on one table of 20000 rows (primes 0..2): the reference's residues for its three
    # long queries, and the short ones equal to the one-workgroup kernel's own tables of their size
    recs = load("fast_wide.json")
    pinned = {}
    for r in recs:
        if r["d"] == 3:
            for i, p, n, M, res in residue_lines(r["stdout"]):
                pinned[(i, n, M)] = res
    long_q = [(20000, 40000), (17000, 20000), (16385, 16386)]
    short_q = [(16000, 16500), (12000, 30000), (8193, 9000), (4000, 4001), (2049, 5000), (1000, 1500), (777, 2000),
               (200, 400), (64, 65), (37, 80)]
    primes = cf.prime_range(0, 3)
    got = eng.run(long_q + short_q, primes, 3)
    assert eng.stats().log2_length == 16
    for b in range(3):
        for j, (n, M) in enumerate(long_q):
            if (b, n, M) in pinned:
                assert got[b][j] == pinned[(b, n, M)], (b, n, M)
    alone = eng.run(short_q, primes, 3)
    assert eng.stats().log2_length == 14
    for b in range(3):
        assert got[b][len(long_q):] == alone[b], b
    assert got[0][-3:] == [918297158731096, 201776960213191, 67342126954205]  # the reference's, prime 0


@pytest.mark.parametrize("device_scope", [False, True])
def test_wide_full_launch_of_16_tasks(eng, monkeypatch, device_scope):
    # 16 primes of (32768, 65536, d = 3): one launch of 16 tasks x 16 workgroups (the whole device),
    # every residue against the reference's; both group-barrier forms
    if device_scope:
        monkeypatch.setenv("CF_WIDE_DEVICE_BARRIERS", "1")
    rec = load("fast_32768_launch.json")
    lines = residue_lines(rec["stdout"])
    assert len(lines) == 16
    print(f"16-task launch, device-scope barriers: {device_scope}", flush=True)
    got = eng.run([(32768, 65536)], [p for _, p, _, _, _ in lines], 3)
    st = eng.stats()
    assert st.log2_length == 16 and st.tasks == 16 and st.launches == 1
    assert [g[0] for g in got] == [r for _, _, _, _, r in lines], ("tasks split over XCDs", st.wide_split)


def test_wide_2_17_against_2_16_and_reference(eng):
    # tables of more than 65536 rows take transforms of 2^17 (16 workgroups of 512 threads); their
    # shorter queries must equal the same queries answered from a 2^16 table, and (32768, 65536)
    # the reference's residues
    P = cf.prime_range(0, 2)
    qs17 = [(65537, 70000), (60000, 100000), (32768, 65536)]
    got = eng.run(qs17, P, 3)
    assert eng.stats().log2_length == 17
    short = eng.run(qs17[1:], P, 3)
    assert eng.stats().log2_length == 16
    for b in range(2):
        assert got[b][1:] == short[b], b
    assert [got[0][2], got[1][2]] == [1051397482096400, 367355341181042]


def test_wide_2_17_reference_prime0(eng):
    try:
        rec = load("fast_70000.json")
    except OSError:
        pytest.skip("fixture not generated")
    lines = residue_lines(rec["stdout"])
    qs = [tuple(q) for q in rec["queries"]]
    got = eng.run(qs, [lines[0][1]], 3)
    assert eng.stats().log2_length == 17
    assert got[0] == [r for _, _, _, _, r in lines]
