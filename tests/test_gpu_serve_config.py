"""The serving workload's own table set: d = 3, rows up to 2048, M up to 8192, primes 0..1535 of
the 50-bit list (about 26 GB resident), against the reference's residues.

tests/golden/serving_2048.json holds, for the first 40 requests of the seed-1 request stream and
for (2048, 8192), what the reference's host build printed over the whole prime set each needs:
the SHA-256 of the residue list, its first residues, and the CRT-combined exact N_bad's bit
length and SHA-256. Checked here: the batched query (cf_query) on all 41 at once and alone, and
the serving loop replaying them as a stream (residues and log2 p_fail)."""
import hashlib
import math

import pytest

import cuckoo_amd as cf
import oracle
import workloads
from golden_data import load

GOLD = load("serving_2048.json")


def _sha_residues(col):
    return hashlib.sha256("".join(f"{r}\n" for r in col).encode()).hexdigest()


@pytest.fixture(scope="module")
def served():
    eng = cf.Engine()
    tabs = cf.Tables(eng, GOLD["d"], GOLD["nmax"], GOLD["Mmax"], cf.prime_range(0, GOLD["primes"]))
    yield eng, tabs
    tabs.close()
    eng.close()


def test_fixture_is_the_stream_prefix():
    n, M, _ = workloads.serving_queries(GOLD["stream_seed"], GOLD["stream_count"], 32, GOLD["nmax"])
    qs = [(q["n"], q["M"]) for q in GOLD["queries"]]
    assert qs[:40] == [(int(n[i]), int(M[i])) for i in range(40)]
    assert qs[40] == (2048, 8192)


@pytest.mark.gpu
def test_batched_queries_equal_reference(served):
    _, tabs = served
    qs = [(q["n"], q["M"]) for q in GOLD["queries"]]
    need = [q["need"] for q in GOLD["queries"]]
    assert [tabs.need(n, M) for n, M in qs] == need
    res = tabs.query(qs, need=need)
    primes = tabs.primes
    for j, q in enumerate(GOLD["queries"]):
        col = [res[b][j] for b in range(q["need"])]
        assert col[:3] == q["head"], (q["n"], q["M"])
        assert _sha_residues(col) == q["residues_sha256"], (q["n"], q["M"])
        assert all(res[b][j] == 0 for b in range(q["need"], len(primes)))
    # the exact N_bad of the largest query, rebuilt from the device residues
    q = GOLD["queries"][-1]
    N, _ = oracle.crt([(primes[b], res[b][40]) for b in range(q["need"])])
    raw = N.to_bytes((N.bit_length() + 7) // 8, "little")
    assert N.bit_length() == q["nbad_bits"] and hashlib.sha256(raw).hexdigest() == q["nbad_sha256"]


@pytest.mark.gpu
def test_query_alone_equals_in_batch(served):
    _, tabs = served
    for q in (GOLD["queries"][0], GOLD["queries"][17], GOLD["queries"][-1]):
        res = tabs.query([(q["n"], q["M"])], need=[q["need"]])
        assert _sha_residues([res[b][0] for b in range(q["need"])]) == q["residues_sha256"]


@pytest.mark.gpu
def test_server_stream_equals_reference(served):
    _, tabs = served
    qs = GOLD["queries"]
    n = [q["n"] for q in qs]
    M = [q["M"] for q in qs]
    arrival = [0.0005 * i for i in range(len(qs))]  # 2000 requests/s: several per batch
    with cf.Server(tabs, slots=2, max_batch=32, crt_threads=4) as srv:
        lat, l2, res, st = srv.replay(n, M, arrival, residues=True)
    assert st.batches >= 2
    for i, q in enumerate(qs):
        col = res[i][:q["need"]]
        assert _sha_residues(col) == q["residues_sha256"], (q["n"], q["M"])
        assert all(r == 0 for r in res[i][q["need"]:])
        assert lat[i] > 0
    # log2 p_fail of the largest request from the exact integers
    q = qs[-1]
    N, _ = oracle.crt([(tabs.primes[b], res[-1][b]) for b in range(q["need"])])
    Qn_bits = q["n"] * math.log2(math.comb(q["M"], 3))
    want = (N.bit_length() - 1 + math.log2(N / (1 << (N.bit_length() - 1)))) - Qn_bits
    assert abs(l2[-1] - want) < 1e-6


@pytest.mark.gpu
def test_tables_beyond_the_query_lds_bound_are_refused():
    # 1/i for i <= max(nmax, Mmax) must fit the query kernel's LDS beside its 512-entry pad
    with cf.Engine() as eng:
        with pytest.raises(cf.CfError) as ex:
            cf.Tables(eng, 3, 100, 20000, cf.prime_range(0, 2))
        assert ex.value.code == cf.CF_E_UNSUPPORTED
        with cf.Tables(eng, 3, 100, 19967, cf.prime_range(0, 2)) as t:
            assert t.query([(100, 19967)])[0][0] == oracle.nbad(3, t.primes[0], [(100, 19967)])[0]
