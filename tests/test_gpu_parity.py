"""Parity of the HIP engine (through the C ABI) with the reference's residues and the oracle.

Bit-exact: every residue is an integer in [0, p) and must equal the reference's / oracle's."""
import random

import pytest

import cuckoo_amd as cf
import oracle
from golden_data import check_instances, fail_case, fail_field, fast_cases, load, residue_lines

pytestmark = pytest.mark.gpu

P_CHECK = (998244353, 1004535809)


@pytest.fixture(scope="module")
def eng():
    e = cf.Engine()
    yield e
    e.close()


def _run_golden(eng, rec):
    """All lines of one reference run, from one engine call with the reference's need mask."""
    lines = residue_lines(rec["stdout"])
    by_prime = {}
    for i, p, n, M, r in lines:
        by_prime.setdefault((i, p), {})[(n, M)] = r
    qs = [tuple(q) for q in rec["queries"]]
    primes = [p for _, p in sorted(by_prime)]
    mask = [[(n, M) in by_prime[k] for n, M in qs] for k in sorted(by_prime)]
    res = eng.run(qs, primes, rec["d"], mask=mask)
    for b, k in enumerate(sorted(by_prime)):
        for j, q in enumerate(qs):
            if mask[b][j]:
                assert res[b][j] == by_prime[k][q], (rec["d"], k, q)
            else:
                assert res[b][j] == 0


@pytest.mark.parametrize("idx", range(len(fast_cases())))
def test_reference_residues(eng, idx):
    _run_golden(eng, fast_cases()[idx])


def test_reference_residues_n16384(eng):
    rec = load("fast_16384.json")
    _run_golden(eng, rec)
    st = eng.stats()
    assert st.log2_length == 14 and st.tasks == rec["count"]


@pytest.mark.parametrize("n,M,d,rep,nbad,total", check_instances())
def test_reference_brute_force_instances(eng, n, M, d, rep, nbad, total):
    res = eng.run([(n, M)], list(P_CHECK), d, model=int(rep))
    assert [r[0] for r in res] == [nbad % P for P in P_CHECK]


@pytest.mark.parametrize("n", [3, 7, 12, 19])
def test_birthday_closed_form(eng, n):
    M = 20
    exact = oracle.birthday_nbad(n, M)
    ps = list(P_CHECK) + cf.prime_range(0, 2)
    for model in (0, 1):
        assert [r[0] for r in eng.run([(n, M)], ps, 1, model=model)] == [exact % p for p in ps]


def test_cross_precision_exact_nbad(eng):
    """97 GPU residues modulo 50-bit primes reconstruct the exact N_bad(200, 400, d=3) printed by the
    reference's 32-bit CLI."""
    ps = cf.prime_range(0, cf.need(200, 400, 3))
    res = eng.run([(200, 400)], ps, 3)
    x, m = oracle.crt([(p, r[0]) for p, r in zip(ps, res)])
    nb = int(fail_field(fail_case(["single", "-n", "200", "-M", "400", "-d", "3", "--algo", "ntt", "--exact"])["stdout"], "N_bad"))
    assert x == nb and m > oracle.q_count(400, 3) ** 200


@pytest.mark.parametrize("seed", range(6))
def test_random_against_oracle(eng, seed):
    rng = random.Random(seed)
    d = rng.randint(1, 5)
    rep = rng.random() < 0.3
    qs = []
    for _ in range(rng.randint(1, 4)):
        n = rng.randint(1, 300)
        qs.append((n, rng.randint(n + 1, 3 * n + 5)))
    ps = cf.prime_range(rng.randint(0, 2000), 2) + [998244353]
    res = eng.run(qs, ps, d, model=int(rep))
    for b, p in enumerate(ps):
        assert res[b] == oracle.nbad(d, p, qs, rep=rep), (seed, p, qs, d, rep)


def test_query_alone_equals_query_in_a_mix(eng):
    ps = cf.prime_range(40, 3)
    mix = [(600, 900), (37, 38), (600, 601), (1000, 1500), (8, 9)]
    joint = eng.run(mix, ps, 3)
    for j, q in enumerate(mix):
        alone = eng.run([q], ps, 3)
        assert [r[0] for r in alone] == [r[j] for r in joint]


def test_mask_and_empty(eng):
    ps = cf.prime_range(0, 3)
    res = eng.run([(50, 60), (20, 25)], ps, 2, mask=[[1, 0], [0, 0], [0, 1]])
    full = eng.run([(50, 60), (20, 25)], ps, 2)
    assert res[0] == [full[0][0], 0] and res[1] == [0, 0] and res[2] == [0, full[2][1]]
    assert eng.run([(5, 9)], [], 2) == []


def test_d_larger_than_n_gives_zero(eng):
    # no item set can fail when every item has d >= n candidates
    assert eng.run([(5, 40)], cf.prime_range(0, 2), 7) == [[0], [0]]


def test_invalid_arguments_raise(eng):
    with pytest.raises(cf.CfError) as ei:
        eng.run([(10, 10)], cf.prime_range(0, 1), 3)
    assert ei.value.code == cf.CF_E_ARGS
    with pytest.raises(cf.CfError):
        eng.run([(10, 20)], [1000003], 0)


def test_primes_below_the_engine_transform(eng):
    # 7681 = 15 * 2^9 + 1 and 40961 = 5 * 2^13 + 1: next_pow2(n) divides p - 1 for n <= 512 but the
    # engine's shortest transform 2^10 does not for 7681, so that prime's tables take the O(n^3)
    # device algorithm; every residue must still be the oracle's
    qs = [(200, 400), (64, 65), (37, 80), (512, 700)]
    primes = [7681, 40961] + cf.prime_range(0, 1)
    got = eng.run(qs, primes, 3, mask=[[1, 1, 1, 0], [1, 1, 1, 1], [1, 1, 1, 1]])
    for b, p in enumerate(primes):
        assert [got[b][j] for j in range(3)] == oracle.nbad(3, p, qs[:3]), p
    assert got[0][3] == 0 and got[1][3] == oracle.nbad(3, 40961, [(512, 700)])[0]
    assert got[2][:3] == [918297158731096, 201776960213191, 67342126954205]
