"""Pins the CPU oracle before anything is checked against it: golden vectors from the reference's
own compiled programs, brute force, closed forms and cross-algorithm identities."""
import math
import random

import pytest

import oracle
from golden_data import check_instances, fail_case, fail_field, fast_cases, load, residue_lines

P_CHECK = (998244353, 1004535809)  # the reference test program's primes


def test_prime_range_matches_reference_lines():
    seen = {}
    for rec in fast_cases() + [load("fast_16384.json")]:
        for i, p, _, _, _ in residue_lines(rec["stdout"]):
            seen[i] = p
    assert len(seen) >= 90
    first = oracle.prime_range(0, max(seen) + 1)
    for i, p in seen.items():
        assert first[i] == p
    # a range starting past 0 is the same list shifted
    assert oracle.prime_range(5, 3) == first[5:8]
    assert all(oracle.lib().or_is_prime_u64(p) for p in first)
    assert all((p - 1) % (1 << 20) == 0 and (1 << 49) < p < (1 << 50) for p in first)


def test_need_decides_printed_lines():
    for rec in fast_cases():
        qs = rec["queries"]
        maxneed = max(oracle.need(n, M, rec["d"]) for n, M in qs)
        want = []
        for i in range(rec["first"], min(rec["first"] + rec["count"], maxneed)):
            for n, M in qs:
                if i < oracle.need(n, M, rec["d"]):
                    want.append((i, n, M))
        got = [(i, n, M) for i, _, n, M, _ in residue_lines(rec["stdout"])]
        assert got == want, rec["queries"]


def test_need_known_values():
    assert oracle.need(200, 400, 3) == 97
    assert oracle.need(16384, 32768, 3) == 14184
    assert oracle.transform_length(16384) == 16384
    assert oracle.transform_length(16385) == 32768
    assert oracle.transform_length(5) == 16


def _golden_residues(max_n):
    for rec in fast_cases():
        for i, p, n, M, r in residue_lines(rec["stdout"]):
            if n <= max_n:
                yield rec["d"], p, n, M, r


def test_oracle_matches_reference_residues_small():
    checked = 0
    cache = {}
    for d, p, n, M, r in _golden_residues(129):
        key = (d, p, n, M)
        if key not in cache:
            cache[key] = oracle.nbad(d, p, [(n, M)])[0]
        assert cache[key] == r, key
        checked += 1
    assert checked > 200


@pytest.mark.slow
def test_oracle_matches_reference_residues_medium():
    groups = {}
    for d, p, n, M, r in _golden_residues(1024):
        if n > 129:
            groups.setdefault((d, p), []).append((n, M, r))
    assert groups
    for (d, p), items in list(groups.items())[:12]:
        got = oracle.nbad(d, p, [(n, M) for n, M, _ in items], algo="horner")
        assert got == [r for _, _, r in items], (d, p)


@pytest.mark.parametrize("n,M,d,rep,nbad,total", check_instances())
def test_oracle_vs_reference_brute_force(n, M, d, rep, nbad, total):
    assert oracle.brute_force(n, M, d, rep) == (nbad, total)
    for P in P_CHECK:
        assert oracle.nbad(d, P, [(n, M)], rep=rep)[0] == nbad % P
        assert oracle.nbad(d, P, [(n, M)], rep=rep, algo="cubic")[0] == nbad % P
        assert oracle.nbad(d, P, [(n, M)], rep=rep, algo="horner")[0] == nbad % P


def test_check_instance_count():
    inst = check_instances()
    assert sum(1 for x in inst if not x[3]) == 23 and sum(1 for x in inst if x[3]) == 14


@pytest.mark.parametrize("n", [3, 7, 12, 19])
def test_birthday_closed_form(n):
    M = 20
    exact = oracle.birthday_nbad(n, M)
    p = oracle.prime_range(0, 1)[0]
    for P in P_CHECK + (p,):
        assert oracle.nbad(1, P, [(n, M)])[0] == exact % P
        assert oracle.nbad(1, P, [(n, M)], rep=True)[0] == exact % P


@pytest.mark.parametrize("n,M,d,rep", [(30, 40, 2, False), (30, 45, 3, False), (24, 60, 4, False), (31, 32, 2, False),
                                       (30, 50, 2, True), (25, 40, 3, True), (60, 100, 3, False)])
def test_ntt_equals_cubic(n, M, d, rep):
    for P in P_CHECK:
        a = oracle.nbad(d, P, [(n, M)], rep=rep)
        b = oracle.nbad(d, P, [(n, M)], rep=rep, algo="cubic")
        c = oracle.nbad(d, P, [(n, M)], rep=rep, algo="horner")
        assert a == b == c


def test_table_reuse_across_queries():
    p = oracle.prime_range(3, 1)[0]
    qs = [(n, M) for n in (5, 12, 24) for M in (n + 1, n + 7, 2 * n + 3)]
    joint = oracle.nbad(3, p, qs)
    alone = [oracle.nbad(3, p, [q])[0] for q in qs]
    assert joint == alone


@pytest.mark.parametrize("n,M,d,args", [
    (200, 400, 3, ["single", "-n", "200", "-M", "400", "-d", "3", "--algo", "ntt", "--exact"]),
    (64, 65, 3, ["single", "-n", "64", "-M", "65", "-d", "3", "--exact"]),
    (12, 20, 1, ["single", "-n", "12", "-M", "20", "-d", "1", "--exact"]),
    (6, 7, 2, ["single", "-n", "6", "-M", "7", "-d", "2", "--exact"]),
])
def test_cross_precision_identity(n, M, d, args):
    """The 50-bit residues of the reference's fast engine reconstruct the exact N_bad printed by its
    32-bit CLI: the two reference paths pin each other, and the oracle is checked against both."""
    rec = next(r for r in fast_cases() if [n, M] in r["queries"] and r["d"] == d and r["first"] == 0
               and r["count"] >= oracle.need(n, M, d))
    x, m = oracle.crt([(p, r) for _, p, nn, MM, r in residue_lines(rec["stdout"]) if (nn, MM) == (n, M)])
    nb = int(fail_field(fail_case(args)["stdout"], "N_bad"))
    Qn = oracle.q_count(M, d) ** n
    assert m > Qn and x == nb and 0 <= nb <= Qn
    if d == 1:
        assert nb == oracle.birthday_nbad(n, M)
    if n == 6:
        assert nb == 23597091  # the reference test program's CRT check
    for p in oracle.prime_range(0, 4):
        assert oracle.nbad(d, p, [(n, M)])[0] == nb % p


def _fail_primes(stdout):
    for line in stdout.splitlines():
        if line.startswith("primes:"):
            k = int(line.split()[1])
            prod = float(line.split("2^")[1].rstrip(")").split(")")[0])
            return k, prod
    return None


@pytest.mark.parametrize("idx", range(16))
def test_oracle_reproduces_32bit_cli(idx):
    """Every exact run of the reference's 32-bit CLI: the prime count loop, and N_bad from the oracle's
    residues modulo those primes."""
    rec = load("fail_cases.json")[idx]
    a = rec["args"]
    if "--exact" not in a or rec["rc"] != 0:
        pytest.skip("not an exact run")
    n, M, d = int(a[a.index("-n") + 1]), int(a[a.index("-M") + 1]), int(a[a.index("-d") + 1])
    rep = "--replacement" in a
    nb = int(fail_field(rec["stdout"], "N_bad"))
    k, prod = _fail_primes(rec["stdout"])
    import math
    L = max(2, 1 << (n - 1).bit_length())
    min_pow2 = L.bit_length() - 1
    bits = n * (d * math.log2(M) if rep else math.log2(math.comb(M, d)))
    count = int(bits / 31) + 2
    while True:
        primes = oracle.ntt_primes32(count, min_pow2)
        have = sum(math.log2(p) for p in primes)
        if have > bits + 2:
            break
        count += int((bits + 2 - have) / 31) + 1
    assert len(primes) == k and abs(have - prod) < 0.06
    if n <= 128:
        res = [(p, oracle.nbad(d, p, [(n, M)], rep=rep)[0]) for p in primes]
        assert oracle.crt(res)[0] == nb


def test_conv_schoolbook_equals_ntt():
    rng = random.Random(5)
    for p, L in ((998244353, 64), (oracle.prime_range(0, 1)[0], 128), (oracle.prime_range(7, 1)[0], 16)):
        a = [rng.randrange(p) for _ in range(L)]
        k = [rng.randrange(p) for _ in range(L)]
        assert oracle.cyclic_conv(p, a, k) == oracle.cyclic_conv(p, a, k, "school")


def test_known_answer_readme():
    out = fail_case(["single", "-n", "20", "-M", "40", "-d", "3"])["stdout"]
    assert "primes: 10 (product 2^311.3)" in out
    assert "p_fail = 2.328850077437917e-07" in out


def test_uneven_batch_instances_pinned():
    """The committed uneven-batch instance list is what the seeded generator draws."""
    import workloads
    from golden_data import load

    g = load("uneven_batch.json")
    assert [tuple(x) for x in g["instances"]] == workloads.uneven_batch(g["seed"], 256)
    for x in g["golden"]:
        assert tuple(g["instances"][x["index"]]) == (x["n"], x["d"], x["M"])
        assert oracle.need(x["n"], x["M"], x["d"]) == x["need"]


def test_forward_oracle_against_reference_forward_runs():
    """The reference CLI's --algo forward --exact N_bad, reduced mod 50-bit primes, equals the
    oracle's O(n^4) recurrence; --mod runs give the residue itself."""
    primes = oracle.prime_range(0, 3)
    seen = 0
    for rec in load("fail_algo_cases.json"):
        a = rec["args"]
        if "forward" not in a:
            continue
        n, M, d = int(a[a.index("-n") + 1]), int(a[a.index("-M") + 1]), int(a[a.index("-d") + 1])
        rep = "--replacement" in a
        nbad = fail_field(rec["stdout"], "N_bad")
        if "--mod" in a:
            P = int(a[a.index("--mod") + 1])
            want = int(rec["stdout"].split(f"N_bad mod {P} = ")[1].split()[0])
            assert oracle.nbad(d, P, [(n, M)], rep=rep, algo="forward")[0] == want
            seen += 1
        elif nbad is not None:
            for p in primes:
                assert oracle.nbad(d, p, [(n, M)], rep=rep, algo="forward")[0] == int(nbad) % p, a
            seen += 1
    assert seen >= 8


def test_forward_oracle_against_brute_force_and_ntt():
    for n, M, d, rep, nbad, _ in check_instances():
        for p in P_CHECK:
            assert oracle.nbad(d, p, [(n, M)], rep=rep, algo="forward")[0] == nbad % p, (n, M, d, rep)
    rng = random.Random(4)
    for _ in range(25):
        n = rng.randint(1, 25)
        M = n + rng.randint(1, 20)
        d, rep = rng.randint(1, 5), bool(rng.randint(0, 1))
        p = rng.choice(P_CHECK)
        assert oracle.nbad(d, p, [(n, M)], rep=rep, algo="forward") == oracle.nbad(d, p, [(n, M)], rep=rep)


def _separate_brute():
    import re
    pat = re.compile(r"\[PASS\] separate\s+n=(\d+) m=\(([\d,]+)\) brute N_bad=(\d+) / (\d+)")
    return [(int(x.group(1)), [int(v) for v in x.group(2).split(",")], int(x.group(3)), int(x.group(4)))
            for x in pat.finditer(load("check_ref.txt"))]


def test_separate_oracle_against_brute_force():
    cases = _separate_brute()
    assert len(cases) >= 20
    for n, m, nbad, total in cases:
        q = 1
        for x in m:
            q *= x
        assert total == q ** n
        for p in P_CHECK + (oracle.prime_range(0, 1)[0],):
            assert oracle.separate_nbad(n, m, p) == nbad % p, (n, m)
            assert oracle.separate_nbad(n, m, p, "forward") == nbad % p, (n, m)


def test_separate_oracle_against_reference_runs():
    """The reference CLI's separate-mode N_bad (--exact) reduced mod 50-bit primes, and its --mod
    residues, equal both oracle algorithms."""
    primes = oracle.prime_range(0, 2)
    seen = 0
    for rec in load("fail_separate_cases.json"):
        a = rec["args"]
        if rec["rc"] != 0:
            continue
        n = int(a[a.index("-n") + 1])
        m = [int(x) for x in a[a.index("-m") + 1].split(",")]
        if "--mod" in a:
            P = int(a[a.index("--mod") + 1])
            want = int(rec["stdout"].split(f"N_bad mod {P} = ")[1].split()[0])
            assert oracle.separate_nbad(n, m, P) == want
            seen += 1
            continue
        nbad = fail_field(rec["stdout"], "N_bad")
        if nbad is None or n > 40:
            continue
        for p in primes:
            assert oracle.separate_nbad(n, m, p) == int(nbad) % p, a
            if n <= 15:
                assert oracle.separate_nbad(n, m, p, "forward") == int(nbad) % p, a
        seen += 1
    assert seen >= 8


def test_separate_with_one_table_is_the_birthday_count():
    # d = 1: an item fails only on a collision, N_bad = m^n - m!/(m-n)!
    p = P_CHECK[0]
    for n, m in [(3, 5), (7, 20), (12, 13)]:
        want = (m ** n - math.factorial(m) // math.factorial(m - n)) % p
        assert oracle.separate_nbad(n, [m], p) == want
        assert oracle.separate_nbad(n, [m], p, "forward") == want
