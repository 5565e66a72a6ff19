"""The exact combiner (bin/cf_post, cf_exact_combine) in place of fast/post_gmp.py.

The oracle (oracle/post.py) is pinned first: by the reference's own exact N_bad, p_fail and log2
values (cuckoo_fail --exact and check's brute-force counts), by the residue files the reference's
cuckoo_fast wrote, and by what the definitions fix (x = r_i mod p_i, truncation of N/D to 40
digits, the 2^-lambda tests). The product is then compared with the oracle's output text."""
import math
import os
import random
import subprocess
from decimal import Decimal
from fractions import Fraction

import pytest

import cuckoo_amd as cf
import oracle
import post
from golden_data import check_instances, fail_case, fail_field, fast_cases

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CF_POST = os.path.join(REPO, "cuckoo-failure_amd", "bin", "cf_post")

# cuckoo_fail single --exact runs of the reference (uniform d-subsets, the model post_gmp.py
# assumes with Q = C(M, d))
EXACT = [(200, 400, 3), (20, 40, 3), (6, 7, 2), (1, 2, 1), (5, 6, 1), (12, 20, 1), (31, 32, 2), (60, 100, 3),
         (64, 65, 3), (100, 150, 4), (128, 1000, 2), (512, 634, 3)]


def reference_exact(n, M, d):
    """(N_bad, p_fail text, log2 text) printed by the reference's cuckoo_fail --exact."""
    args = ["single", "-n", str(n), "-M", str(M), "-d", str(d)]
    args += ["--algo", "ntt", "--exact"] if (n, M, d) == (200, 400, 3) else ["--exact"]
    rec = fail_case(args)
    out = rec["stdout"]
    nbad = int(fail_field(out, "N_bad"))
    return nbad, fail_field(out, "p_fail"), fail_field(out, "log2(p_fail)")


def residue_text(n, M, N, primes, first=0):
    return "".join(f"{first + i} {p} {n} {M} {N % p}\n" for i, p in enumerate(primes))


def run_post(args, files, tmp_path):
    paths = []
    for k, text in enumerate(files):
        p = tmp_path / f"res{k}.txt"
        p.write_text(text)
        paths.append(str(p))
    return subprocess.run([CF_POST, *args[:1], *paths, *args[1:]], capture_output=True, text=True)


# ---------------------------------------------------------------- the oracle, pinned

@pytest.mark.parametrize("n,M,d", EXACT)
def test_oracle_against_reference_exact_counts(n, M, d):
    nbad, pf, lg = reference_exact(n, M, d)
    k = post.need(n, M, d, 49)
    primes = oracle.prime_range(0, k)
    line, x = post.row(n, M, d, {i: (p, nbad % p) for i, p in enumerate(primes)}, pbits=49)
    assert x == nbad
    cols = line.split("\t")
    Qn = math.comb(M, d) ** n
    if nbad == 0:
        assert cols[3] == "-inf" and cols[7] == "0" and pf is None
        return
    # the reference prints p_fail and log2 from long doubles (16 and 6+ digits): the exact
    # values agree to within their last printed place
    assert abs(Decimal(cols[7]) - Decimal(pf)) <= Decimal(pf) * Decimal("2e-15")
    assert abs(Decimal(cols[3]) - Decimal(lg)) <= Decimal("1.5e-6")
    # truncation: the 40 digits are floor(N / D) at that exponent
    lo = Decimal(cols[7])
    ulp = Decimal(10) ** (lo.adjusted() - 39)
    assert Fraction(lo) <= Fraction(nbad, Qn) < Fraction(lo) + Fraction(ulp)


def test_oracle_combines_the_reference_residue_files():
    """Residue lines written by the reference's cuckoo_fast, combined, give the N_bad its
    cuckoo_fail computes on 32-bit primes: two reference programs agree through the oracle."""
    texts = [c["stdout"] for c in fast_cases() if c["count"] in (60, 97)]
    res = post.read(texts)
    for (n, M), d in (((64, 65), 3), ((200, 400), 3)):
        _, x = post.row(n, M, d, res[(n, M)], pbits=49)
        assert x == reference_exact(n, M, d)[0]


def test_oracle_against_brute_force_counts():
    """check's brute-force N_bad over all C(M, d)^n outcomes for tiny shapes."""
    seen = 0
    for n, M, d, rep, nbad, total in check_instances():
        if rep:
            continue
        assert total == math.comb(M, d) ** n
        k = post.need(n, M, d, 49)
        primes = oracle.prime_range(0, k)
        line, x = post.row(n, M, d, {i: (p, nbad % p) for i, p in enumerate(primes)}, pbits=49)
        assert x == nbad
        seen += 1
    assert seen >= 5


def test_oracle_crt_and_digits_by_definition():
    rng = random.Random(3)
    for _ in range(40):
        k = rng.randint(1, 30)
        primes = oracle.prime_range(rng.randint(0, 50), k)
        m = math.prod(primes)
        N = rng.randrange(m)
        x, mm = post.crt([(p, N % p) for p in primes])
        assert mm == m and x == N and all(x % p == N % p for p in primes)
        D = rng.randrange(1, m) + N
        s = post.sci(N, D)
        if N == 0:
            assert s == "0"
            continue
        v = Decimal(s)
        assert len(s.split("e")[0].replace(".", "")) == 40
        assert Fraction(v) <= Fraction(N, D) < Fraction(v) + Fraction(Decimal(10) ** (v.adjusted() - 39))
        lg = Decimal(post.log2_text(N, D))
        assert abs(float(lg) - (math.log2(N) - math.log2(D))) < 1e-9


def test_post_need_matches_engine_need_on_the_reference_shapes():
    # post_gmp.py --pbits 49 reads cuckoo_fast's files: both needs agree on these shapes
    for n, M, d in [(16384, 32768, 3), (200, 400, 3), (64, 65, 3), (2048, 8192, 3), (65536, 131072, 3)]:
        assert post.need(n, M, d, 49) == oracle.need(n, M, d)
    assert post.need(200, 400, 3, 49) == 97


# ---------------------------------------------------------------- the product against the oracle

def test_post_need_abi():
    for n, M, d, pb in [(200, 400, 3, 31), (200, 400, 3, 49), (16384, 32768, 3, 49), (7, 8, 1, 61), (1, 2, 1, 31),
                        (4096, 4097, 5, 49), (1000, 1001, 2, 13)]:
        assert cf.post_need(n, M, d, pb) == post.need(n, M, d, pb), (n, M, d, pb)
    assert cf.post_need(5, 2, 3, 31) == -1


def test_cf_post_reference_residue_files(tmp_path):
    texts = [c["stdout"] for c in fast_cases() if c["count"] in (60, 97)]
    r = run_post(["3", "--pbits", "49"], texts, tmp_path)
    assert r.returncode == 0, r.stderr
    assert r.stdout == post.table(3, texts, pbits=49)
    assert "7.846949813310866398941487621092092415488e-14\t97" in r.stdout


@pytest.mark.parametrize("n,M,d", EXACT)
def test_cf_post_reference_exact_counts(n, M, d, tmp_path):
    nbad = reference_exact(n, M, d)[0]
    k = post.need(n, M, d, 49)
    text = residue_text(n, M, nbad, oracle.prime_range(0, k))
    dump = tmp_path / "dump"
    dump.mkdir()
    r = run_post([str(d), "--pbits", "49", "--dump", str(dump)], [text], tmp_path)
    assert r.returncode == 0, r.stderr
    assert r.stdout == post.table(d, [text], pbits=49)
    assert (dump / f"N_{n}_{M}.txt").read_text() == f"N_bad={nbad}\nQ^n=C({M},{d})^{n}\n"


def test_cf_post_random_and_boundary_values(tmp_path):
    rng = random.Random(7)
    texts = []
    for n, M, d in [(40, 80, 3), (100, 101, 2), (300, 700, 4), (77, 1000, 5), (513, 1100, 3), (9, 10, 1)]:
        Qn = math.comb(M, d) ** n
        k = post.need(n, M, d, 49)
        primes = oracle.prime_range(0, k)
        for N in (0, 1, Qn, Qn - 1, Qn >> 40, (Qn >> 40) + 1, Qn >> 64, (Qn >> 128) + 1, rng.randrange(Qn),
                  rng.randrange((Qn >> 50) + 1)):
            # each value in a file of its own (one (n, M) row per file set)
            t = residue_text(n, M, N, primes)
            r = run_post([str(d), "--pbits", "49"], [t], tmp_path)
            assert r.returncode == 0, r.stderr
            assert r.stdout == post.table(d, [t], pbits=49), (n, M, d, N)
            if N == 1:
                texts.append(t)
    # several shapes in one call, spread over files, in (n, M) order
    r = run_post(["3", "--pbits", "49"], texts[::-1], tmp_path)
    assert r.stdout == post.table(3, texts[::-1], pbits=49)


def test_cf_post_incomplete_duplicates_and_junk(tmp_path):
    n, M, d = 60, 100, 3
    nbad = reference_exact(n, M, d)[0]
    k = post.need(n, M, d, 49)
    primes = oracle.prime_range(0, k + 3)
    good = residue_text(n, M, nbad, primes)           # k + 3 primes: the extra ones are ignored
    wrong = residue_text(n, M, nbad + 12345, primes[:5])  # overridden by the later file
    junk = "# comment\n1 2 3\n\n7 8 9 10 11 12\n"
    r = run_post(["3", "--pbits", "49"], [wrong, junk, good], tmp_path)
    assert r.returncode == 0 and r.stdout == post.table(3, [wrong, junk, good], pbits=49)
    assert r.stdout.splitlines()[1].endswith(f"\t{k}") and "INCOMPLETE" not in r.stdout
    gap = "".join(line + "\n" for i, line in enumerate(good.splitlines()) if i != 3)
    other = residue_text(50, 51, 5, primes[:4])
    r = run_post(["3", "--pbits", "49"], [gap, other], tmp_path)
    assert r.returncode == 0 and r.stdout == post.table(3, [gap, other], pbits=49)
    assert f"INCOMPLETE {k + 2}/{k}" in r.stdout and "INCOMPLETE 4/" in r.stdout
    # default pbits 31: more primes needed than the file holds
    r = run_post(["3"], [good], tmp_path)
    assert r.stdout == post.table(3, [good]) and "INCOMPLETE" in r.stdout


def test_cf_post_errors(tmp_path):
    n, M, d = 60, 100, 3
    Qn = math.comb(M, d) ** n
    k = post.need(n, M, d, 49)
    primes = oracle.prime_range(0, k)
    first = residue_text(20, 40, 0, primes)
    over = residue_text(n, M, Qn + 1, primes)  # more than Q^n outcomes: the reference's assertion
    r = run_post(["3", "--pbits", "49"], [first, over], tmp_path)
    assert r.returncode == 1 and r.stdout == post.HEADER + "\n" + post.table(3, [first], pbits=49).splitlines()[1] + "\n"
    with pytest.raises(AssertionError):
        post.row(n, M, d, {i: (p, (Qn + 1) % p) for i, p in enumerate(primes)}, pbits=49)
    # too few primes for 4 Q^n (a huge --pbits makes need() small)
    r = run_post(["3", "--pbits", "4000"], [residue_text(n, M, 5, primes)], tmp_path)
    assert r.returncode == 1
    for argv in ([], ["x"], ["3", "--pbits"], ["3", "--dump"]):
        assert subprocess.run([CF_POST, *argv], capture_output=True).returncode == 1
    assert subprocess.run([CF_POST, "3", str(tmp_path / "missing.txt")], capture_output=True).returncode == 1
    r = run_post(["3"], ["0 5 6 7 x\n"], tmp_path)
    assert r.returncode == 1


def test_exact_combine_abi():
    n, M, d = 100, 150, 4
    nbad = reference_exact(n, M, d)[0]
    k = post.need(n, M, d, 49)
    primes = oracle.prime_range(0, k)
    res = [nbad % p for p in primes]
    got = cf.exact_combine(d, n, M, primes, res, decimal=True)
    line, _ = post.row(n, M, d, dict(enumerate(zip(primes, res))), pbits=49)
    cols = line.split("\t")
    assert got["log2_pfail"] == cols[3] and got["p_fail"] == cols[7]
    assert ["yes" if x else "no" for x in got["le"]] == cols[4:7]
    assert got["nbad"] == str(nbad) and got["nbad_bits"] == nbad.bit_length()
    assert got["qn_bits"] == (math.comb(M, d) ** n).bit_length()
    # any order of the moduli reconstructs the same integer
    perm = list(range(k))
    random.Random(1).shuffle(perm)
    assert cf.exact_combine(d, n, M, [primes[i] for i in perm], [res[i] for i in perm]) == \
        {key: v for key, v in got.items() if key != "nbad"}
    with pytest.raises(cf.CfError) as ei:
        cf.exact_combine(d, n, M, primes[:5], res[:5])
    assert ei.value.code == cf.CF_E_PRIMES
    for bad in ([4] + primes[1:], [primes[0]] + primes[:-1]):  # an even modulus; a repeated one
        with pytest.raises(cf.CfError) as ei:
            cf.exact_combine(d, n, M, bad, res)
        assert ei.value.code == cf.CF_E_ARGS


def test_cf_post_large_table(tmp_path):
    """n = 4096: 3046 primes, a 25000-bit N_bad."""
    n, M, d = 4096, 8192, 3
    k = post.need(n, M, d, 49)
    Qn = math.comb(M, d) ** n
    N = random.Random(5).randrange(Qn >> 30)
    t = residue_text(n, M, N, oracle.prime_range(0, k))
    r = run_post(["3", "--pbits", "49"], [t], tmp_path)
    assert r.returncode == 0 and r.stdout == post.table(3, [t], pbits=49)


@pytest.mark.gpu
def test_gpu_residues_through_cf_post(tmp_path):
    """cuckoo_fast on the device writes the residue file, cf_post combines it: the rows equal
    the oracle's for the reference's exact N_bad (C1's n = 200 known answer among them)."""
    shapes = [(200, 400), (60, 100), (64, 65), (20, 40)]
    q = tmp_path / "q.txt"
    q.write_text("".join(f"{n} {M}\n" for n, M in shapes))
    out = tmp_path / "res.txt"
    k = max(post.need(n, M, 3, 49) for n, M in shapes)
    print(f"cuckoo_fast 3 over {k} primes, {len(shapes)} queries", flush=True)
    r = subprocess.run([cf.CLI_FAST, "3", str(q), "0", str(k), "--out", str(out)], capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([CF_POST, "3", str(out), "--pbits", "49"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    want = [post.HEADER]
    for n, M in sorted(shapes):
        nbad = reference_exact(n, M, 3)[0]
        kk = post.need(n, M, 3, 49)
        want.append(post.row(n, M, 3, {i: (p, nbad % p) for i, p in enumerate(oracle.prime_range(0, kk))}, 49)[0])
    assert r.stdout == "\n".join(want) + "\n"
    assert "\t-43.534861355228510\tyes\tno\tno\t7.846949813310866398941487621092092415488e-14\t97" in r.stdout
