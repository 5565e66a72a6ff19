"""Residues modulo the reference's 62-bit prime list (its -DFP_RB=64 build: primes c 2^20 + 1 below
2^62, Montgomery R = 2^64), computed by the 64-bit Montgomery kernel.

Pinned by the reference's own lines (tests/golden/fast62_cases.json, written by that build), by
the exact N_bad of its 32-bit CLI (the whole need() set reconstructs it), and by the oracle on
random shapes (both models, multi-query files); the driver bin/cuckoo_fast62 prints the reference's
lines byte for byte. CPU tests check the fixture itself against the oracle and the exact N_bad."""
import os
import random
import subprocess
import tempfile

import pytest

import cuckoo_amd as cf
import oracle
from golden_data import fail_case, fail_field, load, residue_lines

CLI_FAST62 = os.path.join(os.path.dirname(cf.CLI_FAST), "cuckoo_fast62")
P62 = oracle.prime_range(0, 4, 62)


def _cases():
    try:
        return load("fast62_cases.json")
    except OSError:
        return []


def _exact(args):
    return int(fail_field(fail_case(args)["stdout"], "N_bad"))


# ------------------------------------------------------------------ CPU: the fixture itself

def test_fixture_prime_list_is_the_62_bit_list():
    for rec in _cases():
        for i, p, _, _, _ in residue_lines(rec["stdout"]):
            assert oracle.prime_range(i, 1, 62) == [p]
    assert all(2 ** 61 < p < 2 ** 62 and (p - 1) % 2 ** 20 == 0 for p in P62)


def test_fixture_whole_need_sets_give_the_exact_nbad():
    for n, M, args in ((200, 400, ["single", "-n", "200", "-M", "400", "-d", "3", "--algo", "ntt", "--exact"]),
                       (64, 65, ["single", "-n", "64", "-M", "65", "-d", "3", "--exact"])):
        rec = next(r for r in _cases() if r["queries"] == [[n, M]] and r["first"] == 0 and r["count"] > 10)
        lines = residue_lines(rec["stdout"])
        assert len(lines) == oracle.need(n, M, 3, 62)
        N, _ = oracle.crt([(p, r) for _, p, _, _, r in lines])
        assert N == _exact(args)


def test_fixture_small_shapes_equal_the_oracle():
    for rec in _cases():
        qs = [tuple(q) for q in rec["queries"]]
        if max(n for n, _ in qs) > 600:
            continue
        for i, p, n, M, r in residue_lines(rec["stdout"]):
            assert oracle.nbad(rec["d"], p, [(n, M)])[0] == r, (rec["d"], i, n, M)


# ------------------------------------------------------------------ GPU

@pytest.fixture(scope="module")
def eng():
    e = cf.Engine()
    yield e
    e.close()


@pytest.mark.gpu
@pytest.mark.parametrize("idx", range(len(_cases())))
def test_reference_62_bit_lines(eng, idx):
    rec = _cases()[idx]
    lines = residue_lines(rec["stdout"])
    qs = [tuple(q) for q in rec["queries"]]
    primes = sorted({(i, p) for i, p, _, _, _ in lines})
    print(f"d={rec['d']} queries={qs[:4]} primes {primes[0][0]}..{primes[-1][0]}", flush=True)
    mask = [[1 if i < oracle.need(n, M, rec["d"], 62) else 0 for n, M in qs] for i, _ in primes]
    got = eng.run(qs, [p for _, p in primes], rec["d"], mask=mask)
    pos = {i: b for b, (i, _) in enumerate(primes)}
    for i, p, n, M, r in lines:
        assert got[pos[i]][qs.index((n, M))] == r, (i, n, M)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(4))
def test_random_62_bit_against_oracle(eng, seed):
    rng = random.Random(100 + seed)
    d = rng.randint(1, 5)
    model = rng.randint(0, 1)
    qs = []
    for _ in range(rng.randint(1, 11)):
        n = rng.randint(1, 300)
        qs.append((n, rng.randint(n + 1, 3 * n + 10)))
    primes = P62[:2] + [rng.choice(oracle.prime_range(50, 40, 62))]
    got = eng.run(qs, primes, d, model=model)
    for b, p in enumerate(primes):
        assert got[b] == oracle.nbad(d, p, qs, rep=bool(model)), (d, model, p, qs)


@pytest.mark.gpu
def test_mixed_50_and_62_bit_primes_in_one_call(eng):
    qs = [(200, 400), (64, 65), (37, 80)]
    primes = cf.prime_range(0, 2) + P62[:2]
    got = eng.run(qs, primes, 3)
    assert got[0] == [918297158731096, 201776960213191, 67342126954205]
    for b, p in enumerate(primes):
        assert got[b] == oracle.nbad(3, p, qs), p


@pytest.mark.gpu
def test_cuckoo_fast62_prints_the_reference_lines():
    for rec in _cases():
        if max(n for n, _ in rec["queries"]) > 4096:
            continue
        with tempfile.TemporaryDirectory() as td:
            qf = os.path.join(td, "q.txt")
            with open(qf, "w") as f:
                f.write("".join(f"{n} {M}\n" for n, M in rec["queries"]))
            r = subprocess.run([CLI_FAST62, str(rec["d"]), qf, str(rec["first"]), str(rec["count"])], capture_output=True,
                               text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        assert r.stdout == rec["stdout"], (rec["d"], rec["queries"])


def _wide_groups():
    # the reference's lines for tables of more than 16384 rows, grouped by (d, queries)
    try:
        recs = load("fast62_wide.json")
    except OSError:
        return []
    groups = {}
    for r in recs:
        key = (r["d"], tuple(tuple(q) for q in r["queries"]))
        groups.setdefault(key, []).extend(residue_lines(r["stdout"]))
    return sorted(groups.items(), key=lambda g: max(n for n, _ in g[0][1]))


@pytest.mark.gpu
@pytest.mark.parametrize("g", range(7))
def test_62_bit_tables_beyond_16384_rows(eng, g):
    # transforms of 2^15 and 2^16 with the buffer in HBM: every line of the reference's
    # -DFP_RB=64 build for tables of 16385 to 40000 rows, d = 1 to 5
    groups = _wide_groups()
    if g >= len(groups):
        pytest.skip("fixture not generated")
    (d, qs), lines = groups[g]
    idx = sorted({i for i, _, _, _, _ in lines})
    primes = [p for i, p, _, _, _ in sorted(set((i, p, 0, 0, 0) for i, p, _, _, _ in lines))]
    print(f"d={d} queries={qs} primes={idx}", flush=True)
    got = eng.run(list(qs), primes, d)
    for i, p, n, M, r in lines:
        assert got[idx.index(i)][list(qs).index((n, M))] == r, (d, i, n, M)


@pytest.mark.gpu
def test_cuckoo_fast62_beyond_16384_rows_prints_the_reference_lines():
    # the driver end to end on a table of 16400 rows (d = 4, prime 3): the reference's output bytes
    try:
        recs = load("fast62_wide.json")
    except OSError:
        pytest.skip("fixture not generated")
    rec = next(r for r in recs if r["d"] == 4)
    with tempfile.TemporaryDirectory() as td:
        qf = os.path.join(td, "q.txt")
        with open(qf, "w") as f:
            f.write("".join(f"{n} {M}\n" for n, M in rec["queries"]))
        r = subprocess.run([CLI_FAST62, str(rec["d"]), qf, str(rec["first"]), str(rec["count"])], capture_output=True,
                           text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert r.stdout == rec["stdout"]


@pytest.mark.gpu
@pytest.mark.skipif(os.environ.get("CF_SLOW_TESTS") != "1",
                    reason="a table of 65537 rows takes about 126 s on one CU of an MI355X; run with CF_SLOW_TESTS=1")
def test_62_bit_table_of_65537_rows(eng):
    # the longest transform, 2^17 (its twiddle tables fill the LDS): the reference's -DFP_RB=64 lines
    # for two queries on prime 2
    try:
        rec = load("fast62_65537.json")
    except OSError:
        pytest.skip("fixture not generated")
    lines = residue_lines(rec["stdout"])
    qs = [tuple(q) for q in rec["queries"]]
    want = {(n, M): r for _, _, n, M, r in lines}
    print(f"d=3 queries={qs} prime {rec['first']}", flush=True)
    got = eng.run(qs, [lines[0][1]], 3)
    assert got[0] == [want[q] for q in qs]


@pytest.mark.gpu
def test_62_bit_beyond_131072_rows_is_refused(eng):
    with pytest.raises(cf.CfError) as ei:
        eng.run([(131073, 131074)], P62[:1], 3)
    assert ei.value.code == cf.CF_E_UNSUPPORTED
