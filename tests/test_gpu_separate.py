"""The separate-tables model on the device (cf_run_separate, `cuckoo_fail separate`): the CLI
against the reference's own runs of both algorithms, the ABI against the oracle on random
shapes, and the two device algorithms against each other."""
import random
import subprocess

import pytest

import cuckoo_amd as cf
import oracle
from golden_data import load

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = cf.Engine()
    yield e
    e.close()


@pytest.mark.parametrize("idx", [i for i, r in enumerate(load("fail_separate_cases.json")) if r["rc"] == 0])
def test_separate_cli_matches_reference(idx):
    rec = load("fail_separate_cases.json")[idx]
    print("cuckoo_fail", " ".join(rec["args"]), flush=True)
    r = subprocess.run([cf.CLI_FAIL, *rec["args"]], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    want = [l for l in rec["stdout"].splitlines() if "time:" not in l and "(T s)" not in l]
    got = [l for l in r.stdout.splitlines() if "time:" not in l and not l.startswith("N_bad mod")]
    assert got == want
    for wl, gl in zip(rec["stdout"].splitlines(), r.stdout.splitlines()):
        if wl.startswith("primes:"):
            assert gl.split("time:")[0] == wl.split("time:")[0]
        if wl.startswith("N_bad mod"):
            assert gl.split("(")[0] == wl.split("(")[0]


@pytest.mark.parametrize("algo", ["suffix", "forward"])
def test_separate_vs_oracle_random(eng, algo):
    rng = random.Random(23 if algo == "suffix" else 29)
    primes = cf.prime_range(0, 3) + [998244353]
    for _ in range(15):
        d = rng.randint(1, 4)
        m = [rng.randint(1, 9) for _ in range(d)]
        if sum(m) < 2:
            continue
        n = rng.randint(1, min(sum(m) - 1, 14 if algo == "suffix" else 10))
        got = eng.run_separate(n, m, primes, algo)
        for p, g in zip(primes, got):
            assert g == oracle.separate_nbad(n, m, p), (algo, n, m, p)


def test_suffix_equals_forward_on_device(eng):
    primes = cf.prime_range(4, 4)
    for n, m in [(30, [20, 25]), (16, [6, 7, 8]), (40, [41]), (9, [3, 3, 3, 3])]:
        assert eng.run_separate(n, m, primes, "suffix") == eng.run_separate(n, m, primes, "forward"), (n, m)


def test_separate_limits(eng):
    P = cf.prime_range(0, 1)
    with pytest.raises(cf.CfError) as ei:
        eng.run_separate(5, [3, 0], P)
    assert ei.value.code == cf.CF_E_ARGS
    with pytest.raises(cf.CfError) as ei:
        eng.run_separate(10, [4, 5], P)
    assert ei.value.code == cf.CF_E_ARGS
    with pytest.raises(cf.CfError) as ei:
        eng.run_separate(4, [1] * 9, P)
    assert ei.value.code == cf.CF_E_UNSUPPORTED
    with pytest.raises(cf.CfError) as ei:
        eng.run_separate(5, [3, 4], [5])  # prime not above M
    assert ei.value.code == cf.CF_E_ARGS
    assert eng.run_separate(5, [3, 4], []) == []
