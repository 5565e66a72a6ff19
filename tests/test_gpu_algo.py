"""The reference's forward (O(n^4)) and cubic (O(n^3)) one-table algorithms as device kernels
(cf_run_algo), behind `cuckoo_fail single --algo forward|cubic`.

The CLI is compared line by line with the reference's own outputs for both algorithms
(tests/golden/fail_algo_cases.json, written by the reference's cuckoo_fail); the ABI with the
oracle's forward and cubic restatements on random shapes, both models and 50-bit primes."""
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


@pytest.mark.parametrize("idx", range(len(load("fail_algo_cases.json"))))
def test_algo_cli_matches_reference(idx):
    rec = load("fail_algo_cases.json")[idx]
    print("cuckoo_fail", " ".join(rec["args"]), flush=True)
    r = subprocess.run([cf.CLI_FAIL, *rec["args"]], capture_output=True, text=True, timeout=300)
    assert r.returncode == rec["rc"] == 0, r.stderr
    want = [l for l in rec["stdout"].splitlines() if "time:" not in l and "(T s)" not in l]
    got = [l for l in r.stdout.splitlines() if "time:" not in l and not l.startswith("N_bad mod")]
    assert got == want
    for wl, gl in zip(rec["stdout"].splitlines(), r.stdout.splitlines()):
        if wl.startswith("primes:"):
            assert gl.split("time:")[0] == wl.split("time:")[0]
        if wl.startswith("N_bad mod"):
            assert gl.split("(")[0] == wl.split("(")[0]


@pytest.mark.parametrize("algo", ["forward", "cubic"])
def test_algo_vs_oracle_random(eng, algo):
    rng = random.Random(17 if algo == "forward" else 19)
    primes = cf.prime_range(0, 4) + [998244353]
    for _ in range(12):
        n = rng.randint(1, 40)
        M = n + rng.randint(1, 30)
        d, model = rng.randint(1, 5), rng.randint(0, 1)
        got = eng.run_algo(algo, n, M, primes, d, model)
        for p, g in zip(primes, got):
            assert g == oracle.nbad(d, p, [(n, M)], rep=bool(model), algo=algo)[0], (algo, n, M, d, model, p)


def test_algos_agree_with_the_suffix_engine(eng):
    primes = cf.prime_range(0, 8)
    for n, M, d, model in [(100, 130, 3, 0), (64, 65, 2, 1), (150, 400, 4, 0)]:
        ntt = [row[0] for row in eng.run([(n, M)], primes, d, model=model)]
        assert eng.run_algo("cubic", n, M, primes, d, model) == ntt
        if n <= 100:
            assert eng.run_algo("forward", n, M, primes, d, model) == ntt


def test_algo_limits(eng):
    P = cf.prime_range(0, 1)
    for algo, n in (("forward", 257), ("cubic", 1025)):
        with pytest.raises(cf.CfError) as ei:
            eng.run_algo(algo, n, n + 1, P, 3)
        assert ei.value.code == cf.CF_E_UNSUPPORTED
    with pytest.raises(cf.CfError) as ei:
        eng.run_algo("cubic", 50, 60, [59], 3)  # prime not above M
    assert ei.value.code == cf.CF_E_ARGS
    assert eng.run_algo("forward", 5, 9, [], 2) == []
