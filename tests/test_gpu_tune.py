"""The autotuner over the compiled task-kernel variants of each transform-length class
(cf_autotune / cf_set_variant): every variant gives the residues of the built-in one and of
the oracle, and the tuner's choice serves later calls."""
import pytest

import cuckoo_amd as cf
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = cf.Engine()
    yield e
    e.close()


@pytest.mark.parametrize("lg", [10, 11, 12, 13, 14])
def test_autotune_class(eng, lg):
    print(f"autotune 2^{lg}", flush=True)
    t = eng.autotune(lg)
    print(t, flush=True)
    assert t["log2_length"] == lg and 0 <= t["chosen"] < len(t["variants"]) and t["tables"] >= 1
    assert all(v["kernel_ms"] > 0 and v["resident_workgroups"] >= 1 for v in t["variants"])
    assert t["variants"][0]["waves_per_transform"] == (1 << lg) // 16 // 64 // (2 if lg == 14 else 1)
    L = 1 << lg
    queries = [(L // 2 + 37, L + 75), (L // 2 + 1, L // 2 + 2)]
    primes = cf.prime_range(5, 2)
    got = eng.run(queries, primes, 3)
    if lg <= 12:
        for p, r in zip(primes, got):
            assert r == oracle.nbad(3, p, queries), (lg, p)
    eng.set_variant(lg, 0)
    assert eng.run(queries, primes, 3) == got


@pytest.mark.parametrize("lg,variant", [(11, 1), (12, 1), (13, 1), (13, 2), (14, 1)])
def test_each_variant_equals_builtin_and_oracle(eng, lg, variant):
    L = 1 << lg
    cases = [(3, 0, [(L // 2 + 5, L + 11), (L - 3, L - 2), (L // 2 - 100, 3 * L // 2)]),
             (2, 1, [(L - 40, 2 * L)]), (5, 0, [(L // 2 + 900, L + 1000)])]
    primes = cf.prime_range(11, 2) + [cf.prime_range(200, 1)[0]]
    for d, model, queries in cases:
        print(f"2^{lg} variant {variant}: d={d} model={model} {queries}", flush=True)
        eng.set_variant(lg, 0)
        want = eng.run(queries, primes, d, model=model)
        eng.set_variant(lg, variant)
        got = eng.run(queries, primes, d, model=model)
        st = eng.stats()
        eng.set_variant(lg, 0)
        assert st.log2_length == lg
        assert got == want, (lg, variant, d, model)
        if lg <= 12:
            assert got[0] == oracle.nbad(d, primes[0], queries, rep=bool(model))


def test_autotune_on_a_given_size(eng):
    t = eng.autotune(12, n=4000)
    assert t["n"] == 4000 and len(t["variants"]) == 2
    with pytest.raises(cf.CfError):
        eng.autotune(12, n=2048)  # not in the class (2048, 4096]
    eng.set_variant(12, 0)


def test_variant_bounds(eng):
    with pytest.raises(cf.CfError):
        eng.set_variant(10, 1)
    with pytest.raises(cf.CfError):
        eng.set_variant(15, 0)
    eng.set_variant(14, -1)  # back to the built-in choice
