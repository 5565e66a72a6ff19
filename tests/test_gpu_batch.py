"""The uneven batch (one work-queue launch over instances of different n, M, d) through the C ABI.

Pinned against the reference's own residues (cuckoo_fast run once per instance over its whole
prime set) for the batch's instances with n <= 512: the first lines verbatim and the CRT-combined
exact N_bad by hash; against the per-instance engine call for the rest of the shapes; against
the oracle for mixed models and d."""
import hashlib
import random

import pytest

import cuckoo_amd as cf
import oracle
from golden_data import load, residue_lines

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    e = cf.Engine()
    yield e
    e.close()


def _nbad_hash(primes, res):
    N, _ = oracle.crt(zip(primes, res))
    return N.bit_length(), hashlib.sha256(N.to_bytes((N.bit_length() + 7) // 8, "little")).hexdigest()


def test_uneven_small_instances_vs_reference(eng):
    g = load("uneven_batch.json")
    gold = g["golden"]
    insts = [(x["d"], 0, x["n"], x["M"], cf.prime_range(0, x["need"])) for x in gold]
    print(f"uneven batch: {len(insts)} instances, {sum(len(i[4]) for i in insts)} tables", flush=True)
    res = eng.run_batch(insts)
    st = eng.stats()
    # instances of one d share each prime's table
    assert st.launches == 1 and st.tasks == len({(x["d"], p) for x in gold for p in cf.prime_range(0, x["need"])})
    for x, inst, r in zip(gold, insts, res):
        for line in residue_lines("\n".join(x["head"])):
            i, p, n, M, want = line
            assert inst[4][i] == p and r[i] == want, (x["index"], i)
        assert _nbad_hash(inst[4], r) == (x["nbad_bits"], x["nbad_sha256"]), x["index"]
    # one table per (instance, prime): the same residues
    eng.set_batch_sharing(False)
    try:
        assert eng.run_batch(insts) == res
        assert eng.stats().tasks == sum(x["need"] for x in gold)
    finally:
        eng.set_batch_sharing(True)


@pytest.mark.parametrize("share", [True, False])
def test_batch_equals_per_instance_calls(eng, share):
    # every queue length class (L = 2^10, 2^11, 2^12) plus one longer table (2^13) that runs in
    # a launch of its own, models 0 and 1, d from 1 to 5, shared primes across instances (with
    # sharing, primes 0 and 1 of d = 3 are each one table of 5000 rows, a 2^13 transform, that also
    # carries the queries of n = 4000 and 1025)
    P = cf.prime_range(0, 6)
    spec = [(3, 0, 4000, 9000, P[0:3]), (2, 0, 1500, 2000, P[0:4]), (4, 1, 700, 900, P[1:3]),
            (3, 0, 5000, 12000, P[0:2]), (5, 0, 9, 10, P[0:5]), (1, 0, 100, 150, P[2:6]),
            (2, 1, 2048, 2049, P[3:5]), (3, 0, 1025, 3000, P[0:2] + P[4:6]), (4, 0, 64, 65, [998244353, 1004535809])]
    eng.set_batch_sharing(share)
    try:
        res = eng.run_batch(spec)
        st = eng.stats()
    finally:
        eng.set_batch_sharing(True)
    assert st.launches == 2 and st.log2_length == 13
    tables = sum(len(x[4]) for x in spec)
    assert st.tasks == (tables - 4 if share else tables)  # (d = 3, P0) and (d = 3, P1): three instances, one table
    for (d, model, n, M, primes), r in zip(spec, res):
        want = [row[0] for row in eng.run([(n, M)], primes, d, model=model)]
        assert r == want, (d, model, n, M)


def test_batch_vs_oracle_random(eng):
    rng = random.Random(11)
    primes = cf.prime_range(0, 64)
    spec, at = [], 0
    for _ in range(24):
        n = rng.randint(1, 120)
        M = n + rng.randint(1, 2 * n + 2)
        d = rng.randint(1, 5)
        model = rng.randint(0, 1)
        k = rng.randint(1, 3)
        spec.append((d, model, n, M, primes[at:at + k]))
        at = (at + k) % 60
    res = eng.run_batch(spec)
    for (d, model, n, M, ps), r in zip(spec, res):
        for p, got in zip(ps, r):
            assert got == oracle.nbad(d, p, [(n, M)], rep=bool(model))[0], (d, model, n, M, p)


def test_batch_alone_equals_in_mix(eng):
    P = cf.prime_range(10, 4)
    alone = eng.run_batch([(3, 0, 2500, 5000, P[:2])])[0]
    mix = eng.run_batch([(2, 0, 300, 700, P[2:]), (3, 0, 2500, 5000, P[:2]), (4, 1, 90, 91, P[1:3])])
    assert mix[1] == alone


def test_batch_edges(eng):
    assert eng.run_batch([]) == []
    assert eng.run_batch([(3, 0, 50, 60, [])]) == [[]]
    P = cf.prime_range(0, 2)
    with pytest.raises(cf.CfError) as ei:
        eng.run_batch([(0, 0, 50, 60, P)])
    assert ei.value.code == cf.CF_E_ARGS
    with pytest.raises(cf.CfError) as ei:
        eng.run_batch([(3, 0, 60, 60, P)])
    assert ei.value.code == cf.CF_E_ARGS
    with pytest.raises(cf.CfError) as ei:
        eng.run_batch([(3, 0, 20, 60, [59])])  # prime not above M
    assert ei.value.code == cf.CF_E_ARGS
    with pytest.raises(cf.CfError) as ei:
        eng.run_batch([(3, 0, 140000, 280000, P[:1])])  # transform beyond 2^17
    assert ei.value.code == cf.CF_E_UNSUPPORTED


def test_uneven_sample_of_large_instances_vs_reference(eng):
    # instances with n > 512 (every 5th in order of n, and the three largest): the reference's first
    # lines verbatim and the CRT-combined exact N_bad by hash, all in one batch
    try:
        gold = load("uneven_batch_big.json")["golden"]
    except OSError:
        pytest.skip("fixture not generated")
    insts = [(x["d"], 0, x["n"], x["M"], cf.prime_range(0, x["need"])) for x in gold]
    print(f"uneven sample: {len(insts)} instances, {sum(len(i[4]) for i in insts)} tables", flush=True)
    res = eng.run_batch(insts)
    for x, inst, r in zip(gold, insts, res):
        for i, p, n, M, want in residue_lines("\n".join(x["head"])):
            assert inst[4][i] == p and r[i] == want, (x["index"], i)
        assert _nbad_hash(inst[4], r) == (x["nbad_bits"], x["nbad_sha256"]), x["index"]


def test_queue_variants_and_batch_autotuner(eng):
    # every compiled variant of the work-queue kernel gives the per-instance path's residues; the
    # tuner times them on a sample of the batch and keeps the fastest
    P = cf.prime_range(0, 40)
    spec = [(3, 0, 3000, 7000, P[:24]), (2, 0, 1500, 2000, P[:30]), (4, 1, 700, 900, P[5:37]),
            (2, 0, 300, 301, P[:40]), (5, 0, 64, 70, P[:11]), (3, 0, 1025, 2000, P[10:20])]
    want = [[row[0] for row in eng.run([(n, M)], primes, d, model=model)] for d, model, n, M, primes in spec]
    for v in range(3):
        eng.set_batch_variant(v)
        assert eng.run_batch(spec) == want, v
    t = eng.autotune_batch(spec, sample_tables=60)
    print(t, flush=True)
    assert len(t["variants"]) == 3 and 0 <= t["chosen"] < 3 and t["tables"] >= len(spec)
    assert t["variants"][0]["threads"] == 256 and t["variants"][2]["threads"] == 512
    ms = [v["kernel_ms"] for v in t["variants"]]
    assert ms[t["chosen"]] == min(ms)
    assert eng.run_batch(spec) == want
    eng.set_batch_variant(0)
