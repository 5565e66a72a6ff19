"""The C-ABI library: it loads, exports every symbol include/cuckoo_amd.h declares, and its host
logic (no device) agrees with the oracle and the reference's fixtures."""
import os
import re

import pytest

import cuckoo_amd as cf
import oracle
from golden_data import fast_cases, residue_lines

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(REPO, "include", "cuckoo_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 14
    L = cf.lib()
    for s in syms:
        assert hasattr(L, s), s


def test_prime_range_matches_oracle_and_reference():
    assert cf.prime_range(0, 40) == oracle.prime_range(0, 40)
    assert cf.prime_range(13, 5) == oracle.prime_range(13, 5)
    seen = {}
    for rec in fast_cases():
        for i, p, *_ in residue_lines(rec["stdout"]):
            seen[i] = p
    got = cf.prime_range(0, max(seen) + 1)
    assert all(got[i] == p for i, p in seen.items())
    assert cf.prime_range(0, 0) == []
    with pytest.raises(ValueError):
        cf.prime_range(-1, 3)


@pytest.mark.parametrize("n,M,d", [(16384, 32768, 3), (65536, 131072, 3), (200, 400, 3), (2048, 8192, 3),
                                   (8, 9, 1), (4096, 4097, 4), (100, 101, 5), (1, 2, 1), (31, 1000000, 2)])
def test_need_matches_oracle(n, M, d):
    assert cf.need(n, M, d) == oracle.need(n, M, d)
    assert cf.need(n, M, d, 62) == oracle.need(n, M, d, 62)


def test_transform_length_matches_oracle():
    for n in list(range(1, 70)) + [1023, 1024, 1025, 16383, 16384, 16385, 65536]:
        assert cf.transform_length(n) == oracle.transform_length(n)


@pytest.mark.parametrize("count,k", [(1, 1), (10, 5), (152, 8), (400, 10), (30, 20)])
def test_ntt_primes32_matches_oracle(count, k):
    assert cf.ntt_primes32(count, k) == oracle.ntt_primes32(count, k)


def test_mem_per_prime_layout():
    # rectangle-shaped live set: about nm^2/4 words per table instead of nm^2/2
    b = cf.mem_per_prime(16384, 32768)
    assert 1.0e9 < b < 1.2e9
    assert cf.mem_per_prime(1024, 2048) < cf.mem_per_prime(2048, 2048)


def test_engine_open_without_device_fails_loudly():
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK):
        pytest.skip("a HIP device may be present")
    with pytest.raises(cf.CfError):
        cf.Engine()


def test_host_crt_log2_matches_exact():
    """cf_crt_log2 (Garner over 50-bit primes, long double leading digits) against Python's exact CRT."""
    import math
    import random

    rng = random.Random(3)
    primes = cf.prime_range(0, 300)
    for k in (1, 2, 5, 64, 300):
        for _ in range(3):
            res = [rng.randrange(p) for p in primes[:k]]
            N, _ = oracle.crt(zip(primes[:k], res))
            shift = max(N.bit_length() - 60, 0)
            want = math.log2(N >> shift) + shift if N else -math.inf
            got = cf.crt_log2(primes[:k], res)
            assert abs(got - want) < 1e-12 * max(1.0, want), (k, got, want)
    # small values, including zero and values below the first prime
    assert cf.crt_log2(primes[:3], [0, 0, 0]) == -math.inf
    assert cf.crt_log2(primes[:3], [5, 5, 5]) == math.log2(5)
