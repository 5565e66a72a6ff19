"""Python face of the CPU oracle (oracle.c) plus exact-integer helpers.

TEST INFRASTRUCTURE ONLY: the checker for the HIP engine. Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may import this module; the product never does.

The exact-integer helpers use Python's unbounded integers, so the Chinese remaindering here is
independent of the product's own C++ reconstruction.
"""
import ctypes
import math
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(HERE, "liboracle.so")
_lib = None

u64 = ctypes.c_uint64


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.or_is_prime_u64.argtypes = [u64]
        L.or_prime_range.restype = ctypes.c_long
        L.or_prime_range.argtypes = [ctypes.c_int, ctypes.c_long, ctypes.c_long, ctypes.POINTER(u64)]
        L.or_log2_q.restype = ctypes.c_double
        L.or_log2_q.argtypes = [ctypes.c_int, ctypes.c_int]
        L.or_need.restype = ctypes.c_long
        L.or_need.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.or_transform_length.restype = ctypes.c_size_t
        L.or_transform_length.argtypes = [ctypes.c_size_t]
        L.or_ntt_primes32.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32)]
        for f in (L.or_suffix_table_ntt, L.or_suffix_table_cubic):
            f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, u64, ctypes.POINTER(u64)]
        L.or_nbad.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, u64, ctypes.c_int,
                              ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(u64)]
        L.or_nbad_horner.argtypes = [ctypes.c_int, ctypes.c_int, u64, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(u64)]
        L.or_nbad_forward.argtypes = [ctypes.c_int, ctypes.c_int, u64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(u64)]
        L.or_brute_force.argtypes = [ctypes.c_int] * 4 + [ctypes.POINTER(ctypes.c_ulonglong)] * 2
        for f in (L.or_separate_nbad, L.or_separate_forward):
            f.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int, u64, ctypes.POINTER(u64)]
        for f in (L.or_cyclic_conv_schoolbook, L.or_cyclic_conv_ntt):
            f.argtypes = [u64, ctypes.c_size_t, ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.POINTER(u64)]
        L.or_cyclic_conv_ntt.restype = ctypes.c_int
        _lib = L
    return _lib


def _arr(t, xs):
    return (t * len(xs))(*xs)


def prime_range(first, count, prime_bits=50):
    out = (u64 * max(count, 1))()
    k = lib().or_prime_range(prime_bits, first, count, out)
    return [out[i] for i in range(k)]


def need(n, M, d, prime_bits=50):
    return lib().or_need(n, M, d, prime_bits)


def log2_q(M, d):
    return lib().or_log2_q(M, d)


def transform_length(nmax):
    return lib().or_transform_length(nmax)


def ntt_primes32(count, min_pow2):
    out = (ctypes.c_uint32 * count)()
    if lib().or_ntt_primes32(count, min_pow2, out) < 0:
        raise RuntimeError("not enough NTT-friendly primes")
    return list(out)


def suffix_table(nm, d, p, rep=False, algo="ntt"):
    T = nm * (nm + 1) // 2
    V = (u64 * T)()
    f = lib().or_suffix_table_ntt if algo == "ntt" else lib().or_suffix_table_cubic
    if f(nm, d, int(rep), p, V):
        raise RuntimeError("oracle suffix table failed")
    return list(V)


def nbad(d, p, queries, rep=False, algo="ntt"):
    """N_bad mod p for each (n, M) by the paper's double sum (algo 'ntt', 'cubic'), by the fast
    engine's Horner form (algo 'horner') or by the O(n^4) forward recurrence (algo 'forward')."""
    ns = _arr(ctypes.c_int, [q[0] for q in queries])
    Ms = _arr(ctypes.c_int, [q[1] for q in queries])
    out = (u64 * len(queries))()
    if algo == "horner":
        rc = lib().or_nbad_horner(d, int(rep), p, len(queries), ns, Ms, out)
    elif algo == "forward":
        rc = 0
        for j, (n, M) in enumerate(queries):
            x = u64()
            rc = rc or lib().or_nbad_forward(d, int(rep), p, n, M, ctypes.byref(x))
            out[j] = x.value
    else:
        rc = lib().or_nbad(0 if algo == "ntt" else 1, d, int(rep), p, len(queries), ns, Ms, out)
    if rc:
        raise RuntimeError("oracle nbad failed")
    return list(out)


def separate_nbad(n, m, p, algo="suffix"):
    """N_bad mod p for n items in separate tables of lengths m (list), by the suffix table's
    defining sums (algo 'suffix') or the forward recurrence (algo 'forward')."""
    x = u64()
    f = lib().or_separate_nbad if algo == "suffix" else lib().or_separate_forward
    if f(n, _arr(ctypes.c_int, m), len(m), p, ctypes.byref(x)):
        raise RuntimeError("oracle separate-table count failed")
    return x.value


def brute_force(n, M, d, rep=False):
    a, b = ctypes.c_ulonglong(), ctypes.c_ulonglong()
    if lib().or_brute_force(n, M, d, int(rep), ctypes.byref(a), ctypes.byref(b)):
        raise RuntimeError("brute force out of range")
    return a.value, b.value


def cyclic_conv(p, a, k, method="ntt"):
    L = len(a)
    A, K, O = _arr(u64, a), _arr(u64, k), (u64 * L)()
    if method == "ntt":
        if lib().or_cyclic_conv_ntt(p, L, A, K, O):
            raise RuntimeError("bad length")
    else:
        lib().or_cyclic_conv_schoolbook(p, L, A, K, O)
    return list(O)


# ---- exact integers ----

def crt(pairs):
    """The unique x in [0, prod p) with x = r (mod p) for every (p, r)."""
    x, m = 0, 1
    for p, r in pairs:
        t = ((r - x) * pow(m, -1, p)) % p
        x, m = x + m * t, m * p
    return x, m


def q_count(M, d, rep=False):
    return M ** d if rep else math.comb(M, d)


def birthday_nbad(n, M):
    """d = 1 closed form: N_bad = M^n - M!/(M-n)!."""
    return M ** n - math.perm(M, n)
