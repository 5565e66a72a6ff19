"""Python binding of the MI355X one-table failure-count engine (include/cuckoo_amd.h).

The engine is the HIP library lib/libcuckoo_amd.so built by `make` in this directory. There is
no fallback: if the library is missing or no HIP device is present, the calls raise.

    from cuckoo_amd import Engine, prime_range, need
    with Engine() as e:
        res = e.run([(16384, 32768)], prime_range(0, 8), d=3)   # res[b][j] = N_bad mod p_b

Host-side functions (prime_range, need, transform_length, ntt_primes32) mirror the reference's
host logic (fast/cuckoo_fast.cu:423-472, include/cuckoo/modint.hpp:183-196) and need no device.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libcuckoo_amd.so")
CLI_FAST = os.path.join(HERE, "bin", "cuckoo_fast")
CLI_FAIL = os.path.join(HERE, "bin", "cuckoo_fail")

CF_OK, CF_E_DEVICE, CF_E_ARGS, CF_E_NOMEM, CF_E_PRIMES, CF_E_UNSUPPORTED = 0, 1, 2, 3, 4, 5
CF_ALGO_FORWARD, CF_ALGO_CUBIC, CF_ALGO_SUFFIX = 1, 2, 3

_lib = None
u64 = ctypes.c_uint64


class CfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} (code {code})")
        self.code = code


class Stats(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("wall_ms", ctypes.c_double), ("launches", ctypes.c_int),
                ("tasks", ctypes.c_int), ("batches", ctypes.c_int), ("log2_length", ctypes.c_int),
                ("device_bytes", ctypes.c_size_t), ("graph", ctypes.c_int),
                ("wide_split", ctypes.c_int)]


class Instance(ctypes.Structure):
    _fields_ = [("d", ctypes.c_int32), ("model", ctypes.c_int32), ("n", ctypes.c_int32), ("M", ctypes.c_int32),
                ("prime_off", ctypes.c_int64), ("prime_cnt", ctypes.c_int64)]


class Exact(ctypes.Structure):
    _fields_ = [("log2_pfail", ctypes.c_char * 48), ("le", ctypes.c_int32 * 3), ("p_fail", ctypes.c_char * 64),
                ("nbad_bits", ctypes.c_int64), ("qn_bits", ctypes.c_int64), ("error", ctypes.c_char * 128)]


class Tune(ctypes.Structure):
    _fields_ = [("log2_length", ctypes.c_int), ("variants", ctypes.c_int), ("chosen", ctypes.c_int),
                ("tables", ctypes.c_int), ("n", ctypes.c_int), ("waves", ctypes.c_int * 4), ("deferred", ctypes.c_int * 4),
                ("resident", ctypes.c_int * 4), ("ms", ctypes.c_double * 4)]


class BatchTune(ctypes.Structure):
    _fields_ = [("variants", ctypes.c_int), ("chosen", ctypes.c_int), ("tables", ctypes.c_int),
                ("threads", ctypes.c_int * 4), ("per_cu", ctypes.c_int * 4), ("resident", ctypes.c_int * 4),
                ("ms", ctypes.c_double * 4)]


class ServeStats(ctypes.Structure):
    _fields_ = [("wall_s", ctypes.c_double), ("queries_per_s", ctypes.c_double), ("p50_ms", ctypes.c_double),
                ("p99_ms", ctypes.c_double), ("mean_batch", ctypes.c_double), ("batches", ctypes.c_long)]


def lib():
    """The loaded engine library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is missing: build it with `make -C {HERE}`")
        L = ctypes.CDLL(LIB_PATH)
        L.cf_open.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.cf_close.argtypes = [ctypes.c_void_p]
        L.cf_close.restype = None
        L.cf_strerror.restype = ctypes.c_char_p
        L.cf_strerror.argtypes = [ctypes.c_int]
        L.cf_last_error.restype = ctypes.c_char_p
        L.cf_last_error.argtypes = [ctypes.c_void_p]
        L.cf_last_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(Stats)]
        L.cf_device_name.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
        L.cf_set_graphs.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.cf_run_separate.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int32),
                                      ctypes.c_int, ctypes.POINTER(u64), ctypes.c_size_t, ctypes.POINTER(u64)]
        L.cf_autotune.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Tune)]
        L.cf_set_variant.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.cf_set_batch_variant.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.cf_set_batch_sharing.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.cf_autotune_batch.argtypes = [ctypes.c_void_p, ctypes.POINTER(Instance), ctypes.c_size_t, ctypes.POINTER(u64),
                                        ctypes.c_size_t, ctypes.POINTER(BatchTune)]
        L.cf_max_batch62.restype = ctypes.c_long
        L.cf_max_batch62.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_size_t]
        L.cf_mem_per_prime62.restype = ctypes.c_size_t
        L.cf_mem_per_prime62.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_size_t]
        L.cf_run_fused.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int32),
                                   ctypes.POINTER(ctypes.c_int32), ctypes.c_size_t, ctypes.POINTER(u64), ctypes.c_size_t,
                                   ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(u64)]
        L.cf_run_batch.argtypes = [ctypes.c_void_p, ctypes.POINTER(Instance), ctypes.c_size_t, ctypes.POINTER(u64),
                                   ctypes.POINTER(u64)]
        L.cf_run_algo.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 5 + [ctypes.POINTER(u64), ctypes.c_size_t,
                                                                           ctypes.POINTER(u64)]
        vp = ctypes.c_void_p
        L.cf_tables_build.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(u64),
                                      ctypes.c_size_t, ctypes.POINTER(vp)]
        L.cf_tables_free.argtypes = [vp]
        L.cf_tables_free.restype = None
        L.cf_tables_bytes.argtypes = [vp]
        L.cf_tables_bytes.restype = ctypes.c_size_t
        L.cf_tables_need.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.cf_tables_need.restype = ctypes.c_long
        L.cf_query.argtypes = [vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), ctypes.c_size_t,
                               ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(u64)]
        L.cf_query_noalloc.argtypes = L.cf_query.argtypes
        L.cf_crt_log2.argtypes = [ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.c_size_t]
        L.cf_crt_log2.restype = ctypes.c_double
        L.cf_exact_combine.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(u64), ctypes.POINTER(u64),
                                       ctypes.c_size_t, ctypes.POINTER(Exact), ctypes.c_char_p, ctypes.c_size_t]
        L.cf_post_need.argtypes = [ctypes.c_int] * 4
        L.cf_post_need.restype = ctypes.c_long
        L.cf_server_open.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
        L.cf_server_close.argtypes = [vp]
        L.cf_server_close.restype = None
        L.cf_server_slots.argtypes = [vp]
        dp = ctypes.POINTER(ctypes.c_double)
        L.cf_server_replay.argtypes = [vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                       dp, dp, dp, ctypes.POINTER(u64), ctypes.POINTER(ServeStats)]
        L.cf_max_batch.restype = ctypes.c_long
        L.cf_max_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_size_t]
        L.cf_mem_per_prime.restype = ctypes.c_size_t
        L.cf_mem_per_prime.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_size_t]
        L.cf_prime_range.restype = ctypes.c_long
        L.cf_prime_range.argtypes = [ctypes.c_int, ctypes.c_long, ctypes.c_long, ctypes.POINTER(u64)]
        L.cf_need.restype = ctypes.c_long
        L.cf_need.argtypes = [ctypes.c_int] * 4
        L.cf_transform_length.restype = ctypes.c_size_t
        L.cf_transform_length.argtypes = [ctypes.c_size_t]
        L.cf_ntt_primes32.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_uint32)]
        L.cf_engine_log2_length.argtypes = [ctypes.c_int, u64]
        _lib = L
    return _lib


# ---- host logic ----

def prime_range(first, count, prime_bits=50):
    out = (u64 * max(count, 1))()
    k = lib().cf_prime_range(prime_bits, first, count, out)
    if k < 0:
        raise ValueError("bad prime range")
    return [out[i] for i in range(k)]


def need(n, M, d, prime_bits=50):
    return lib().cf_need(n, M, d, prime_bits)


def transform_length(nmax):
    return lib().cf_transform_length(nmax)


def ntt_primes32(count, min_pow2):
    out = (ctypes.c_uint32 * max(count, 1))()
    if lib().cf_ntt_primes32(count, min_pow2, out) != count:
        raise RuntimeError("not enough NTT-friendly primes")
    return list(out)[:count]


def crt_log2(primes, res):
    """log2 of the CRT combination of res modulo primes (host Garner); -inf for 0."""
    k = len(primes)
    return lib().cf_crt_log2((u64 * max(k, 1))(*primes), (u64 * max(k, 1))(*res), k)


def exact_combine(d, n, M, primes, res, decimal=False):
    """The exact columns of one post_gmp.py row for N_bad = CRT(res mod primes), Q = C(M, d):
    a dict with log2_pfail, le (three bools), p_fail, nbad_bits, qn_bits (and nbad, the decimal
    digits, with decimal=True)."""
    k = len(primes)
    out = Exact()
    buf = ctypes.create_string_buffer(20 * k + 2) if decimal else None
    rc = lib().cf_exact_combine(d, n, M, (u64 * max(k, 1))(*primes), (u64 * max(k, 1))(*res), k, ctypes.byref(out),
                                buf, 20 * k + 2 if decimal else 0)
    if rc:
        raise CfError(rc, out.error.decode())
    r = {"log2_pfail": out.log2_pfail.decode(), "le": [bool(x) for x in out.le], "p_fail": out.p_fail.decode(),
         "nbad_bits": out.nbad_bits, "qn_bits": out.qn_bits}
    if decimal:
        r["nbad"] = buf.value.decode()
    return r


def post_need(n, M, d, pbits=31):
    """post_gmp.py's need(n, M) (double precision)."""
    return lib().cf_post_need(n, M, d, pbits)


def mem_per_prime(nmax, Mmax, J=1):
    return lib().cf_mem_per_prime(nmax, Mmax, J)


# ---- the device engine ----

class Engine:
    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        rc = lib().cf_open(device, ctypes.byref(self._h))
        if rc != CF_OK:
            raise CfError(rc, f"cannot open HIP device {device}: {lib().cf_strerror(rc).decode()}")

    def close(self):
        if self._h:
            lib().cf_close(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_name(self):
        buf = ctypes.create_string_buffer(256)
        lib().cf_device_name(self._h, buf, 256)
        return buf.value.decode()

    def set_graphs(self, on=True):
        """Graph mode: each run / run_batch call is one captured hipGraph, replayed while the launch
        plan repeats."""
        lib().cf_set_graphs(self._h, 1 if on else 0)

    def autotune(self, log2_length, n=0):
        """Times every compiled kernel variant of one transform-length class on tables of n rows
        (0: the class's smallest) and keeps the fastest for this engine; returns the table."""
        t = Tune()
        rc = lib().cf_autotune(self._h, log2_length, n, ctypes.byref(t))
        if rc != CF_OK:
            raise CfError(rc, lib().cf_last_error(self._h).decode())
        k = t.variants
        return {"log2_length": t.log2_length, "chosen": t.chosen, "tables": t.tables, "n": t.n,
                "variants": [{"waves_per_transform": t.waves[v], "deferred_flush": bool(t.deferred[v]),
                              "resident_workgroups": t.resident[v], "kernel_ms": round(t.ms[v], 2)} for v in range(k)]}

    def set_variant(self, log2_length, variant):
        rc = lib().cf_set_variant(self._h, log2_length, variant)
        if rc != CF_OK:
            raise CfError(rc, "bad variant")

    def _instances(self, instances):
        k = len(instances)
        inst = (Instance * max(k, 1))()
        flat = []
        for i, (d, model, n, M, primes) in enumerate(instances):
            inst[i] = Instance(d, model, n, M, len(flat), len(primes))
            flat.extend(primes)
        return inst, k, flat

    def autotune_batch(self, instances, sample_tables=0):
        """Times the work-queue kernel's variants on a sample of this batch (instances as for
        run_batch; about sample_tables tables, 0 = all) and keeps the fastest; returns the table."""
        inst, k, flat = self._instances(instances)
        P = (u64 * max(len(flat), 1))(*flat)
        t = BatchTune()
        rc = lib().cf_autotune_batch(self._h, inst, k, P, sample_tables, ctypes.byref(t))
        if rc != CF_OK:
            raise CfError(rc, lib().cf_last_error(self._h).decode())
        return {"chosen": t.chosen, "tables": t.tables,
                "variants": [{"threads": t.threads[v], "workgroups_per_cu": t.per_cu[v], "resident_workgroups": t.resident[v],
                              "kernel_ms": round(t.ms[v], 2)} for v in range(t.variants)]}

    def set_batch_variant(self, variant):
        rc = lib().cf_set_batch_variant(self._h, variant)
        if rc != CF_OK:
            raise CfError(rc, "bad variant")

    def set_batch_sharing(self, on):
        """run_batch with one table per (d, model, prime) serving every instance on it (True, the
        default) or one table per (instance, prime) (False); the residues are the same."""
        rc = lib().cf_set_batch_sharing(self._h, 1 if on else 0)
        if rc != CF_OK:
            raise CfError(rc, "bad engine")

    def run_separate(self, n, m, primes, algo="suffix"):
        """N_bad(n, m) mod p for separate tables of lengths m by the device suffix table (algo
        "suffix") or forward recurrence (algo "forward")."""
        code = {"suffix": CF_ALGO_SUFFIX, "forward": CF_ALGO_FORWARD}[algo]
        np_ = len(primes)
        P = (u64 * max(np_, 1))(*primes)
        res = (u64 * max(np_, 1))()
        mm = (ctypes.c_int32 * max(len(m), 1))(*m)
        rc = lib().cf_run_separate(self._h, code, n, mm, len(m), P, np_, res)
        if rc != CF_OK:
            raise CfError(rc, lib().cf_last_error(self._h).decode())
        return list(res)[:np_]

    def run_algo(self, algo, n, M, primes, d, model=0):
        """N_bad(n, M) mod p for every prime by the device forward (algo "forward", O(n^4)) or cubic
        (algo "cubic", O(n^3)) algorithm."""
        code = {"forward": CF_ALGO_FORWARD, "cubic": CF_ALGO_CUBIC}[algo]
        np_ = len(primes)
        P = (u64 * max(np_, 1))(*primes)
        res = (u64 * max(np_, 1))()
        rc = lib().cf_run_algo(self._h, code, d, model, n, M, P, np_, res)
        if rc != CF_OK:
            raise CfError(rc, lib().cf_last_error(self._h).decode())
        return list(res)[:np_]

    def run(self, queries, primes, d, model=0, mask=None):
        """N_bad(n, M) mod p for every prime and query: returns res[b][j] (0 where mask[b][j] is 0)."""
        J, np_ = len(queries), len(primes)
        n = (ctypes.c_int32 * max(J, 1))(*[q[0] for q in queries])
        M = (ctypes.c_int32 * max(J, 1))(*[q[1] for q in queries])
        P = (u64 * max(np_, 1))(*primes)
        res = (u64 * max(np_ * J, 1))()
        mk = None
        if mask is not None:
            flat = [1 if mask[b][j] else 0 for b in range(np_) for j in range(J)]
            mk = (ctypes.c_uint8 * max(len(flat), 1))(*flat)
        rc = lib().cf_run_fused(self._h, d, model, n, M, J, P, np_, mk, res)
        if rc != CF_OK:
            raise CfError(rc, lib().cf_last_error(self._h).decode())
        return [[res[b * J + j] for j in range(J)] for b in range(np_)]

    def run_batch(self, instances):
        """Uneven batch in one work-queue launch. instances: (d, model, n, M, primes) tuples;
        returns, per instance, N_bad(n, M) mod each of its primes."""
        k = len(instances)
        inst = (Instance * max(k, 1))()
        flat = []
        for i, (d, model, n, M, primes) in enumerate(instances):
            inst[i] = Instance(d, model, n, M, len(flat), len(primes))
            flat.extend(primes)
        P = (u64 * max(len(flat), 1))(*flat)
        res = (u64 * max(len(flat), 1))()
        rc = lib().cf_run_batch(self._h, inst, k, P, res)
        if rc != CF_OK:
            raise CfError(rc, lib().cf_last_error(self._h).decode())
        return [[res[inst[i].prime_off + b] for b in range(inst[i].prime_cnt)] for i in range(k)]

    def stats(self):
        s = Stats()
        lib().cf_last_stats(self._h, ctypes.byref(s))
        return s

    def max_batch(self, nmax, Mmax, J=1):
        return lib().cf_max_batch(self._h, nmax, Mmax, J)


class Tables:
    """A resident table set: suffix tables of every prime for rows 0..nmax, kept in device memory."""

    def __init__(self, engine, d, nmax, Mmax, primes, model=0):
        self.engine, self.primes = engine, list(primes)
        self._h = ctypes.c_void_p()
        P = (u64 * max(len(self.primes), 1))(*self.primes)
        rc = lib().cf_tables_build(engine._h, d, model, nmax, Mmax, P, len(self.primes), ctypes.byref(self._h))
        if rc != CF_OK:
            raise CfError(rc, lib().cf_last_error(engine._h).decode())

    def close(self):
        if self._h:
            lib().cf_tables_free(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def bytes(self):
        return lib().cf_tables_bytes(self._h)

    def need(self, n, M):
        return lib().cf_tables_need(self._h, n, M)

    def query(self, queries, need=None, noalloc=False):
        """res[b][j] = N_bad(n_j, M_j) mod p_b for b < need[j] (None: all primes), else 0. With
        noalloc, through cf_query_noalloc (no device allocation; fixed per-device scratch)."""
        J, np_ = len(queries), len(self.primes)
        n = (ctypes.c_int32 * max(J, 1))(*[q[0] for q in queries])
        M = (ctypes.c_int32 * max(J, 1))(*[q[1] for q in queries])
        nd = None if need is None else (ctypes.c_int64 * max(J, 1))(*need)
        res = (u64 * max(np_ * J, 1))()
        rc = (lib().cf_query_noalloc if noalloc else lib().cf_query)(self._h, n, M, J, nd, res)
        if rc != CF_OK:
            raise CfError(rc, lib().cf_last_error(self.engine._h).decode())
        return [[res[b * J + j] for j in range(J)] for b in range(np_)]


class Server:
    """The serving loop over a resident table set (continuous batching, host CRT)."""

    def __init__(self, tables, slots=2, max_batch=64, crt_threads=8):
        self.tables = tables
        self._h = ctypes.c_void_p()
        rc = lib().cf_server_open(tables._h, slots, max_batch, crt_threads, ctypes.byref(self._h))
        if rc != CF_OK:
            raise CfError(rc, lib().cf_last_error(tables.engine._h).decode())

    def close(self):
        if self._h:
            lib().cf_server_close(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def slots(self):
        return lib().cf_server_slots(self._h)

    def replay(self, n, M, arrival, residues=False):
        """Replays requests in real time. Returns (latency_ms, log2_pfail, residues or None, stats)."""
        k, np_ = len(n), len(self.tables.primes)
        N = (ctypes.c_int32 * max(k, 1))(*[int(x) for x in n])
        MM = (ctypes.c_int32 * max(k, 1))(*[int(x) for x in M])
        A = (ctypes.c_double * max(k, 1))(*[float(x) for x in arrival])
        lat = (ctypes.c_double * max(k, 1))()
        l2 = (ctypes.c_double * max(k, 1))()
        res = (u64 * max(k * np_, 1))() if residues else None
        st = ServeStats()
        rc = lib().cf_server_replay(self._h, k, N, MM, A, lat, l2, res, ctypes.byref(st))
        if rc != CF_OK:
            raise CfError(rc, lib().cf_last_error(self.tables.engine._h).decode())
        rr = [[res[i * np_ + b] for b in range(np_)] for i in range(k)] if residues else None
        return list(lat)[:k], list(l2)[:k], rr, st
