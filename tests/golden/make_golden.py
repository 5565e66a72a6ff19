#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ by running the reference's own programs.

The programs are the reference's C++ sources compiled unmodified by `make -C oracle ref`
(into oracle/_ref/, never committed). Nothing from the reference is imported or copied: only
the programs' outputs are kept, as data.

    python tests/golden/make_golden.py            # everything except the n = 65536 prime
    python tests/golden/make_golden.py --big      # also n = 65536 (about 32 GB RAM, ~3 min)
    python tests/golden/make_golden.py --only uneven   # just the uneven-batch fixture
    python tests/golden/make_golden.py --only wide     # just the n > 16384 fixture (~2 min, 8 cores)
    python tests/golden/make_golden.py --only algos    # just the --algo forward|cubic fixture
    python tests/golden/make_golden.py --only separate # just the separate-tables fixture
    python tests/golden/make_golden.py --only headline   # n = 16384, primes 0..511 (~8 min, 8 cores)
    python tests/golden/make_golden.py --only widelaunch # n = 32768, primes 0..15 (one 16-task launch)
    python tests/golden/make_golden.py --only serving    # the serving table set's 41 queries, 1536 primes
    python tests/golden/make_golden.py --only uneven_big # a sample of the uneven batch with n > 512
    python tests/golden/make_golden.py --only big        # n = 65536: primes 1..3, and 126 queries on 0..1
    python tests/golden/make_golden.py --only fast62     # the reference's -DFP_RB=64 build (62-bit primes)
    python tests/golden/make_golden.py --only fast62wide # the same build on tables beyond 16384 rows
    python tests/golden/make_golden.py --only fast62big  # and on one table of 65537 rows (2^17)
    python tests/golden/make_golden.py --only big17      # n = 70000 (transforms of 2^17), prime 0 (~39 GB, ~4 min)
"""
import json
import os
import re
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(REPO, "oracle", "_ref")
FAST = os.path.join(REF, "cuckoo_fast_ifma")
if not os.path.exists(FAST) or "avx512ifma" not in open("/proc/cpuinfo").read():
    FAST = os.path.join(REF, "cuckoo_fast_cpu")
FAIL = os.path.join(REF, "cuckoo_fail")
CHECK = os.path.join(REF, "check")


def run(argv, cwd=None):
    r = subprocess.run(argv, capture_output=True, text=True, cwd=cwd)
    return r.returncode, r.stdout, r.stderr


def strip_fast_stderr(s):
    # progress lines carry wall-clock times; keep only their structure
    return re.sub(r"build done at [0-9.]+ s, queries at [0-9.]+ s", "build done at T s, queries at T s", s)


def fast_run(d, queries, first, count, extra=(), out_init=None):
    with tempfile.TemporaryDirectory() as td:
        qf = os.path.join(td, "q.txt")
        with open(qf, "w") as f:
            f.write("".join(f"{n} {M}\n" for n, M in queries))
        argv = [FAST, str(d), qf, str(first), str(count), *extra]
        outf = None
        if out_init is not None:
            outf = os.path.join(td, "out.txt")
            with open(outf, "w") as f:
                f.write(out_init)
            argv += ["--out", outf]
        rc, so, se = run(argv)
        rec = {"d": d, "queries": queries, "first": first, "count": count, "extra": list(extra),
               "rc": rc, "stdout": so, "stderr": strip_fast_stderr(se)}
        if outf is not None:
            rec["out_init"] = out_init
            rec["out_final"] = open(outf).read()
        return rec


def fast_cases():
    """Residue lines of the 50-bit engine over many shapes: every n class around the powers of
    two (the c[nmax] -> c[0] wrap when L = nmax), d = 1..5, multi-query files with different
    need, and prime ranges that start past 0."""
    cases = []
    for d in (1, 2, 3, 4, 5):
        for n in (8, 9, 15, 16, 17, 31, 32, 33, 64, 100, 127, 128, 129):
            cases.append((d, [(n, n + 1), (n, 2 * n + 3)], 0, 2))
    cases.append((3, [(37, 80), (64, 65), (200, 400), (512, 634)], 0, 3))
    cases.append((3, [(1000, 1500), (777, 2000), (37, 38)], 0, 2))
    cases.append((2, [(255, 256), (256, 257), (257, 258), (1024, 1025), (1023, 3000)], 5, 3))
    cases.append((4, [(2048, 4096)], 0, 2))
    cases.append((3, [(4096, 8192)], 0, 2))
    # query lists whose need straddles the range: lines only for i < need(n, M)
    cases.append((3, [(10, 11), (8, 200), (40, 41), (12, 13)], 0, 12))
    cases.append((2, [(30, 31), (9, 10)], 3, 9))
    # full prime sets, for exact reconstruction by CRT
    cases.append((3, [(64, 65)], 0, 60))
    cases.append((3, [(200, 400)], 0, 97))
    cases.append((1, [(12, 20)], 0, 8))
    cases.append((2, [(6, 7), (8, 9)], 0, 3))
    with ThreadPoolExecutor(max_workers=os.cpu_count()) as ex:
        return list(ex.map(lambda c: fast_run(*c), cases))


def fast_errors():
    """Exit codes and messages of the argument/input checks (cuckoo_fast.cu:654-695, 724-732)."""
    recs = []
    recs.append(fast_run(3, [(7, 20)], 0, 2))  # max n < 8
    recs.append(fast_run(3, [(20, 20)], 0, 2))  # M <= n
    recs.append(fast_run(3, [(20, 40)], 0, 2, extra=("--bogus",)))
    recs.append(fast_run(3, [(20, 40)], 500, 2))  # range past need: nothing to do
    recs.append(fast_run(0, [(20, 40)], 0, 2))  # d < 1
    rc, so, se = run([FAST, "3"])
    recs.append({"argv": ["3"], "rc": rc, "stdout": so, "stderr": se})
    rc, so, se = run([FAST, "3", "/nonexistent/q.txt", "0", "1"])
    recs.append({"argv": ["3", "/nonexistent/q.txt", "0", "1"], "rc": rc, "stdout": so, "stderr": se})
    return recs


def fast_resume():
    """--out resume semantics (cuckoo_fast.cu:697-727): complete, partial, foreign and truncated
    records in a pre-existing file."""
    recs = []
    q = [(20, 40), (12, 13)]
    full = fast_run(3, q, 0, 6)["stdout"].splitlines(keepends=True)
    recs.append(fast_run(3, q, 0, 6, out_init=""))  # fresh file
    recs.append(fast_run(3, q, 0, 6, out_init="".join(full[:4])))  # two primes complete
    recs.append(fast_run(3, q, 0, 6, out_init="".join(full[:3])))  # prime 1 partial
    recs.append(fast_run(3, q, 0, 6, out_init="".join(full)))  # all done
    recs.append(fast_run(3, q, 0, 6, out_init="".join(full[:4]) + "2 1125899846025217 20"))  # truncated tail
    recs.append(fast_run(3, q, 0, 6, out_init="0 5 9 9 9\n0 5 9 9 9\n1 5 9 9 9\n"))  # foreign (n, M)
    recs.append(fast_run(3, q, 2, 3, out_init="".join(full[:6])))  # range starting past done primes
    return recs


def fast_big(n, M, first, count):
    with ThreadPoolExecutor(max_workers=min(count, os.cpu_count())) as ex:
        recs = list(ex.map(lambda i: fast_run(3, [(n, M)], i, 1), range(first, first + count)))
    return {"d": 3, "queries": [[n, M]], "first": first, "count": count,
            "stdout": "".join(r["stdout"] for r in recs)}


def fast_wide():
    """Tables longer than 16384 rows (transforms of 2^15 and 2^16): a few primes each."""
    cases = [(3, [(16385, 16386)], 0, 2), (3, [(20000, 40000), (17000, 20000)], 0, 3), (2, [(24000, 30000)], 5, 2),
             (3, [(32768, 65536)], 0, 2), (4, [(40000, 41000)], 1, 1)]
    jobs = [(d, qs, i) for d, qs, first, count in cases for i in range(first, first + count)]
    with ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count())) as ex:
        recs = list(ex.map(lambda j: fast_run(j[0], j[1], j[2], 1), jobs))
    return [{"d": d, "queries": [list(q) for q in qs], "prime": i, "stdout": r["stdout"]} for (d, qs, i), r in zip(jobs, recs)]


def fail_cases():
    """The reference's 32-bit CLI (src/cuckoo_fail.cpp), ntt algorithm."""
    args = [
        ["single", "-n", "200", "-M", "400", "-d", "3", "--algo", "ntt", "--exact"],
        ["single", "-n", "20", "-M", "40", "-d", "3"],
        ["single", "-n", "20", "-M", "40", "-d", "3", "--exact"],
        ["single", "-n", "6", "-M", "7", "-d", "2", "--exact"],
        ["single", "-n", "1", "-M", "2", "-d", "1", "--exact"],
        ["single", "-n", "5", "-M", "6", "-d", "1", "--exact"],
        ["single", "-n", "12", "-M", "20", "-d", "1", "--exact"],
        ["single", "-n", "31", "-M", "32", "-d", "2", "--exact"],
        ["single", "-n", "60", "-M", "100", "-d", "3", "--exact"],
        ["single", "-n", "64", "-M", "65", "-d", "3", "--exact"],
        ["single", "-n", "100", "-M", "150", "-d", "4", "--exact"],
        ["single", "-n", "128", "-M", "1000", "-d", "2", "--exact"],
        ["single", "-n", "300", "-M", "301", "-d", "5"],
        ["single", "-n", "512", "-M", "634", "-d", "3", "--exact"],
        ["single", "-n", "30", "-M", "50", "-d", "2", "--replacement", "--exact"],
        ["single", "-n", "25", "-M", "40", "-d", "3", "--replacement", "--exact"],
        ["single", "-n", "200", "-M", "400", "-d", "3", "--mod", "998244353"],
        ["single", "-n", "200", "-M", "400", "-d", "3", "--mod", "1004535809"],
        ["single", "-n", "50", "-M", "51", "-d", "3", "--mod", "469762049"],
        ["single", "-n", "10", "-M", "10", "-d", "3"],
        ["single", "-n", "10", "-M", "20", "-d", "3", "--algo", "bogus"],
        ["single", "-n", "10", "-M", "20", "-d", "3", "--mod", "1000"],
        ["bogus"],
        [],
    ]
    with ThreadPoolExecutor(max_workers=os.cpu_count()) as ex:
        outs = list(ex.map(lambda a: run([FAIL, *a]), args))
    recs = []
    for a, (rc, so, se) in zip(args, outs):
        so = re.sub(r"time: [0-9.]+ s", "time: T s", so)
        so = re.sub(r"\([0-9.]+ s\)", "(T s)", so)
        recs.append({"args": a, "rc": rc, "stdout": so, "stderr": se})
    return recs


def fail_algo_cases():
    """The reference's 32-bit CLI with its O(n^4) forward and O(n^3) cubic algorithms."""
    args = []
    for algo in ("forward", "cubic"):
        args += [
            ["single", "-n", "20", "-M", "40", "-d", "3", "--algo", algo, "--exact"],
            ["single", "-n", "6", "-M", "7", "-d", "2", "--algo", algo, "--exact"],
            ["single", "-n", "1", "-M", "2", "-d", "1", "--algo", algo, "--exact"],
            ["single", "-n", "12", "-M", "20", "-d", "1", "--algo", algo, "--exact"],
            ["single", "-n", "31", "-M", "32", "-d", "2", "--algo", algo, "--exact"],
            ["single", "-n", "60", "-M", "100", "-d", "3", "--algo", algo, "--exact"],
            ["single", "-n", "64", "-M", "65", "-d", "3", "--algo", algo],
            ["single", "-n", "30", "-M", "50", "-d", "2", "--replacement", "--algo", algo, "--exact"],
            ["single", "-n", "25", "-M", "40", "-d", "3", "--replacement", "--algo", algo, "--exact"],
            ["single", "-n", "40", "-M", "41", "-d", "5", "--algo", algo, "--mod", "998244353"],
        ]
    args += [
        ["single", "-n", "200", "-M", "400", "-d", "3", "--algo", "cubic", "--exact"],
        ["single", "-n", "100", "-M", "150", "-d", "4", "--algo", "cubic"],
        ["single", "-n", "128", "-M", "1000", "-d", "2", "--replacement", "--algo", "cubic", "--mod", "1004535809"],
        ["single", "-n", "80", "-M", "90", "-d", "4", "--algo", "forward"],
    ]
    with ThreadPoolExecutor(max_workers=os.cpu_count()) as ex:
        outs = list(ex.map(lambda a: run([FAIL, *a]), args))
    recs = []
    for a, (rc, so, se) in zip(args, outs):
        so = re.sub(r"time: [0-9.]+ s", "time: T s", so)
        so = re.sub(r"\([0-9.]+ s\)", "(T s)", so)
        recs.append({"args": a, "rc": rc, "stdout": so, "stderr": se})
    return recs


def fail_separate_cases():
    """The reference's 32-bit CLI in separate-tables mode (suffix and forward algorithms)."""
    args = [
        ["separate", "-n", "10", "-m", "6,6", "--exact"],
        ["separate", "-n", "10", "-m", "6,6", "--algo", "forward", "--exact"],
        ["separate", "-n", "30", "-m", "20,25", "--exact"],
        ["separate", "-n", "40", "-m", "30,30"],
        ["separate", "-n", "12", "-m", "5,5,5", "--exact"],
        ["separate", "-n", "15", "-m", "6,7,8", "--algo", "forward"],
        ["separate", "-n", "20", "-m", "25", "--exact"],
        ["separate", "-n", "8", "-m", "3,3,3,3", "--exact"],
        ["separate", "-n", "60", "-m", "40,40", "--mod", "998244353"],
        ["separate", "-n", "25", "-m", "9,9,9", "--algo", "suffix", "--exact"],
        ["separate", "-n", "3", "-m", "1,3", "--exact"],
        ["separate", "-n", "100", "-m", "70,70", "--exact"],
        ["separate", "-n", "40", "-m", "20,20,20", "--exact"],
        ["separate", "-n", "10", "-m", "4,5"],
        ["separate", "-n", "5", "-m", "3,0"],
        ["separate", "-n", "5", "-m", "3,4", "--algo", "cubic"],
    ]
    with ThreadPoolExecutor(max_workers=os.cpu_count()) as ex:
        outs = list(ex.map(lambda a: run([FAIL, *a]), args))
    recs = []
    for a, (rc, so, se) in zip(args, outs):
        so = re.sub(r"time: [0-9.]+ s", "time: T s", so)
        so = re.sub(r"\([0-9.]+ s\)", "(T s)", so)
        recs.append({"args": a, "rc": rc, "stdout": so, "stderr": se})
    return recs


def uneven_golden(n_max=512):
    """The uneven-batch instance list (seed 0) and, for its instances with n <= n_max, the
    reference's residues over the instance's whole prime set (one cuckoo_fast run per instance,
    FIRST = 0, COUNT = need): the first three lines, and the CRT-combined exact N_bad as its
    bit length and the SHA-256 of its little-endian bytes."""
    import hashlib
    sys.path.insert(0, os.path.join(REPO, "cuckoo-failure_amd"))
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    import workloads

    inst = workloads.uneven_batch(0, 256)
    small = [(i, n, d, M) for i, (n, d, M) in enumerate(inst) if n <= n_max]

    def one(a):
        i, n, d, M = a
        k = oracle.need(n, M, d)
        rec = fast_run(d, [(n, M)], 0, k)
        assert rec["rc"] == 0, rec["stderr"]
        lines = rec["stdout"].splitlines()
        assert len(lines) == k
        pairs = [(int(t.split()[1]), int(t.split()[4])) for t in lines]
        N, _ = oracle.crt(pairs)
        raw = N.to_bytes((N.bit_length() + 7) // 8, "little")
        return {"index": i, "n": n, "d": d, "M": M, "need": k, "head": lines[:3],
                "nbad_bits": N.bit_length(), "nbad_sha256": hashlib.sha256(raw).hexdigest()}

    with ThreadPoolExecutor(max_workers=os.cpu_count()) as ex:
        gold = list(ex.map(one, small))
    return {"seed": 0, "instances": [list(x) for x in inst], "golden": gold}


def _lines_by_prime(stdout):
    """'i p n M r' lines -> {i: [(p, n, M, r), ...]} in output order."""
    out = {}
    for line in stdout.splitlines():
        t = line.split()
        if len(t) == 5:
            out.setdefault(int(t[0]), []).append(tuple(int(x) for x in t[1:]))
    return out


# Successful reference runs are cached by their arguments (the long fixtures take tens of minutes
# of CPU, so an interrupted generation resumes where it stopped). Set CF_GOLDEN_CACHE="" to disable.
CACHE = os.environ.get("CF_GOLDEN_CACHE", os.path.join(tempfile.gettempdir(), "cf_golden_cache"))


def cached_run(d, queries, first, count):
    """stdout of a successful `cuckoo_fast d queries first count` run, from the cache if present."""
    import hashlib
    key = hashlib.sha256(json.dumps([os.path.basename(FAST), d, [list(q) for q in queries], first, count]).encode()).hexdigest()
    path = os.path.join(CACHE, key[:2], key) if CACHE else None
    if path and os.path.exists(path):
        with open(path) as f:
            return f.read()
    rec = fast_run(d, queries, first, count)
    assert rec["rc"] == 0, rec["stderr"]
    if path:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path + ".tmp", "w") as f:
            f.write(rec["stdout"])
        os.replace(path + ".tmp", path)
    return rec["stdout"]


def fast_slice(d, queries, first, count, workers):
    """One reference process per prime over [first, first + count) (the host build runs one prime
    per batch anyway), `workers` at a time; returns the concatenated stdout in prime order."""
    with ThreadPoolExecutor(max_workers=workers) as ex:
        outs = list(ex.map(lambda i: cached_run(d, queries, i, 1), range(first, first + count)))
    return "".join(outs)


def headline_slice():
    """The whole 512-prime slice of the n = 16384, M = 32768, d = 3 workload (about 8 minutes of
    the IFMA build on 8 cores, 2 GB per process)."""
    return {"d": 3, "queries": [[16384, 32768]], "first": 0, "count": 512,
            "stdout": fast_slice(3, [(16384, 32768)], 0, 512, os.cpu_count())}


def wide_launch_slice():
    """16 primes of n = 32768, M = 65536, d = 3: one full 16-task launch of the 2^16 kernel
    (8.6 GB per process, so at most 5 at once in 62 GB)."""
    return {"d": 3, "queries": [[32768, 65536]], "first": 0, "count": 16,
            "stdout": fast_slice(3, [(32768, 65536)], 0, 16, 5)}


def queries_126():
    """126 queries on one n = 65536 table (the query count of the reference's published timing,
    fast/README.md:80-86): 63 at n = 65536 with M = 65537 + 1040 k, and 63 at n = 1040 (k + 1)
    with M = 2 n, k = 0..62."""
    qs = [(65536, 65537 + 1040 * k) for k in range(63)]
    qs += [(1040 * (k + 1), 2 * 1040 * (k + 1)) for k in range(63)]
    return qs


def big_tables(primes_131072=(1, 2, 3), primes_126=(0, 1)):
    """n = 65536 tables, one process at a time (32 GB each): more primes of (65536, 131072), and
    the 126-query file on primes 0 and 1 (about 5 minutes per prime)."""
    out = {"d": 3, "queries": [[65536, 131072]], "primes": list(primes_131072),
           "stdout": "".join(cached_run(3, [(65536, 131072)], i, 1) for i in primes_131072)}
    qs = queries_126()
    out126 = {"d": 3, "queries": [list(q) for q in qs], "primes": list(primes_126),
              "stdout": "".join(cached_run(3, qs, i, 1) for i in primes_126)}
    return out, out126


def serving_golden(n_queries=40, count=1536):
    """The serving workload's table set (d = 3, n_max = 2048, M_max = 8192, primes 0..1535): the
    reference's residues for the first 40 requests of the seed-1 stream (100k requests) and for
    (2048, 8192), the largest shape the set answers. One reference run over the prime slice with
    all 41 queries in one file (lines for i < need(n, M), file order). Per query: need, the first
    residues, the SHA-256 of the residue list (decimal, one per line, primes ascending), and the
    CRT-combined exact N_bad's bit length and SHA-256."""
    import hashlib
    sys.path.insert(0, os.path.join(REPO, "cuckoo-failure_amd"))
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    import workloads

    n, M, _ = workloads.serving_queries(1, 100000, 32, 2048)
    qs = [(int(n[i]), int(M[i])) for i in range(n_queries)] + [(2048, 8192)]
    # one process per 48-prime sub-range (each builds its own table per prime)
    ranges = [(f, min(48, count - f)) for f in range(0, count, 48)]
    with ThreadPoolExecutor(max_workers=os.cpu_count()) as ex:
        outs = list(ex.map(lambda r: cached_run(3, qs, r[0], r[1]), ranges))
    by_prime = {}
    for out in outs:
        by_prime.update(_lines_by_prime(out))
    gold = []
    for j, (qn, qM) in enumerate(qs):
        k = oracle.need(qn, qM, 3)
        rows = []
        for i in range(min(k, count)):
            hits = [x for x in by_prime.get(i, []) if x[1] == qn and x[2] == qM]
            assert len(hits) >= 1, (i, qn, qM)
            rows.append((hits[0][0], hits[0][3]))
        text = "".join(f"{r}\n" for _, r in rows)
        N, _ = oracle.crt(rows)
        raw = N.to_bytes((N.bit_length() + 7) // 8, "little")
        gold.append({"n": qn, "M": qM, "need": k, "head": [r for _, r in rows[:3]],
                     "residues_sha256": hashlib.sha256(text.encode()).hexdigest(),
                     "nbad_bits": N.bit_length(), "nbad_sha256": hashlib.sha256(raw).hexdigest()})
    return {"d": 3, "nmax": 2048, "Mmax": 8192, "primes": count, "stream_seed": 1, "stream_count": 100000,
            "queries": gold}


def uneven_sample_golden():
    """The reference's residues for a sample of the uneven batch's instances with n > 512: every
    5th of them in order of n, and the three largest (one cuckoo_fast run per instance over its
    whole prime set, 8 at once). Same record as uneven_golden()."""
    import hashlib
    sys.path.insert(0, os.path.join(REPO, "cuckoo-failure_amd"))
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    import workloads

    inst = workloads.uneven_batch(0, 256)
    big = sorted((n, i) for i, (n, d, M) in enumerate(inst) if n > 512)
    pick = sorted(set([i for _, i in big[::5]] + [i for _, i in big[-3:]]))

    def one(i):
        n, d, M = inst[i]
        k = oracle.need(n, M, d)
        stdout = fast_slice(d, [(n, M)], 0, k, 1)
        lines = stdout.splitlines()
        assert len(lines) == k
        pairs = [(int(t.split()[1]), int(t.split()[4])) for t in lines]
        N, _ = oracle.crt(pairs)
        raw = N.to_bytes((N.bit_length() + 7) // 8, "little")
        return {"index": i, "n": n, "d": d, "M": M, "need": k, "head": lines[:3],
                "nbad_bits": N.bit_length(), "nbad_sha256": hashlib.sha256(raw).hexdigest()}

    # one instance per worker: each runs its need() primes in one process after the other
    with ThreadPoolExecutor(max_workers=os.cpu_count()) as ex:
        gold = list(ex.map(one, pick))
    return {"seed": 0, "golden": gold}


def fast62_cases():
    """The reference's -DFP_RB=64 build (62-bit primes c 2^20 + 1 below 2^62, Montgomery R = 2^64):
    residue lines over several shapes, the whole need() set of (200, 400, 3) and (64, 65, 3) for
    exact reconstruction, a query file straddling need(), and one prime at n = 16384."""
    fast62 = os.path.join(REF, "cuckoo_fast62")
    cases = []
    for d in (1, 2, 3, 4, 5):
        for n in (8, 15, 16, 17, 100, 129):
            cases.append((d, [(n, n + 1), (n, 2 * n + 3)], 0, 2))
    cases.append((3, [(37, 80), (64, 65), (200, 400), (512, 634)], 0, 3))
    cases.append((2, [(1024, 1025), (1023, 3000), (255, 256)], 4, 2))
    cases.append((3, [(4096, 8192)], 0, 1))
    cases.append((3, [(10, 11), (8, 200), (40, 41), (12, 13)], 0, 12))
    cases.append((3, [(64, 65)], 0, 60))
    cases.append((3, [(200, 400)], 0, 97))
    cases.append((3, [(16384, 32768)], 0, 1))

    def one(c):
        d, qs, first, count = c
        with tempfile.TemporaryDirectory() as td:
            qf = os.path.join(td, "q.txt")
            with open(qf, "w") as f:
                f.write("".join(f"{n} {M}\n" for n, M in qs))
            rc, so, se = run([fast62, str(d), qf, str(first), str(count)])
        assert rc == 0, se
        return {"d": d, "queries": qs, "first": first, "count": count, "stdout": so}

    with ThreadPoolExecutor(max_workers=os.cpu_count()) as ex:
        return list(ex.map(one, cases))


def fast62_wide_cases():
    """The reference's -DFP_RB=64 build on tables of more than 16384 rows (transforms of 2^15 and
    2^16), one prime per process: several queries per table, d = 1 to 5."""
    fast62 = os.path.join(REF, "cuckoo_fast62")
    cases = [(3, [(16385, 16386), (20000, 40000), (17000, 20000)], 0, 1),
             (3, [(16385, 16386), (20000, 40000), (17000, 20000)], 1, 1),
             (2, [(16500, 17000)], 2, 1),
             (4, [(16400, 16401)], 3, 1),
             (3, [(32768, 65536)], 0, 1),
             (3, [(40000, 80000)], 1, 1),
             (1, [(16390, 16391), (16400, 20000)], 0, 1),
             (5, [(16400, 16500)], 3, 1)]

    def one(c):
        d, qs, first, count = c
        with tempfile.TemporaryDirectory() as td:
            qf = os.path.join(td, "q.txt")
            with open(qf, "w") as f:
                f.write("".join(f"{n} {M}\n" for n, M in qs))
            rc, so, se = run([fast62, str(d), qf, str(first), str(count)])
        assert rc == 0, se
        return {"d": d, "queries": qs, "first": first, "count": count, "stdout": so}

    with ThreadPoolExecutor(max_workers=len(cases)) as ex:
        return list(ex.map(one, cases))


def _qfile(qs):
    """A query file in a temporary directory that lives until the process ends."""
    td = tempfile.mkdtemp()
    qf = os.path.join(td, "q.txt")
    with open(qf, "w") as f:
        f.write("".join(f"{n} {M}\n" for n, M in qs))
    return qf


def _write(name, data):
    """Writes one fixture file after its data is complete (an interrupted run leaves the old file)."""
    tmp = os.path.join(HERE, name + ".tmp")
    with open(tmp, "w") as f:
        json.dump(data, f, indent=0)
    os.replace(tmp, os.path.join(HERE, name))


def main():
    os.makedirs(HERE, exist_ok=True)
    only = sys.argv[sys.argv.index("--only") + 1] if "--only" in sys.argv else None
    if only == "headline":
        _write("fast_16384.json", headline_slice())
        return
    if only == "widelaunch":
        _write("fast_32768_launch.json", wide_launch_slice())
        return
    if only == "serving":
        _write("serving_2048.json", serving_golden())
        return
    if only == "uneven_big":
        _write("uneven_batch_big.json", uneven_sample_golden())
        return
    if only == "big":
        more, q126 = big_tables()
        _write("fast_65536_more.json", more)
        _write("fast_65536_q126.json", q126)
        return
    if only == "uneven":
        _write("uneven_batch.json", uneven_golden())
        return
    if only == "big17":
        _write("fast_70000.json", {"d": 3, "queries": [[70000, 140000], [66000, 70001]], "first": 0, "count": 1,
                                   "stdout": cached_run(3, [(70000, 140000), (66000, 70001)], 0, 1)})
        return
    if only == "fast62":
        _write("fast62_cases.json", fast62_cases())
        return
    if only == "fast62wide":
        _write("fast62_wide.json", fast62_wide_cases())
        return
    if only == "fast62big":  # one table of 65537 rows: a transform of 2^17
        _write("fast62_65537.json", {"d": 3, "queries": [[65537, 65540], [40000, 80000]], "first": 2, "count": 1,
                                     "stdout": run([os.path.join(REF, "cuckoo_fast62"), "3", _qfile([(65537, 65540), (40000, 80000)]),
                                                    "2", "1"])[1]})
        return
    if only == "separate":
        _write("fail_separate_cases.json", fail_separate_cases())
        return
    if only == "algos":
        _write("fail_algo_cases.json", fail_algo_cases())
        return
    if only == "wide":
        _write("fast_wide.json", fast_wide())
        return
    rc, so, _ = run([CHECK])
    assert rc == 0, "reference check program failed"
    with open(os.path.join(HERE, "check_ref.txt"), "w") as f:
        f.write(re.sub(r"\([0-9.]+ s\)", "(T s)", so))
    _write("fast_cases.json", fast_cases())
    _write("fast_errors.json", fast_errors())
    _write("fast_resume.json", fast_resume())
    _write("fail_cases.json", fail_cases())
    _write("fail_algo_cases.json", fail_algo_cases())
    _write("fail_separate_cases.json", fail_separate_cases())
    _write("fast_16384.json", headline_slice())
    _write("uneven_batch.json", uneven_golden())
    _write("fast_wide.json", fast_wide())
    _write("fast_32768_launch.json", wide_launch_slice())
    _write("serving_2048.json", serving_golden())
    _write("uneven_batch_big.json", uneven_sample_golden())
    if "--big" in sys.argv:
        _write("fast_65536.json", fast_big(65536, 131072, 0, 1))
        more, q126 = big_tables()
        _write("fast_65536_more.json", more)
        _write("fast_65536_q126.json", q126)


if __name__ == "__main__":
    main()
