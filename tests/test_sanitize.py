"""The engine's host code built with AddressSanitizer and UndefinedBehaviorSanitizer
(tests/sanitize): the bignum, the CRT digits and log2, the exact combiner, the prime lists, need(),
the tile-slot plan and the per-prime constants, each answer compared with Python's exact integers;
and the two command-line drivers on their argument, input and resume paths (which end before any
device call), against the reference's recorded outputs. Any sanitizer report fails the run."""
import math
import os
import random
import subprocess
import tempfile

import pytest

import oracle
from golden_data import load

HERE = os.path.dirname(os.path.abspath(__file__))
SAN = os.path.join(HERE, "sanitize")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")


@pytest.fixture(scope="module")
def built():
    subprocess.run(["make", "-s", "-C", SAN], check=True)
    return SAN


def host(built, cmds):
    r = subprocess.run([os.path.join(built, "host_check")], input="\n".join(cmds) + "\n", capture_output=True, text=True,
                       env=ENV, timeout=300)
    assert r.returncode == 0 and "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-3000:]
    return r.stdout.splitlines()


def test_bignum_crt_and_exact_combine(built):
    rng = random.Random(3)
    cmds, want = [], []
    for _ in range(12):
        x = rng.getrandbits(rng.randint(1, 600))
        k = 22
        cmds.append(f"crt32 {k} {x}")
        want.append(("crt", x))
        a, e = rng.getrandbits(rng.randint(1, 200)), rng.randint(0, 40)
        cmds.append(f"pow {a} {e}")
        want.append(("int", a ** e))
        b = rng.getrandbits(rng.randint(1, 300)) + 1
        c = rng.getrandbits(rng.randint(1, 700))
        cmds.append(f"div {c} {b}")
        want.append(("int", c // b))
    # exact combine: a value below Q^n over the whole need() set of 50-bit primes
    for d, n, M in ((3, 64, 65), (2, 30, 31), (3, 200, 400)):
        k = oracle.need(n, M, d)
        Qn = math.comb(M, d) ** n
        x = rng.randrange(Qn)
        cmds.append(f"exact {d} {n} {M} {k} {x}")
        want.append(("exact", (x, Qn)))
    out = host(built, cmds)
    assert len(out) == len(want)
    for line, (kind, w) in zip(out, want):
        t = line.split()
        if kind == "crt":
            assert int(t[0]) == w
            if w:
                assert abs(float(t[1]) - math.log2(w)) < 1e-6  # printed to 6 decimals
        elif kind == "int":
            assert int(t[0]) == w
        else:
            x, Qn = w
            assert int(t[0]) == x and int(t[1]) == x.bit_length()
            if x:
                assert abs(float(t[2]) - (math.log2(x) - math.log2(Qn))) < 1e-9
            assert [int(v) for v in t[4:7]] == [int(x * 2 ** lam <= Qn) for lam in (40, 64, 128)]


def test_prime_lists_need_slots_constants(built):
    cmds = ["primes 50 0 6", "primes 62 3 4", "need 16384 32768 3 50", "need 200 400 3 62", "need 65536 131072 3 50"]
    cmds += [f"slots {nm}" for nm in (8, 17, 255, 256, 257, 4096, 16384, 40000, 65536)]
    P = oracle.prime_range(0, 2)
    cmds += [f"consts {P[0]} {lg}" for lg in (10, 11, 12, 13, 14, 16, 17)] + [f"consts {P[1]} 16"]
    out = host(built, cmds)
    assert [int(x) for x in out[0].split()] == oracle.prime_range(0, 6)
    assert [int(x) for x in out[1].split()] == oracle.prime_range(3, 4, 62)
    assert int(out[2]) == oracle.need(16384, 32768, 3) and int(out[3]) == oracle.need(200, 400, 3, 62)
    assert int(out[4]) == oracle.need(65536, 131072, 3)
    for line in out[5:14]:
        pool, nseg, nb, segs = (int(v) for v in line.split())
        assert pool > 0 and nseg > 0 and segs == nb * nseg
    assert out[14:] == ["1"] * 8


@pytest.mark.parametrize("idx", range(7))
def test_fast_driver_argument_paths(built, idx):
    rec = load("fast_errors.json")[idx]
    with tempfile.TemporaryDirectory() as td:
        if "argv" in rec:
            argv = rec["argv"]
        else:
            qf = os.path.join(td, "q.txt")
            with open(qf, "w") as f:
                f.write("".join(f"{n} {M}\n" for n, M in rec["queries"]))
            argv = [str(rec["d"]), qf, str(rec["first"]), str(rec["count"]), *rec["extra"]]
        r = subprocess.run([os.path.join(built, "cuckoo_fast_san"), *argv], capture_output=True, text=True, env=ENV,
                           timeout=60)
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
    assert r.returncode == rec["rc"] and r.stderr.strip() == rec["stderr"].strip()


def test_fast_driver_resume_all_done(built):
    rec = load("fast_resume.json")[3]
    with tempfile.TemporaryDirectory() as td:
        qf, of = os.path.join(td, "q.txt"), os.path.join(td, "out.txt")
        with open(qf, "w") as f:
            f.write("".join(f"{n} {M}\n" for n, M in rec["queries"]))
        with open(of, "w") as f:
            f.write(rec["out_init"])
        r = subprocess.run([os.path.join(built, "cuckoo_fast_san"), str(rec["d"]), qf, str(rec["first"]), str(rec["count"]),
                            "--out", of], capture_output=True, text=True, env=ENV, timeout=60)
        final = open(of).read()
    assert "Sanitizer" not in r.stderr and r.returncode == 0 and final == rec["out_final"]


@pytest.mark.parametrize("idx", [19, 20, 21, 22, 23])
def test_fail_driver_argument_paths(built, idx):
    rec = load("fail_cases.json")[idx]
    r = subprocess.run([os.path.join(built, "cuckoo_fail_san"), *rec["args"]], capture_output=True, text=True, env=ENV,
                       timeout=60)
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
    assert r.returncode == rec["rc"] == 2 and r.stderr.splitlines()[0] == rec["stderr"].splitlines()[0]
