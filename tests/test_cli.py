"""The two drivers against the reference's own outputs: argument and input errors, resume
decisions that stop before any device work (CPU), and full runs (GPU)."""
import os
import subprocess
import tempfile

import pytest

import cuckoo_amd as cf
from golden_data import load, residue_lines


def run_fast(d, queries, first, count, extra=(), out_init=None, argv=None):
    with tempfile.TemporaryDirectory() as td:
        if argv is None:
            qf = os.path.join(td, "q.txt")
            with open(qf, "w") as f:
                f.write("".join(f"{n} {M}\n" for n, M in queries))
            argv = [str(d), qf, str(first), str(count), *extra]
        outf = None
        if out_init is not None:
            outf = os.path.join(td, "out.txt")
            with open(outf, "w") as f:
                f.write(out_init)
            argv = argv + ["--out", outf]
        r = subprocess.run([cf.CLI_FAST, *argv], capture_output=True, text=True, timeout=600)
        final = open(outf).read() if outf else None
        return r.returncode, r.stdout, r.stderr, final


@pytest.mark.parametrize("idx", range(7))
def test_fast_argument_errors(idx):
    rec = load("fast_errors.json")[idx]
    if "argv" in rec:
        argv = rec["argv"]
        rc, so, se, _ = run_fast(None, None, None, None, argv=argv)
    else:
        rc, so, se, _ = run_fast(rec["d"], rec["queries"], rec["first"], rec["count"], extra=rec["extra"])
    assert rc == rec["rc"]
    assert so == rec["stdout"]
    assert se.strip() == rec["stderr"].strip()


def test_fast_resume_all_done_stops_before_the_device():
    rec = load("fast_resume.json")[3]
    rc, so, se, final = run_fast(rec["d"], rec["queries"], rec["first"], rec["count"], out_init=rec["out_init"])
    assert rc == rec["rc"] == 0
    assert se.strip() == rec["stderr"].strip()
    assert final == rec["out_final"]


@pytest.mark.parametrize("idx", [19, 20, 21, 22, 23])
def test_fail_argument_errors(idx):
    rec = load("fail_cases.json")[idx]
    r = subprocess.run([cf.CLI_FAIL, *rec["args"]], capture_output=True, text=True, timeout=60)
    assert r.returncode == rec["rc"] == 2
    assert r.stderr.splitlines()[0] == rec["stderr"].splitlines()[0]
    assert r.stdout.startswith("usage:")


@pytest.mark.parametrize("idx", [i for i, r in enumerate(load("fail_separate_cases.json")) if r["rc"] != 0])
def test_separate_argument_errors(idx):
    rec = load("fail_separate_cases.json")[idx]
    r = subprocess.run([cf.CLI_FAIL, *rec["args"]], capture_output=True, text=True, timeout=60)
    assert r.returncode == rec["rc"] == 2
    assert r.stderr.splitlines()[0] == rec["stderr"].splitlines()[0]
    assert r.stdout.startswith("usage:")


# ------------------------------------------------------------------ GPU

@pytest.mark.gpu
@pytest.mark.parametrize("idx", [0, 1, 2, 4, 5, 6])
def test_fast_resume_scenarios(idx):
    rec = load("fast_resume.json")[idx]
    rc, so, se, final = run_fast(rec["d"], rec["queries"], rec["first"], rec["count"], out_init=rec["out_init"])
    assert rc == rec["rc"] == 0, se
    assert final == rec["out_final"]
    assert so == ""


@pytest.mark.gpu
@pytest.mark.parametrize("idx", range(0, len(load("fast_cases.json")), 7))
def test_fast_cli_matches_reference_lines(idx):
    rec = load("fast_cases.json")[idx]
    rc, so, se, _ = run_fast(rec["d"], rec["queries"], rec["first"], rec["count"])
    assert rc == rec["rc"] == 0, se
    assert so == rec["stdout"]


@pytest.mark.gpu
@pytest.mark.parametrize("idx", [1, 3])
def test_fast_cli_tables_beyond_16384_rows(idx):
    # the reference's own lines for n = 16385 and for two queries of 20000 / 17000 rows
    rec = load("fast_wide.json")[idx]
    rc, so, se, _ = run_fast(rec["d"], rec["queries"], rec["prime"], 1)
    assert rc == 0, se
    assert so == rec["stdout"]


@pytest.mark.gpu
def test_fast_cli_batches_do_not_change_output():
    rec = next(r for r in load("fast_cases.json") if r["queries"] == [[10, 11], [8, 200], [40, 41], [12, 13]])
    for b in ("1", "3", "50"):
        rc, so, se, _ = run_fast(rec["d"], rec["queries"], rec["first"], rec["count"], extra=("--batch", b))
        assert rc == 0, se
        assert so == rec["stdout"]


@pytest.mark.gpu
@pytest.mark.parametrize("idx", range(19))
def test_fail_cli_matches_reference(idx):
    rec = load("fail_cases.json")[idx]
    r = subprocess.run([cf.CLI_FAIL, *rec["args"]], capture_output=True, text=True, timeout=600)
    assert r.returncode == rec["rc"], r.stderr
    want = [l for l in rec["stdout"].splitlines() if "time:" not in l and "(T s)" not in l]
    got = [l for l in r.stdout.splitlines() if "time:" not in l and not l.startswith("N_bad mod")]
    assert got == want
    for wl, gl in zip(rec["stdout"].splitlines(), r.stdout.splitlines()):
        if wl.startswith("primes:"):
            assert gl.split("time:")[0] == wl.split("time:")[0]
        if wl.startswith("N_bad mod"):
            assert gl.split("(")[0] == wl.split("(")[0]


@pytest.mark.gpu
def test_fail_cli_mod_prime_below_the_engine_transform():
    # 7681 = 15 * 2^9 + 1: next_pow2(200) = 2^8 divides p - 1 (the reference's requirement), the
    # engine's shortest transform 2^10 does not; the residue is the exact N_bad's
    from golden_data import fail_case, fail_field
    exact = int(fail_field(fail_case(["single", "-n", "200", "-M", "400", "-d", "3", "--algo", "ntt", "--exact"])["stdout"],
                           "N_bad"))
    r = subprocess.run([cf.CLI_FAIL, "single", "-n", "200", "-M", "400", "-d", "3", "--mod", "7681"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.splitlines()[-1].startswith(f"N_bad mod 7681 = {exact % 7681}  (")


@pytest.mark.gpu
def test_fast_cli_not_enough_device_memory_exits_3():
    # one prime of a table of 2^20 - 1 rows needs terabytes: no batch fits (fast/cuckoo_fast.cu:751-753)
    rc, so, se, _ = run_fast(3, [(1048575, 2000000)], 0, 1)
    assert rc == 3 and so == ""
    assert se.strip().splitlines()[-1].startswith("not enough device memory:"), se


@pytest.mark.gpu
def test_fast_cli_killed_mid_run_then_resumed():
    # a run cut off after its first batches reached the file (SIGKILL), then the same command again:
    # every prime of the range ends up in the file once with the uninterrupted run's residue (a
    # line cut in half by the kill is followed by the resumed output on the same line, as with the
    # reference; such a line has more than five fields and is not a record)
    import signal
    import time
    queries = [(8192, 16384), (5000, 9000)]
    with tempfile.TemporaryDirectory() as td:
        qf, of = os.path.join(td, "q.txt"), os.path.join(td, "out.txt")
        with open(qf, "w") as f:
            f.write("".join(f"{n} {M}\n" for n, M in queries))
        argv = [cf.CLI_FAST, "3", qf, "0", "24", "--batch", "2", "--out", of]
        p = subprocess.Popen(argv, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        t0 = time.time()
        while time.time() - t0 < 120 and p.poll() is None:
            if os.path.exists(of) and open(of).read().count("\n") >= 4:
                break
            time.sleep(0.05)
        killed = p.poll() is None
        if killed:
            p.send_signal(signal.SIGKILL)
        p.wait(timeout=60)
        print(f"killed mid-run: {killed}; lines before resume: {open(of).read().count(chr(10))}", flush=True)
        r = subprocess.run(argv, capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stderr
        text = open(of).read()
        clean = subprocess.run([cf.CLI_FAST, "3", qf, "0", "24"], capture_output=True, text=True, timeout=600)
        assert clean.returncode == 0
    want = {(t[0], t[2], t[3]): t[4] for t in (l.split() for l in clean.stdout.splitlines())}
    recs = [t for t in (l.split() for l in text.splitlines()) if len(t) == 5]
    got = {}
    for t in recs:
        assert want[(t[0], t[2], t[3])] == t[4], t
        got[(t[0], t[2], t[3])] = t[4]
    assert got == want
