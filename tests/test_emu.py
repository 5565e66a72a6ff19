"""The device kernels' own phase sequences run on the CPU (tests/emu/cf_emu: the kernel headers
compiled for the host, every modular product checked exactly with 128-bit integers), against the
oracle: each transform-length class of the one-workgroup kernel (both tile-flush variants), the
16-workgroup kernel of tables beyond 16384 rows with more than 8 queries, both models, and every
phase's threads run in reverse order (a phase must not depend on its own writes). The 16-workgroup
kernels also run with their workgroups one at a time between group barriers, in ascending and in
descending order: a read of another workgroup's data that no group barrier orders after its write
then gives different answers in the two orders."""
import os
import subprocess

import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
EMU_DIR = os.path.join(HERE, "emu")
EMU = os.path.join(EMU_DIR, "cf_emu")
P0, P1 = 1125899865948161, 1125899861753857  # primes 0 and 1 of the 50-bit list


@pytest.fixture(scope="module")
def emu():
    subprocess.run(["make", "-s", "-C", EMU_DIR], check=True)
    return EMU


def run_emu(emu, d, model, p, queries, lg=None, defer=None, reverse=False, serial=0, poison=False):
    argv = [emu]
    if lg:
        argv += ["--lg", str(lg)]
    if defer is not None:
        argv += ["--defer", str(defer)]
    if reverse:
        argv += ["--reverse"]
    if serial:
        argv += ["--serial-wg", str(serial)]
    if poison:
        argv += ["--poison"]
    argv += [str(d), str(model), str(p)] + [str(x) for q in queries for x in q]
    r = subprocess.run(argv, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    stats = lines[-1].split()
    assert stats[0] == "products" and int(stats[1]) > 0
    return [int(l.split()[2]) for l in lines[:-1]], float(stats[3])


@pytest.mark.parametrize("lg,defer", [(10, None), (11, None), (12, None), (13, None), (13, 0), (14, None), (14, 0)])
def test_each_length_class(emu, lg, defer):
    # tables far shorter than L are valid (any L >= n); n crosses several 16-row blocks and the
    # queries straddle the block and segment boundaries
    qs = [(300, 900), (300, 301), (257, 600), (129, 130), (40, 400), (17, 18), (16, 50), (5, 6), (250, 251)]
    got, ratio = run_emu(emu, 3, 0, P0, qs, lg=lg, defer=defer)
    assert got == oracle.nbad(3, P0, qs)
    assert ratio < 2.0  # every product result well inside the 2^52 bound (p < 2^50)


@pytest.mark.parametrize("d,model", [(1, 0), (2, 0), (5, 0), (2, 1), (3, 1)])
def test_models_and_d(emu, d, model):
    qs = [(150, 300), (150, 151), (33, 34), (d, d + 1)]
    got, _ = run_emu(emu, d, model, P1, qs)
    assert got == oracle.nbad(d, P1, qs, rep=bool(model))


def test_engine_choice_at_the_class_boundaries(emu):
    # n just above a power of two picks the next length; n = 1024 runs its c[nm] -> c[0] wrap
    for qs in ([(1024, 2048), (1000, 1001)], [(1025, 1100)], [(2048, 2049)]):
        got, _ = run_emu(emu, 3, 0, P0, qs)
        assert got == oracle.nbad(3, P0, qs)


def test_reverse_thread_order_changes_nothing(emu):
    qs = [(200, 400), (64, 65), (199, 1000)]
    for lg in (10, 14):
        fwd, _ = run_emu(emu, 3, 0, P0, qs, lg=lg)
        rev, _ = run_emu(emu, 3, 0, P0, qs, lg=lg, reverse=True)
        assert fwd == rev == oracle.nbad(3, P0, qs)


@pytest.mark.slow
def test_sixteen_workgroup_kernel_with_many_queries(emu):
    # the 2^16 kernel (16 workgroups, host barriers) on a short table: 20 queries, so 12 of them
    # take the extra query phases; and in reverse thread order
    qs = [(300, 900), (300, 301), (299, 500), (250, 700), (200, 400), (180, 181), (150, 600), (120, 240),
          (100, 800), (90, 91), (64, 65), (50, 70), (40, 400), (30, 31), (20, 40), (10, 11), (9, 10), (8, 200),
          (300, 450), (5, 6)]
    want = oracle.nbad(3, P0, qs)
    got, _ = run_emu(emu, 3, 0, P0, qs, lg=16)
    assert got == want
    got, _ = run_emu(emu, 3, 0, P0, qs[:9], lg=16, reverse=True)
    assert got == want[:9]


@pytest.mark.slow
@pytest.mark.parametrize("serial", [1, 2])
def test_sixteen_workgroup_kernel_in_fixed_workgroup_orders(emu, serial):
    # the hand-overs across workgroups 0 -> 1 (row 256), the per-workgroup query slots (more than
    # 8 queries, and row nm right after the last step), from a poisoned start: arena NaN, LDS and
    # the threads' register arrays all-ones (none of them is cleared between launches)
    qs = [(300, 900), (300, 301), (299, 500), (250, 700), (200, 400), (180, 181), (150, 600), (120, 240),
          (100, 800), (90, 91), (64, 65), (50, 70), (40, 400), (30, 31), (20, 40), (10, 11), (9, 10), (8, 200),
          (300, 450), (5, 6)]
    got, _ = run_emu(emu, 3, 0, P0, qs, lg=16, serial=serial, poison=True)
    assert got == oracle.nbad(3, P0, qs)


def test_poisoned_start_changes_nothing(emu):
    # one-workgroup kernels of each flush variant from a poisoned arena, LDS and registers
    qs = [(300, 900), (300, 301), (64, 65)]
    for lg, defer in ((12, None), (14, None), (14, 0)):
        got, _ = run_emu(emu, 3, 0, P1, qs, lg=lg, defer=defer, poison=True)
        assert got == oracle.nbad(3, P1, qs), (lg, defer)


@pytest.mark.slow
def test_sixteen_workgroup_kernel_of_2_17(emu):
    # the 2^17 form (16 workgroups of 512 threads, inner transforms of 2^13) on a short table
    qs = [(300, 900), (300, 301), (200, 400), (64, 65), (20, 40), (9, 10), (250, 251), (100, 800), (90, 91), (30, 31)]
    got, ratio = run_emu(emu, 3, 0, P0, qs, lg=17)
    assert got == oracle.nbad(3, P0, qs) and ratio < 2.0
