"""Readers for the fixtures under tests/golden/ (written by tests/golden/make_golden.py from the
reference's own compiled programs)."""
import json
import os
import re

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f) if name.endswith(".json") else f.read()


def residue_lines(stdout):
    """'i p n M residue' lines -> list of int tuples, in output order."""
    out = []
    for line in stdout.splitlines():
        t = line.split()
        if len(t) == 5:
            out.append(tuple(int(x) for x in t))
    return out


def fast_cases():
    return load("fast_cases.json")


def check_instances():
    """(n, M, d, replacement, nbad, total) for every one-table brute-force line of the reference's
    test program output."""
    pat = re.compile(r"\[PASS\] one table\s+n=(\d+) M=(\d+) d=(\d+) (distinct|replacement)\s+brute N_bad=(\d+) / (\d+)")
    out = []
    for line in load("check_ref.txt").splitlines():
        m = pat.search(line)
        if m:
            n, M, d = int(m.group(1)), int(m.group(2)), int(m.group(3))
            out.append((n, M, d, m.group(4) == "replacement", int(m.group(5)), int(m.group(6))))
    return out


def fail_case(args):
    for rec in load("fail_cases.json"):
        if rec["args"] == args:
            return rec
    raise KeyError(args)


def fail_field(stdout, key):
    """Value printed after 'key = ' in the 32-bit CLI's output."""
    for line in stdout.splitlines():
        if line.startswith(key + " ="):
            return line.split("=", 1)[1].strip().split()[0].rstrip(",")
    return None
