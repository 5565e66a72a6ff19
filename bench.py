#!/usr/bin/env python3
"""Throughput of the one-table suffix-table engine on one MI355X.

Workload (the reference's fast-path batch): d = 3, one table, n = 16384, M = 32768, the 512-prime
slice 0..511 of the 50-bit list. One step = one engine call that builds the suffix table of every
prime in the slice and answers the query (512 residues of N_bad). The metric is prime-residues/s.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-baseline ifma|scalar|none]

Prints one JSON line. `roofline` is for the dominant kernel (the L = 2^14 task kernel):
algorithmic exact modular multiplies per launch (the work formula below) over the launch's
average duration, measured with HIP events on the engine's stream, expressed as FP64 VALU
FMA-equivalent TFLOP/s (one exact modmul = 6 FP64 VALU instructions) against the MI355X FP64
vector peak of 78.6 TFLOP/s. `cpu_baseline` times the reference's own host build (compiled from
its unmodified sources into oracle/_ref) on this machine's cores, on a bounded sample of the
same workload. `traffic` is the HBM bytes of one launch from rocprofv3 FETCH_SIZE / WRITE_SIZE
passes scaled by the counters' readings for a plain copy of known size; it is read from the
committed profile `profiles/k_tasks14_n16384_traffic.json` (null if that file is absent);
`hbm_bytes_per_s` and `hbm_frac` are that traffic over the measured launch time, against 8 TB/s.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "cuckoo-failure_amd"))

N, M, D, NPRIMES = 16384, 32768, 3, 512
PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 vector, dense
PEAK_HBM_BYTES_PER_S = 8e12  # MI355X HBM3E
FLOP_PER_MODMUL = 12         # 6 FP64 VALU instructions, counted as FMA-equivalents (2 flops each)


def work_modmul(n, d, J=1):
    """Algorithmic exact modular multiplies per residue: 2n-2d-1 convolutions of length L, each two
    radix-2 NTTs plus a pointwise product, the stage pre/post products and the query rows."""
    L = max(16, 1 << (n - 1).bit_length())
    C = 2 * n - 2 * d - 1
    return C * (L * int(math.log2(L)) + L) + 6.5 * n * n + 1.5 * n * n * J


def table_work(d, ns):
    """Algorithmic exact modular multiplies of one table of max(ns) rows answering queries with
    rows ns (one per entry): the table's build once, each query's row sums."""
    nm = max(ns)
    return work_modmul(nm, d, 0) + 1.5 * sum(n * n for n in ns)


REF_BUILDS = {"ifma": ("cuckoo_fast_ifma", "AVX-512 IFMA host build"), "scalar": ("cuckoo_fast_cpu", "scalar host build"),
              "fast62": ("cuckoo_fast62", "-DFP_RB=64 host build (62-bit primes, scalar)")}


def progress(msg):
    """One line on stderr as each stage of the run ends (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(kind, procs=None, residues=None):
    """The reference host build on `procs` cores: one process per prime (the reference's own
    scale-out by disjoint prime ranges), wall time from first start to last exit. When `residues`
    is a dict, the printed residues are put in it by prime index."""
    name, what = REF_BUILDS[kind]
    exe = os.path.join(REPO, "oracle", "_ref", name)
    if not os.path.exists(exe):
        return None
    if kind == "ifma" and "avx512ifma" not in open("/proc/cpuinfo").read():
        return None
    cores = len(os.sched_getaffinity(0))
    procs = procs or min(cores, 16)
    with tempfile.TemporaryDirectory() as td:
        qf = os.path.join(td, "q.txt")
        with open(qf, "w") as f:
            f.write(f"{N} {M}\n")
        t0 = time.perf_counter()
        ps = [subprocess.Popen([exe, str(D), qf, str(i), "1"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL)
              for i in range(procs)]
        outs = [p.communicate()[0] for p in ps]
        wall = time.perf_counter() - t0
    if any(p.returncode for p in ps) or sum(len(o.splitlines()) for o in outs) != procs:
        return None
    if residues is not None:
        for o in outs:
            t = o.split()
            residues[int(t[0])] = int(t[4])
    return {"value": procs / wall, "unit": "prime-residues/s", "cores": procs, "kind": "reference",
            "sample": f"{procs} primes (i=0..{procs - 1}) of n={N}, M={M}, d={D}, one process per prime, "
                      f"{what}, wall {wall:.1f} s"}


def reference_residues(name):
    """{prime index: residue} from the reference host build's lines in tests/golden/<name> (empty
    if absent). Every file the bench compares against is data written by that build."""
    out = {}
    for fname in (name,) if isinstance(name, str) else name:
        try:
            with open(os.path.join(REPO, "tests", "golden", fname)) as f:
                rec = json.load(f)
        except (OSError, ValueError):
            continue
        for line in rec["stdout"].splitlines():
            t = line.split()
            if len(t) == 5:
                out[int(t[0])] = int(t[4])
    return out


def serving_leg(a):
    """The serving workload inside the default run: tables of d = 3, n <= 2048, M <= 8192 for primes
    0..1535 built once and kept in HBM; `serve_queries` requests of the seed-1 Poisson stream
    replayed in real time through the continuous-batching server, then a prefix through the same
    server with one request per launch and one slot. The 41 requests with reference residues
    (tests/golden/serving_2048.json) are checked through the batched query first."""
    import hashlib

    import cuckoo_amd as cf
    import workloads

    nmax, Mmax, np_ = 2048, 8192, 1536
    n, M, t = workloads.serving_queries(1, 100000, 32, nmax, a.rate)
    k_srv, k_seq = min(a.serve_queries, len(n)), min(a.seq_queries, a.serve_queries)
    eng = cf.Engine()
    t0 = time.perf_counter()
    tabs = cf.Tables(eng, 3, nmax, Mmax, cf.prime_range(0, np_))
    build_s = time.perf_counter() - t0
    progress(f"serving: tables built in {build_s:.2f} s; replaying {k_srv} requests")
    checked = 0
    try:
        with open(os.path.join(REPO, "tests", "golden", "serving_2048.json")) as f:
            gold = json.load(f)["queries"]
    except (OSError, ValueError):
        gold = []
    if gold:
        res = tabs.query([(q["n"], q["M"]) for q in gold], need=[q["need"] for q in gold])
        for j, q in enumerate(gold):
            col = "".join(f"{res[b][j]}\n" for b in range(q["need"]))
            assert hashlib.sha256(col.encode()).hexdigest() == q["residues_sha256"], ("serving residues", q)
            checked += q["need"]
    with cf.Server(tabs, slots=a.slots, max_batch=a.max_batch, crt_threads=a.crt_threads) as srv:
        srv.replay(n[:64], M[:64], t[:64] - t[0])  # warm-up
        _, l2, _, st = srv.replay(n[:k_srv], M[:k_srv], t[:k_srv] - t[0])
    progress(f"serving: {k_srv} requests replayed; {k_seq} more one per launch")
    with cf.Server(tabs, slots=1, max_batch=1, crt_threads=a.crt_threads) as seq:
        _, l2s, _, sst = seq.replay(n[:k_seq], M[:k_seq], t[:k_seq] - t[0])
    assert l2s == l2[:k_seq], "served and sequential answers differ"
    tabs.close()
    eng.close()
    return {"metric": "nbad queries/s with p50/p99 latency (arrival to CRT'd p_fail)",
            "queries_per_s": round(st.queries_per_s, 1), "p50_ms": round(st.p50_ms, 3), "p99_ms": round(st.p99_ms, 3),
            "requests": k_srv, "rate": a.rate, "mean_batch": round(st.mean_batch, 2),
            "sequential": {"requests": k_seq, "queries_per_s": round(sst.queries_per_s, 1),
                           "p50_ms": round(sst.p50_ms, 3), "p99_ms": round(sst.p99_ms, 3)},
            "speedup_vs_sequential": round(st.queries_per_s / sst.queries_per_s, 2) if sst.queries_per_s else None,
            "workload": f"tables d=3, n_max={nmax}, M_max={Mmax}, primes 0..{np_ - 1} resident; seed-1 Poisson "
                        f"stream, n uniform [32, {nmax}], M uniform [n+1, 4n]",
            "build_s": round(build_s, 2), "reference_residues_checked": checked,
            "identical_answers": "served = sequential log2 p_fail for every request of the prefix"}


def measured_modmul_peak():
    """The exact FP64 modular product's peak rate on this GPU (bin/cf_modmul_rate: 8 independent
    chains per thread, 2048 x 256 threads), or None if the microbenchmark is not built."""
    import cuckoo_amd as cf
    exe = os.path.join(os.path.dirname(cf.CLI_FAST), "cf_modmul_rate")
    if not os.path.exists(exe):
        return None
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    if r.returncode:
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


def profiled_traffic():
    path = os.path.join(REPO, "profiles", "k_tasks14_n16384_traffic.json")
    try:
        with open(path) as f:
            return float(json.load(f)["traffic_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def bench_uneven(a):
    """The uneven batch: 256 instances (seed 0) of n in [64, 4096], d in {2, 3, 4}, M = ceil(alpha n),
    each over its whole prime set, as one work-queue launch in which instances of one d share each
    prime's table; the same batch with one table per (instance, prime); and the plain path, one
    engine call per instance. All three must give identical residues."""
    import cuckoo_amd as cf
    import workloads

    inst = workloads.uneven_batch(0, a.instances)
    spec = [(d, 0, n, M, cf.prime_range(0, cf.need(n, M, d))) for n, d, M in inst]
    tables = sum(len(x[4]) for x in spec)
    mm = sum(len(x[4]) * work_modmul(x[2], x[0]) for x in spec)
    shared = {}  # (d, prime) -> the rows of the queries on that table
    for d, _, n, M, primes in spec:
        for p in primes:
            shared.setdefault((d, p), []).append(n)
    mm_shared = sum(table_work(d, ns) for (d, _), ns in shared.items())
    eng = cf.Engine()
    # the work-queue kernel's variant, chosen by the autotuner on a 1/16 sample of this batch
    tune = eng.autotune_batch(spec, sample_tables=tables // 16)
    eng.run_batch([x[:4] + (x[4][:1],) for x in spec[:8]])  # warm-up: allocation, code load
    t0 = time.perf_counter()
    got = eng.run_batch(spec)
    wall_q = time.perf_counter() - t0
    st = eng.stats()
    # the work queue alone: one table per (instance, prime)
    eng.set_batch_sharing(False)
    t0 = time.perf_counter()
    got_ns = eng.run_batch(spec)
    wall_ns = time.perf_counter() - t0
    st_ns = eng.stats()
    eng.set_batch_sharing(True)
    assert got_ns == got, "shared and unshared batch residues differ"
    t0 = time.perf_counter()
    plain, plain_kernel_ms, plain_launches = [], 0.0, 0
    for d, model, n, M, primes in spec:
        plain.append([r[0] for r in eng.run([(n, M)], primes, d, model=model)])
        plain_kernel_ms += eng.stats().kernel_ms
        plain_launches += eng.stats().launches
    wall_p = time.perf_counter() - t0
    eng.close()
    assert got == plain, "work-queue and per-instance residues differ"
    # the reference's own usage: one `cuckoo_fast d q_i.txt 0 need_i` process per instance, timed
    # on every 16th instance (process start, device set-up and run included)
    sample = list(range(0, len(spec), 16))
    wall_c, tables_c = 0.0, 0
    with tempfile.TemporaryDirectory() as td:
        for i in sample:
            d, _, n, M, primes = spec[i]
            qf, of = os.path.join(td, f"q{i}.txt"), os.path.join(td, f"o{i}.txt")
            with open(qf, "w") as f:
                f.write(f"{n} {M}\n")
            t0 = time.perf_counter()
            subprocess.run([cf.CLI_FAST, str(d), qf, "0", str(len(primes)), "--out", of], check=True,
                           capture_output=True)
            wall_c += time.perf_counter() - t0
            tables_c += len(primes)
            lines = [l.split() for l in open(of).read().splitlines()]
            assert [int(t[4]) for t in lines] == got[i], "per-instance CLI residues differ"
    achieved = mm_shared * FLOP_PER_MODMUL / (st.kernel_ms * 1e-3) / 1e12
    achieved_ns = mm * FLOP_PER_MODMUL / (st_ns.kernel_ms * 1e-3) / 1e12
    print(json.dumps({
        "metric": "uneven-batch prime-residues/s (256 instances, seed 0, bit-exact)",
        "value": round(tables / wall_q, 2), "unit": "prime-residues/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f64 (exact integer residues < 2^50)", "data": "synthetic (seeded instance list)",
        "config": {"workload": f"{a.instances} instances, n log-uniform in [64, 4096], d in {{2,3,4}}, "
                               f"M = ceil(alpha n), alpha in [1.05, 3]; every instance over primes 0..need-1",
                   "residues": tables},
        "work_queue": {"wall_s": round(wall_q, 3), "kernel_ms": round(st.kernel_ms, 1), "launches": st.launches,
                       "tables": st.tasks, "device_bytes": st.device_bytes,
                       "note": "instances of one d share each prime's table (max n rows, every query on it)"},
        "work_queue_unshared": {"wall_s": round(wall_ns, 3), "kernel_ms": round(st_ns.kernel_ms, 1),
                                "launches": st_ns.launches, "tables": st_ns.tasks, "value": round(tables / wall_ns, 2),
                                "frac": round(achieved_ns / PEAK_FP64_TFLOPS, 4)},
        "plain_per_instance": {"wall_s": round(wall_p, 3), "kernel_ms": round(plain_kernel_ms, 1),
                               "launches": plain_launches, "value": round(tables / wall_p, 2)},
        "speedup_vs_plain": round(wall_p / wall_q, 3),
        "speedup_unshared_vs_plain": round(wall_p / wall_ns, 3),
        "autotune": tune,
        "cli_per_instance": {"instances": len(sample), "tables": tables_c, "wall_s": round(wall_c, 3),
                             "value": round(tables_c / wall_c, 2),
                             "speedup_of_batch": round((tables / wall_q) / (tables_c / wall_c), 3),
                             "note": "one cuckoo_fast process per instance (every 16th instance), as the reference "
                                     "runs an uneven batch; residues identical to the batch's"},
        "roofline": {"bound": "valu", "kernel": "k_queue", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": None},
        "identical_residues": True,
    }))


def bench_serve(a):
    """Serving: suffix tables of d = 3, n <= 2048, for 1536 primes (M <= 8192) built once and kept
    in HBM; a Poisson stream of nbad(n, M) requests (seed 1, n uniform in [32, 2048], M uniform in
    [n+1, 4n]) replayed in real time through the continuous-batching server, then a prefix of
    the same stream through the same server with one request per launch and one slot."""
    import cuckoo_amd as cf
    import workloads

    nmax, Mmax, np_ = 2048, 8192, 1536
    n, M, t = workloads.serving_queries(1, a.queries, 32, nmax, a.rate)
    eng = cf.Engine()
    primes = cf.prime_range(0, np_)
    t0 = time.perf_counter()
    tabs = cf.Tables(eng, 3, nmax, Mmax, primes)
    build_s = time.perf_counter() - t0
    with cf.Server(tabs, slots=a.slots, max_batch=a.max_batch, crt_threads=a.crt_threads) as srv:
        srv.replay(n[:64], M[:64], t[:64] - t[0])  # warm-up
        lat, l2, _, st = srv.replay(n, M, t - t[0])
        slots = srv.slots()
    k = min(a.seq_queries, a.queries)
    with cf.Server(tabs, slots=1, max_batch=1, crt_threads=a.crt_threads) as seq:
        _, l2s, _, sst = seq.replay(n[:k], M[:k], t[:k] - t[0])
    assert l2s == l2[:k], "served and sequential answers differ"
    tabs.close()
    eng.close()
    print(json.dumps({
        "metric": "served nbad queries/s (d=3, n<=2048, 1536 primes, CRT to p_fail, bit-exact)",
        "value": round(st.queries_per_s, 1), "unit": "queries/s", "n_gpus": 1, "higher_is_better": True,
        "p50_ms": round(st.p50_ms, 3), "p99_ms": round(st.p99_ms, 3),
        "dtype": "f64 (exact integer residues < 2^50)", "data": "synthetic (seeded Poisson request stream)",
        "config": {"workload": f"tables d=3, n_max={nmax}, M_max={Mmax}, primes 0..{np_ - 1} resident "
                               f"({tabs_bytes_gb(np_, nmax, Mmax):.1f} GB); {a.queries} requests, Poisson rate "
                               f"{a.rate}/s, n uniform [32, {nmax}], M uniform [n+1, 4n]",
                   "slots": slots, "max_batch": a.max_batch, "crt_threads": a.crt_threads},
        "build_s": round(build_s, 2), "mean_batch": round(st.mean_batch, 2), "batches": st.batches,
        "sequential": {"queries": k, "queries_per_s": round(sst.queries_per_s, 1), "p50_ms": round(sst.p50_ms, 3),
                       "p99_ms": round(sst.p99_ms, 3)},
        "speedup_vs_sequential": round(st.queries_per_s / sst.queries_per_s, 2) if sst.queries_per_s else None,
        "identical_answers": True,
    }))


def bench_wide(a):
    """One table of n = 65536, M = 131072, d = 3 per prime over a slice of the 50-bit list: transforms
    of 2^16 spread over 16 workgroups per prime, 16 primes per launch (device memory)."""
    import cuckoo_amd as cf

    n, M, d = a.wide_n, 2 * a.wide_n, 3
    count = a.primes if a.primes != NPRIMES else 256
    primes = cf.prime_range(0, count)
    # --wide-queries 126: the query count of the reference's published n = 65536 timing
    # (fast/README.md:80-86): 63 queries at n with M = n + 1 + 1040 k, and 63 at n = 1040 (k + 1),
    # M = 2 n (k = 0..62); otherwise the one query (n, 2 n)
    if a.wide_queries == 126:
        qs = [(n, n + 1 + 1040 * k) for k in range(63)] + [(1040 * (k + 1), 2 * 1040 * (k + 1)) for k in range(63)]
    else:
        qs = [(n, M)]
    eng = cf.Engine()
    eng.run(qs, primes[:min(count, 16)], d)  # warm-up: one launch, the device arena allocated
    t0 = time.perf_counter()
    got = eng.run(qs, primes, d)
    wall = time.perf_counter() - t0
    st = eng.stats()
    dev = eng.device_name()
    eng.close()
    pinned, bad = [], []
    if len(qs) == 1:
        fixtures = {65536: ["fast_65536.json", "fast_65536_more.json"], 32768: ["fast_32768_launch.json"]}
        ref = reference_residues(fixtures.get(n, []))
        pinned = [i for i in range(count) if i in ref]
        bad = [i for i in pinned if got[i][0] != ref[i]]
    elif n == 65536:
        try:
            with open(os.path.join(REPO, "tests", "golden", "fast_65536_q126.json")) as f:
                lines = [l.split() for l in json.load(f)["stdout"].splitlines()]
        except (OSError, ValueError):
            lines = []
        for t in lines:
            i, j = int(t[0]), qs.index((int(t[2]), int(t[3])))
            if i < count:
                pinned.append(i)
                if got[i][j] != int(t[4]):
                    bad.append(i)
        pinned = sorted(set(pinned))
    assert not bad, f"residues differ from the reference at prime indices {bad}"
    per_launch_ms = st.kernel_ms / st.launches
    mm_per_launch = work_modmul(n, d, len(qs)) * st.tasks / st.launches
    achieved = mm_per_launch * FLOP_PER_MODMUL / (per_launch_ms * 1e-3) / 1e12
    print(json.dumps({
        "metric": f"prime-residues/s at n={n},d=3 (bit-exact)",
        "value": round(count * len(qs) / wall, 4), "unit": "prime-residues/s", "n_gpus": 1, "higher_is_better": True,
        "primes_per_s": round(count / wall, 4),
        "ms_per_step": round(wall * 1e3, 1), "dtype": "f64 (exact integer residues < 2^50)",
        "data": "synthetic (deterministic: the reference's 50-bit prime list)",
        "config": {"workload": f"one-table suffix-table build + nbad, n={n}, d={d}, {len(qs)} "
                               f"quer{'y (M=' + str(M) + ')' if len(qs) == 1 else 'ies'}, {count}-prime slice 0..{count - 1}",
                   "device": dev, "launches": st.launches, "primes_per_launch": st.tasks // max(st.launches, 1),
                   "device_bytes": st.device_bytes},
        "roofline": {"bound": "valu", "kernel": "k_wide", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": None,
                     "launch_ms": round(per_launch_ms, 1), "modmul_per_launch": mm_per_launch},
        "seconds_per_prime": round(wall / count, 4),
        "reference_check": f"residues of primes {pinned} equal the reference host build's",
    }))


def bench_fast62(a):
    """The headline shape over the reference's 62-bit prime list (its -DFP_RB=64 build): n = 16384,
    M = 32768, d = 3, primes 0..count-1 of c 2^20 + 1 below 2^62, on the 64-bit Montgomery kernel.
    A 62-bit residue carries 61 bits of the CRT product against 49 for a 50-bit one, so
    information-equivalent throughput is also reported."""
    import cuckoo_amd as cf

    count = a.primes if a.primes != NPRIMES else 256
    primes = cf.prime_range(0, count, prime_bits=62)
    eng = cf.Engine()
    eng.run([(N, M)], primes[:8], D)  # warm-up
    t0 = time.perf_counter()
    got = eng.run([(N, M)], primes, D)
    wall = time.perf_counter() - t0
    st = eng.stats()
    dev = eng.device_name()
    eng.close()
    ref = {}
    try:
        with open(os.path.join(REPO, "tests", "golden", "fast62_cases.json")) as f:
            for rec in json.load(f):
                if rec["queries"] == [[N, M]]:
                    for line in rec["stdout"].splitlines():
                        t = line.split()
                        ref[int(t[0])] = int(t[4])
    except (OSError, ValueError):
        pass
    base = None
    if a.cpu_baseline != "none":  # the reference's 62-bit build on the host cores, its residues checked too
        timed = {}
        base = cpu_baseline("fast62", residues=timed)
        ref.update(timed)
    pinned = [i for i in ref if i < count]
    assert all(got[i][0] == ref[i] for i in pinned), "62-bit residues differ from the reference's"
    mm = work_modmul(N, D) * count
    mpeak = measured_modmul_peak()
    mont = mpeak.get("mont64_modmul_per_s") if mpeak else None
    print(json.dumps({
        "metric": "prime-residues/s at n=16384,d=3, 62-bit primes (bit-exact)",
        "value": round(count / wall, 3), "unit": "prime-residues/s", "n_gpus": 1, "higher_is_better": True,
        "ms_per_step": round(wall * 1e3, 1), "dtype": "u64 (Montgomery, R = 2^64)",
        "data": "synthetic (deterministic: the reference's 62-bit prime list)",
        "config": {"workload": f"one-table suffix-table build + nbad, n={N}, M={M}, d={D}, {count} primes of the 62-bit list",
                   "device": dev, "launches": st.launches, "kernel_ms": round(st.kernel_ms, 1)},
        "equivalent_50_bit_residues_per_s": round(count / wall * 61 / 49, 3),
        "modmul_per_s": round(mm / (st.kernel_ms * 1e-3), 1),
        "roofline": {"bound": "valu", "kernel": "k_p62<14>", "achieved": round(mm / (st.kernel_ms * 1e-3), 1),
                     "peak": mont, "unit": "Montgomery products/s (measured peak, bin/cf_modmul_rate)",
                     "frac": round(mm / (st.kernel_ms * 1e-3) / mont, 4) if mont else None, "traffic": None},
        "cpu_baseline": base,
        "reference_check": f"residues of primes {sorted(pinned)} equal the reference's 62-bit build's",
    }))


def tabs_bytes_gb(np_, nmax, Mmax):
    F1 = max(nmax, Mmax) + 1
    return np_ * 8 * (3 * F1 + nmax * (nmax + 1) // 2) / 1e9


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--primes", type=int, default=NPRIMES)
    ap.add_argument("--cpu-baseline", default="ifma", choices=["ifma", "scalar", "none"])
    ap.add_argument("--workload", default="fast", choices=["fast", "uneven", "serve", "wide", "fast62"],
                    help="fast: the headline n=16384 batch; uneven: the 256-instance work-queue batch; "
                         "serve: resident tables and a replayed request stream; wide: n=65536 over 256 primes; "
                         "fast62: the headline shape over the 62-bit prime list")
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--wide-n", type=int, default=65536, help="wide: table rows (M = 2 n)")
    ap.add_argument("--wide-queries", type=int, default=1, choices=[1, 126],
                    help="wide: 1 query, or the 126 of the reference's published n = 65536 timing")
    ap.add_argument("--queries", type=int, default=100000)
    ap.add_argument("--rate", type=float, default=1500.0)
    ap.add_argument("--seq-queries", type=int, default=1000)
    ap.add_argument("--serve-queries", type=int, default=100000,
                    help="fast: requests of the serving stream replayed after the headline (0: no serving figures)")
    ap.add_argument("--slots", type=int, default=2)
    ap.add_argument("--max-batch", type=int, default=32)
    ap.add_argument("--crt-threads", type=int, default=8)
    ap.add_argument("--autotune", action="store_true",
                    help="fast: time the 2^14 kernel variants on the workload's own tables first and run the fastest")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (primes shard across GPUs by "
              f"disjoint FIRST/COUNT ranges with no communication)", file=sys.stderr)
        sys.exit(2)

    if a.workload == "uneven":
        return bench_uneven(a)
    if a.workload == "serve":
        return bench_serve(a)
    if a.workload == "wide":
        return bench_wide(a)
    if a.workload == "fast62":
        return bench_fast62(a)
    import cuckoo_amd as cf
    eng = cf.Engine()
    primes = cf.prime_range(0, a.primes)
    tune = eng.autotune(14, n=N) if a.autotune else None
    for _ in range(a.warmup):
        eng.run([(N, M)], primes, D)
    progress(f"headline: {a.warmup} warm-up step(s) done")
    kernel_ms, launches, tasks = 0.0, 0, 0
    outs = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        outs.append(eng.run([(N, M)], primes, D))
        st = eng.stats()
        kernel_ms += st.kernel_ms
        launches += st.launches
        tasks += st.tasks
    wall = time.perf_counter() - t0
    progress(f"headline: {a.steps} timed step(s) in {wall:.2f} s")
    dev = eng.device_name()
    eng.close()
    # every residue of every timed call against the reference host build's lines for the slice
    ref = reference_residues("fast_16384.json")
    pinned = [i for i in range(a.primes) if i in ref]
    for got in outs:
        bad = [i for i in pinned if got[i][0] != ref[i]]
        assert not bad, f"residues differ from the reference at prime indices {bad[:8]}"

    residues = a.primes * a.steps
    value = residues / wall
    per_launch_ms = kernel_ms / launches
    mm_per_launch = work_modmul(N, D) * tasks / launches
    achieved = mm_per_launch * FLOP_PER_MODMUL / (per_launch_ms * 1e-3) / 1e12
    base = None if a.cpu_baseline == "none" else cpu_baseline(a.cpu_baseline)
    progress(f"cpu baseline: {base['sample'] if base else 'not run'}")
    mpeak = measured_modmul_peak()
    progress("modmul peak measured")
    traffic = profiled_traffic()
    out = {
        "metric": "prime-residues/s at n=16384,d=3 (bit-exact); nbad queries/s + p99 latency",
        "value": round(value, 3),
        "unit": "prime-residues/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (exact integer residues < 2^50)",
        "data": "synthetic (deterministic: the reference's 50-bit prime list)",
        "config": {"workload": f"one-table suffix-table build + nbad, n={N}, M={M}, d={D}, "
                               f"{a.primes}-prime slice 0..{a.primes - 1}", "n": N, "M": M, "d": D,
                   "primes": a.primes, "queries_per_prime": 1, "device": dev},
        "roofline": {"bound": "valu", "kernel": "k_tasks<14>", "achieved": round(achieved, 3),
                     "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
                     "traffic": traffic, "launch_ms": round(per_launch_ms, 2),
                     "hbm_bytes_per_s": round(traffic / (per_launch_ms * 1e-3)) if traffic else None,
                     "hbm_frac": round(traffic / (per_launch_ms * 1e-3) / PEAK_HBM_BYTES_PER_S, 4) if traffic else None,
                     "modmul_per_launch": mm_per_launch,
                     "measured_modmul_peak_per_s": mpeak["fp64_exact_modmul_per_s"] if mpeak else None,
                     "frac_of_measured_modmul_peak": round(mm_per_launch / (per_launch_ms * 1e-3) /
                                                           mpeak["fp64_exact_modmul_per_s"], 4) if mpeak else None},
        "cpu_baseline": base,
        "bit_exact": {"residues_per_step_equal_to_reference": len(pinned), "steps_checked": a.steps,
                      "reference": "the reference's host build (fast/cuckoo_fast.cu) on primes 0..%d" % (max(pinned) if pinned else -1)},
    }
    if a.serve_queries > 0:
        out["serving"] = serving_leg(a)
    if tune:
        out["autotune"] = tune
        if tune["chosen"]:  # the profiled traffic is the built-in variant's
            out["roofline"]["kernel"] = f"k_tasks<14> variant {tune['chosen']}"
            out["roofline"]["traffic"] = out["roofline"]["hbm_bytes_per_s"] = out["roofline"]["hbm_frac"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
