"""The exact combiner of fast/post_gmp.py restated with Python integers (no gmpy2 here).

TEST INFRASTRUCTURE ONLY: the checker for cf_post / cf_exact_combine. Only tests/ may import it.

Each function follows the reference file:line it restates:
  need()  fast/post_gmp.py:46-48   (double precision, pbits default 31)
  crt()   fast/post_gmp.py:51-60   (pairwise merge with a modular inverse)
  sci()   fast/post_gmp.py:63-76   (40 significant digits by exact integer division)
  row()   fast/post_gmp.py:79-93   (one TSV row, INCOMPLETE rows, the assertion)
  table() fast/post_gmp.py:36-43, 79  (reading the residue files, header, (n, M) order)
log2 is taken with the decimal module at 120 digits from the top 400 bits of each integer (the
reference uses 320-bit MPFR); both round to the same 15 printed decimals.
"""
import decimal
import math
from collections import defaultdict

HEADER = "n\tM\tratio\tlog2_pfail\tle_2^-40\tle_2^-64\tle_2^-128\tp_fail\tprimes"


def need(n, M, d, pbits=31):
    b = sum(math.log2(M - i) - math.log2(i + 1) for i in range(d))
    return math.ceil((n * b + 3) / float(pbits)) + 1


def crt(pairs):
    items = [(r, p) for p, r in pairs]
    while len(items) > 1:
        nxt = []
        for k in range(0, len(items) - 1, 2):
            (x1, m1), (x2, m2) = items[k], items[k + 1]
            nxt.append((x1 + m1 * ((x2 - x1) * pow(m1, -1, m2) % m2), m1 * m2))
        if len(items) % 2:
            nxt.append(items[-1])
        items = nxt
    return items[0]


def sci(N, D, digits=40):
    if N == 0:
        return "0"
    e = int((N.bit_length() - D.bit_length()) * 0.30103)  # a start; the loop settles e (as :68-76)
    while True:
        s = digits - 1 - e
        q = (N * 10 ** s) // D if s >= 0 else N // (D * 10 ** (-s))
        t = str(q)
        if len(t) == digits:
            return t[0] + "." + t[1:] + "e%d" % e
        e += 1 if len(t) > digits else -1


def _log2(x, ctx):
    sh = max(0, x.bit_length() - 400)
    return ctx.add(decimal.Decimal(sh), ctx.divide(ctx.ln(decimal.Decimal(x >> sh)), ctx.ln(decimal.Decimal(2))))


def log2_text(x, Qn):
    """log2(x / Qn) printed as post_gmp.py:85,88 prints it ("%.15f", or -inf for x = 0)."""
    if x == 0:
        return "-inf"
    ctx = decimal.Context(prec=120)
    lg = ctx.subtract(_log2(x, ctx), _log2(Qn, ctx))
    return format(lg.quantize(decimal.Decimal("1e-15"), rounding=decimal.ROUND_HALF_EVEN, context=ctx), "f")


def row(n, M, d, got, pbits=31):
    """got: {prime index: (p, r)}. Returns (line, N_bad or None); raises AssertionError as the
    reference does when the primes do not cover 4 Q^n or the residues exceed Q^n."""
    k = need(n, M, d, pbits)
    if any(i not in got for i in range(k)):
        return f"{n}\t{M}\t{M / n:.6f}\tINCOMPLETE {len(got)}/{k}", None
    x, m = crt([got[i] for i in range(k)])
    Qn = math.comb(M, d) ** n
    assert m > 4 * Qn and x <= Qn, (n, M)
    tests = "\t".join("yes" if (x << lam) <= Qn else "no" for lam in (40, 64, 128))
    return f"{n}\t{M}\t{M / n:.6f}\t{log2_text(x, Qn)}\t{tests}\t{sci(x, Qn)}\t{k}", x


def read(texts):
    res = defaultdict(dict)
    for text in texts:
        for line in text.splitlines():
            t = line.split()
            if len(t) == 5:
                res[(int(t[2]), int(t[3]))][int(t[0])] = (int(t[1]), int(t[4]))
    return res


def table(d, texts, pbits=31):
    """The whole stdout of post_gmp.py for residue files with these contents."""
    res = read(texts)
    out = [HEADER]
    for (n, M) in sorted(res):
        out.append(row(n, M, d, res[(n, M)], pbits)[0])
    return "\n".join(out) + "\n"
