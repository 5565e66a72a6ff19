"""The register passes of the 62-bit kernel (cuckoo-failure_amd/csrc/cf_p62.hip, Work62::dif / dit
and the pass lists fwd_mid / inv_mid / conv_mid), restated over Python integers: radix-16 DIF passes
from the top and one partial pass at the bottom, the inverse DIT passes in the opposite order, each
group's twiddles made from one base w^e (e < L/16) and the roots of order 16. They must give the
radix-2 transforms word for word (the prologue's fwd() and the textbook DIT inverse), and the
inverse of the forward must be L times the input. A wrong group index, twiddle exponent or pass
order fails here for some L."""
import random

import pytest



def _prime():
    c = (1 << 61) // (1 << 20)
    while True:
        p = c * (1 << 20) + 1
        if all(pow(a, p - 1, p) == 1 for a in (2, 3, 5, 7, 11, 13)):
            return p
        c -= 1


def _root(p, L):
    for x in range(2, 1000):
        if pow(x, (p - 1) // 2, p) == p - 1:
            return pow(x, (p - 1) // L, p)


def _tables(p, L):
    # tw[h + j] = w_2h^j and its inverse, as the host builds them (cf_p62_host.cpp prime_tables)
    w = _root(p, L)
    wi = pow(w, p - 2, p)
    tw, twi = [0] * L, [0] * L
    h = 1
    while h < L:
        s, si = pow(w, L // (2 * h), p), pow(wi, L // (2 * h), p)
        f = fi = 1
        for j in range(h):
            tw[h + j], twi[h + j] = f, fi
            f, fi = f * s % p, fi * si % p
        h *= 2
    return tw, twi


def _dif2(x, tw, p):
    x, L, h = x[:], len(x), len(x) // 2
    while h:
        for b in range(0, L, 2 * h):
            for j in range(h):
                u, v = x[b + j], x[b + j + h]
                x[b + j], x[b + j + h] = (u + v) % p, (u - v) * tw[h + j] % p
        h //= 2
    return x


def _dit2(x, twi, p):
    x, L, h = x[:], len(x), 1
    while h < L:
        for b in range(0, L, 2 * h):
            for j in range(h):
                u, o = x[b + j], x[b + j + h] * twi[h + j] % p
                x[b + j], x[b + j + h] = (u + o) % p, (u - o) % p
        h *= 2
    return x


def _passes(lg):
    full, r = lg // 4, lg % 4
    fwd, h = [], 1 << (lg - 1)
    for _ in range(full):
        fwd.append((16, h))
        h >>= 4
    if r:
        fwd.append((1 << r, (1 << r) >> 1))
    inv, h = [], 1
    if r:
        inv.append((1 << r, 1))
        h = 1 << r
    for _ in range(full):
        inv.append((16, h))
        h <<= 4
    return fwd, inv


def _dif_pass(x, R, h, tw, p):
    L, s = len(x), 2 * h // R
    for k in range(L // R):
        j = k % s
        i0 = (k - j) * 16 + j if R == 16 else k * R  # the kernel's group index (partial pass: j = 0)
        b = tw[L // 2 + j * (L // (2 * h))]  # the LDS base table w^e, e < L/16
        assert j * (L // (2 * h)) < max(1, L // 16)
        g = [x[i0 + m * s] for m in range(R)]
        l, e = 0, R // 2
        while e:
            tws = [b if r == 0 else b * tw[8 + ((16 // R) << l) * r] % p for r in range(e)]
            for k0 in range(0, R, 2 * e):
                for r in range(e):
                    u, v = g[k0 + r], g[k0 + r + e]
                    g[k0 + r], g[k0 + r + e] = (u + v) % p, (u - v) * tws[r] % p
            b, l, e = b * b % p, l + 1, e // 2
        for m in range(R):
            x[i0 + m * s] = g[m]


def _dit_pass(x, R, h, twi, p):
    L = len(x)
    for k in range(L // R):
        j = k % h
        i0 = (k - j) * 16 + j if R == 16 else k * R
        c = twi[L // 2 + j * (L // (R * h))]
        assert j * (L // (R * h)) < max(1, L // 16)
        g = [x[i0 + m * h] for m in range(R)]
        lgr = R.bit_length() - 1
        pw = [0] * lgr
        pw[lgr - 1] = c
        for l in range(lgr - 2, -1, -1):
            pw[l] = pw[l + 1] * pw[l + 1] % p
        for l in range(lgr):
            e = 1 << l
            tws = [pw[l] if r == 0 else pw[l] * twi[8 + (8 >> l) * r] % p for r in range(e)]
            for k0 in range(0, R, 2 * e):
                for r in range(e):
                    u, o = g[k0 + r], g[k0 + r + e] * tws[r] % p
                    g[k0 + r], g[k0 + r + e] = (u + o) % p, (u - o) % p
        for m in range(R):
            x[i0 + m * h] = g[m]


@pytest.mark.parametrize("lg", range(4, 15))
def test_register_passes_equal_the_radix_2_transforms(lg):
    p = _prime()
    L = 1 << lg
    tw, twi = _tables(p, L)
    rng = random.Random(lg)
    x = [rng.randrange(p) for _ in range(L)]
    want = _dif2(x, tw, p)
    fwd, inv = _passes(lg)
    y = x[:]
    for R, h in fwd:
        _dif_pass(y, R, h, tw, p)
    assert y == want
    back = _dit2(want, twi, p)
    for R, h in inv:
        _dit_pass(y, R, h, twi, p)
    assert y == back
    assert back == [v * L % p for v in x]


def test_swizzle_is_a_bijection():
    # the LDS word of entry i (Work62::sw)
    for lg in range(4, 15):
        L = 1 << lg
        assert sorted(i ^ ((i >> 4) & 31) for i in range(L)) == list(range(L))
