"""The register passes of the 62-bit kernel (cuckoo-failure_amd/csrc/cf_p62.hip, Work62::dif_c /
dit_c, twiddle_chain / twiddle_tab and the pass lists first_pass / fwd_mid / conv_mid / inv_mid /
last_pass), restated over Python integers: radix-16 DIF passes from the top and one partial pass at
the bottom, the inverse DIT passes in the opposite order; each pass a constant network of roots of
order 16 and one twiddle per entry, from the LDS tables W[x] = w^(16x), Wi[x] = w^-(16x) or, for
the outermost passes, a chain of powers of w^(+-j). They must give the radix-2 transforms word for
word (the prologue's fwd() and the textbook DIT inverse), and the inverse of the forward must be L
times the input. A wrong group index, table index, twiddle exponent or pass order fails here for
some L."""
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


def _wexp(tab, e, p):
    # w^e from tw (or w^-e from twi), e < L: tab[L/2 + e] below L/2, and w^(L/2) = -1 (Work62::wexp)
    L = len(tab)
    return tab[L // 2 + e] if e < L // 2 else (p - tab[e]) % p


def _rev(k, bits):
    return int(format(k, f"0{bits}b")[::-1], 2) if bits else 0


def _dif_c(g, R, tw, p):
    # the constant DIF network of R entries (Work62::dif_c): level l pairs (k, k + e), e = R/2^(l+1),
    # the difference times w_16^((16/R) 2^l r), r = k mod e
    l, e = 0, R // 2
    while e:
        for k0 in range(0, R, 2 * e):
            for r in range(e):
                u, v = g[k0 + r], g[k0 + r + e]
                g[k0 + r], g[k0 + r + e] = (u + v) % p, (u - v) * (tw[8 + ((16 // R) << l) * r] if r else 1) % p
        l, e = l + 1, e // 2


def _dit_c(g, R, twi, p):
    # the constant DIT network (Work62::dit_c): level l pairs (k, k + 2^l), the second entry times
    # w_16^-((8 >> l) r), r = k mod 2^l
    l, e = 0, 1
    while e < R:
        for k0 in range(0, R, 2 * e):
            for r in range(e):
                u, o = g[k0 + r], g[k0 + r + e] * (twi[8 + (8 >> l) * r] if r else 1) % p
                g[k0 + r], g[k0 + r + e] = (u + o) % p, (u - o) % p
        l, e = l + 1, e * 2


def _chain(g, b, p):
    # entry k times b^rev4(k) (Work62::twiddle_chain)
    for k in range(1, 16):
        g[k] = g[k] * pow(b, _rev(k, 4), p) % p


def _dif_pass(x, R, h, tw, p, P):
    # forward pass P (Work62::first_pass for P = 0, fwd_mid, and the shared middle phase for the
    # partial pass): constant network, then entry k times w_2h^(j rev(k))
    L, s = len(x), 2 * h // R
    W = [_wexp(tw, 16 * i, p) for i in range(max(1, L // 16))]
    for k in range(L // R):
        j = k % s
        i0 = (k - j) * 16 + j if R == 16 else k * R
        g = [x[i0 + m * s] for m in range(R)]
        _dif_c(g, R, tw, p)
        if R == 16 and j:
            if P == 0:  # base w^j = w^(j mod 16) W[j / 16]
                _chain(g, _wexp(tw, j & 15, p) * W[j >> 4] % p, p)
            else:  # W[16^(P-1) j rev4(k)]
                for m in range(1, 16):
                    g[m] = g[m] * W[(1 << (4 * (P - 1))) * j * _rev(m, 4)] % p
        for m in range(R):
            x[i0 + m * s] = g[m]


def _dit_pass(x, R, h, twi, p):
    # inverse pass (Work62::inv_mid, last_pass, the shared middle phase): entry k times
    # w_(Rh)^-(j rev(k)), then the constant network
    L = len(x)
    Wi = [_wexp(twi, 16 * i, p) for i in range(max(1, L // 16))]
    for k in range(L // R):
        j = k % h
        i0 = (k - j) * 16 + j if R == 16 else k * R
        g = [x[i0 + m * h] for m in range(R)]
        if R == 16 and j:
            if h == L // 16:  # the last pass: base w^-j
                _chain(g, _wexp(twi, j & 15, p) * Wi[j >> 4] % p, p)
            else:  # Wi[(L / 256h) j rev4(k)]
                for m in range(1, 16):
                    g[m] = g[m] * Wi[(L // (256 * h)) * j * _rev(m, 4)] % p
        _dit_c(g, R, twi, p)
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
    for P, (R, h) in enumerate(fwd):
        _dif_pass(y, R, h, tw, p, P)
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
