"""The device field / curve arithmetic (csrc/field.hpp, csrc/curve.hpp),
compiled for the CPU as libzkhost.so, against the oracle.  CPU only."""
import ctypes as ct
import os
import random

import pytest

from oracle import curve as C
from oracle.field import Q_MOD, R_MOD, from_mont, to_mont, to_words, from_words

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "pyzkp_amd", "libzkhost.so")


@pytest.fixture(scope="module")
def host():
    if not os.path.exists(LIB):
        pytest.skip("libzkhost.so not built")
    lib = ct.CDLL(LIB)
    lib.zkhost_field.argtypes = [ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_uint64]
    lib.zkhost_xyzz.argtypes = [ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_uint32, ct.c_void_p]
    return lib


def _arr(vals):
    a = (ct.c_uint64 * (4 * len(vals)))()
    for i, v in enumerate(vals):
        for j, w in enumerate(to_words(v)):
            a[4 * i + j] = w
    return a


def _vals(a, n):
    return [from_words(a[4 * i:4 * i + 4]) for i in range(n)]


@pytest.mark.parametrize("field,p", [(0, R_MOD), (1, Q_MOD)])
def test_field_ops(host, field, p):
    rng = random.Random(field)
    n = 200
    xs = [rng.randrange(p) for _ in range(n - 6)] + [0, 1, p - 1, p - 2, 2, (p + 1) // 2]
    ys = [rng.randrange(p) for _ in range(n - 6)] + [p - 1, p - 1, p - 1, 1, 0, (p - 1) // 2]
    out = (ct.c_uint64 * (4 * n))()
    mx, my = [to_mont(x, p) for x in xs], [to_mont(y, p) for y in ys]
    host.zkhost_field(field, 0, _arr(mx), _arr(my), out, n)
    assert [from_mont(v, p) for v in _vals(out, n)] == [x * y % p for x, y in zip(xs, ys)]
    host.zkhost_field(field, 1, _arr(xs), _arr(ys), out, n)
    assert _vals(out, n) == [(x + y) % p for x, y in zip(xs, ys)]
    host.zkhost_field(field, 2, _arr(xs), _arr(ys), out, n)
    assert _vals(out, n) == [(x - y) % p for x, y in zip(xs, ys)]
    host.zkhost_field(field, 3, _arr(xs), None, out, n)
    assert _vals(out, n) == [(-x) % p for x in xs]
    # to_mont accepts any 256-bit input and reduces it
    big = [rng.randrange(1 << 256) for _ in range(n - 2)] + [(1 << 256) - 1, p]
    host.zkhost_field(field, 4, _arr(big), None, out, n)
    assert _vals(out, n) == [to_mont(x % p, p) for x in big]
    host.zkhost_field(field, 5, _arr(mx), None, out, n)
    assert _vals(out, n) == xs


def _xyzz_from_affine(a):
    if a is None:
        return [to_mont(1, Q_MOD), to_mont(1, Q_MOD), 0, 0]
    return [to_mont(a[0], Q_MOD), to_mont(a[1], Q_MOD), to_mont(1, Q_MOD), to_mont(1, Q_MOD)]


def _xyzz_to_affine(v):
    X, Y, ZZ, ZZZ = (from_mont(x, Q_MOD) for x in v)
    if ZZ == 0:
        return None
    assert pow(ZZ, 3, Q_MOD) == pow(ZZZ, 2, Q_MOD)
    return (X * pow(ZZ, -1, Q_MOD) % Q_MOD, Y * pow(ZZZ, -1, Q_MOD) % Q_MOD)


def _scaled(a, lam):
    """Same point in XYZZ with ZZ = lam^2, ZZZ = lam^3 (non-trivial Z)."""
    if a is None:
        return _xyzz_from_affine(None)
    x, y = a
    return [to_mont(x * lam * lam, Q_MOD), to_mont(y * pow(lam, 3, Q_MOD), Q_MOD),
            to_mont(lam * lam, Q_MOD), to_mont(pow(lam, 3, Q_MOD), Q_MOD)]


def _call(host, op, p, q, k=0):
    out = (ct.c_uint64 * 16)()
    host.zkhost_xyzz(op, _arr(p), _arr(q) if q is not None else None, k, out)
    return _xyzz_to_affine(_vals(out, 4))


def test_xyzz_ops(host):
    rng = random.Random(21)
    G = C.G1_GENERATOR
    for _ in range(12):
        a, b = rng.randrange(1, R_MOD), rng.randrange(1, R_MOD)
        P, Qp = C.g1_mul_affine(G, a), C.g1_mul_affine(G, b)
        lp, lq = rng.randrange(2, Q_MOD), rng.randrange(2, Q_MOD)
        sp, sq = _scaled(P, lp), _scaled(Qp, lq)
        assert _call(host, 0, sp, sq) == C.aff_add(P, Qp)
        assert _call(host, 0, sp, _scaled(P, lq)) == C.aff_add(P, P)          # P + P
        assert _call(host, 0, sp, _scaled(C.aff_neg(P), lq)) is None          # P - P
        assert _call(host, 0, sp, _xyzz_from_affine(None)) == P               # P + O
        assert _call(host, 0, _xyzz_from_affine(None), sq) == Qp
        assert _call(host, 1, sp, None) == C.aff_add(P, P)
        aq = [to_mont(Qp[0], Q_MOD), to_mont(Qp[1], Q_MOD), 0, 0]
        assert _call(host, 2, sp, aq) == C.aff_add(P, Qp)                     # madd
        ap = [to_mont(P[0], Q_MOD), to_mont(P[1], Q_MOD), 0, 0]
        assert _call(host, 2, sp, ap) == C.aff_add(P, P)                      # madd doubling
        an = [to_mont(P[0], Q_MOD), to_mont((-P[1]) % Q_MOD, Q_MOD), 0, 0]
        assert _call(host, 2, sp, an) is None                                 # madd to identity
        assert _call(host, 2, _xyzz_from_affine(None), aq) == Qp              # madd into identity
        k = rng.randrange(0, 1 << 20)
        assert _call(host, 3, sp, None, k) == C.g1_mul_affine(P, k)
