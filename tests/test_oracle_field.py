"""Pins the oracle's Goldilocks / F_{p^2} arithmetic (oracle/field.hpp) against Python big
integers and the field's structure (reference README.md:65-66: p = 2^64 - 2^32 + 1,
BinomialExtensionField<Goldilocks, 2> with X^2 = 7)."""
import numpy as np

import oracle

P = 2**64 - 2**32 + 1
EDGE = [0, 1, 2, P - 1, P - 2, 2**32, 2**32 - 1, 2**63, 2**48, P - 2**32, 0xFFFFFFFF00000000]


def _rand(n, seed):
    r = np.random.default_rng(seed)
    v = r.integers(0, 2**63, size=n, dtype=np.uint64) * np.uint64(2) + r.integers(0, 2, size=n, dtype=np.uint64)
    return np.array([int(x) % P for x in v], dtype=np.uint64)


def test_base_ops_match_bigints():
    a = np.concatenate([np.array(EDGE, np.uint64), _rand(2000, 1)])
    b = np.concatenate([np.array(EDGE[::-1], np.uint64), _rand(2000, 2)])
    add, sub, mul = oracle.field_ops(a, b)
    for x, y, s, d, m in zip(a.tolist(), b.tolist(), add.tolist(), sub.tolist(), mul.tolist()):
        assert s == (x + y) % P
        assert d == (x - y) % P
        assert m == (x * y) % P


def test_reduce128_matches_bigints():
    r = np.random.default_rng(3)
    lo = r.integers(0, 2**64 - 1, size=3000, dtype=np.uint64, endpoint=True)
    hi = r.integers(0, 2**64 - 1, size=3000, dtype=np.uint64, endpoint=True)
    lo[:4] = [0, 2**64 - 1, 2**64 - 1, 0]
    hi[:4] = [0, 2**64 - 1, 0, 2**64 - 1]
    out = oracle.reduce128(lo, hi)
    for l, h, o in zip(lo.tolist(), hi.tolist(), out.tolist()):
        assert o == ((h << 64) | l) % P


def _ef_mul_big(a, b):
    # (a0 + a1 X)(b0 + b1 X) in F_p[X]/(X^2 - 7), computed with Python integers
    return ((a[0] * b[0] + 7 * a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)


def test_ext_mul_and_inverse():
    a = np.stack([_rand(500, 4), _rand(500, 5)], axis=1)
    b = np.stack([_rand(500, 6), _rand(500, 7)], axis=1)
    mul, inv = oracle.ef_ops(a, b)
    for x, y, m in zip(a.tolist(), b.tolist(), mul.tolist()):
        assert tuple(m) == _ef_mul_big(x, y)
    # a * a^-1 = 1
    one, _ = oracle.ef_ops(a, inv)
    assert (one[:, 0] == 1).all() and (one[:, 1] == 0).all()


def test_seven_is_a_nonresidue_and_frobenius_conjugates():
    # X^2 = 7 defines a field only if 7 is not a square mod p (Euler's criterion)
    assert pow(7, (P - 1) // 2, P) == P - 1
    # Frobenius: (a0 + a1 X)^p = a0 - a1 X, checked through the oracle's multiplication
    a = np.array([[123456789, 987654321]], np.uint64)
    acc = np.array([[1, 0]], np.uint64)
    base = a.copy()
    e = P
    while e:
        if e & 1:
            acc, _ = oracle.ef_ops(acc, base)
        base, _ = oracle.ef_ops(base, base)
        e >>= 1
    assert acc[0, 0] == 123456789 and int(acc[0, 1]) == P - 987654321


def test_two_adic_generators():
    # closed forms of the small-order roots and primitivity of every g_k
    assert oracle.two_adic_generator(32) == 1753635133440165772
    assert oracle.two_adic_generator(1) == P - 1
    assert oracle.two_adic_generator(2) == 2**48
    assert oracle.two_adic_generator(3) == P - 2**24
    for k in range(1, 33):
        g = oracle.two_adic_generator(k)
        assert pow(g, 2**k, P) == 1
        assert pow(g, 2 ** (k - 1), P) == P - 1
        if k > 1:
            assert pow(g, 2, P) == oracle.two_adic_generator(k - 1)
