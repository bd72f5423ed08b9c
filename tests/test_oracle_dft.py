"""Pins the oracle DFT (oracle/dft.hpp) to the closed form out[j] = sum_k c_k g^(jk) used by the
reference's RS encoding (pcs/src/ligesis.rs:222-235) via an O(n^2) big-integer evaluation."""
import numpy as np
import pytest

import oracle
from ligesis_amd import synth

P = synth.P


def _dft_brute(c):
    n = len(c)
    k = n.bit_length() - 1
    g = oracle.two_adic_generator(k)
    return [sum(int(c[t]) * pow(g, j * t, P) for t in range(n)) % P for j in range(n)]


@pytest.mark.parametrize("logn", [0, 1, 2, 3, 5, 7])
def test_dft_closed_form(logn):
    c = synth.field_elems(logn + 10, 1 << logn)
    assert oracle.dft(c).tolist() == _dft_brute(c.tolist())


def test_dft_ext_is_componentwise():
    c = synth.ext_elems(9, 64)
    out = oracle.dft_ext(c)
    assert out[:, 0].tolist() == oracle.dft(c[:, 0]).tolist()
    assert out[:, 1].tolist() == oracle.dft(c[:, 1]).tolist()


def test_rs_encode_is_evaluation_of_row_polynomial():
    # row r encodes to r(g^j) for j < 2n, g of order 2n; the first n evals at even j are the
    # size-n DFT (the LDE contains the plain transform)
    n = 32
    f = synth.field_elems(5, 4 * n)
    enc = oracle.rs_encode_rows(f, 4, n)
    g = oracle.two_adic_generator(6)
    for r in range(4):
        row = [int(x) for x in f[r * n:(r + 1) * n]]
        for j in (0, 1, 7, 33, 63):
            assert int(enc[r, j]) == sum(row[t] * pow(g, j * t, P) for t in range(n)) % P
        assert enc[r, ::2].tolist() == oracle.dft(f[r * n:(r + 1) * n]).tolist()


def test_dft_inverse_relation():
    # DFT(DFT(x))[j] = n * x[-j mod n]
    n = 256
    x = synth.field_elems(77, n)
    y = oracle.dft(oracle.dft(x))
    for j in range(n):
        assert int(y[j]) == (n * int(x[(-j) % n])) % P
