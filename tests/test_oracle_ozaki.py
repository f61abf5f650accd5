"""Pins the Ozaki split DGEMM oracle against exact arithmetic."""
from fractions import Fraction

import numpy as np
import pytest

import oracle


def wide(rng, shape):
    return np.ldexp(rng.uniform(-1, 1, shape), rng.integers(-20, 21, shape))


@pytest.mark.parametrize("bits", [oracle.OZ_INT8_BITS, oracle.OZ_BF16_BITS])
def test_slices_reconstruct_exactly(bits):
    rng = np.random.default_rng(0)
    x = wide(rng, (7, 33))
    e = oracle.oz_exponents(x, 1)
    s = 6
    d = oracle.oz_slices(x, e, 1, s, bits)
    assert np.all(np.abs(d) <= (1 << bits) - 1)
    assert np.all(d == np.trunc(d))
    for i in range(7):
        amax = np.max(np.abs(x[i]))
        assert amax < 2.0 ** e[i] and amax >= 2.0 ** (e[i] - 1)
        for k in range(33):
            approx = sum(Fraction(int(d[t, i, k])) * Fraction(2) ** (int(e[i]) - bits * (t + 1)) for t in range(s))
            resid = Fraction(x[i, k]) - approx
            assert abs(resid) < Fraction(2) ** (int(e[i]) - bits * s)
            # truncation: the residual has the sign of x (or is zero)
            assert resid == 0 or (resid > 0) == (x[i, k] > 0)


def test_zero_rows_and_columns():
    a = np.zeros((3, 5)); a[1] = [1, -2, 3, 0.5, 0]
    b = np.ones((5, 4)); b[:, 2] = 0
    c = oracle.ozaki_dgemm(a, b, 4)
    assert np.array_equal(c, a @ b)


@pytest.mark.parametrize("bits,s", [(7, 1), (7, 2), (5, 2), (5, 3)])
def test_exact_for_short_integers(bits, s):
    # entries with at most bits*s significant bits below the row scale come out exact
    rng = np.random.default_rng(1)
    lim = 2 ** (bits * s - 1)
    a = rng.integers(-lim, lim, (9, 40)).astype(float)
    b = rng.integers(-lim, lim, (40, 7)).astype(float)
    a[:, 0] = lim - 1   # pin each row's scale
    b[0, :] = -(lim - 1)
    got = oracle.ozaki_dgemm(a, b, s, bits=bits)
    # the product set t+u <= s+1 drops only terms with t,u > ... ; with s=1/2 digits
    # every pair satisfies t+u <= s+1 only for s=1, so compare with the bound instead
    exact = a @ b   # exact: integers below 2^53
    if s == 1:
        assert np.array_equal(got, exact)
    else:
        e = oracle.oz_exponents(a, 1); f = oracle.oz_exponents(b, 0)
        bound = 40 * (s + 3) * 2.0 ** (-bits * s) * np.ldexp(1.0, e.reshape(-1, 1) + f.reshape(1, -1))
        assert np.all(np.abs(got - exact) <= bound)


@pytest.mark.parametrize("s", [3, 5, 8])
def test_error_within_a_priori_bound(s):
    rng = np.random.default_rng(s)
    a = wide(rng, (6, 50))
    b = wide(rng, (50, 5))
    got = oracle.ozaki_dgemm(a, b, s)
    e = oracle.oz_exponents(a, 1); f = oracle.oz_exponents(b, 0)
    for i in range(6):
        for j in range(5):
            exact = sum(Fraction(a[i, k]) * Fraction(b[k, j]) for k in range(50))
            bound = Fraction(50 * (s + 3)) * Fraction(2) ** (int(e[i] + f[j]) - 7 * s)
            # plus the FP64 rounding of the recombination (a few ulps of the result)
            slack = Fraction(abs(float(exact))) * Fraction(2) ** -50 + Fraction(2) ** (int(e[i] + f[j]) - 60)
            assert abs(Fraction(got[i, j]) - exact) <= bound + slack


def test_error_decreases_with_slices():
    rng = np.random.default_rng(3)
    a = rng.uniform(-1, 1, (32, 4096))
    b = rng.uniform(-1, 1, (4096, 32))
    ref = a @ b
    errs = [oracle.rel_frobenius(oracle.ozaki_dgemm(a, b, s), ref) for s in range(3, 9)]
    assert all(errs[i + 1] < errs[i] / 16 for i in range(len(errs) - 2))
    assert errs[-1] < 2e-15


def test_flush_schedule_long_k():
    # K = 70000: one INT8 slice product already fills half of int32, so every
    # product is flushed on its own; the result must still be exact for short integers
    rng = np.random.default_rng(4)
    a = rng.integers(-100, 101, (4, 70000)).astype(float)
    b = rng.integers(-100, 101, (70000, 3)).astype(float)
    assert oracle.oz_flush_every(70000, 7) == 1
    got = oracle.ozaki_dgemm(a, b, 2)   # 14 bits >= 8 bits of the entries: exact digits
    assert np.array_equal(got, a @ b)


def test_bf16_chunking():
    assert oracle.oz_kchunk(16384, 5) == 16384
    assert oracle.oz_kchunk(20000, 5) == 17408
    rng = np.random.default_rng(5)
    a = rng.integers(-30, 31, (3, 20000)).astype(float)
    b = rng.integers(-30, 31, (20000, 2)).astype(float)
    got = oracle.ozaki_dgemm(a, b, 2, bits=5)   # 10 bits >= 5 bits of the entries
    assert np.array_equal(got, a @ b)


# ---------------------------------------------------------------- Ozaki scheme II (CRT)

def test_crt_moduli_pairwise_coprime():
    from math import gcd
    ps = oracle.CRT_MODULI[:16]
    assert all(p <= 256 for p in ps)
    assert all(gcd(a, b) == 1 for i, a in enumerate(ps) for b in ps[i + 1:])


@pytest.mark.parametrize("nmod", [2, 5, 10, 16])
def test_crt_reconstruct_exact(nmod):
    import random
    P = 1
    for p in oracle.CRT_MODULI[:nmod]:
        P *= p
    rnd = random.Random(nmod)
    vals = [rnd.randrange(-(P // 4), P // 4) for _ in range(200)] + [0, 1, -1, P // 4 - 1, -(P // 4)]
    res = [np.array([v % p for v in vals], dtype=np.int64) for p in oracle.CRT_MODULI[:nmod]]
    got = oracle.crt_reconstruct(res, nmod)
    assert list(got) == vals


def test_crt_q_bound():
    for K in (1, 100, 4096, 16384, 100000):
        for nm in range(2, 17):
            q = oracle.crt_q(K, nm)
            P = 1
            for p in oracle.CRT_MODULI[:nm]:
                P *= p
            if q > 0:
                assert K * 4 ** q * 4 <= P
                assert q == 53 or K * 4 ** (q + 1) * 4 > P


def test_crt_scale_rounds_to_nearest_even():
    # row exponent 0 and q = 3: the scaled value is the entry itself
    X = np.array([[0.7, -0.7, 0.5, 1.5, 2.5, -2.5, 0.49, 7.5]]) * 2.0 ** -3
    out = oracle.crt_scale(X, np.array([0]), 1, 3)
    assert out.tolist() == [[1, -1, 0, 2, 2, -2, 0, 8]]
    # column scaling (axis 0) uses one exponent per column
    Y = np.array([[0.3, 3.0], [-0.3, -1.75]])
    out = oracle.crt_scale(Y, np.array([-1, 2]), 0, 2)  # column 0 by 2^3, column 1 by 2^0
    assert out.tolist() == [[2, 3], [-2, -2]]


def test_crt_exact_for_short_integers():
    rng = np.random.default_rng(11)
    a = rng.integers(-1000, 1000, (7, 50)).astype(float)
    b = rng.integers(-1000, 1000, (50, 9)).astype(float)
    assert np.array_equal(oracle.ozaki2_dgemm(a, b, 16), a @ b)


@pytest.mark.parametrize("nmod", [8, 12, 16])
def test_crt_error_within_bound(nmod):
    rng = np.random.default_rng(nmod)
    a = wide(rng, (5, 40))
    b = wide(rng, (40, 6))
    got = oracle.ozaki2_dgemm(a, b, nmod)
    q = oracle.crt_q(40, nmod)
    e = oracle.oz_exponents(a, 1)
    f = oracle.oz_exponents(b, 0)
    for i in range(5):
        for j in range(6):
            exact = sum(Fraction(a[i, k]) * Fraction(b[k, j]) for k in range(40))
            bound = Fraction(40 * 2) * Fraction(2) ** (int(e[i] + f[j]) - q)
            slack = Fraction(abs(float(exact))) * Fraction(2) ** -52
            assert abs(Fraction(got[i, j]) - exact) <= bound + slack
