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


def test_crt_Q_bound():
    import math
    for K in (1, 100, 4096, 16384, 100000):
        for nm in range(2, 17):
            P = math.prod(oracle.CRT_MODULI[:nm])
            Q = oracle.crt_Q(K, nm)
            rk = math.isqrt(K) + (0 if math.isqrt(K) ** 2 == K else 1)
            fits = lambda q: 4 * (2 ** (2 * q) + rk * 2 ** q + K) <= P
            if Q is None:
                assert not fits(1)
            else:
                assert fits(Q) and not fits(Q + 1)


def test_crt_bound_sums_known_answers():
    # row max 0.75 = 0.75 * 2^0 (e = 0): 256*0.75 = 192 -> 192^2 * 2^16
    X = np.array([[0.75, -0.375, 0.0, 2.0 ** -9 * 0.6]])
    e = oracle.oz_exponents(X, 1)
    assert e.tolist() == [0]
    S = oracle.crt_bound_sums(X, e, 1)
    # -0.375 = 0.75 * 2^-1: d = 1 -> 192^2 * 2^14; 0.6 * 2^-9: m = 0.6, ex = -9, d = 9 ->
    # ceil(153.6) = 154, ceil(154^2 / 2^2); zero -> 0
    assert S.tolist() == [192 ** 2 * 2 ** 16 + 192 ** 2 * 2 ** 14 + -(-154 ** 2 // 4)]
    # d = 10: ceil(M^2 / 2^4); d = 30: 1
    Y = np.array([[0.5, 0.5 * 2.0 ** -10], [0.5, 0.99 * 2.0 ** -30]]).T   # columns
    f = oracle.oz_exponents(Y, 0)
    S2 = oracle.crt_bound_sums(Y, f, 0)
    assert S2.tolist() == [128 ** 2 * 2 ** 16 + -(-128 ** 2 // 2 ** 4), 128 ** 2 * 2 ** 16 + 1]


def test_crt_bound_sums_is_a_tight_upper_bound():
    rng = np.random.default_rng(3)
    X = wide(rng, (20, 300))
    e = oracle.oz_exponents(X, 1)
    S = oracle.crt_bound_sums(X, e, 1)
    for i in range(20):
        exact = sum((Fraction(float(x)) * Fraction(2) ** (-int(e[i]))) ** 2 for x in X[i]) * 2 ** 32
        assert Fraction(int(S[i])) >= exact
        assert Fraction(int(S[i])) <= exact * Fraction(102, 100) + 300


def test_crt_line_bounds_one_chunk_is_the_line_sum():
    # K <= the chunk: the chunk exponent is the line's and nothing is rescaled
    rng = np.random.default_rng(4)
    X = wide(rng, (9, oracle.CRT_KCHUNK))
    X[3] = 0.0
    e, S = oracle.crt_line_bounds(X, 1)
    assert e.tolist() == oracle.oz_exponents(X, 1).tolist()
    assert S.tolist() == oracle.crt_bound_sums(X, oracle.oz_exponents(X, 1), 1).tolist()
    e0, S0 = oracle.crt_line_bounds(X.T, 0)
    assert e0.tolist() == e.tolist() and S0.tolist() == S.tolist()


def test_crt_line_bounds_known_answers():
    kc = oracle.CRT_KCHUNK
    # chunk 0: 0.75 (e_c = 0); chunk 1: 0.1 = 0.8 * 2^-3 (e_c = -3): ceil(204.8) = 205,
    # 205^2 2^16 rescaled by 4^-3 exactly
    x = np.zeros((1, 2 * kc))
    x[0, 5] = 0.75
    x[0, kc + 7] = 0.1
    e, S = oracle.crt_line_bounds(x, 1)
    assert e.tolist() == [0] and S.tolist() == [192 ** 2 * 2 ** 16 + 205 ** 2 * 2 ** 10]
    # chunk 1 holds two entries 0.6 * 2^-40 (e_c = -40): 2 * 154^2 * 2^16 / 4^40 rounds up to 1
    # (the per-entry bound at the line exponent counts each as 1)
    x[0, kc + 7] = 0.6 * 2.0 ** -40
    x[0, kc + 9] = -0.6 * 2.0 ** -40
    e, S = oracle.crt_line_bounds(x, 1)
    assert S.tolist() == [192 ** 2 * 2 ** 16 + 1]
    assert oracle.crt_bound_sums(x, e, 1).tolist() == [192 ** 2 * 2 ** 16 + 2]
    # a zero chunk does not set the line exponent; an all-zero line is (0, 0)
    y = np.zeros((2, 3 * kc))
    y[0, 2 * kc + 1] = 0.3 * 2.0 ** -12
    e, S = oracle.crt_line_bounds(y, 1)
    _, ex = np.frexp(0.3 * 2.0 ** -12)
    assert e.tolist() == [ex, 0] and S.tolist() == [int(np.ceil(0.6 * 256)) ** 2 * 2 ** 16, 0]


def test_crt_line_bounds_tight_upper_bound():
    rng = np.random.default_rng(8)
    X = wide(rng, (12, 300))
    X[:, 64:128] *= 2.0 ** -30       # a chunk far below the line maximum
    e, S = oracle.crt_line_bounds(X, 1)
    assert e.tolist() == oracle.oz_exponents(X, 1).tolist()
    nch = -(-300 // oracle.CRT_KCHUNK)
    for i in range(12):
        exact = sum((Fraction(float(x)) * Fraction(2) ** (-int(e[i]))) ** 2 for x in X[i]) * 2 ** 32
        assert Fraction(int(S[i])) >= exact
        assert Fraction(int(S[i])) <= exact * Fraction(102, 100) + nch


@pytest.mark.parametrize("nmod", [6, 10, 14, 16])
def test_crt_shifts_bound_the_product(nmod):
    # |A'B'| <= P/4 with exact integers, and each shift is the largest the norm bound allows
    import math
    rng = np.random.default_rng(nmod)
    K = 64
    a = wide(rng, (6, K))
    b = rng.uniform(-1, 1, (K, 5))
    Q = oracle.crt_Q(K, nmod)
    ea = oracle.crt_shifts(*oracle.crt_line_bounds(a, 1), Q)
    fb = oracle.crt_shifts(*oracle.crt_line_bounds(b, 0), Q)
    A = oracle.crt_scale(a, ea, 1, Q)
    B = oracle.crt_scale(b, fb, 0, Q)
    P = math.prod(oracle.CRT_MODULI[:nmod])
    V = [[sum(int(A[i, k]) * int(B[k, j]) for k in range(K)) for j in range(5)] for i in range(6)]
    assert max(abs(v) for r in V for v in r) <= P // 4
    e, S = oracle.crt_line_bounds(a, 1)
    for i in range(6):
        sig = Q - int(ea[i])
        assert 2 ** (2 * (sig + int(e[i]))) * int(S[i]) <= 2 ** (2 * Q + 32) or sig == 53 - int(e[i])
        assert sig == 53 - int(e[i]) or 2 ** (2 * (sig + 1 + int(e[i]))) * int(S[i]) > 2 ** (2 * Q + 32)


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
    # |a - a' 2^-s| <= 2^(-s-1) per entry with s = Q - e', so per product
    # term the error is at most 2^(e+f'-Q-1) + 2^(f+e'-Q-1) + 2^(e'+f'-2Q-2)
    rng = np.random.default_rng(nmod)
    a = wide(rng, (5, 40))
    b = wide(rng, (40, 6))
    got = oracle.ozaki2_dgemm(a, b, nmod)
    Q = oracle.crt_Q(40, nmod)
    e = oracle.oz_exponents(a, 1)
    f = oracle.oz_exponents(b, 0)
    ea = oracle.crt_shifts(*oracle.crt_line_bounds(a, 1), Q)
    fb = oracle.crt_shifts(*oracle.crt_line_bounds(b, 0), Q)
    for i in range(5):
        for j in range(6):
            exact = sum(Fraction(a[i, k]) * Fraction(b[k, j]) for k in range(40))
            two = Fraction(2)
            term = (two ** (int(e[i]) + int(fb[j]) - Q - 1) + two ** (int(f[j]) + int(ea[i]) - Q - 1)
                    + two ** (int(ea[i]) + int(fb[j]) - 2 * Q - 2))
            slack = Fraction(abs(float(exact))) * two ** -52
            assert abs(Fraction(got[i, j]) - exact) <= 40 * term + slack