"""Pins the CPU oracle: every restated function is checked against a
published known answer, the reference's own compiled code, brute-force
enumeration or a closed form that a plausible slip (swapped index, dropped
term, off-by-one) would break."""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle

# ------------------------------------------------------------------ SplitMix64

M64 = (1 << 64) - 1


def py_splitmix(seed, n):
    s, out = seed, []
    for _ in range(n):
        s = (s + 0x9E3779B97F4A7C15) & M64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        out.append(z ^ (z >> 31))
    return out


def test_splitmix64_published_vectors():
    # Vigna's splitmix64.c reference outputs
    assert oracle.splitmix64_stream(0, 3) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    assert oracle.splitmix64_stream(1234567, 5) == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                                    4593380528125082431, 16408922859458223821]
    assert py_splitmix(1234567, 5) == oracle.splitmix64_stream(1234567, 5)


@pytest.mark.parametrize("seed", [0, 1, 2, 3, 0xDEADBEEF])
def test_generate_matches_exact_rationals(seed):
    n = 257
    zs = py_splitmix(seed, n)
    ze = py_splitmix(seed ^ 0xD1B54A32D192ED03, n)
    u = oracle.generate(seed, 0, n)
    w = oracle.generate(seed, 1, n)
    z01 = oracle.generate(seed, 2, n)
    for i in range(n):
        frac = Fraction(zs[i] >> 11, 1 << 53)
        assert Fraction(z01[i]) == frac
        assert Fraction(u[i]) == 2 * frac - 1
        e = (ze[i] >> 32) % 41 - 20
        assert Fraction(w[i]) == (2 * frac - 1) * Fraction(2) ** e
    assert np.all(u >= -1.0) and np.all(u < 1.0)


def nearest_f32_brute(x):
    """The FP32 value nearest to the Fraction x, ties to the even significand,
    by search over the two FP32 neighbours of a truncated candidate."""
    lo = np.float32(float(x))  # any FP32 value within a few ulps
    cands = {np.float32(lo)}
    for _ in range(3):
        cands |= {np.nextafter(c, np.float32(np.inf)) for c in list(cands)}
        cands |= {np.nextafter(c, np.float32(-np.inf)) for c in list(cands)}
    best = min(abs(Fraction(float(c)) - x) for c in cands)
    ties = [c for c in cands if abs(Fraction(float(c)) - x) == best]
    if len(ties) == 1:
        return ties[0]
    return [c for c in ties if (int(np.array(c).view(np.uint32)) & 1) == 0][0]


@pytest.mark.parametrize("dist", [0, 1, 2])
def test_generate_f32_is_nearest_even(dist):
    got = oracle.generate_f32(4242, dist, 300)
    exact = oracle.generate(4242, dist, 300)
    for g, x in zip(got, exact):
        assert g.view(np.uint32) == nearest_f32_brute(Fraction(float(x))).view(np.uint32)


def test_generate_f32_ties_known_answers():
    # a halfway case rounds to the even neighbour, a value past it away
    assert nearest_f32_brute(Fraction(1) + Fraction(1, 2 ** 24)) == np.float32(1.0)
    assert nearest_f32_brute(Fraction(1) + Fraction(3, 2 ** 24)) == np.float32(1.0 + 2.0 ** -22)
    assert np.float32(np.float64(1.0 + 2.0 ** -24)) == np.float32(1.0)
    assert np.float32(np.float64(1.0 + 3 * 2.0 ** -24)) == np.float32(1.0 + 2.0 ** -22)


def test_generate_golden_fixture():
    import json, os
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "generate.json")))
    for case in fx["cases"]:
        got = oracle.generate(case["seed"], case["dist"], len(case["bits"]))
        assert [int(v) for v in got.view(np.uint64)] == [int(b, 16) for b in case["bits"]]


# ------------------------------------------------------------------ numroc / grids

def numroc_brute(n, nb, iproc, isrc, nprocs):
    # block b of the dimension lives on process (isrc + b) mod nprocs
    return sum(1 for g in range(n) if ((g // nb) + isrc) % nprocs == iproc)


@pytest.mark.parametrize("n,nb,nprocs", [(0, 3, 2), (1, 1, 1), (23, 3, 7), (17, 2, 5), (16384, 512, 4), (100, 7, 3),
                                         (5, 8, 2), (64, 8, 4)])
def test_numroc_brute_force(n, nb, nprocs):
    tot = 0
    for ip in range(nprocs):
        for isrc in range(nprocs):
            assert oracle.numroc(n, nb, ip, isrc, nprocs) == numroc_brute(n, nb, ip, isrc, nprocs)
        tot += oracle.numroc(n, nb, ip, 0, nprocs)
    assert tot == n


def test_numroc_config2():
    assert [oracle.numroc(16384, 512, p, 0, 4) for p in range(4)] == [4096] * 4


def grid_brute(size):
    best = 1
    for d in range(1, size + 1):
        if size % d == 0 and d * d <= size:
            best = d
    return size // best, best


def test_virtual_grid():
    for size in range(1, 200):
        assert oracle.virtual_grid(size) == grid_brute(size)
    assert oracle.virtual_grid(35) == (7, 5)
    assert oracle.virtual_grid(16) == (4, 4)
    assert oracle.virtual_grid(32) == (8, 4)


def test_block_splits_remainder_to_last():
    assert oracle.block_splits(20000, 7) == [0, 2857, 5714, 8571, 11428, 14285, 17142, 20000]
    assert oracle.block_splits(20000, 5) == [0, 4000, 8000, 12000, 16000, 20000]
    assert oracle.block_splits(2048, 2) == [0, 1024, 2048]


# ------------------------------------------------------------------ layouts

def test_pack_block2d_definition():
    rows, cols, ldg = 23, 17, 25
    ro, co = oracle.block_splits(rows, 7), oracle.block_splits(cols, 5)
    G = np.arange(ldg * cols, dtype=np.float64)
    P = oracle.pack_block2d(G, ldg, rows, cols, ro, co)
    # definition: rank r = i*5 + j holds tile (i, j) column-major, ranks concatenated
    expect = []
    for i in range(7):
        for j in range(5):
            for c in range(co[j], co[j + 1]):
                for r in range(ro[i], ro[i + 1]):
                    expect.append(G[r + c * ldg])
    assert np.array_equal(P, np.array(expect))
    G2 = np.full(ldg * cols, -1.0)
    oracle.unpack_block2d(P, ro, co, G2, ldg, rows, cols)
    g, g2 = G.reshape(cols, ldg), G2.reshape(cols, ldg)
    assert np.array_equal(g2[:, :rows], g[:, :rows]) and np.all(g2[:, rows:] == -1.0)


@pytest.mark.parametrize("rows,cols,mb,nb,P,Q", [(23, 17, 3, 2, 7, 5), (23, 17, 3, 2, 2, 3), (16, 16, 4, 4, 2, 2),
                                                 (10, 7, 16, 16, 3, 2), (31, 29, 5, 6, 3, 4)])
def test_pack_block_cyclic_definition(rows, cols, mb, nb, P, Q):
    ldg = rows + 2
    G = np.arange(ldg * cols, dtype=np.float64)
    Pk = oracle.pack_block_cyclic(G, ldg, rows, cols, mb, nb, P, Q)
    lr = [sum(1 for I in range(rows) if (I // mb) % P == p) for p in range(P)]
    lc = [sum(1 for J in range(cols) if (J // nb) % Q == q) for q in range(Q)]
    base, acc = {}, 0
    for r in range(P * Q):
        base[r] = acc
        acc += lr[r // Q] * lc[r % Q]
    assert Pk.size == acc
    for J in range(cols):
        for I in range(rows):
            pr, pc = (I // mb) % P, (J // nb) % Q
            li = (I // mb // P) * mb + I % mb
            lj = (J // nb // Q) * nb + J % nb
            assert Pk[base[pr * Q + pc] + li + lj * lr[pr]] == G[I + J * ldg]
    G2 = np.full(ldg * cols, -1.0)
    oracle.unpack_block_cyclic(Pk, mb, nb, P, Q, G2, ldg, rows, cols)
    assert np.array_equal(G2.reshape(cols, ldg)[:, :rows], G.reshape(cols, ldg)[:, :rows])


def test_layer_reduce_order():
    C = np.array([1.0, 1e16])
    part = np.array([1e16, 1.0, -1e16, -1e16])
    oracle.layer_reduce(C, part, 2, 2)
    # (1 + 1e16) - 1e16 == 0 in layer order; (1e16 + 1) - 1e16 == 0
    assert list(C) == [0.0, 0.0]


# ------------------------------------------------------------------ Cannon schedule

def triangular_closed_form(p, c, i, j, k, r):
    return (i + j - k * (p // c) - r * (r + 1) // 2) % p


@pytest.mark.parametrize("p,c", [(1, 1), (2, 1), (2, 2), (4, 1), (4, 2), (4, 4), (8, 1), (8, 2), (8, 4), (16, 4),
                                 (3, 1), (3, 3), (5, 1), (6, 2), (6, 3), (7, 1), (9, 3)])
def test_cannon_simulation(p, c):
    mA, mB = oracle.cannon_simulate(p, c)
    steps = p // c
    # the A tile's column index always equals the B tile's row index
    assert np.array_equal(mA, mB)
    for k in range(c):
        for i in range(p):
            for j in range(p):
                for r in range(steps):
                    assert mA[k, i, j, r] == triangular_closed_form(p, c, i, j, k, r)
    covered = all(sorted(mA[:, i, j, :].ravel()) == list(range(p)) for i in range(p) for j in range(p))
    offs = sorted(((-k * steps - r * (r + 1) // 2) % p) for k in range(c) for r in range(steps))
    assert covered == (offs == list(range(p)))
    if p & (p - 1) == 0 or steps <= 2:
        assert covered
    if (p, c) in [(3, 1), (5, 1), (7, 1), (6, 2)]:
        assert not covered


# ------------------------------------------------------------------ DGEMM

@pytest.mark.parametrize("opA,opB", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_dgemm_loops_vs_blas(opA, opB):
    rng = np.random.default_rng(0)
    m, n, k = 13, 9, 21
    a = rng.uniform(-1, 1, (m, k) if opA == "N" else (k, m))
    b = rng.uniform(-1, 1, (k, n) if opB == "N" else (n, k))
    c = rng.uniform(-1, 1, (m, n))
    C = np.asfortranarray(c).ravel(order="F").copy()
    oracle.dgemm_loops(opA, opB, m, n, k, 1.5, np.asfortranarray(a).ravel(order="F"), a.shape[0],
                       np.asfortranarray(b).ravel(order="F"), b.shape[0], -0.5, C, m)
    ref = oracle.host_dgemm(opA, opB, 1.5, a, b, -0.5, c)
    assert oracle.rel_frobenius(C.reshape(n, m).T, ref) < 1e-15


def test_dgemm_loops_beta_zero_ignores_nan():
    A = np.ones(4)
    B = np.ones(4)
    C = np.full(4, np.nan)
    oracle.dgemm_loops("N", "N", 2, 2, 1, 1.0, A, 2, B, 1, 0.0, C, 2)
    assert list(C) == [1.0, 1.0, 1.0, 1.0]


@pytest.mark.parametrize("opA,opB,M,N,K,p,c,std", [("N", "N", 8, 8, 8, 2, 1, False), ("N", "T", 16, 8, 24, 4, 2, False),
                                                   ("T", "T", 9, 6, 12, 3, 1, True), ("T", "N", 8, 8, 16, 4, 4, False)])
def test_cannon25d_restatement(opA, opB, M, N, K, p, c, std):
    rng = np.random.default_rng(1)
    ra, ca = (M, K) if opA == "N" else (K, M)
    rb, cb = (K, N) if opB == "N" else (N, K)
    A = rng.integers(-3, 4, ra * ca).astype(float)
    B = rng.integers(-3, 4, rb * cb).astype(float)
    C = rng.integers(-3, 4, M * N).astype(float)
    sp = oracle.block_splits
    Ap = oracle.pack_block2d(A, ra, ra, ca, sp(ra, p), sp(ca, p))
    Bp = oracle.pack_block2d(B, rb, rb, cb, sp(rb, p), sp(cb, p))
    Cp = oracle.pack_block2d(C, M, M, N, sp(M, p), sp(N, p))
    oracle.cannon25d(opA, opB, M, N, K, 3.0, Ap, Bp, -2.0, Cp, p, c, std)
    G = np.zeros(M * N)
    oracle.unpack_block2d(Cp, sp(M, p), sp(N, p), G, M, M, N)
    ref = oracle.host_dgemm(opA, opB, 3.0, A.reshape(ca, ra).T, B.reshape(cb, rb).T, -2.0, C.reshape(N, M).T)
    assert np.array_equal(G.reshape(N, M).T, ref)


def test_summa_restatement():
    rng = np.random.default_rng(2)
    M, N, K, mb, nb, kb, P, Q = 23, 17, 19, 3, 2, 4, 3, 2
    A = rng.integers(-3, 4, M * K).astype(float)
    B = rng.integers(-3, 4, K * N).astype(float)
    C = rng.integers(-3, 4, M * N).astype(float)
    Ap = oracle.pack_block_cyclic(A, M, M, K, mb, kb, P, Q)
    Bp = oracle.pack_block_cyclic(B, K, K, N, kb, nb, P, Q)
    Cp = oracle.pack_block_cyclic(C, M, M, N, mb, nb, P, Q)
    oracle.summa(M, N, K, 1.0, Ap, Bp, 0.5, Cp, mb, nb, kb, P, Q)
    G = np.zeros(M * N)
    oracle.unpack_block_cyclic(Cp, mb, nb, P, Q, G, M, M, N)
    ref = oracle.host_dgemm("N", "N", 1.0, A.reshape(K, M).T, B.reshape(N, K).T, 0.5, C.reshape(N, M).T)
    assert np.array_equal(G.reshape(N, M).T, ref)


@pytest.mark.parametrize("opA,opB", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("P,Q", [(3, 2), (2, 2), (1, 3)])
def test_summa_op_restatement(opA, opB, P, Q):
    """or_summa_op against the textbook product on exact integers; with
    'N' operands it equals or_summa bit for bit."""
    rng = np.random.default_rng(P * 10 + Q + ord(opA) + 3 * ord(opB))
    M, N, K, mb, nb, kb = 23, 17, 19, 3, 2, 4
    rA, cA = (M, K) if opA == "N" else (K, M)
    rB, cB = (K, N) if opB == "N" else (N, K)
    A = rng.integers(-3, 4, rA * cA).astype(float)
    B = rng.integers(-3, 4, rB * cB).astype(float)
    C = rng.integers(-3, 4, M * N).astype(float)
    Ap = oracle.pack_block_cyclic(A, rA, rA, cA, *((mb, kb) if opA == "N" else (kb, mb)), P, Q)
    Bp = oracle.pack_block_cyclic(B, rB, rB, cB, *((kb, nb) if opB == "N" else (nb, kb)), P, Q)
    Cp = oracle.pack_block_cyclic(C, M, M, N, mb, nb, P, Q)
    got = oracle.lib()
    out = Cp.copy()
    got.or_summa_op(opA.encode(), opB.encode(), M, N, K, 1.0, oracle._p(Ap), oracle._p(Bp), -0.5, oracle._p(out),
                    mb, nb, kb, P, Q)
    G = np.zeros(M * N)
    oracle.unpack_block_cyclic(out, mb, nb, P, Q, G, M, M, N)
    ref = oracle.host_dgemm(opA, opB, 1.0, A.reshape(cA, rA).T, B.reshape(cB, rB).T, -0.5, C.reshape(N, M).T)
    assert np.array_equal(G.reshape(N, M).T, ref)
    if opA == opB == "N":
        assert np.array_equal(out, oracle.summa(M, N, K, 1.0, Ap, Bp, -0.5, Cp.copy(), mb, nb, kb, P, Q))


@pytest.mark.parametrize("col_major", [True, False])
def test_block_seq_tiles_definition(col_major):
    """Every rank's tile element equals the global element it stands for
    (uneven 3x2 grid, remainders in the last row and column)."""
    M, N, K, size = 23, 17, 5, 6
    gA = np.arange(M * K, dtype=float)
    gB = np.arange(K * N, dtype=float) + 0.5
    gC = np.arange(M * N, dtype=float) + 0.25
    tA, tB, tC, sizes = oracle.block_seq_tiles(col_major, M, N, K, size, gA, gB, gC)
    dRow, dCol = oracle.virtual_grid(size)
    assert (dRow, dCol) == (3, 2)
    A = gA.reshape(K, M).T if col_major else gA.reshape(M, K)
    B = gB.reshape(N, K).T if col_major else gB.reshape(K, N)
    C = gC.reshape(N, M).T if col_major else gC.reshape(M, N)
    oa = ob = oc = 0
    for r, (tm, tn) in enumerate(sizes):
        i, j = divmod(r, dCol)
        r0, c0 = (M // dRow) * i, (N // dCol) * j
        assert tm == (M // dRow if i < dRow - 1 else M - r0) and tn == (N // dCol if j < dCol - 1 else N - c0)
        def tile(flat, rows, cols):
            return flat.reshape(cols, rows).T if col_major else flat.reshape(rows, cols)
        assert np.array_equal(tile(tA[oa: oa + tm * K], tm, K), A[r0: r0 + tm, :])
        assert np.array_equal(tile(tB[ob: ob + K * tn], K, tn), B[:, c0: c0 + tn])
        assert np.array_equal(tile(tC[oc: oc + tm * tn], tm, tn), C[r0: r0 + tm, c0: c0 + tn])
        oa, ob, oc = oa + tm * K, ob + K * tn, oc + tm * tn


# ------------------------------------------------------------------ validation

def planted(base, digits_err):
    """base with relative errors 3*10^-d planted on the positive entries."""
    x = base.copy()
    for idx, d in digits_err:
        x[idx] = base[idx] * (1 - 3 * 10.0 ** (-d))
    return x


def test_dtest_equality_known_answers():
    ref = np.linspace(1.0, 2.0, 50)
    assert oracle.Dtest_equality(ref, planted(ref, [(3, 11)])) == 11
    assert oracle.Dtest_equality(ref, planted(ref, [(3, 11), (7, 13)])) == 11
    assert oracle.Dtest_equality(ref, planted(ref, [(9, 6)])) == 0     # fails already at 1e-8
    assert oracle.Dtest_equality(ref, ref.copy()) == 30                # never fails: capped at 30
    neg = -ref
    assert oracle.Dtest_equality(neg, neg * 1.5) == 30                 # signed reference: negatives never fail
    z = np.zeros(4)
    assert oracle.Dtest_equality(z, np.array([0, 0, 3e-12, 0.0])) == 12  # zero reference: absolute error


def test_dtest_equality_matches_reference_build():
    ref_lib = oracle.ref_lib()
    if ref_lib is None:
        pytest.skip("reference Validation.cpp not built (oracle/_ref)")
    rng = np.random.default_rng(5)
    for trial in range(40):
        a = rng.uniform(-2, 2, 200)
        b = a * (1 + rng.choice([0, 1e-15, 3e-12, 1e-9, 1e-6], size=200) * rng.choice([-1, 1], size=200))
        b[rng.integers(0, 200)] = 0.0 if trial % 5 == 0 else b[0]
        mine = oracle.Dtest_equality(a, b)
        theirs = ref_lib.ref_Dtest_equality(oracle._p(a), oracle._p(b), a.size)
        assert mine == theirs
        for eps in (1e-8, 1e-12, 1e-15):
            assert oracle.Dvec_diff(a, b, eps) == ref_lib.ref_Dvec_diff(oracle._p(a), oracle._p(b), a.size, eps)
