"""Virtual-grid executors on the GPU: 2D/2.5D Cannon (dg_cannon25d) and
SUMMA (dg_summa), checked against the oracle's CPU restatement on exact
integer data (bit-exact) and against the host dgemm at the configuration
shapes (relative Frobenius error <= 1e-13)."""
import numpy as np
import pytest
import torch

import dgamd
import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-13


def to_dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def packed2d(G, rows, cols, p):
    return oracle.pack_block2d(G, rows, rows, cols, oracle.block_splits(rows, p), oracle.block_splits(cols, p))


def unpacked2d(P, rows, cols, p):
    G = np.zeros(rows * cols)
    oracle.unpack_block2d(P, oracle.block_splits(rows, p), oracle.block_splits(cols, p), G, rows, rows, cols)
    return G


def stored_shapes(opA, opB, M, N, K):
    ra, ca = (M, K) if opA == "N" else (K, M)
    rb, cb = (K, N) if opB == "N" else (N, K)
    return ra, ca, rb, cb


@pytest.mark.parametrize("opA,opB,M,N,K,p,c,sched", [
    ("N", "N", 64, 64, 64, 2, 1, dgamd.DG_SCHED_REFERENCE),
    ("N", "N", 128, 64, 96, 4, 2, dgamd.DG_SCHED_REFERENCE),
    ("N", "T", 128, 64, 96, 4, 2, dgamd.DG_SCHED_AUTO),
    ("T", "N", 96, 96, 192, 4, 1, dgamd.DG_SCHED_REFERENCE),
    ("T", "T", 60, 90, 30, 3, 1, dgamd.DG_SCHED_STANDARD),
    ("N", "N", 64, 64, 64, 8, 4, dgamd.DG_SCHED_AUTO),
    # grids where the reference's rotation offsets miss products: standard Cannon
    ("N", "N", 72, 48, 60, 6, 2, dgamd.DG_SCHED_STANDARD),
    ("T", "T", 72, 48, 60, 6, 2, dgamd.DG_SCHED_AUTO),
    ("T", "N", 50, 40, 35, 5, 1, dgamd.DG_SCHED_STANDARD),
    ("N", "T", 42, 84, 126, 7, 1, dgamd.DG_SCHED_STANDARD),
    # p = c: one step per layer, every product on its own layer
    ("N", "T", 64, 64, 64, 4, 4, dgamd.DG_SCHED_REFERENCE),
])
def test_cannon_exact_vs_oracle(dev, opA, opB, M, N, K, p, c, sched):
    rng = np.random.default_rng(M + N + K + p + c)
    ra, ca, rb, cb = stored_shapes(opA, opB, M, N, K)
    A = rng.integers(-4, 5, ra * ca).astype(np.float64)
    B = rng.integers(-4, 5, rb * cb).astype(np.float64)
    C = rng.integers(-4, 5, M * N).astype(np.float64)
    Ap, Bp, Cp = packed2d(A, ra, ca, p), packed2d(B, rb, cb, p), packed2d(C, M, N, p)
    # DG_SCHED_AUTO takes the reference rotations where they cover every k tile
    ref_covers = dgamd.lib().dg_cannon_m_index(p, c, dgamd.DG_SCHED_REFERENCE, 0, 0, 0, 0) >= 0
    std = sched == dgamd.DG_SCHED_STANDARD or (sched == dgamd.DG_SCHED_AUTO and not ref_covers)
    ref = oracle.cannon25d(opA, opB, M, N, K, 2.0, Ap, Bp, 0.5, Cp.copy(), p, c, sched_std=std)
    Ad, Bd, Cd = to_dev(Ap, dev), to_dev(Bp, dev), to_dev(Cp, dev)
    ws = torch.full((max(1, dgamd.cannon25d_workspace(M, N, p, c)),), float("nan"), dtype=torch.float64, device=dev)
    dgamd.cannon25d_t(opA, opB, M, N, K, 2.0, Ad, Bd, 0.5, Cd, p, c, sched, ws)
    assert np.array_equal(Cd.cpu().numpy(), ref)
    # and the unpacked result is the textbook product
    a2 = A.reshape(ca, ra).T
    b2 = B.reshape(cb, rb).T
    full = oracle.host_dgemm(opA, opB, 2.0, a2, b2, 0.5, C.reshape(N, M).T)
    assert np.array_equal(unpacked2d(ref, M, N, p).reshape(N, M).T, full)


def test_config1_cannon_2048(dev):
    """M=N=K=2048 on a 2x2x1 virtual grid, 1024^2 tiles, alpha=1, beta=0 with a
    NaN-filled C, validated against the host dgemm."""
    M = N = K = 2048
    p, c = 2, 1
    A = torch.empty(M * K, dtype=torch.float64, device=dev)
    B = torch.empty(K * N, dtype=torch.float64, device=dev)
    dgamd.generate_t(1, dgamd.DG_DIST_UNIFORM_PM1, A)
    dgamd.generate_t(2, dgamd.DG_DIST_UNIFORM_PM1, B)
    ro = dgamd.block_splits(M, p)
    Ap, Bp = torch.empty_like(A), torch.empty_like(B)
    dgamd.pack_block2d_t(A, M, M, K, ro, ro, Ap)
    dgamd.pack_block2d_t(B, K, K, N, ro, ro, Bp)
    Cp = torch.full((M * N,), float("nan"), dtype=torch.float64, device=dev)
    dgamd.cannon25d_t("N", "N", M, N, K, 1.0, Ap, Bp, 0.0, Cp, p, c, dgamd.DG_SCHED_REFERENCE)
    C = torch.empty_like(Cp)
    dgamd.unpack_block2d_t(Cp, ro, ro, C, M, M, N)
    a = A.cpu().numpy().reshape(K, M).T
    b = B.cpu().numpy().reshape(N, K).T
    ref = oracle.host_dgemm("N", "N", 1.0, a, b, 0.0, None)
    got = C.cpu().numpy().reshape(N, M).T
    err = oracle.rel_frobenius(got, ref)
    assert err <= TOL, err


def test_config3_shape_scaled(dev):
    """2.5D Cannon on a 4x4x2 grid, rectangular M != N != K, opB = 'T', at a
    quarter of the 16384 x 8192 x 32768 shape (same tile aspect)."""
    M, N, K, p, c = 4096, 2048, 8192, 4, 2
    for opB in ("N", "T"):
        ra, ca, rb, cb = stored_shapes("N", opB, M, N, K)
        A = torch.empty(ra * ca, dtype=torch.float64, device=dev)
        B = torch.empty(rb * cb, dtype=torch.float64, device=dev)
        C0 = torch.empty(M * N, dtype=torch.float64, device=dev)
        dgamd.generate_t(11, 0, A)
        dgamd.generate_t(12, 0, B)
        dgamd.generate_t(13, 0, C0)
        Ap, Bp, Cp = torch.empty_like(A), torch.empty_like(B), torch.empty_like(C0)
        dgamd.pack_block2d_t(A, ra, ra, ca, dgamd.block_splits(ra, p), dgamd.block_splits(ca, p), Ap)
        dgamd.pack_block2d_t(B, rb, rb, cb, dgamd.block_splits(rb, p), dgamd.block_splits(cb, p), Bp)
        dgamd.pack_block2d_t(C0, M, M, N, dgamd.block_splits(M, p), dgamd.block_splits(N, p), Cp)
        ws = torch.empty(dgamd.cannon25d_workspace(M, N, p, c), dtype=torch.float64, device=dev)
        dgamd.cannon25d_t("N", opB, M, N, K, 1.0, Ap, Bp, 0.5, Cp, p, c, dgamd.DG_SCHED_AUTO, ws)
        C = torch.empty_like(C0)
        dgamd.unpack_block2d_t(Cp, dgamd.block_splits(M, p), dgamd.block_splits(N, p), C, M, M, N)
        a = A.cpu().numpy().reshape(ca, ra).T
        b = B.cpu().numpy().reshape(cb, rb).T
        ref = oracle.host_dgemm("N", opB, 1.0, a, b, 0.5, C0.cpu().numpy().reshape(N, M).T)
        err = oracle.rel_frobenius(C.cpu().numpy().reshape(N, M).T, ref)
        assert err <= TOL, (opB, err)


@pytest.mark.parametrize("M,N,K,mb,nb,kb,P,Q", [
    (64, 64, 64, 8, 8, 8, 4, 4),
    (100, 70, 50, 16, 8, 7, 3, 2),
    (33, 65, 129, 4, 16, 32, 2, 4),
    (40, 40, 40, 64, 64, 64, 2, 2),   # blocks larger than the matrix: ranks with no data
    # whole 128x128 rank tiles: one k-step per panel (a segment change at every
    # step, also inside the prologue), 2 to 26 steps, and 3-step panels
    (512, 512, 16, 64, 64, 8, 2, 2),
    (512, 512, 80, 64, 64, 8, 2, 2),
    (512, 256, 208, 64, 64, 8, 2, 2),
    (512, 256, 144, 128, 64, 24, 2, 2),
])
def test_summa_exact_vs_oracle(dev, M, N, K, mb, nb, kb, P, Q):
    rng = np.random.default_rng(M * N + K)
    A = rng.integers(-4, 5, M * K).astype(np.float64)
    B = rng.integers(-4, 5, K * N).astype(np.float64)
    C = rng.integers(-4, 5, M * N).astype(np.float64)
    Ap = oracle.pack_block_cyclic(A, M, M, K, mb, kb, P, Q)
    Bp = oracle.pack_block_cyclic(B, K, K, N, kb, nb, P, Q)
    Cp = oracle.pack_block_cyclic(C, M, M, N, mb, nb, P, Q)
    ref = oracle.summa(M, N, K, -1.0, Ap, Bp, 0.5, Cp.copy(), mb, nb, kb, P, Q)
    Ad, Bd, Cd = to_dev(Ap, dev), to_dev(Bp, dev), to_dev(Cp, dev)
    dgamd.summa_t("N", "N", M, N, K, -1.0, Ad, Bd, 0.5, Cd, mb, nb, kb, P, Q)
    assert np.array_equal(Cd.cpu().numpy(), ref)
    # unpacked, it is the textbook product
    G = np.zeros(M * N)
    oracle.unpack_block_cyclic(ref, mb, nb, P, Q, G, M, M, N)
    full = oracle.host_dgemm("N", "N", -1.0, A.reshape(K, M).T, B.reshape(N, K).T, 0.5, C.reshape(N, M).T)
    assert np.array_equal(G.reshape(N, M).T, full)


@pytest.mark.parametrize("opA,opB", [("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("M,N,K,mb,nb,kb,P,Q", [
    (64, 64, 64, 8, 8, 8, 4, 4),
    (100, 70, 50, 16, 8, 7, 3, 2),      # P != Q: a transposed panel spans several owners
    (33, 65, 129, 4, 16, 32, 2, 4),
    (40, 40, 40, 64, 64, 64, 2, 2),     # ranks with no data
    (512, 256, 144, 128, 64, 24, 2, 2),
])
def test_summa_transposed_exact_vs_oracle(dev, opA, opB, M, N, K, mb, nb, kb, P, Q):
    """op(A) = A^T with A stored K x M (blocks kb x mb), op(B) = B^T with B
    stored N x K (blocks nb x kb), as cublasMpGemm's TransA / TransB."""
    rng = np.random.default_rng(M * N + K + ord(opA) + 2 * ord(opB))
    rA, cA = (M, K) if opA == "N" else (K, M)
    rB, cB = (K, N) if opB == "N" else (N, K)
    A = rng.integers(-4, 5, rA * cA).astype(np.float64)
    B = rng.integers(-4, 5, rB * cB).astype(np.float64)
    C = rng.integers(-4, 5, M * N).astype(np.float64)
    bA = (mb, kb) if opA == "N" else (kb, mb)
    bB = (kb, nb) if opB == "N" else (nb, kb)
    Ap = oracle.pack_block_cyclic(A, rA, rA, cA, bA[0], bA[1], P, Q)
    Bp = oracle.pack_block_cyclic(B, rB, rB, cB, bB[0], bB[1], P, Q)
    Cp = oracle.pack_block_cyclic(C, M, M, N, mb, nb, P, Q)
    ref = oracle.summa(M, N, K, -1.0, Ap, Bp, 0.5, Cp.copy(), mb, nb, kb, P, Q, opA, opB)
    Ad, Bd, Cd = to_dev(Ap, dev), to_dev(Bp, dev), to_dev(Cp, dev)
    dgamd.summa_t(opA, opB, M, N, K, -1.0, Ad, Bd, 0.5, Cd, mb, nb, kb, P, Q)
    assert np.array_equal(Cd.cpu().numpy(), ref)


def test_config2_summa_scaled(dev):
    """SUMMA 4x4, nb=512, alpha=1, beta=0.5 at M=N=K=4096 vs host dgemm."""
    n, nb, P = 4096, 512, 4
    A = torch.empty(n * n, dtype=torch.float64, device=dev)
    B = torch.empty_like(A)
    C0 = torch.empty_like(A)
    for s, t in ((1, A), (2, B), (3, C0)):
        dgamd.generate_t(s, 0, t)
    Ap, Bp, Cp = torch.empty_like(A), torch.empty_like(A), torch.empty_like(A)
    dgamd.pack_block_cyclic_t(A, n, n, n, nb, nb, P, P, Ap)
    dgamd.pack_block_cyclic_t(B, n, n, n, nb, nb, P, P, Bp)
    dgamd.pack_block_cyclic_t(C0, n, n, n, nb, nb, P, P, Cp)
    dgamd.summa_t("N", "N", n, n, n, 1.0, Ap, Bp, 0.5, Cp, nb, nb, nb, P, P)
    C = torch.empty_like(A)
    dgamd.unpack_block_cyclic_t(Cp, nb, nb, P, P, C, n, n, n)
    h = lambda t: t.cpu().numpy().reshape(n, n).T
    ref = oracle.host_dgemm("N", "N", 1.0, h(A), h(B), 0.5, h(C0))
    err = oracle.rel_frobenius(h(C), ref)
    assert err <= TOL, err


def test_grid_argument_errors(dev):
    L = dgamd.lib()
    # p=3, c=1 under the reference rotations does not visit every k tile
    st = L.dg_cannon25d(b"N", b"N", 9, 9, 9, 1.0, None, None, 0.0, None, 3, 1, dgamd.DG_SCHED_REFERENCE, None, 0, None)
    assert st == dgamd.DG_ERR_SCHEDULE
    st = L.dg_cannon25d(b"N", b"N", 10, 9, 9, 1.0, None, None, 0.0, None, 3, 1, dgamd.DG_SCHED_STANDARD, None, 0, None)
    assert st == dgamd.DG_ERR_INVALID
    st = L.dg_summa(b"T", b"N", 8, 8, 8, 1.0, None, None, 0.0, None, 2, 2, 2, 2, 2, None)
    assert st == dgamd.DG_ERR_INVALID   # null operands
    st = L.dg_summa(b"X", b"N", 8, 8, 8, 1.0, None, None, 0.0, None, 2, 2, 2, 2, 2, None)
    assert st == dgamd.DG_ERR_INVALID
