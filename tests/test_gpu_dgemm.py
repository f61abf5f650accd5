"""FP64 MFMA DGEMM (dg_dgemm) against the host BLAS dgemm oracle.

Tolerance: relative Frobenius error <= 1e-13 against the host dgemm (the
acceptance bar of the reference's validation against cblas_dgemm); inputs
whose products and partial sums are exact integers must match bit for bit,
whatever the summation order.
"""
import numpy as np
import pytest
import torch

import dgamd
import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-13


def colmajor(rows, cols, ld, rng, dev, integer=False, fill=None, offset=0):
    """A (ld x cols) column-major device buffer whose top rows x cols block is
    random, starting `offset` doubles into its allocation (NaN before it);
    returns (device tensor view, host 2-D view of the logical matrix)."""
    buf = np.full((cols, ld), np.nan if fill is None else fill, dtype=np.float64)
    if integer:
        vals = rng.integers(-8, 9, size=(cols, rows)).astype(np.float64)
    else:
        vals = rng.uniform(-1.0, 1.0, size=(cols, rows))
    buf[:, :rows] = vals
    flat = np.concatenate([np.full(offset, np.nan), buf.reshape(-1)])
    t = torch.from_numpy(flat).to(dev)[offset:]
    return t, vals.T.copy()


def run(dev, opA, opB, m, n, k, alpha, beta, pad=(0, 0, 0), integer=False, seed=0, offset=(0, 0)):
    rng = np.random.default_rng(seed)
    ra, ca = (m, k) if opA == "N" else (k, m)
    rb, cb = (k, n) if opB == "N" else (n, k)
    lda, ldb, ldc = max(ra, 1) + pad[0], max(rb, 1) + pad[1], max(m, 1) + pad[2]
    A, a = colmajor(ra, ca, lda, rng, dev, integer, offset=offset[0])
    B, b = colmajor(rb, cb, ldb, rng, dev, integer, offset=offset[1])
    C, c = colmajor(m, n, ldc, rng, dev, integer, fill=7.0)
    if beta == 0.0:
        # beta == 0 must not read C: poison the logical block with NaN
        C.view(n, ldc)[:, :m] = float("nan")
    dgamd.dgemm_t(opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc)
    torch.cuda.synchronize()
    out = C.cpu().numpy().reshape(n, ldc)
    got = out[:, :m].T
    ref = oracle.host_dgemm(opA, opB, alpha, a, b, beta, c)
    # rows of the padding must be untouched
    if pad[2]:
        assert np.all(out[:, m:] == 7.0)
    return got, ref


def test_mfma_layout_exact(dev):
    # one 16x16 output block, k = 4: checks the MFMA operand/result lane layout
    got, ref = run(dev, "N", "N", 16, 16, 4, 1.0, 0.0, integer=True)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("opA,opB", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_ops_exact_integer(dev, opA, opB):
    # interior tiles, edge tiles and a partial k-step, all exact
    got, ref = run(dev, opA, opB, 300, 200, 77, 1.0, 0.0, integer=True, seed=1)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("opA,opB", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_ops_random(dev, opA, opB):
    got, ref = run(dev, opA, opB, 512, 384, 640, 1.5, 0.5, seed=2)
    assert oracle.rel_frobenius(got, ref) <= TOL


def test_beta_zero_nan_c(dev):
    got, ref = run(dev, "N", "N", 256, 256, 256, 1.0, 0.0, seed=3)
    assert np.all(np.isfinite(got))
    assert oracle.rel_frobenius(got, ref) <= TOL


@pytest.mark.parametrize("opA,opB", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_all_interior_tiles_exact_integer(dev, opA, opB):
    # m, n multiples of 128 and an even number of 8-deep k-steps: the
    # direct-to-LDS kernel, every operand layout, exact
    got, ref = run(dev, opA, opB, 384, 256, 208, 1.0, 0.0, integer=True, seed=4)
    assert np.array_equal(got, ref)
    got, ref = run(dev, opA, opB, 128, 128, 16, -2.0, 0.5, integer=True, seed=5)
    assert np.array_equal(got, ref)


# Edge tiles of the direct-to-LDS kernel: even leading dimensions keep the
# 128-bit path; the NaN padding of A and B (rows past the matrix) fails the
# test if any load reads past a row or past a segment's K; odd row counts of
# the row-contiguous operand (A for opA = N, B for opB = T) send the last row
# or column of C to the remainder kernel.
@pytest.mark.parametrize("opA,opB,m,n,k", [
    ("N", "N", 301, 203, 86),    # odd m: last row split off; partial last k-step (6 of 8)
    ("N", "T", 255, 129, 77),    # odd m and n: last row and column split off; odd K
    ("T", "N", 129, 131, 94),    # k-contiguous A and B: clamped rows, partial k-step
    ("T", "T", 130, 257, 62),    # odd n with opB = T: last column split off
    ("N", "N", 1, 300, 18),      # a single row: all of it in the remainder kernel
    ("N", "T", 300, 1, 40),      # a single column: likewise
    ("T", "N", 5, 7, 2),         # one partial k-step, padded to two
])
def test_dma_edge_tiles_exact_integer(dev, opA, opB, m, n, k):
    ra = m if opA == "N" else k
    rb = k if opB == "N" else n
    got, ref = run(dev, opA, opB, m, n, k, 1.0, 0.0, pad=(ra % 2 + 2, rb % 2 + 2, 3), integer=True, seed=8)
    assert np.array_equal(got, ref)
    got, ref = run(dev, opA, opB, m, n, k, -0.5, 2.0, pad=(ra % 2, rb % 2, 1), integer=True, seed=9)
    assert np.array_equal(got, ref)


# A row-contiguous operand (A for opA = N, B for opB = T) that starts 8 bytes
# past a 16-byte boundary: the direct-to-LDS loads start one row earlier
# (NaN there must not reach C).
@pytest.mark.parametrize("opA,opB,m,n,k", [
    ("N", "N", 301, 200, 96),
    ("N", "T", 128, 256, 64),
    ("T", "T", 200, 129, 40),
    ("N", "T", 255, 257, 33),
])
def test_dma_shifted_row_operand_exact_integer(dev, opA, opB, m, n, k):
    ra = m if opA == "N" else k
    rb = k if opB == "N" else n
    off = (1 if opA == "N" else 0, 1 if opB == "T" else 0)
    got, ref = run(dev, opA, opB, m, n, k, 1.0, 0.0, pad=(ra % 2, rb % 2, 0), integer=True, seed=10, offset=off)
    assert np.array_equal(got, ref)
    got, ref = run(dev, opA, opB, m, n, k, 0.5, -1.0, pad=(ra % 2, rb % 2, 2), integer=True, seed=11, offset=off)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("pad", [(1, 0, 0), (0, 1, 0), (1, 1, 5)])
def test_odd_leading_dims(dev, pad):
    # odd lda (257) or ldb (131) forces the 64-bit load path; ldc padding must stay untouched
    got, ref = run(dev, "N", "T", 256, 131, 129, -0.75, 2.0, pad=pad, seed=4)
    assert oracle.rel_frobenius(got, ref) <= TOL


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (1, 300, 17), (300, 1, 33), (15, 17, 1), (129, 129, 16)])
def test_small_and_ragged(dev, m, n, k):
    got, ref = run(dev, "N", "N", m, n, k, 1.0, 0.25, seed=5)
    assert oracle.rel_frobenius(got, ref) <= TOL


def test_k_zero_scales_c(dev):
    got, ref = run(dev, "N", "N", 64, 64, 0, 1.0, 0.5, seed=6)
    assert np.array_equal(got, ref)


def test_config1_shape_2048(dev):
    # the tile product of the 2048^3 Cannon configuration at full size
    got, ref = run(dev, "N", "N", 2048, 2048, 2048, 1.0, 0.0, seed=7)
    err = oracle.rel_frobenius(got, ref)
    assert err <= TOL, err
    # the reference's digit-count report (Validation.cpp:57-83), informative only:
    # its per-element relative error is unbounded near zero outputs
    digits = oracle.Dtest_equality(np.ascontiguousarray(ref.T).ravel(), np.ascontiguousarray(got.T).ravel())
    print("Dtest_equality digits:", digits, "rel Frobenius:", err)


def test_dtest_equality_on_gpu_result(dev):
    """The reference's digit count (Validation.cpp:57-83) on a GPU DGEMM
    result whose entries are bounded away from zero (A, B in [0.5, 1)):
    the oracle's restatement, the reference's own Validation.cpp built from
    its sources (oracle/_ref, when built) and the host API's drop-in
    Dtest_equality (libdghost.so) report the same digit count, and the FP64
    result agrees with the host dgemm to at least 13 digits everywhere."""
    import ctypes
    import os
    m, n, k = 2048, 1536, 2048
    rng = np.random.default_rng(17)
    a = rng.uniform(0.5, 1.0, (m, k))
    b = rng.uniform(0.5, 1.0, (k, n))
    up = lambda x: torch.from_numpy(np.asfortranarray(x).ravel(order="F").copy()).to(dev)
    A, B = up(a), up(b)
    C = torch.empty(m * n, dtype=torch.float64, device=dev)
    dgamd.dgemm_t("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
    got = np.ascontiguousarray(C.cpu().numpy())
    ref = np.ascontiguousarray((a @ b).T).ravel()    # column-major
    digits = oracle.Dtest_equality(ref, got)
    assert digits >= 13, digits
    R = oracle.ref_lib()
    if R is not None:
        dp = ctypes.POINTER(ctypes.c_double)
        assert R.ref_Dtest_equality(ref.ctypes.data_as(dp), got.ctypes.data_as(dp), ref.size) == digits
    host = ctypes.CDLL(os.path.join(os.path.dirname(dgamd.__file__), "libdghost.so"))
    f = host._Z14Dtest_equalityPdS_x   # short Dtest_equality(double*, double*, long long)
    f.restype = ctypes.c_short
    f.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), ctypes.c_longlong]
    dp = ctypes.POINTER(ctypes.c_double)
    assert f(ref.ctypes.data_as(dp), got.ctypes.data_as(dp), ref.size) == digits


def test_invalid_arguments(dev):
    L = dgamd.lib()
    assert L.dg_dgemm(b"X", b"N", 4, 4, 4, 1.0, None, 4, None, 4, 0.0, None, 4, None) == dgamd.DG_ERR_INVALID
    assert L.dg_dgemm(b"N", b"N", 4, 4, 4, 1.0, None, 3, None, 4, 0.0, None, 4, None) == dgamd.DG_ERR_INVALID


@pytest.mark.parametrize("opA,opB", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("steps", [2, 4, 6, 8, 10, 12, 14, 16, 18, 26, 52])
def test_interior_step_counts_exact(dev, opA, opB, steps):
    """Whole 128x128 tiles (A held in registers, three k-steps ahead, in
    turns of six steps; row pairs for opA = N, k pairs for opA = T): k-step
    counts from the shortest prologue
    through whole turns and tails of 4, 6 and 8 steps, bit-exact on integer
    data.  The second shape has beta != 0 and an odd ldc (element stores);
    the padding rows of C stay untouched."""
    k = 8 * steps
    for m, n, beta, padc in [(256, 384, 0.0, 0), (384, 128, 0.5, 1)]:
        got, ref = run(dev, opA, opB, m, n, k, 1.0, beta, pad=(0, 0, padc), integer=True, seed=steps)
        assert np.array_equal(got, ref), (m, n, k, beta, padc)


@pytest.mark.parametrize("mode", [0, 1, 2])  # autotune, one launch per tile, persistent
def test_interior_grouped_modes_exact(dev, mode):
    """A grouped batch of interior opA = N problems with different k-step
    counts (and so different turn and tail splits), in every launch mode."""
    rng = np.random.default_rng(40 + mode)
    shapes = [(128, 256, 16), (256, 128, 96), (384, 256, 208), (128, 128, 144)]
    descs, outs, keep = [], [], []  # descriptors hold raw addresses: keep A and B alive
    for i, (m, n, k) in enumerate(shapes):
        opB = "N" if i % 2 == 0 else "T"
        rb, cb = (k, n) if opB == "N" else (n, k)
        a = rng.integers(-8, 9, (m, k)).astype(np.float64)
        b = rng.integers(-8, 9, (rb, cb)).astype(np.float64)
        c = rng.integers(-8, 9, (m, n)).astype(np.float64)
        A = torch.from_numpy(np.asfortranarray(a).ravel(order="F")).to(dev)
        B = torch.from_numpy(np.asfortranarray(b).ravel(order="F")).to(dev)
        C = torch.from_numpy(np.asfortranarray(c).ravel(order="F")).to(dev)
        keep.append((A, B))
        descs.append(dgamd.gemm_desc("N", opB, m, n, k, 2.0, A, m, B, rb, -1.0, C, m))
        ref = 2.0 * (a @ (b if opB == "N" else b.T)) - c
        outs.append((C, m, n, ref))
    plan = dgamd.GroupedPlan(descs, mode)
    plan.execute()
    torch.cuda.synchronize()
    bad = [(m, n, int((C.cpu().numpy().reshape(n, m).T != ref).sum())) for C, m, n, ref in outs
           if not np.array_equal(C.cpu().numpy().reshape(n, m).T, ref)]
    assert not bad, bad
