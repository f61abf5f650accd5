"""Ozaki split DGEMM on the GPU (dg_ozaki_dgemm).

The slice extraction, the exact 32-bit slice products and the FP64
recombination follow the oracle's operation order, so the GPU result must
equal oracle.ozaki_dgemm bit for bit.  At larger sizes the error against
the host FP64 dgemm is checked per slice count."""
import numpy as np
import pytest
import torch

import dgamd
import oracle

pytestmark = pytest.mark.gpu


def upload(a2d, dev):
    return torch.from_numpy(np.asfortranarray(a2d).ravel(order="F").copy()).to(dev)


def wide(rng, shape):
    return np.ldexp(rng.uniform(-1, 1, shape), rng.integers(-20, 21, shape))


def run(dev, opA, opB, m, n, k, s, kind, alpha=1.0, beta=0.0, gen=None, seed=0):
    rng = np.random.default_rng(seed)
    gen = gen or (lambda shape: rng.uniform(-1, 1, shape))
    a = gen((m, k) if opA == "N" else (k, m))
    b = gen((k, n) if opB == "N" else (n, k))
    c = rng.uniform(-1, 1, (m, n))
    A, B = upload(a, dev), upload(b, dev)
    C = upload(c, dev)
    if beta == 0.0:
        C.fill_(float("nan"))
    dgamd.ozaki_dgemm_t(opA, opB, m, n, k, alpha, A, a.shape[0], B, b.shape[0], beta, C, m, s, kind)
    got = C.cpu().numpy().reshape(n, m).T
    bits = oracle.OZ_BF16_BITS if kind == dgamd.DG_OZAKI_BF16 else oracle.OZ_INT8_BITS
    aa = a if opA == "N" else a.T
    bb = b if opB == "N" else b.T
    ref = oracle.ozaki_dgemm(np.ascontiguousarray(aa), np.ascontiguousarray(bb), s, alpha, beta, c, bits=bits)
    return got, ref, aa, bb, c


KINDS = [dgamd.DG_OZAKI_INT8, dgamd.DG_OZAKI_BF16]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("opA,opB", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_bit_exact_vs_oracle_ops(dev, kind, opA, opB):
    got, ref, *_ = run(dev, opA, opB, 200, 150, 300, 6, kind, alpha=1.5, beta=-0.5)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("s", [1, 5, 7, 9])
def test_bit_exact_vs_oracle_slices(dev, kind, s):
    rng = np.random.default_rng(s)
    got, ref, *_ = run(dev, "N", "N", 256, 256, 512, s, kind, gen=lambda sh: wide(rng, sh))
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (17, 3, 129), (130, 260, 1), (128, 128, 128)])
def test_ragged_shapes(dev, m, n, k):
    got, ref, *_ = run(dev, "N", "N", m, n, k, 8, dgamd.DG_OZAKI_INT8, beta=2.0)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


def test_long_k_flush_points(dev):
    # INT8: K = 40000 -> three products share one int32 accumulator;
    # BF16: K = 20000 -> the K loop is cut into 17408 + 2592 exact chunks
    for kind, k in ((dgamd.DG_OZAKI_INT8, 40000), (dgamd.DG_OZAKI_INT8, 70000), (dgamd.DG_OZAKI_BF16, 20000)):
        got, ref, *_ = run(dev, "N", "T", 64, 48, k, 5, kind, seed=k)
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), (kind, k)


def test_zero_lines(dev):
    m = n = 64
    k = 100
    a = np.zeros((m, k)); a[3, 5] = 1.0; a[10] = np.linspace(-1, 1, k)
    b = np.zeros((k, n)); b[5] = 2.0; b[:, 7] = 1.0
    A, B = upload(a, dev), upload(b, dev)
    C = torch.zeros(m * n, dtype=torch.float64, device=dev)
    dgamd.ozaki_dgemm_t("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m, 8)
    got = C.cpu().numpy().reshape(n, m).T
    assert np.array_equal(got, oracle.ozaki_dgemm(a, b, 8))
    assert np.all(np.isfinite(got))


@pytest.mark.parametrize("dist", ["uniform", "wide"])
def test_error_per_slice_count(dev, dist):
    """Relative Frobenius error vs the host FP64 dgemm at 1024 x 1024 x 16384
    for s = 5..9: it falls ~2^-7 per slice and is within 2x the native FP64
    MFMA DGEMM's own error from s = 8 on."""
    m = n = 1024
    k = 16384
    rng = np.random.default_rng(7)
    gen = (lambda sh: rng.uniform(-1, 1, sh)) if dist == "uniform" else (lambda sh: wide(rng, sh))
    a, b = gen((m, k)), gen((k, n))
    A, B = upload(a, dev), upload(b, dev)
    ref = a @ b
    C = torch.empty(m * n, dtype=torch.float64, device=dev)
    dgamd.dgemm_t("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
    native = oracle.rel_frobenius(C.cpu().numpy().reshape(n, m).T, ref)
    errs = {}
    for s in range(5, 10):
        dgamd.ozaki_dgemm_t("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m, s)
        errs[s] = oracle.rel_frobenius(C.cpu().numpy().reshape(n, m).T, ref)
    print(dist, "native %.3e" % native, {s: "%.3e" % e for s, e in errs.items()})
    assert native < 1e-14
    # each slice adds 7 bits until the FP64 floor of the reference is reached at s = 8
    for s in range(5, 8):
        assert errs[s + 1] < errs[s] / 10
    assert errs[8] <= 2 * native
    assert errs[9] <= 2 * native


# ---------------------------------------------------------------- Ozaki scheme II (CRT)

def run_crt(dev, opA, opB, m, n, k, nmod, alpha=1.0, beta=0.0, gen=None, seed=0):
    rng = np.random.default_rng(seed)
    gen = gen or (lambda shape: rng.uniform(-1, 1, shape))
    a = gen((m, k) if opA == "N" else (k, m))
    b = gen((k, n) if opB == "N" else (n, k))
    c = rng.uniform(-1, 1, (m, n))
    A, B, C = upload(a, dev), upload(b, dev), upload(c, dev)
    if beta == 0.0:
        C.fill_(float("nan"))
    bound = dgamd.ozaki_dgemm_t(opA, opB, m, n, k, alpha, A, a.shape[0], B, b.shape[0], beta, C, m, nmod,
                                dgamd.DG_OZAKI_INT8_CRT)
    got = C.cpu().numpy().reshape(n, m).T
    aa = np.ascontiguousarray(a if opA == "N" else a.T)
    bb = np.ascontiguousarray(b if opB == "N" else b.T)
    ref = oracle.ozaki2_dgemm(aa, bb, nmod, alpha, beta, c)
    return got, ref, bound


@pytest.mark.parametrize("opA,opB", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_crt_bit_exact_ops(dev, opA, opB):
    got, ref, _ = run_crt(dev, opA, opB, 130, 90, 200, 14, alpha=1.5, beta=-0.5)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("nmod", [2, 8, 12, 15, 16])
def test_crt_bit_exact_moduli(dev, nmod):
    rng = np.random.default_rng(nmod)
    got, ref, _ = run_crt(dev, "N", "N", 96, 64, 300, nmod, gen=lambda sh: wide(rng, sh), seed=nmod)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


def test_crt_ragged_and_zero_lines(dev):
    got, ref, _ = run_crt(dev, "N", "N", 1, 300, 17, 16)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    m = n = 40
    k = 33
    a = np.zeros((m, k)); a[5] = np.linspace(-3, 3, k)
    b = np.zeros((k, n)); b[:, 2] = 1.0
    A, B = upload(a, dev), upload(b, dev)
    C = torch.zeros(m * n, dtype=torch.float64, device=dev)
    dgamd.ozaki_dgemm_t("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m, 16, dgamd.DG_OZAKI_INT8_CRT)
    got = C.cpu().numpy().reshape(n, m).T
    assert np.array_equal(got, oracle.ozaki2_dgemm(a, b, 16))
    assert np.array_equal(got, a @ b)


def upload_padded(x2d, ld, offset, dev):
    """x2d stored column-major with leading dimension ld >= rows, starting
    `offset` doubles into a NaN-filled device buffer; returns (buffer, view)."""
    r, c = x2d.shape
    buf = np.full(offset + ld * c + 1, np.nan)
    for j in range(c):
        buf[offset + j * ld: offset + j * ld + r] = x2d[:, j]
    t = torch.from_numpy(buf).to(dev)
    return t, t[offset:]


@pytest.mark.parametrize("opA,opB,m,n,k,pad,offset", [
    ("T", "N", 70, 50, 201, 1, 0),    # k-contiguous lines, odd k, even ld (16-byte pair loads, scalar tail)
    ("T", "N", 70, 50, 201, 1, 1),    # base 8 bytes off 16-byte alignment: a wave per line
    ("N", "T", 65, 33, 130, 3, 1),    # strided lines, partial last chunk, odd ld
    ("T", "T", 40, 90, 64, 2, 0),     # one whole chunk
    ("N", "N", 33, 47, 257, 0, 1),    # a single entry in the last chunk
])
def test_crt_scan_layouts_bit_exact(dev, opA, opB, m, n, k, pad, offset):
    rng = np.random.default_rng(m * n + k + pad + offset)
    a = wide(rng, (m, k) if opA == "N" else (k, m))
    b = wide(rng, (k, n) if opB == "N" else (n, k))
    a[:, 1] *= 2.0 ** -35   # a column / row far below the others
    lda, ldb = a.shape[0] + pad, b.shape[0] + pad
    _, A = upload_padded(a, lda, offset, dev)
    _, B = upload_padded(b, ldb, offset, dev)
    C = torch.full((m * n,), float("nan"), dtype=torch.float64, device=dev)
    dgamd.ozaki_dgemm_t(opA, opB, m, n, k, 1.0, A, lda, B, ldb, 0.0, C, m, 14, dgamd.DG_OZAKI_INT8_CRT)
    got = C.cpu().numpy().reshape(n, m).T
    aa = np.ascontiguousarray(a if opA == "N" else a.T)
    bb = np.ascontiguousarray(b if opB == "N" else b.T)
    ref = oracle.ozaki2_dgemm(aa, bb, 14)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("dist", ["uniform", "wide"])
def test_crt_error_per_moduli(dev, dist):
    """1024 x 1024 x 16384: error vs the host FP64 dgemm falls ~2^-8 per
    modulus; with 16 moduli it is within 2x the native FP64 MFMA error."""
    m = n = 1024
    k = 16384
    rng = np.random.default_rng(9)
    gen = (lambda sh: rng.uniform(-1, 1, sh)) if dist == "uniform" else (lambda sh: wide(rng, sh))
    a, b = gen((m, k)), gen((k, n))
    A, B = upload(a, dev), upload(b, dev)
    ref = a @ b
    C = torch.empty(m * n, dtype=torch.float64, device=dev)
    dgamd.dgemm_t("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
    native = oracle.rel_frobenius(C.cpu().numpy().reshape(n, m).T, ref)
    errs = {}
    for nm in (12, 13, 14, 15, 16):
        dgamd.ozaki_dgemm_t("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m, nm, dgamd.DG_OZAKI_INT8_CRT)
        errs[nm] = oracle.rel_frobenius(C.cpu().numpy().reshape(n, m).T, ref)
    print(dist, "native %.3e" % native, {s: "%.3e" % e for s, e in errs.items()})
    for nm in (12, 13):
        assert errs[nm + 1] < errs[nm] / 8
    assert errs[16] <= 2 * native
