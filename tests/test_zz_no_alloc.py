"""dg_grouped_dgemm_noalloc through the C ABI: the allocation-free grouped
GEMM (descriptor tables in the library's static device scratch array) gives
the results of dg_grouped_dgemm bit for bit, on two calls in a row (the
second with a smaller batch), and rejects a batch whose tables exceed the
scratch array without touching any output."""
import ctypes

import numpy as np
import pytest
import torch

import dgamd

pytestmark = pytest.mark.gpu


def noalloc_fn():
    return dgamd.lib().dg_grouped_dgemm_noalloc   # argtypes from dgamd's signature table


def make_batch(dev, count, seed, max_dim):
    rng = np.random.default_rng(seed)
    f64 = dict(dtype=torch.float64, device=dev)
    keep, descs, outs = [], [], []
    for t in range(count):
        opA, opB = "NT"[rng.integers(2)], "NT"[rng.integers(2)]
        m, n, k = (int(rng.integers(1, max_dim + 1)) for _ in range(3))
        ra, ca = (m, k) if opA == "N" else (k, m)
        rb, cb = (k, n) if opB == "N" else (n, k)
        lda, ldb, ldc = ra + int(rng.integers(0, 3)), rb + int(rng.integers(0, 3)), m + int(rng.integers(0, 3))
        A = torch.empty(lda * ca, **f64)
        B = torch.empty(ldb * cb, **f64)
        C = torch.empty(ldc * n, **f64)
        dgamd.generate_t(1000 + 3 * t + seed, dgamd.DG_DIST_UNIFORM_PM1, A)
        dgamd.generate_t(1001 + 3 * t + seed, dgamd.DG_DIST_UNIFORM_PM1, B)
        dgamd.generate_t(1002 + 3 * t + seed, dgamd.DG_DIST_UNIFORM_PM1, C)
        alpha = float(rng.uniform(-2, 2))
        beta = 0.0 if t % 5 == 0 else float(rng.uniform(-1, 1))
        descs.append(dgamd.gemm_desc(opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc))
        keep += [A, B, C]
        outs.append(C)
    return descs, outs, keep


def run_both(dev, descs, outs):
    """(outputs of dg_grouped_dgemm, outputs of dg_grouped_dgemm_noalloc)
    from the same initial C."""
    init = [c.clone() for c in outs]
    dgamd.grouped_dgemm_t(descs)
    torch.cuda.synchronize()
    ref = [c.clone() for c in outs]
    for c, c0 in zip(outs, init):
        c.copy_(c0)
    arr = (dgamd.GemmDesc * len(descs))(*descs)
    rc = noalloc_fn()(arr, len(descs), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == dgamd.DG_OK, dgamd.last_error()
    torch.cuda.synchronize()
    return ref, [c.clone() for c in outs]


def test_noalloc_matches_grouped_twice(dev):
    descs, outs, keep = make_batch(dev, 48, 1, 300)
    ref, got = run_both(dev, descs, outs)
    for i, (r, g) in enumerate(zip(ref, got)):
        assert torch.equal(r, g), i
    # second call, smaller batch: the scratch array holds the new tables
    descs2, outs2, keep2 = make_batch(dev, 7, 2, 520)
    ref2, got2 = run_both(dev, descs2, outs2)
    for i, (r, g) in enumerate(zip(ref2, got2)):
        assert torch.equal(r, g), i
    del keep, keep2


def test_noalloc_rejects_oversized_batch(dev):
    count = 4000
    A = torch.ones(1, dtype=torch.float64, device=dev)
    B = torch.ones(1, dtype=torch.float64, device=dev)
    C = torch.full((count,), 5.0, dtype=torch.float64, device=dev)
    base = C.data_ptr()
    descs = [dgamd.gemm_desc("N", "N", 1, 1, 1, 1.0, A, 1, B, 1, 0.0, base + 8 * i, 1) for i in range(count)]
    arr = (dgamd.GemmDesc * count)(*descs)
    rc = noalloc_fn()(arr, count, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == dgamd.DG_ERR_WORKSPACE
    assert "scratch" in dgamd.last_error()
    torch.cuda.synchronize()
    assert torch.all(C == 5.0)
