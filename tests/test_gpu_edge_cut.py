"""Uneven FP64 problems are cut into an aligned interior block (A-in-registers
kernel) and strips around it: edge-handling direct-to-LDS tiles, or, for
strips past the interior at most 64 rows or columns wide, the register-staged
kernel's 64x128 and 128x32 tiles.  The cut must not change a single bit:
dg_dgemm (which cuts) against dg_grouped_dgemm_noalloc (which runs every
problem whole on the edge-handling kernel), for every op combination, bases
8 bytes off 16-byte alignment, ragged sizes, both sides of the narrow-strip
width limit and K-step counts the cut accepts; and both within 1e-13 of the
host dgemm."""
import ctypes
import itertools

import numpy as np
import pytest
import torch

import dgamd

pytestmark = pytest.mark.gpu

SHAPES = [
    (1000, 900, 512),    # one interior block and two strips (rows after, columns after)
    (1031, 777, 272),    # 34 k-steps
    (2857, 600, 512),    # a rank tile of the 7x5 grid over 20000
    (384, 384, 256),     # exactly 9 interior tiles when aligned
    (1088, 832, 256),    # 64 rows and 64 columns past the interior: the widest narrow strips
    (1089, 833, 256),    # 65 of each: strips on the 128x128 edge tiles again
]


def operand(dev, rows, cols, off, seed):
    """rows x cols column-major matrix at element offset `off` of a buffer (ld = rows + 2)."""
    ld = rows + 2
    buf = torch.empty(off + ld * cols + 1, dtype=torch.float64, device=dev)
    dgamd.generate_t(seed, dgamd.DG_DIST_UNIFORM_PM1, buf)
    host = buf.cpu().numpy()[off:off + ld * cols].reshape(cols, ld).T[:rows, :]
    return buf, buf.data_ptr() + 8 * off, ld, host


@pytest.mark.parametrize("opA,opB", list(itertools.product("NT", "NT")))
@pytest.mark.parametrize("m,n,k", SHAPES)
@pytest.mark.parametrize("off", [0, 1])
def test_cut_is_bit_identical(dev, opA, opB, m, n, k, off):
    print("edge cut case", opA, opB, m, n, k, off, flush=True)
    ra, ca = (m, k) if opA == "N" else (k, m)
    rb, cb = (k, n) if opB == "N" else (n, k)
    bufa, pa, lda, a = operand(dev, ra, ca, off, 11 + m)
    bufb, pb, ldb, b = operand(dev, rb, cb, off, 12 + n)
    bufc, pc, ldc, c = operand(dev, m, n, off, 13 + k)
    c0 = bufc.clone()
    alpha, beta = 1.25, -0.5
    L = dgamd.lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = L.dg_dgemm(opA.encode(), opB.encode(), m, n, k, alpha, ctypes.c_void_p(pa), lda, ctypes.c_void_p(pb), ldb,
                    beta, ctypes.c_void_p(pc), ldc, stream)
    assert rc == dgamd.DG_OK, dgamd.last_error()
    torch.cuda.synchronize()
    cut = bufc.clone()
    bufc.copy_(c0)
    desc = dgamd.gemm_desc(opA, opB, m, n, k, alpha, pa, lda, pb, ldb, beta, pc, ldc)
    arr = (dgamd.GemmDesc * 1)(desc)
    rc = L.dg_grouped_dgemm_noalloc(arr, 1, stream)
    assert rc == dgamd.DG_OK, dgamd.last_error()
    torch.cuda.synchronize()
    assert torch.equal(cut.view(torch.int64), bufc.view(torch.int64))
    got = cut.cpu().numpy()[off:off + ldc * n].reshape(n, ldc).T[:m, :]
    aa = a if opA == "N" else a.T
    bb = b if opB == "N" else b.T
    ref = alpha * (aa @ bb) + beta * c
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err <= 1e-13, err
    # the padding rows between columns are untouched
    whole = cut.cpu().numpy()[off:off + ldc * n].reshape(n, ldc).T
    assert np.array_equal(whole[m:, :], c0.cpu().numpy()[off:off + ldc * n].reshape(n, ldc).T[m:, :])


@pytest.mark.parametrize("opA,opB", [("N", "N"), ("T", "T")])
@pytest.mark.parametrize("mode", ["nonpersistent", "persistent"])
def test_cut_plans_and_graph_replays_bit_identical(dev, opA, opB, mode):
    """A batch whose problems are cut (narrow strips included, bases on and
    off 16-byte alignment) through a plan in either scheduling mode, then
    through two replays of the plan's graph with the outputs poisoned in
    between: every run gives the bits of the uncut no-allocation path."""
    shapes = [(1031, 777, 272, 0), (2857, 600, 512, 1), (1088, 832, 256, 0), (1000, 900, 512, 1)]
    L = dgamd.lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    keep, descs, outs = [], [], []
    for i, (m, n, k, off) in enumerate(shapes):
        print("plan cut case", opA, opB, mode, m, n, k, off, flush=True)
        ra, ca = (m, k) if opA == "N" else (k, m)
        rb, cb = (k, n) if opB == "N" else (n, k)
        bufa, pa, lda, _ = operand(dev, ra, ca, off, 31 + i)
        bufb, pb, ldb, _ = operand(dev, rb, cb, off, 41 + i)
        bufc, pc, ldc, _ = operand(dev, m, n, off, 51 + i)
        keep += [bufa, bufb]
        outs.append((bufc, bufc.clone()))
        descs.append(dgamd.gemm_desc(opA, opB, m, n, k, 0.75, pa, lda, pb, ldb, 0.5, pc, ldc))
    arr = (dgamd.GemmDesc * len(descs))(*descs)
    assert L.dg_grouped_dgemm_noalloc(arr, len(descs), stream) == dgamd.DG_OK, dgamd.last_error()
    torch.cuda.synchronize()
    ref = [c.clone() for c, _ in outs]
    for c, c0 in outs:
        c.copy_(c0)
    plan = dgamd.GroupedPlan(descs, dgamd.DG_PLAN_PERSISTENT if mode == "persistent" else dgamd.DG_PLAN_NONPERSISTENT)
    plan.execute()
    torch.cuda.synchronize()
    for (c, _), r in zip(outs, ref):
        assert torch.equal(c.view(torch.int64), r.view(torch.int64))
    for rep in range(2):
        for c, c0 in outs:
            c.copy_(c0)   # beta != 0: each replay starts from the same C
        torch.cuda.synchronize()
        plan.graph_launch()
        torch.cuda.synchronize()
        for (c, _), r in zip(outs, ref):
            assert torch.equal(c.view(torch.int64), r.view(torch.int64)), rep
    plan.close()
