"""The configurations at full size on one GPU.

Checks that need no second implementation of the whole product:
* sampled rows against the host dgemm (relative Frobenius <= 1e-13);
* SUMMA and the grouped batch sum every output element over k in the same
  order as a plain dg_dgemm of the same operands, so they must agree with it
  bit for bit;
* pack/unpack round trips."""
import numpy as np
import pytest
import torch

import dgamd
import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-13


def gen(n_elems, seed, dev, dist=dgamd.DG_DIST_UNIFORM_PM1):
    t = torch.empty(n_elems, dtype=torch.float64, device=dev)
    dgamd.generate_t(seed, dist, t)
    return t


def host_rows(Gt, rows, nrows_total, ncols):
    """Rows of a column-major (nrows_total x ncols) device matrix, as a host array."""
    v = Gt.view(ncols, nrows_total)
    return v[:, rows].cpu().numpy().T


def test_config2_summa_16384_full(dev):
    n, nb, P = 16384, 512, 4
    A, B, C0 = gen(n * n, 1, dev), gen(n * n, 2, dev), gen(n * n, 3, dev)
    Ap, Bp, Cp = torch.empty_like(A), torch.empty_like(B), torch.empty_like(C0)
    for G, Pk in ((A, Ap), (B, Bp), (C0, Cp)):
        dgamd.pack_block_cyclic_t(G, n, n, n, nb, nb, P, P, Pk)
    dgamd.summa_t("N", "N", n, n, n, 1.0, Ap, Bp, 0.5, Cp, nb, nb, nb, P, P)
    del Ap, Bp
    C = torch.empty_like(C0)
    dgamd.unpack_block_cyclic_t(Cp, nb, nb, P, P, C, n, n, n)
    del Cp
    # plain DGEMM of the same operands: same k order per element -> identical bits
    D = C0.clone()
    dgamd.dgemm_t("N", "N", n, n, n, 1.0, A, n, B, n, 0.5, D, n)
    assert torch.equal(C, D)
    del D
    rows = [0, 1, 511, 512, 4095, 4096, 9000, 16383]
    a = host_rows(A, rows, n, n)
    b = B.cpu().numpy().reshape(n, n).T
    ref = a @ b + 0.5 * host_rows(C0, rows, n, n)
    err = oracle.rel_frobenius(host_rows(C, rows, n, n), ref)
    assert err <= TOL, err


@pytest.mark.parametrize("opB", ["N", "T"])
def test_config3_cannon_full(dev, opB):
    M, N, K, p, c = 16384, 8192, 32768, 4, 2
    rb, cb = (K, N) if opB == "N" else (N, K)
    A, B = gen(M * K, 11, dev), gen(rb * cb, 12, dev)
    sp = dgamd.block_splits
    Ap, Bp = torch.empty_like(A), torch.empty_like(B)
    dgamd.pack_block2d_t(A, M, M, K, sp(M, p), sp(K, p), Ap)
    dgamd.pack_block2d_t(B, rb, rb, cb, sp(rb, p), sp(cb, p), Bp)
    Cp = torch.full((M * N,), float("nan"), dtype=torch.float64, device=dev)
    ws = torch.empty(dgamd.cannon25d_workspace(M, N, p, c), dtype=torch.float64, device=dev)
    dgamd.cannon25d_t("N", opB, M, N, K, 1.0, Ap, Bp, 0.0, Cp, p, c, dgamd.DG_SCHED_AUTO, ws)
    del Ap, Bp, ws
    C = torch.empty_like(Cp)
    dgamd.unpack_block2d_t(Cp, sp(M, p), sp(N, p), C, M, M, N)
    del Cp
    D = torch.empty_like(C)
    dgamd.dgemm_t("N", opB, M, N, K, 1.0, A, M, B, rb, 0.0, D, M)
    err_gpu = (torch.linalg.vector_norm(C - D) / torch.linalg.vector_norm(D)).item()
    assert err_gpu <= TOL, err_gpu
    del D
    rows = [0, 4095, 4096, 16383]
    a = host_rows(A, rows, M, K)
    bh = B.cpu().numpy().reshape(cb, rb).T
    ref = a @ (bh if opB == "N" else bh.T)
    err = oracle.rel_frobenius(host_rows(C, rows, M, N), ref)
    assert err <= TOL, err


def test_config5_grouped_full(dev):
    """Tiles of the 7x5 grid over M=N=K=20000 (rows 6x2857 + 2858, columns
    5x4000) plus 256 GEMMs with log-uniform M, N, K in [64, 4096]: one plan,
    bitwise equal to per-problem dg_dgemm, sampled against the host dgemm."""
    n = 20000
    drow, dcol = dgamd.virtual_grid(35)
    ro, co = dgamd.block_splits(n, drow), dgamd.block_splits(n, dcol)
    assert (drow, dcol) == (7, 5) and ro[-1] - ro[-2] == 2858 and co[1] == 4000
    A, B = gen(n * n, 21, dev), gen(n * n, 22, dev)
    C = torch.full((n * n,), float("nan"), dtype=torch.float64, device=dev)
    descs, checks = [], []
    for i in range(drow):
        for j in range(dcol):
            m_, n_ = ro[i + 1] - ro[i], co[j + 1] - co[j]
            Ai = A[ro[i]:]           # row panel view: element (r, k) at ro[i] + r + k*n
            Bj = B[co[j] * n:]       # column panel: element (k, c) at (co[j] + c)*n + k
            Cij = C[ro[i] + co[j] * n:]
            descs.append(dgamd.gemm_desc("N", "N", m_, n_, n, 1.0, Ai, n, Bj, n, 0.0, Cij, n))
    rng = np.random.default_rng(5)
    extras = []
    for t in range(256):
        m_, n_, k_ = (int(round(np.exp(rng.uniform(np.log(64), np.log(4096))))) for _ in range(3))
        Ae, Be = gen(m_ * k_, 100 + 3 * t, dev), gen(k_ * n_, 101 + 3 * t, dev)
        Ce = torch.full((m_ * n_,), float("nan"), dtype=torch.float64, device=dev)
        extras.append((m_, n_, k_, Ae, Be, Ce))
        descs.append(dgamd.gemm_desc("N", "N", m_, n_, k_, 1.0, Ae, m_, Be, k_, 0.0, Ce, m_))
    plan = dgamd.GroupedPlan(descs, dgamd.DG_PLAN_AUTOTUNE)
    plan.graph_launch()
    torch.cuda.synchronize()
    # bitwise against per-problem DGEMM, tile by tile
    D = torch.full_like(C, float("nan"))
    for i in range(drow):
        for j in range(dcol):
            m_, n_ = ro[i + 1] - ro[i], co[j + 1] - co[j]
            dgamd.dgemm_t("N", "N", m_, n_, n, 1.0, A[ro[i]:], n, B[co[j] * n:], n, 0.0, D[ro[i] + co[j] * n:], n)
    assert torch.equal(C, D)
    del D
    for (m_, n_, k_, Ae, Be, Ce) in extras[::16]:
        De = torch.empty_like(Ce)
        dgamd.dgemm_t("N", "N", m_, n_, k_, 1.0, Ae, m_, Be, k_, 0.0, De, m_)
        assert torch.equal(Ce, De)
        ref = Ae.cpu().numpy().reshape(k_, m_).T @ Be.cpu().numpy().reshape(n_, k_).T
        assert oracle.rel_frobenius(Ce.cpu().numpy().reshape(n_, m_).T, ref) <= TOL
    rows = [0, 2856, 2857, 19999]
    ref = host_rows(A, rows, n, n) @ B.cpu().numpy().reshape(n, n).T
    err = oracle.rel_frobenius(host_rows(C, rows, n, n), ref)
    assert err <= TOL, err
    plan.close()


def test_block2d_full_round_trip_20000(dev):
    n = 20000
    X = gen(n * n, 31, dev)
    ro, co = dgamd.block_splits(n, 7), dgamd.block_splits(n, 5)
    Pk = torch.empty_like(X)
    dgamd.pack_block2d_t(X, n, n, n, ro, co, Pk)
    Y = torch.empty_like(X)
    dgamd.unpack_block2d_t(Pk, ro, co, Y, n, n, n)
    assert torch.equal(X, Y)
    # tile (6, 4) (the 2858 x 4000 corner) sits last, column-major with ld 2858
    base = ro[6] * n + (n - ro[6]) * co[4]
    assert Pk[base].item() == X[ro[6] + co[4] * n].item()
    assert Pk[base + 2857 + 3999 * 2858].item() == X[n - 1 + (n - 1) * n].item()


def test_config5_grouped_ozaki_form_full(dev):
    """The configuration-5 batch (7x5 uneven tiles of M=N=K=20000 plus 256
    log-uniform GEMMs) as one plan in the Ozaki-INT8 form: the autotuner runs
    the large problems as INT8 CRT emulation with the FP64-level moduli count,
    and every output agrees with the FP64 plan's to the FP64 level."""
    n = 20000
    drow, dcol = dgamd.virtual_grid(35)
    ro, co = dgamd.block_splits(n, drow), dgamd.block_splits(n, dcol)
    A, B = gen(n * n, 21, dev), gen(n * n, 22, dev)
    C = torch.full((n * n,), float("nan"), dtype=torch.float64, device=dev)
    descs, outs = [], [C]
    for i in range(drow):
        for j in range(dcol):
            m_, n_ = ro[i + 1] - ro[i], co[j + 1] - co[j]
            descs.append(dgamd.gemm_desc("N", "N", m_, n_, n, 1.0, A[ro[i]:], n, B[co[j] * n:], n, 0.0,
                                         C[ro[i] + co[j] * n:], n))
    rng = np.random.default_rng(5)
    keep = []
    for t in range(256):
        m_, n_, k_ = (int(round(np.exp(rng.uniform(np.log(64), np.log(4096))))) for _ in range(3))
        Ae, Be = gen(m_ * k_, 100 + 3 * t, dev), gen(k_ * n_, 101 + 3 * t, dev)
        Ce = torch.full((m_ * n_,), float("nan"), dtype=torch.float64, device=dev)
        keep += [Ae, Be]
        outs.append(Ce)
        descs.append(dgamd.gemm_desc("N", "N", m_, n_, k_, 1.0, Ae, m_, Be, k_, 0.0, Ce, m_))
    plan = dgamd.GroupedPlan(descs, dgamd.DG_PLAN_AUTOTUNE)
    plan.graph_launch()
    torch.cuda.synchronize()
    ref = [o.clone() for o in outs]
    plan.close()
    for o in outs:
        o.fill_(float("nan"))
    oz = dgamd.GroupedPlan(descs, dgamd.DG_PLAN_AUTOTUNE, form=dgamd.DG_FORM_OZAKI_INT8)
    oz.graph_launch()
    torch.cuda.synchronize()
    forms = [oz.problem_form(i) for i in range(len(descs))]
    n_oz = sum(1 for f, _ in forms if f == dgamd.DG_FORM_OZAKI_INT8)
    assert n_oz >= 1, forms[:40]
    # the FP64-level count for each problem's k: 16 moduli at k = 20000, 15 at the extras' smaller k
    assert all(m in (15, 16) for f, m in forms if f == dgamd.DG_FORM_OZAKI_INT8)
    worst = max((torch.linalg.vector_norm(o - r) / torch.linalg.vector_norm(r)).item() for o, r in zip(outs, ref))
    print("Ozaki-form problems: %d of %d; worst relative error vs the FP64 plan %.3e" % (n_oz, len(descs), worst))
    assert worst <= 1e-13, worst
    oz.close()
    del keep
