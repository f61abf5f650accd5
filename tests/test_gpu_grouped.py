"""Grouped uneven batches (dg_grouped_dgemm and plans): the tiles of a
non-divisible 7x5 grid plus extra random GEMMs, in one launch, against the
host dgemm (relative Frobenius error <= 1e-13 per problem) and bit for bit
across scheduling variants and HIP-graph replays."""
import numpy as np
import pytest
import torch

import dgamd
import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-13


def make_batch(dev, n_grid=700, extras=24, seed=0, beta=0.0):
    """7x5 BlockSequential tiles over an n_grid^3 problem (last row/column
    absorb the remainder) plus `extras` GEMMs with log-uniform sizes."""
    rng = np.random.default_rng(seed)
    drow, dcol = dgamd.virtual_grid(35)
    assert (drow, dcol) == (7, 5)
    ro, co = dgamd.block_splits(n_grid, drow), dgamd.block_splits(n_grid, dcol)
    K = n_grid
    shapes = [(ro[i + 1] - ro[i], co[j + 1] - co[j], K) for i in range(drow) for j in range(dcol)]
    for _ in range(extras):
        shapes.append(tuple(int(round(np.exp(rng.uniform(np.log(16), np.log(400))))) for _ in range(3)))
    probs = []
    for (m, n, k) in shapes:
        a = rng.uniform(-1, 1, (m, k))
        b = rng.uniform(-1, 1, (k, n))
        c = rng.uniform(-1, 1, (m, n))
        A = torch.from_numpy(np.asfortranarray(a).ravel(order="F").copy()).to(dev)
        B = torch.from_numpy(np.asfortranarray(b).ravel(order="F").copy()).to(dev)
        C = torch.from_numpy(np.asfortranarray(c).ravel(order="F").copy()).to(dev)
        probs.append(dict(m=m, n=n, k=k, a=a, b=b, c=c, A=A, B=B, C=C, C0=C.clone()))
    descs = [dgamd.gemm_desc("N", "N", p["m"], p["n"], p["k"], 1.0, p["A"], p["m"], p["B"], p["k"], beta, p["C"],
                             p["m"]) for p in probs]
    return probs, descs


def check(probs, beta):
    for p in probs:
        got = p["C"].cpu().numpy().reshape(p["n"], p["m"]).T
        ref = oracle.host_dgemm("N", "N", 1.0, p["a"], p["b"], beta, p["c"])
        err = oracle.rel_frobenius(got, ref)
        assert err <= TOL, (p["m"], p["n"], p["k"], err)


def test_grouped_one_call(dev):
    probs, descs = make_batch(dev, beta=0.5)
    dgamd.grouped_dgemm_t(descs)
    check(probs, 0.5)


@pytest.mark.parametrize("mode", [dgamd.DG_PLAN_NONPERSISTENT, dgamd.DG_PLAN_PERSISTENT, dgamd.DG_PLAN_AUTOTUNE])
def test_plan_modes_bit_identical(dev, mode):
    probs, descs = make_batch(dev, seed=1)
    dgamd.grouped_dgemm_t(descs)
    ref = [p["C"].clone() for p in probs]
    for p in probs:
        p["C"].copy_(p["C0"])
    plan = dgamd.GroupedPlan(descs, mode)
    plan.execute()
    torch.cuda.synchronize()
    assert plan.variant() in (dgamd.DG_PLAN_NONPERSISTENT, dgamd.DG_PLAN_PERSISTENT)
    for p, r in zip(probs, ref):
        assert torch.equal(p["C"], r)
    check(probs, 0.0)
    plan.close()


def test_autotune_leaves_outputs_untouched_with_beta(dev):
    probs, descs = make_batch(dev, seed=2, beta=1.0, extras=8)
    plan = dgamd.GroupedPlan(descs, dgamd.DG_PLAN_AUTOTUNE)
    plan.execute()   # tunes (several trial runs), then runs once
    torch.cuda.synchronize()
    check(probs, 1.0)
    plan.close()


def test_graph_replays_match_eager(dev):
    probs, descs = make_batch(dev, seed=3)
    plan = dgamd.GroupedPlan(descs, dgamd.DG_PLAN_PERSISTENT)
    plan.execute()
    torch.cuda.synchronize()
    eager = [p["C"].clone() for p in probs]
    for rep in range(4):
        for p in probs:
            p["C"].fill_(float("nan"))   # eager work between the replays
        torch.cuda.synchronize()
        plan.graph_launch()
        torch.cuda.synchronize()
        for p, e in zip(probs, eager):
            assert torch.equal(p["C"], e), rep
    plan.close()


def test_mixed_ops_and_empty(dev):
    rng = np.random.default_rng(4)
    descs, probs = [], []
    for opA, opB, m, n, k in [("N", "T", 100, 70, 50), ("T", "N", 64, 64, 64), ("T", "T", 33, 17, 129),
                              ("N", "N", 0, 10, 10), ("N", "N", 20, 30, 0)]:
        a = rng.uniform(-1, 1, (m, k) if opA == "N" else (k, m))
        b = rng.uniform(-1, 1, (k, n) if opB == "N" else (n, k))
        c = rng.uniform(-1, 1, (m, n))
        up = lambda x: torch.from_numpy(np.asfortranarray(x).ravel(order="F").copy()).to(dev)
        A, B, C = up(a), up(b), up(c)
        descs.append(dgamd.gemm_desc(opA, opB, m, n, k, 2.0, A, max(1, a.shape[0]), B, max(1, b.shape[0]), -1.0, C,
                                     max(1, m)))
        probs.append((opA, opB, m, n, a, b, c, C, A, B))
    dgamd.grouped_dgemm_t(descs)
    for opA, opB, m, n, a, b, c, C, *_ in probs:
        if m == 0 or n == 0:
            continue
        got = C.cpu().numpy().reshape(n, m).T
        ref = oracle.host_dgemm(opA, opB, 2.0, a, b, -1.0, c)
        assert oracle.rel_frobenius(got, ref) <= TOL


def test_plan_rejects_bad_input(dev):
    L = dgamd.lib()
    bad = dgamd.gemm_desc("N", "N", 10, 10, 10, 1.0, 0, 5, 0, 10, 0.0, 0, 10)
    arr = (dgamd.GemmDesc * 1)(bad)
    assert L.dg_grouped_dgemm(arr, 1, None) == dgamd.DG_ERR_INVALID


def test_ozaki_form_plan(dev):
    """DG_FORM_OZAKI_INT8: the autotuner runs the 8192^3 problem as INT8 CRT
    emulation (16 moduli, FP64-level scaling), keeps problems under 2^32
    flops in FP64, and every output stays within 1e-13 of the FP64 MFMA
    result; a graph replay reproduces the eager outputs bit for bit."""
    shapes = [(8192, 8192, 8192), (3000, 2000, 5000), (300, 200, 100), (64, 64, 64)]
    probs = []
    for t, (m, n, k) in enumerate(shapes):
        A = torch.empty(m * k, dtype=torch.float64, device=dev)
        B = torch.empty(k * n, dtype=torch.float64, device=dev)
        C = torch.empty(m * n, dtype=torch.float64, device=dev)
        dgamd.generate_t(500 + 3 * t, dgamd.DG_DIST_WIDE, A)
        dgamd.generate_t(501 + 3 * t, dgamd.DG_DIST_WIDE, B)
        dgamd.generate_t(502 + 3 * t, dgamd.DG_DIST_UNIFORM_PM1, C)
        ref = C.clone()
        dgamd.dgemm_t("N", "N", m, n, k, 1.0, A, m, B, k, 0.5, ref, m)
        probs.append(dict(m=m, n=n, k=k, A=A, B=B, C=C, C0=C.clone(), ref=ref))
    descs = [dgamd.gemm_desc("N", "N", p["m"], p["n"], p["k"], 1.0, p["A"], p["m"], p["B"], p["k"], 0.5, p["C"],
                             p["m"]) for p in probs]
    plan = dgamd.GroupedPlan(descs, dgamd.DG_PLAN_AUTOTUNE, form=dgamd.DG_FORM_OZAKI_INT8)
    assert plan.problem_form(0) == (-1, 0)  # chosen on first use
    plan.execute()
    torch.cuda.synchronize()
    forms = [plan.problem_form(i) for i in range(len(probs))]
    assert forms[0] == (dgamd.DG_FORM_OZAKI_INT8, 16), forms
    assert forms[2] == (dgamd.DG_FORM_FP64, 0) and forms[3] == (dgamd.DG_FORM_FP64, 0), forms
    assert forms[1][0] in (dgamd.DG_FORM_FP64, dgamd.DG_FORM_OZAKI_INT8)
    eager = []
    for p in probs:
        err = (torch.linalg.vector_norm(p["C"] - p["ref"]) / torch.linalg.vector_norm(p["ref"])).item()
        assert err <= TOL, (p["m"], p["n"], p["k"], err)
        eager.append(p["C"].clone())
    for p in probs:
        p["C"].copy_(p["C0"])
    for _ in range(2):
        for p in probs:
            p["C"].copy_(p["C0"])
        torch.cuda.synchronize()
        plan.graph_launch()
        torch.cuda.synchronize()
        for p, e in zip(probs, eager):
            assert torch.equal(p["C"], e)
    plan.close()


def test_ozaki_form_fixed_moduli(dev):
    """An explicit moduli count: the large problem uses it and stays within
    the scheme's a-priori bound (13 moduli: q = 44 bits at k = 4096)."""
    m = n = k = 4096
    A = torch.empty(m * k, dtype=torch.float64, device=dev)
    B = torch.empty(k * n, dtype=torch.float64, device=dev)
    C = torch.zeros(m * n, dtype=torch.float64, device=dev)
    dgamd.generate_t(61, dgamd.DG_DIST_UNIFORM_PM1, A)
    dgamd.generate_t(62, dgamd.DG_DIST_UNIFORM_PM1, B)
    ref = torch.empty_like(C)
    dgamd.dgemm_t("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, ref, m)
    plan = dgamd.GroupedPlan([dgamd.gemm_desc("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)],
                             dgamd.DG_PLAN_NONPERSISTENT, form=dgamd.DG_FORM_OZAKI_INT8, moduli=13)
    plan.execute()
    torch.cuda.synchronize()
    f, mod = plan.problem_form(0)
    err = (torch.linalg.vector_norm(C - ref) / torch.linalg.vector_norm(ref)).item()
    if f == dgamd.DG_FORM_OZAKI_INT8:
        assert mod == 13
        assert err <= 1e-9, err
    else:
        assert (f, mod) == (dgamd.DG_FORM_FP64, 0) and err <= TOL
    plan.close()


def test_ozaki_form_shared_operands(dev):
    """Ozaki-form problems that read the same lines (the 2x2 tiles of one
    8192^3 product share tile rows of A and tile columns of B; a fifth
    problem reads a block of A transposed) are prepared once per operand and
    still give, bit for bit, what dg_ozaki_dgemm gives each problem alone;
    a graph replay reproduces them."""
    n, h = 8192, 4096
    A = torch.empty(n * n, dtype=torch.float64, device=dev)
    B = torch.empty(n * n, dtype=torch.float64, device=dev)
    dgamd.generate_t(701, dgamd.DG_DIST_WIDE, A)
    dgamd.generate_t(702, dgamd.DG_DIST_WIDE, B)
    outs, descs, args = [], [], []
    for i in range(2):
        for j in range(2):
            C = torch.empty(h * h, dtype=torch.float64, device=dev)
            dgamd.generate_t(710 + 2 * i + j, dgamd.DG_DIST_UNIFORM_PM1, C)
            a = ("N", "N", h, h, n, 1.0, A[h * i:], n, B[h * j * n:], n, 0.5, C, h)
            outs.append(C)
            args.append(a)
    # op(A) = A(0:4096, 0:4096)^T, op(B) = the first tile column of B: not the same lines as any tile's
    C5 = torch.empty(h * h, dtype=torch.float64, device=dev)
    dgamd.generate_t(720, dgamd.DG_DIST_UNIFORM_PM1, C5)
    args.append(("T", "N", h, h, h, -2.0, A, n, B, n, 0.25, C5, h))
    outs.append(C5)
    descs = [dgamd.gemm_desc(*a) for a in args]
    starts = [o.clone() for o in outs]
    plan = dgamd.GroupedPlan(descs, dgamd.DG_PLAN_AUTOTUNE, form=dgamd.DG_FORM_OZAKI_INT8)
    plan.execute()
    torch.cuda.synchronize()
    forms = [plan.problem_form(i) for i in range(len(descs))]
    assert all(f == dgamd.DG_FORM_OZAKI_INT8 for f, _ in forms[:4]), forms
    got = [o.clone() for o in outs]
    for t, (a, (f, mod)) in enumerate(zip(args, forms)):
        ref = starts[t].clone()
        opA, opB, m_, n_, k_, alpha, Aa, lda, Bb, ldb, beta, _, ldc = a
        if f == dgamd.DG_FORM_OZAKI_INT8:
            dgamd.ozaki_dgemm_t(opA, opB, m_, n_, k_, alpha, Aa, lda, Bb, ldb, beta, ref, ldc, mod,
                                kind=dgamd.DG_OZAKI_INT8_CRT)
        else:
            dgamd.dgemm_t(opA, opB, m_, n_, k_, alpha, Aa, lda, Bb, ldb, beta, ref, ldc)
        torch.cuda.synchronize()
        assert torch.equal(got[t], ref), (t, f, mod)
    for o, s0 in zip(outs, starts):
        o.copy_(s0)
    torch.cuda.synchronize()
    plan.graph_launch()
    torch.cuda.synchronize()
    for o, g in zip(outs, got):
        assert torch.equal(o, g)
    plan.close()
