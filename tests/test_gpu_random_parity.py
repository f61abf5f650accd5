"""Seeded random parity sweep, every result bit-exact.

Covers dg_dgemm, dg_sgemm, the grouped batch and the CRT Ozaki DGEMM over
random ops, shapes, leading-dimension padding, base offsets, alpha and beta.
Shapes are log-uniform, with a share of whole tiles and whole k-steps so the
lean instantiations run too.

The GEMM data are small integers, whose products and sums are exact in FP32
and FP64 whatever the summation order, so they are compared with the exact
host product. The Ozaki results are compared with the oracle's scheme II
restatement, on wide-range inputs.
"""
import math

import numpy as np
import pytest
import torch

import dgamd
import oracle

pytestmark = pytest.mark.gpu


class Gen:
    def __init__(self, seed, maxd):
        self.rng = np.random.default_rng(seed)
        self.maxd = maxd

    def logu(self, lo, hi):
        return int(round(math.exp(self.rng.uniform(math.log(lo), math.log(hi)))))

    def problem(self, dev, single):
        rng = self.rng
        opA, opB = str(rng.choice(["N", "T"])), str(rng.choice(["N", "T"]))
        m, n = self.logu(1, self.maxd), self.logu(1, self.maxd)
        k = self.logu(1, 500 if single else self.maxd)
        if rng.random() < 0.3:  # whole 128 x 128 tiles and whole k-steps
            m, n, k = 128 * max(1, m // 128), 128 * max(1, n // 128), 16 * max(1, k // 16)
        ra, ca = (m, k) if opA == "N" else (k, m)
        rb, cb = (k, n) if opB == "N" else (n, k)
        pads = rng.integers(0, 6, 3) if rng.random() < 0.5 else np.zeros(3, int)
        offs = rng.integers(0, 4, 3) if rng.random() < 0.5 else np.zeros(3, int)
        lda, ldb, ldc = ra + int(pads[0]), rb + int(pads[1]), m + int(pads[2])
        alpha = float(rng.choice([1.0, 2.0, -1.0, 0.5]))
        beta = float(rng.choice([0.0, 1.0, -2.0, 0.25]))
        dt = np.float32 if single else np.float64
        a = rng.integers(-8, 9, (ra, ca)).astype(dt)
        b = rng.integers(-8, 9, (rb, cb)).astype(dt)
        c = rng.integers(-8, 9, (m, n)).astype(dt)

        def buf(rows, cols, ld, off, vals):
            x = np.full((cols, ld), np.nan, dtype=dt)
            x[:, :rows] = vals.T
            flat = np.concatenate([np.full(off, np.nan, dtype=dt), x.reshape(-1)])
            return torch.from_numpy(flat).to(dev)[off:]

        A, B, C = buf(ra, ca, lda, int(offs[0]), a), buf(rb, cb, ldb, int(offs[1]), b), buf(m, n, ldc, int(offs[2]), c)
        if beta == 0.0:
            C.view(n, ldc)[:, :m] = float("nan")  # beta == 0 must not read C
        opa = a.astype(np.float64) if opA == "N" else a.T.astype(np.float64)
        opb = b.astype(np.float64) if opB == "N" else b.T.astype(np.float64)
        ref = alpha * (opa @ opb) + (beta * c.astype(np.float64) if beta != 0.0 else 0.0)
        return dict(opA=opA, opB=opB, m=m, n=n, k=k, A=A, B=B, C=C, lda=lda, ldb=ldb, ldc=ldc, alpha=alpha,
                    beta=beta, ref=ref)


def result(p):
    torch.cuda.synchronize()
    return p["C"].cpu().numpy().reshape(p["n"], p["ldc"])[:, :p["m"]].T.astype(np.float64)


def args(p):
    return (p["opA"], p["opB"], p["m"], p["n"], p["k"], p["alpha"], p["A"], p["lda"], p["B"], p["ldb"], p["beta"],
            p["C"], p["ldc"])


def shape(p):
    return {x: p[x] for x in ("opA", "opB", "m", "n", "k", "lda", "ldb", "ldc", "alpha", "beta")}


@pytest.mark.parametrize("seed", [11, 12])
def test_random_dgemm_sgemm_grouped(dev, seed):
    g = Gen(seed, 900)
    for it in range(30):
        p = g.problem(dev, single=False)
        dgamd.dgemm_t(*args(p))
        assert np.array_equal(result(p), p["ref"]), ("dgemm", it, shape(p))
        p = g.problem(dev, single=True)
        dgamd.sgemm_t(*args(p))
        assert np.array_equal(result(p), p["ref"]), ("sgemm", it, shape(p))
    probs = [g.problem(dev, single=False) for _ in range(12)]
    dgamd.grouped_dgemm_t([dgamd.gemm_desc(*args(p)) for p in probs])
    for p in probs:
        assert np.array_equal(result(p), p["ref"]), ("grouped", shape(p))


@pytest.mark.parametrize("seed", [21, 22])
def test_random_crt_ozaki_vs_oracle(dev, seed):
    g = Gen(seed, 300)
    rng = g.rng
    for it in range(4):
        m, n, k = g.logu(1, 300), g.logu(1, 300), g.logu(1, 600)
        nm = int(rng.integers(2, 17))
        a = np.ldexp(rng.uniform(-1, 1, (m, k)), rng.integers(-20, 21, (m, k)))
        b = np.ldexp(rng.uniform(-1, 1, (k, n)), rng.integers(-20, 21, (k, n)))
        A = torch.from_numpy(np.asfortranarray(a).ravel(order="F")).to(dev)
        B = torch.from_numpy(np.asfortranarray(b).ravel(order="F")).to(dev)
        C = torch.zeros(m * n, dtype=torch.float64, device=dev)
        if oracle.crt_Q(k, nm) is None:
            with pytest.raises(dgamd.DGError):
                dgamd.ozaki_dgemm_t("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m, nm, dgamd.DG_OZAKI_INT8_CRT)
            continue
        dgamd.ozaki_dgemm_t("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m, nm, dgamd.DG_OZAKI_INT8_CRT)
        torch.cuda.synchronize()
        got = C.cpu().numpy().reshape(n, m).T
        assert np.array_equal(got, oracle.ozaki2_dgemm(a, b, nm)), (it, m, n, k, nm)
