"""Uneven FP32 problems with a k-contiguous operand are cut into an interior
block (the register-operand kernel) and the strips after it (the
direct-to-LDS edge kernel).  Both kernels sum every element in the same
order, so the cut must not change a single bit: every case equals the uncut
problem, computed in a child process with DG_GEMM_NO_AREG=1 set (there the
whole problem runs on the direct-to-LDS kernel).  Cases cover every op
combination (NT has no k-contiguous operand and is never cut), row and
column strips, odd step counts (not cut) and C with an odd ldc."""
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

CASES = [
    # (opA, opB, m, n, k, ldc pad)
    ("N", "N", 1000, 900, 512, 0),    # rows and columns after the interior
    ("T", "N", 1031, 777, 320, 0),
    ("T", "T", 2200, 1300, 256, 1),   # odd ldc: element stores in every launch
    ("N", "T", 1000, 900, 512, 0),    # no k-contiguous operand: never cut
    ("N", "N", 1156, 1024, 272, 0),   # 17 k-steps: not cut
    ("N", "N", 1024, 1111, 96, 0),    # columns only
]


def compute(dev):
    """Runs every case through dg_sgemm on `dev`; returns the C buffers."""
    import torch
    import dgamd
    outs = {}
    for i, (opA, opB, m, n, k, padc) in enumerate(CASES):
        g = torch.Generator(device="cpu").manual_seed(1000 + i)
        ra, ca = (m, k) if opA == "N" else (k, m)
        rb, cb = (k, n) if opB == "N" else (n, k)
        A = (torch.rand(ca, ra, generator=g) * 2 - 1).to(dev)
        B = (torch.rand(cb, rb, generator=g) * 2 - 1).to(dev)
        ldc = m + padc
        C = (torch.rand(n, ldc, generator=g) * 2 - 1).to(dev)
        dgamd.sgemm_t(opA, opB, m, n, k, 1.25, A, ra, B, rb, -0.5, C, ldc)
        torch.cuda.synchronize()
        outs["c%d" % i] = C.cpu().numpy()
    return outs


pytestmark = pytest.mark.gpu


def test_cut_is_bit_identical(tmp_path):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    out = tmp_path / "uncut.npz"
    env = dict(os.environ, DG_GEMM_NO_AREG="1")
    print("sgemm cut: child computes the uncut cases", flush=True)
    subprocess.run([sys.executable, os.path.abspath(__file__), str(out)], env=env, check=True, timeout=240)
    ref = np.load(out)
    got = compute(torch.device("cuda", 0))
    for i, case in enumerate(CASES):
        print("sgemm cut case", case, flush=True)
        assert np.array_equal(got["c%d" % i].view(np.uint32), ref["c%d" % i].view(np.uint32)), case


if __name__ == "__main__":
    # child: the uncut results (DG_GEMM_NO_AREG set by the parent)
    sys.path.insert(0, os.path.join(ROOT, "distributed-gemm-thesis_amd"))
    import torch
    np.savez(sys.argv[1], **compute(torch.device("cuda", 0)))
