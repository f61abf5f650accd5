"""Throughput of the per-rank GEMM hot path on one MI355X.

Default workload: FP64 SUMMA on a virtual 4x4 process grid, M=N=K=16384,
block-cyclic 512x512 blocks, A, B, C uniform[-1,1) from fixed seeds,
alpha=1, beta=0.5.  One step = one dg_summa call over the packed
block-cyclic operands already resident in HBM (every virtual rank's
local GEMM, all 32 panels, one launch).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload summa|cannon|dgemm|grouped|cannon2x2|sgemm]

Prints one JSON line.  `roofline` is the dominant kernel against the FP64
matrix peak (78.6 TFLOP/s; the FP32 matrix peak, 157.3, for sgemm), timed with HIP events on the launch stream;
`hbm` times the HBM-bound pack/unpack, copy, layer-reduction and
generation kernels at the configurations' shapes against the 8 TB/s peak;
`cpu_baseline` is the host BLAS dgemm of the oracle on the workload's own
global product and data, on this host's cores (median of 3).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "distributed-gemm-thesis_amd"))

FP64_PEAK = 78.6  # TFLOP/s, MI355X dense FP64 matrix
FP32_PEAK = 157.3  # TFLOP/s, MI355X dense FP32 matrix
HBM_PEAK = 8000.0  # GB/s

WORKLOADS = {
    "summa": "FP64 SUMMA 4x4 virtual grid, M=N=K=16384, block-cyclic nb=512, alpha=1, beta=0.5",
    "cannon": "FP64 2.5D Cannon 4x4x2 virtual grid, M=16384 N=8192 K=32768, opA=N opB={opB}, layer reduce in HBM",
    "dgemm": "FP64 DGEMM NN M=N=K=16384, alpha=1, beta=0",
    "grouped": "FP64 grouped batch: 7x5 uneven tiles of M=N=K=20000 + 256 GEMMs with log-uniform M,N,K in "
               "[64,4096], one autotuned plan replayed as a HIP graph",
    "cannon2x2": "FP64 2D Cannon 2x2x1 virtual grid, M=N=K=2048, 1024^2 tiles, alpha=1, beta=0",
    "sgemm": "FP32 SGEMM NN M=N=K=16384, alpha=1, beta=0 (FP32 matrix cores)",
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="summa")
    ap.add_argument("--opB", choices=["N", "T"], default="N",
                    help="cannon workload: B stored K x N (N) or N x K with ldb = N (T)")
    ap.add_argument("--no-hbm", action="store_true", help="skip the HBM leg (pack, reduce, copy, generate)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ozaki-slices", default="5,6,7,8,9",
                    help="INT8 slice counts of the Ozaki leg (empty: skip the leg)")
    ap.add_argument("--ozaki-bf16-slices", default="11", help="BF16 slice counts of the Ozaki leg")
    ap.add_argument("--ozaki-crt-moduli", default="13,14,15,16",
                    help="moduli counts of the Ozaki scheme II (INT8, CRT) leg")
    return ap.parse_args()


def load_traffic(kernel_tag):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE),
    committed as profiles/traffic_<tag>.json; None when not measured."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % kernel_tag)
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


def setup(workload, dev, opB="N"):
    import torch
    import dgamd

    f64 = dict(dtype=torch.float64, device=dev)
    if workload == "summa":
        n, nb, P = 16384, 512, 4
        M = N = K = n
        G = torch.empty(n * n, **f64)
        packed = []
        for seed in (1, 2, 3):
            dgamd.generate_t(seed, dgamd.DG_DIST_UNIFORM_PM1, G)
            Pk = torch.empty(n * n, **f64)
            dgamd.pack_block_cyclic_t(G, n, n, n, nb, nb, P, P, Pk)
            packed.append(Pk)
        del G
        Ap, Bp, Cp = packed

        def step():
            dgamd.summa_t("N", "N", M, N, K, 1.0, Ap, Bp, 0.5, Cp, nb, nb, nb, P, P)

        cfg = {"workload": WORKLOADS[workload], "M": M, "N": N, "K": K, "grid": "4x4", "block": nb,
               "alpha": 1.0, "beta": 0.5, "layout": "block-cyclic, column-major local matrices"}
        return step, (M, N, K), cfg, "summa16384"
    if workload == "cannon":
        M, N, K, p, c = 16384, 8192, 32768, 4, 2
        bufs = []
        shapes = ((M, K), (K, N) if opB == "N" else (N, K), (M, N))
        for seed, (r, cc) in zip((1, 2, 3), shapes):
            G = torch.empty(r * cc, **f64)
            dgamd.generate_t(seed, dgamd.DG_DIST_UNIFORM_PM1, G)
            Pk = torch.empty_like(G)
            dgamd.pack_block2d_t(G, r, r, cc, dgamd.block_splits(r, p), dgamd.block_splits(cc, p), Pk)
            del G
            bufs.append(Pk)
        Ap, Bp, Cp = bufs
        ws = torch.empty(dgamd.cannon25d_workspace(M, N, p, c), **f64)

        def step():
            dgamd.cannon25d_t("N", opB, M, N, K, 1.0, Ap, Bp, 0.0, Cp, p, c, dgamd.DG_SCHED_AUTO, ws)

        cfg = {"workload": WORKLOADS[workload].format(opB=opB), "M": M, "N": N, "K": K, "grid": "4x4x2",
               "alpha": 1.0, "beta": 0.0, "opA": "N", "opB": opB}
        return step, (M, N, K), cfg, "cannon16384x8192x32768" + ("nt" if opB == "T" else "")
    if workload == "grouped":
        import numpy as np
        n = 20000
        drow, dcol = dgamd.virtual_grid(35)
        ro, co = dgamd.block_splits(n, drow), dgamd.block_splits(n, dcol)
        A = torch.empty(n * n, **f64)
        B = torch.empty(n * n, **f64)
        C = torch.empty(n * n, **f64)
        dgamd.generate_t(21, 0, A)
        dgamd.generate_t(22, 0, B)
        descs, keep = [], [A, B, C]
        flops = 2.0 * n ** 3
        for i in range(drow):
            for j in range(dcol):
                descs.append(dgamd.gemm_desc("N", "N", ro[i + 1] - ro[i], co[j + 1] - co[j], n, 1.0,
                                             A.data_ptr() + 8 * ro[i], n, B.data_ptr() + 8 * co[j] * n, n, 0.0,
                                             C.data_ptr() + 8 * (ro[i] + co[j] * n), n))
        rng = np.random.default_rng(5)
        for t in range(256):
            m_, n_, k_ = (int(round(np.exp(rng.uniform(np.log(64), np.log(4096))))) for _ in range(3))
            Ae, Be, Ce = torch.empty(m_ * k_, **f64), torch.empty(k_ * n_, **f64), torch.empty(m_ * n_, **f64)
            dgamd.generate_t(100 + 3 * t, 0, Ae)
            dgamd.generate_t(101 + 3 * t, 0, Be)
            keep += [Ae, Be, Ce]
            descs.append(dgamd.gemm_desc("N", "N", m_, n_, k_, 1.0, Ae, m_, Be, k_, 0.0, Ce, m_))
            flops += 2.0 * m_ * n_ * k_
        plan = dgamd.GroupedPlan(descs, dgamd.DG_PLAN_AUTOTUNE)
        plan.execute()  # autotunes once

        def step():
            plan.graph_launch()

        step.keep = (keep, plan)
        step.descs = descs
        step.outs = [C] + keep[5::3]  # the 7x5 grid's C, then every extra problem's C
        step.flops = flops
        cfg = {"workload": WORKLOADS[workload], "problems": len(descs), "flop_per_step": flops}
        return step, (flops / 2.0, 1, 1), cfg, "grouped20000"
    if workload == "cannon2x2":
        M = N = K = 2048
        p = 2
        bufs = []
        for seed in (1, 2):
            G = torch.empty(M * M, **f64)
            dgamd.generate_t(seed, 0, G)
            Pk = torch.empty_like(G)
            dgamd.pack_block2d_t(G, M, M, M, dgamd.block_splits(M, p), dgamd.block_splits(M, p), Pk)
            bufs.append(Pk)
        Cp = torch.empty(M * N, **f64)

        def step():
            dgamd.cannon25d_t("N", "N", M, N, K, 1.0, bufs[0], bufs[1], 0.0, Cp, p, 1, dgamd.DG_SCHED_REFERENCE)

        return step, (M, N, K), {"workload": WORKLOADS[workload], "M": M, "N": N, "K": K, "grid": "2x2x1"}, "cannon2048"
    if workload == "sgemm":
        n = 16384
        f32 = dict(dtype=torch.float32, device=dev)
        A = torch.empty(n * n, **f32)
        B = torch.empty(n * n, **f32)
        C = torch.empty(n * n, **f32)
        dgamd.generate_t(1, 0, A)
        dgamd.generate_t(2, 0, B)

        def step():
            dgamd.sgemm_t("N", "N", n, n, n, 1.0, A, n, B, n, 0.0, C, n)

        return step, (n, n, n), {"workload": WORKLOADS[workload], "M": n, "N": n, "K": n}, "sgemm16384"
    n = 16384
    A = torch.empty(n * n, **f64)
    B = torch.empty(n * n, **f64)
    C = torch.empty(n * n, **f64)
    dgamd.generate_t(1, 0, A)
    dgamd.generate_t(2, 0, B)

    def step():
        dgamd.dgemm_t("N", "N", n, n, n, 1.0, A, n, B, n, 0.0, C, n)

    return step, (n, n, n), {"workload": WORKLOADS["dgemm"], "M": n, "N": n, "K": n}, "dgemm16384"


INT8_PEAK = 5030.0   # TOPS, MI355X dense INT8 matrix
BF16_PEAK = 2516.0   # TFLOP/s, MI355X dense BF16 matrix


def ozaki_leg(dev, int8_slices, bf16_slices, crt_moduli=(), reps=2):
    """Ozaki split DGEMM at M=N=K=16384 on wide-range inputs u*2^e
    (u uniform[-1,1), e uniform in [-20,20]): FP64-equivalent TFLOP/s
    (2MNK/t), utilisation of the low-precision matrix peak counting the
    s(s+1)/2 slice products, and relative Frobenius error against the
    native FP64 MFMA DGEMM of the same inputs.  One call includes the
    exponent scan, the slicing and the recombination."""
    import torch
    import dgamd
    n = 16384
    f64 = dict(dtype=torch.float64, device=dev)
    A = torch.empty(n * n, **f64)
    B = torch.empty(n * n, **f64)
    dgamd.generate_t(4, dgamd.DG_DIST_WIDE, A)
    dgamd.generate_t(5, dgamd.DG_DIST_WIDE, B)
    ref = torch.empty(n * n, **f64)
    dgamd.dgemm_t("N", "N", n, n, n, 1.0, A, n, B, n, 0.0, ref, n)
    C = torch.empty(n * n, **f64)
    flops = 2.0 * n ** 3
    ref_norm = torch.linalg.vector_norm(ref).item()
    # Near-exact reference on a sample: 4 rows x 2048 columns of C from the
    # same inputs, summed in 64-bit-mantissa long double on the host (error
    # ~1e-18 relative, far below the FP64 errors compared).
    import numpy as np
    rows = torch.linspace(0, n - 1, 4, device=dev).long()
    cols = torch.linspace(0, n - 1, 2048, device=dev).long()
    a_s = A.view(n, n)[:, rows].T.cpu().numpy().astype(np.longdouble)      # op(A) rows, 4 x K
    b_s = B.view(n, n)[cols, :].T.cpu().numpy().astype(np.longdouble)      # op(B) columns, K x 2048
    exact = a_s @ b_s
    exact_norm = np.sqrt(np.sum(exact * exact))

    def err_exact(Ct):
        got = Ct.view(n, n)[cols][:, rows].T.cpu().numpy().astype(np.longdouble)
        d = got - exact
        return float(np.sqrt(np.sum(d * d)) / exact_norm)

    native_err = err_exact(ref)
    out = {"config": "M=N=K=16384, A,B = u*2^e wide-range (seeds 4,5), alpha=1, beta=0; errors: relative Frobenius "
                     "vs the native FP64 MFMA result (whole C) and vs a long-double host reference on 4 rows x "
                     "2048 columns",
           "native_fp64_rel_err_vs_exact": float("%.3e" % native_err),
           "int8": {}, "bf16": {}, "int8_crt": {}}
    legs = [("int8", dgamd.DG_OZAKI_INT8, int8_slices, INT8_PEAK), ("bf16", dgamd.DG_OZAKI_BF16, bf16_slices, BF16_PEAK),
            ("int8_crt", dgamd.DG_OZAKI_INT8_CRT, crt_moduli, INT8_PEAK)]
    for name, kind, slices, peak in legs:
        for s in slices:
            ws = torch.empty(dgamd.ozaki_workspace(n, n, n, s, kind), dtype=torch.uint8, device=dev)
            dgamd.ozaki_dgemm_t("N", "N", n, n, n, 1.0, A, n, B, n, 0.0, C, n, s, kind, ws)  # warm-up
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                dgamd.ozaki_dgemm_t("N", "N", n, n, n, 1.0, A, n, B, n, 0.0, C, n, s, kind, ws)
            e1.record()
            torch.cuda.synchronize()
            t = e0.elapsed_time(e1) / 1e3 / reps
            err = torch.linalg.vector_norm(C - ref).item() / ref_norm
            nprod = s if kind == dgamd.DG_OZAKI_INT8_CRT else s * (s + 1) // 2
            ee = err_exact(C)
            out[name][str(s)] = {"tflops_fp64_equiv": round(flops / t / 1e12, 2), "ms": round(t * 1e3, 2),
                                 "products": nprod, "lowp_frac": round(nprod * flops / t / 1e12 / peak, 4),
                                 "rel_err_vs_fp64": float("%.3e" % err), "rel_err_vs_exact": float("%.3e" % ee),
                                 "err_ratio_vs_native": round(ee / native_err, 2)}
            del ws
    return out


def grouped_ozaki_leg(step, reps=3):
    """The grouped batch in the Ozaki form (DG_FORM_OZAKI_INT8, FP64-level
    moduli count per problem, form chosen per problem by the autotuner), one
    graph launch per step: FP64-equivalent TFLOP/s, how many problems run as
    INT8 CRT emulation, and the largest relative Frobenius error of any
    output against the FP64 plan's."""
    import torch
    import dgamd
    refs = [t.clone() for t in step.outs]
    plan = dgamd.GroupedPlan(step.descs, dgamd.DG_PLAN_AUTOTUNE, form=dgamd.DG_FORM_OZAKI_INT8)
    plan.graph_launch()  # chooses the forms, tunes, captures
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        plan.graph_launch()
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) / 1e3 / reps
    forms = [plan.problem_form(i) for i in range(len(step.descs))]
    n_oz = sum(1 for f, _ in forms if f == dgamd.DG_FORM_OZAKI_INT8)
    err = max((torch.linalg.vector_norm(o - r) / torch.linalg.vector_norm(r)).item() for o, r in zip(step.outs, refs))
    plan.close()
    return {"tflops_fp64_equiv": round(step.flops / t / 1e12, 2), "ms": round(t * 1e3, 2), "problems": len(forms),
            "problems_ozaki_int8": n_oz, "moduli": sorted({m for f, m in forms if f == dgamd.DG_FORM_OZAKI_INT8}),
            "max_rel_err_vs_fp64": float("%.3e" % err)}


DATA = {
    "summa": "synthetic: SplitMix64-seeded uniform[-1,1) A, B, C (seeds 1, 2, 3), generated and packed on the GPU",
    "cannon": "synthetic: SplitMix64-seeded uniform[-1,1) A, B, C (seeds 1, 2, 3), generated and packed on the GPU",
    "dgemm": "synthetic: SplitMix64-seeded uniform[-1,1) A, B (seeds 1, 2), generated on the GPU",
    "grouped": "synthetic: SplitMix64-seeded uniform[-1,1); 20000^2 A, B seeds 21, 22; extra problem t seeds "
               "100+3t, 101+3t; shapes from numpy default_rng(5)",
    "cannon2x2": "synthetic: SplitMix64-seeded uniform[-1,1) A, B (seeds 1, 2), generated and packed on the GPU",
    "sgemm": "synthetic: SplitMix64-seeded uniform[-1,1) A, B (seeds 1, 2) rounded to FP32, generated on the GPU",
}

# The host product each workload's CPU baseline times: the workload's own
# global product, on the same generated data (the oracle's generator is the
# GPU's, bit for bit): (opB, M, N, K, seeds of A and B, beta with C's seed).
CPU_PRODUCT = {
    "summa": ("N", 16384, 16384, 16384, (1, 2), (0.5, 3)),
    "dgemm": ("N", 16384, 16384, 16384, (1, 2), (0.0, None)),
    "cannon": (None, 16384, 8192, 32768, (1, 2), (0.0, None)),
    "cannon2x2": ("N", 2048, 2048, 2048, (1, 2), (0.0, None)),
    "grouped": ("N", 20000, 20000, 20000, (21, 22), (0.0, None)),
    "sgemm": ("N", 16384, 16384, 16384, (1, 2), (0.0, None)),
}


def cpu_baseline(workload, opB="N", runs=3):
    """Host BLAS dgemm (the oracle's numpy BLAS; sgemm for the FP32
    workload) on the workload's own global product and data: the operands
    come from the oracle's C restatement of the generator (the GPU's values
    bit for bit), one warm-up product of 2048^3, then the median of `runs`
    timed products.  For the grouped batch the sample is the 7x5 grid's
    product (20000^3), without the 256 extra problems."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle

    ob, M, N, K, (sa, sb), (beta, sc) = CPU_PRODUCT[workload]
    ob = ob or opB
    single = workload == "sgemm"
    dt = np.float32 if single else np.float64
    info = {}
    try:
        from threadpoolctl import threadpool_info
        for i in threadpool_info():
            if i.get("user_api") == "blas":
                info = i
                break
    except Exception:
        pass
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    # column-major flat arrays; .reshape(cols, rows).T is the matrix (a view)
    a = oracle.generate(sa, 0, M * K).astype(dt).reshape(K, M).T
    b = oracle.generate(sb, 0, K * N).astype(dt)
    # op(B) as K x N: the stored K x N matrix (N), or the transpose of the stored N x K one (T)
    b = b.reshape(N, K).T if ob == "N" else b.reshape(K, N)
    c = oracle.generate(sc, 0, M * N).astype(dt).reshape(N, M).T if sc is not None else None
    w = np.ones((2048, 2048), dtype=dt)
    w @ w  # warm-up: thread pool and code paths
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        out = a @ b
        if c is not None:
            out += dt(beta) * c
        times.append(time.perf_counter() - t0)
        del out
    t = sorted(times)[len(times) // 2]
    lib = "%s %s" % (info.get("internal_api", "numpy BLAS"), info.get("version", "")) if info else "numpy BLAS"
    return {"value": round(2.0 * M * N * K / t / 1e12, 4), "unit": "TFLOP/s",
            "cores": info.get("num_threads") or os.cpu_count(), "kind": "port",
            "nproc": os.cpu_count(), "cpu_model": model, "library": lib.strip(),
            "sample": "host %s %s%s %dx%dx%d on the workload's generated data (seeds %d, %d%s), median of %d "
                      "(%s s)" % ("sgemm" if single else "dgemm", "N", ob, M, N, K, sa, sb,
                                   ", C seed %d, beta %g" % (sc, beta) if sc is not None else "", runs,
                                   ", ".join("%.2f" % x for x in times))}


def hbm_leg(dev, reps=5):
    """The HBM-bound kernels at the configurations' shapes, each timed with
    HIP events on the launch stream over `reps` calls after one warm-up:
    bytes moved per call (16 B per element copied, 8 (c+1) B per element of a
    (c-1)-partial layer reduction, 8 B per generated element), GB/s and the
    fraction of the 8 TB/s HBM peak."""
    import torch
    import dgamd
    f64 = dict(dtype=torch.float64, device=dev)
    out = {}

    def timed(name, bytes_moved, fn, shape):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1) / 1e3 / reps
        gbs = bytes_moved / t / 1e9
        out[name] = {"shape": shape, "ms": round(t * 1e3, 4), "bytes": int(bytes_moved), "GB_s": round(gbs, 1),
                     "frac": round(gbs / HBM_PEAK, 4)}

    n, nb, P = 16384, 512, 4
    G = torch.empty(n * n, **f64)
    Pk = torch.empty(n * n, **f64)
    timed("generate", 8 * n * n, lambda: dgamd.generate_t(1, dgamd.DG_DIST_UNIFORM_PM1, G), "16384^2 FP64")
    timed("pack_block_cyclic", 16 * n * n, lambda: dgamd.pack_block_cyclic_t(G, n, n, n, nb, nb, P, P, Pk),
          "16384^2, 512x512 blocks on 4x4")
    timed("unpack_block_cyclic", 16 * n * n, lambda: dgamd.unpack_block_cyclic_t(Pk, nb, nb, P, P, G, n, n, n),
          "16384^2, 512x512 blocks on 4x4")
    timed("copy_block", 16 * (n - 8) * (n - 8),
          lambda: dgamd.copy_block_t(n - 8, n - 8, G, n, Pk, n), "16376^2 of ld 16384")
    timed("torch_copy (calibration)", 16 * n * n, lambda: Pk.copy_(G), "16384^2, torch's device copy")
    del G, Pk
    M, N, K, p = 16384, 8192, 32768, 4
    G = torch.empty(M * K, **f64)
    Pk = torch.empty(M * K, **f64)
    dgamd.generate_t(1, dgamd.DG_DIST_UNIFORM_PM1, G)
    ro, ko = dgamd.block_splits(M, p), dgamd.block_splits(K, p)
    timed("pack_block2d", 16 * M * K, lambda: dgamd.pack_block2d_t(G, M, M, K, ro, ko, Pk), "16384x32768 on 4x4")
    timed("unpack_block2d", 16 * M * K, lambda: dgamd.unpack_block2d_t(Pk, ro, ko, G, M, M, K),
          "16384x32768 on 4x4")
    del G, Pk
    C = torch.empty(M * N, **f64)
    W = torch.empty(M * N, **f64)
    dgamd.generate_t(3, dgamd.DG_DIST_UNIFORM_PM1, C)
    dgamd.generate_t(4, dgamd.DG_DIST_UNIFORM_PM1, W)
    timed("layer_reduce", 8 * 3 * M * N, lambda: dgamd.layer_reduce_t(C, W, M * N, 1),
          "16384x8192, c=2 (one partial layer)")
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU transport "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    import dgamd

    dev = torch.device("cuda", 0)
    dgamd.lib()
    if dgamd.lib().dg_device_ok() != 1:
        raise SystemExit("no gfx950 device: " + dgamd.last_error())
    step, (M, N, K), cfg, tag = setup(args.workload, dev, args.opB)
    flops = 2.0 * M * N * K
    fp32 = args.workload == "sgemm"
    peak = FP32_PEAK if fp32 else FP64_PEAK

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s, e in ev:
        s.record(stream)
        step()
        e.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = [s.elapsed_time(e) / 1e3 for s, e in ev]  # seconds per step (events on the launch stream)
    avg = sum(per) / len(per)
    value = flops * args.steps / wall / 1e12
    achieved = flops / avg / 1e12

    out = {
        "metric": "FP64-equiv TFLOP/s @16384\u00b3 (native vs Ozaki-INT8), % MFMA peak, rel err vs FP64",
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if fp32 else "f64",
        "data": DATA[args.workload],
        "config": cfg,
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4), "traffic": load_traffic(tag)},
    }
    if args.workload == "grouped":
        out["ozaki_form"] = grouped_ozaki_leg(step)
    del step  # the workload's buffers
    torch.cuda.empty_cache()
    int8_s = [int(x) for x in args.ozaki_slices.split(",") if x.strip()]
    bf16_s = [int(x) for x in args.ozaki_bf16_slices.split(",") if x.strip()]
    crt_m = [int(x) for x in args.ozaki_crt_moduli.split(",") if x.strip()]
    if int8_s or bf16_s or crt_m:
        out["ozaki"] = ozaki_leg(dev, int8_s, bf16_s, crt_m)
        torch.cuda.empty_cache()
    if not args.no_hbm:
        out["hbm"] = hbm_leg(dev)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.workload, args.opB)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
