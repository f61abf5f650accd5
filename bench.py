"""Throughput of the per-rank GEMM hot path on one MI355X.

Default workload: FP64 SUMMA on a virtual 4x4 process grid, M=N=K=16384,
block-cyclic 512x512 blocks, A, B, C uniform[-1,1) from fixed seeds,
alpha=1, beta=0.5.  One step = one dg_summa call over the packed
block-cyclic operands already resident in HBM (every virtual rank's
local GEMM, all 32 panels, one launch).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload summa|cannon|dgemm|grouped|cannon2x2|sgemm]

Prints one JSON line.  `roofline` is the dominant kernel against the FP64
matrix peak (78.6 TFLOP/s; the FP32 matrix peak, 157.3, for sgemm), timed with HIP events on the launch stream;
`cpu_baseline` is the host BLAS dgemm of the oracle on a bounded sample of
the same work, on this host's cores.
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
    "cannon": "FP64 2.5D Cannon 4x4x2 virtual grid, M=16384 N=8192 K=32768, opA=N opB=N, layer reduce in HBM",
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
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of the CPU baseline sample")
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


def setup(workload, dev):
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
        for seed, (r, cc) in zip((1, 2, 3), ((M, K), (K, N), (M, N))):
            G = torch.empty(r * cc, **f64)
            dgamd.generate_t(seed, dgamd.DG_DIST_UNIFORM_PM1, G)
            Pk = torch.empty_like(G)
            dgamd.pack_block2d_t(G, r, r, cc, dgamd.block_splits(r, p), dgamd.block_splits(cc, p), Pk)
            del G
            bufs.append(Pk)
        Ap, Bp, Cp = bufs
        ws = torch.empty(dgamd.cannon25d_workspace(M, N, p, c), **f64)

        def step():
            dgamd.cannon25d_t("N", "N", M, N, K, 1.0, Ap, Bp, 0.0, Cp, p, c, dgamd.DG_SCHED_AUTO, ws)

        cfg = {"workload": WORKLOADS[workload], "M": M, "N": N, "K": K, "grid": "4x4x2", "alpha": 1.0, "beta": 0.0}
        return step, (M, N, K), cfg, "cannon16384x8192x32768"
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


def cpu_baseline(budget_s, single=False):
    """Host BLAS dgemm (the oracle's numpy/OpenBLAS; sgemm with `single`) on
    4096^3 products, repeated for about `budget_s` seconds;
    TFLOP/s = 2*4096^3*reps / time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle  # noqa: F401  (the checker's host dgemm is numpy's BLAS)

    threads = None
    try:
        from threadpoolctl import threadpool_info
        for info in threadpool_info():
            if info.get("user_api") == "blas":
                threads = info.get("num_threads")
                break
    except Exception:
        pass
    n = 4096
    rng = np.random.default_rng(1)
    dt = np.float32 if single else np.float64
    a = rng.uniform(-1, 1, (n, n)).astype(dt)
    b = rng.uniform(-1, 1, (n, n)).astype(dt)
    c = np.empty((n, n), dtype=dt)
    np.matmul(a, b, out=c)  # warm-up
    reps, t0 = 0, time.perf_counter()
    while True:
        np.matmul(a, b, out=c)
        reps += 1
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
    tf = 2.0 * n ** 3 * reps / el / 1e12
    return {"value": round(tf, 4), "unit": "TFLOP/s", "cores": threads or os.cpu_count(), "kind": "port",
            "sample": "%d host-BLAS (numpy OpenBLAS) %s calls of 4096x4096x4096 in %.1f s; "
                      "the 16384^3 step is 64 such products" % (reps, "sgemm" if single else "dgemm", el)}


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
    step, (M, N, K), cfg, tag = setup(args.workload, dev)
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
        "data": "synthetic: SplitMix64-seeded uniform[-1,1) (seeds 1,2,3), generated and packed on the GPU",
        "config": cfg,
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4), "traffic": load_traffic(tag)},
    }
    if args.workload == "grouped":
        out["ozaki_form"] = grouped_ozaki_leg(step)
    int8_s = [int(x) for x in args.ozaki_slices.split(",") if x.strip()]
    bf16_s = [int(x) for x in args.ozaki_bf16_slices.split(",") if x.strip()]
    crt_m = [int(x) for x in args.ozaki_crt_moduli.split(",") if x.strip()]
    if int8_s or bf16_s or crt_m:
        del step
        torch.cuda.empty_cache()
        out["ozaki"] = ozaki_leg(dev, int8_s, bf16_s, crt_m)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds, single=fp32)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
