"""The descriptor-table cache under a caller's HIP graph: tables that a
captured call reads stay resident however many new tables follow, and new
tables reuse only buffers whose last use has completed.

A graph holding dg_dgemm, dg_summa and a block-cyclic pack is captured,
then 300 calls with distinct tables (every alpha a different power of two)
push the cache far past its 256 resident entries, with synchronisations in
between so that retired buffers are reused.  Every one of those calls is
checked bit for bit (scaling by a power of two is exact), and the graph is
replayed several times afterwards, each replay bit-exact against the eager
result."""
import pytest
import torch

import dgamd

pytestmark = pytest.mark.gpu


def test_captured_tables_survive_cache_turnover(dev):
    f64 = dict(dtype=torch.float64, device=dev)
    m, n, k = 256, 192, 320
    A = torch.empty(m * k, **f64)
    B = torch.empty(k * n, **f64)
    dgamd.generate_t(81, dgamd.DG_DIST_UNIFORM_PM1, A)
    dgamd.generate_t(82, dgamd.DG_DIST_UNIFORM_PM1, B)
    C0 = torch.empty(m * n, **f64)
    dgamd.generate_t(83, dgamd.DG_DIST_UNIFORM_PM1, C0)
    M = 384
    nb, P = 64, 2
    G = [torch.empty(M * M, **f64) for _ in range(3)]
    for s, t in zip((84, 85, 86), G):
        dgamd.generate_t(s, dgamd.DG_DIST_UNIFORM_PM1, t)
    Pk = [torch.empty_like(t) for t in G]
    for t, p in zip(G, Pk):
        dgamd.pack_block_cyclic_t(t, M, M, M, nb, nb, P, P, p)
    C1, C2, Pg = torch.empty_like(C0), torch.empty_like(Pk[2]), torch.empty_like(Pk[2])

    def work():
        C1.copy_(C0)
        C2.copy_(Pk[2])
        dgamd.dgemm_t("N", "T", m, n, k, 1.25, A, m, B, n, -0.5, C1, m)
        dgamd.summa_t("N", "N", M, M, M, 0.75, Pk[0], Pk[1], 0.5, C2, nb, nb, nb, P, P)
        dgamd.pack_block_cyclic_t(G[0], M, M, M, nb, nb, P, P, Pg)

    work()
    torch.cuda.synchronize()
    eager = [C1.clone(), C2.clone(), Pg.clone()]

    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            work()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()

    # 300 distinct tables: C = 2^e * A B exactly, checked against e = 0
    mm, nn, kk = 96, 80, 128
    a = torch.empty(mm * kk, **f64)
    b = torch.empty(kk * nn, **f64)
    dgamd.generate_t(87, dgamd.DG_DIST_UNIFORM_PM1, a)
    dgamd.generate_t(88, dgamd.DG_DIST_UNIFORM_PM1, b)
    ref = torch.empty(mm * nn, **f64)
    dgamd.dgemm_t("N", "N", mm, nn, kk, 1.0, a, mm, b, kk, 0.0, ref, mm)
    outs = [torch.empty(mm * nn, **f64) for _ in range(300)]
    for i, c in enumerate(outs):
        dgamd.dgemm_t("N", "N", mm, nn, kk, 2.0 ** (i - 150), a, mm, b, kk, 0.0, c, mm)
        if i % 50 == 49:
            torch.cuda.synchronize()   # earlier tables complete: their buffers become reusable
    torch.cuda.synchronize()
    for i, c in enumerate(outs):
        assert torch.equal(c, ref * 2.0 ** (i - 150)), i

    for rep in range(3):
        for t in (C1, C2, Pg):
            t.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for got, want in zip((C1, C2, Pg), eager):
            assert torch.equal(got, want), rep
    # the eager path still works after the turnover
    work()
    torch.cuda.synchronize()
    for got, want in zip((C1, C2, Pg), eager):
        assert torch.equal(got, want)
    del g
