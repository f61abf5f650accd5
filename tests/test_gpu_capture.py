"""Single-call entry points captured into a caller's HIP graph (through
torch.cuda.graph): after one eager call has made the descriptor tables
resident, dg_dgemm, dg_summa and dg_ozaki_dgemm (scheme II) record only
kernel launches, and every replay reproduces the eager result bit for bit.
The graph is replayed several times with eager work (NaN fills) between the
replays, and each replay is checked."""
import pytest
import torch

import dgamd

pytestmark = pytest.mark.gpu


def test_dgemm_summa_ozaki_replays_match_eager(dev):
    f64 = dict(dtype=torch.float64, device=dev)
    m, n, k = 384, 256, 512
    A = torch.empty(m * k, **f64)
    B = torch.empty(k * n, **f64)
    dgamd.generate_t(71, dgamd.DG_DIST_UNIFORM_PM1, A)
    dgamd.generate_t(72, dgamd.DG_DIST_WIDE, B)
    C0 = torch.empty(m * n, **f64)
    dgamd.generate_t(73, dgamd.DG_DIST_UNIFORM_PM1, C0)

    # SUMMA operands on a 2x2 block-cyclic grid
    M = N = K = 512
    nb, P = 64, 2
    G = [torch.empty(M * M, **f64) for _ in range(3)]
    for s, t in zip((74, 75, 76), G):
        dgamd.generate_t(s, dgamd.DG_DIST_UNIFORM_PM1, t)
    Pk = [torch.empty_like(t) for t in G]
    for t, p in zip(G, Pk):
        dgamd.pack_block_cyclic_t(t, M, M, M, nb, nb, P, P, p)

    nmod = 14
    ws = torch.empty(max(1, dgamd.ozaki_workspace(m, n, k, nmod, dgamd.DG_OZAKI_INT8_CRT)), dtype=torch.uint8,
                     device=dev)
    C1, C2, C3 = torch.empty_like(C0), torch.empty_like(C0), torch.empty_like(Pk[2])

    def work():
        C1.copy_(C0)
        C2.copy_(C0)
        C3.copy_(Pk[2])
        dgamd.dgemm_t("N", "N", m, n, k, 1.5, A, m, B, k, 0.5, C1, m)
        dgamd.ozaki_dgemm_t("N", "N", m, n, k, 1.5, A, m, B, k, 0.5, C2, m, nmod, dgamd.DG_OZAKI_INT8_CRT, ws)
        dgamd.summa_t("N", "N", M, N, K, -1.0, Pk[0], Pk[1], 0.25, C3, nb, nb, nb, P, P)

    # eager call: uploads the descriptor tables, gives the reference outputs
    work()
    torch.cuda.synchronize()
    eager = [C1.clone(), C2.clone(), C3.clone()]
    assert all(torch.isfinite(e).all() for e in eager)

    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            work()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    for rep in range(3):
        for t in (C1, C2, C3):
            t.fill_(float("nan"))   # eager work between the replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for got, ref in zip((C1, C2, C3), eager):
            assert torch.equal(got, ref), rep
    del g
