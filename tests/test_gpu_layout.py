"""Pack/unpack, strided copy, layer reduction and generation on the GPU,
bit-exact against the CPU oracle (data movement and fixed-order sums)."""
import numpy as np
import pytest
import torch

import dgamd
import oracle

pytestmark = pytest.mark.gpu


def dev_copy(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


@pytest.mark.parametrize("dist", [dgamd.DG_DIST_UNIFORM_PM1, dgamd.DG_DIST_WIDE, dgamd.DG_DIST_UNIFORM_01])
def test_generate_bit_exact(dev, dist):
    n = 100003
    t = torch.empty(n, dtype=torch.float64, device=dev)
    dgamd.generate_t(12345, dist, t)
    got = t.cpu().numpy()
    ref = oracle.generate(12345, dist, n)
    assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("n,off", [(1, 0), (2, 1), (3, 3), (5, 1), (100003, 1), (100003, 2), (100003, 3), (4097, 0)])
def test_generate_offsets_and_tails(dev, dtype, n, off):
    """Generation into buffers that start 4, 8 or 12 bytes past a 16-byte
    boundary and whose lengths leave a tail: the single-element head and
    tail and the 16-byte body agree with the oracle bit for bit, and the
    elements around the buffer stay untouched."""
    dt = torch.float64 if dtype == "f64" else torch.float32
    buf = torch.full((n + off + 4,), 7.0, dtype=dt, device=dev)
    dgamd.generate_t(4321, dgamd.DG_DIST_WIDE, buf[off:off + n])
    got = buf.cpu().numpy()
    ref = oracle.generate(4321, dgamd.DG_DIST_WIDE, n) if dtype == "f64" else oracle.generate_f32(4321, dgamd.DG_DIST_WIDE, n)
    u = np.uint64 if dtype == "f64" else np.uint32
    assert np.array_equal(got[off:off + n].view(u), ref.view(u))
    assert np.all(got[:off] == 7.0) and np.all(got[off + n:] == 7.0)


@pytest.mark.parametrize("rows,cols,ldg,pr,pc", [(2048, 2048, 2048, 2, 2), (23, 17, 29, 7, 5), (20000, 40, 20003, 7, 5),
                                                 (5, 3, 5, 1, 1), (9, 9, 9, 3, 3)])
def test_block2d_pack_unpack(dev, rows, cols, ldg, pr, pc):
    rng = np.random.default_rng(rows * 31 + cols)
    G = rng.standard_normal(ldg * cols)
    ro, co = oracle.block_splits(rows, pr), oracle.block_splits(cols, pc)
    ref = oracle.pack_block2d(G, ldg, rows, cols, ro, co)
    Gd = dev_copy(G, dev)
    P = torch.full((rows * cols,), float("nan"), dtype=torch.float64, device=dev)
    dgamd.pack_block2d_t(Gd, ldg, rows, cols, ro, co, P)
    assert np.array_equal(P.cpu().numpy(), ref)
    G2 = torch.full((ldg * cols,), 3.5, dtype=torch.float64, device=dev)
    dgamd.unpack_block2d_t(P, ro, co, G2, ldg, rows, cols)
    g2 = G2.cpu().numpy().reshape(cols, ldg)
    assert np.array_equal(g2[:, :rows], G.reshape(cols, ldg)[:, :rows])
    assert np.all(g2[:, rows:] == 3.5)  # rows beyond the matrix untouched


@pytest.mark.parametrize("col_major", [True, False])
@pytest.mark.parametrize("M,N,K,size", [(23, 17, 5, 6), (300, 260, 90, 35), (16, 16, 3, 4), (9, 30, 1, 7)])
def test_block_sequential_tiles(dev, col_major, M, N, K, size):
    """BlockSequentialDecomposer's layouts, as its host code packs them
    (row-major globals through their column-major transposes), hold every
    rank's deliverMatrix tiles bit for bit (reference branch per storage order)."""
    rng = np.random.default_rng(M + N + K + size + int(col_major))
    gA, gB, gC = (rng.standard_normal(n) for n in (M * K, K * N, M * N))
    tA, tB, tC, sizes = oracle.block_seq_tiles(col_major, M, N, K, size, gA, gB, gC)
    dRow, dCol = oracle.virtual_grid(size)
    ro, co, ko = oracle.block_splits(M, dRow), oracle.block_splits(N, dCol), [0, K]
    PA = torch.empty(M * K, dtype=torch.float64, device=dev)
    PB = torch.empty(K * N, dtype=torch.float64, device=dev)
    PC = torch.empty(M * N, dtype=torch.float64, device=dev)
    if col_major:
        dgamd.pack_block2d_t(dev_copy(gA, dev), M, M, K, ro, ko, PA)
        dgamd.pack_block2d_t(dev_copy(gB, dev), K, K, N, ko, co, PB)
        dgamd.pack_block2d_t(dev_copy(gC, dev), M, M, N, ro, co, PC)
    else:
        dgamd.pack_block2d_t(dev_copy(gA, dev), K, K, M, ko, ro, PA)
        dgamd.pack_block2d_t(dev_copy(gB, dev), N, N, K, co, ko, PB)
        dgamd.pack_block2d_t(dev_copy(gC, dev), N, N, M, co, ro, PC)
    pa, pb, pc = PA.cpu().numpy(), PB.cpu().numpy(), PC.cpu().numpy()
    oa = ob = oc = 0
    for r, (tm, tn) in enumerate(sizes):
        i, j = divmod(r, dCol)
        a = pa[ro[i] * K: ro[i] * K + tm * K]
        b = pb[co[j] * K: co[j] * K + K * tn]
        cb = (ro[i] * N + tm * co[j]) if col_major else (co[j] * M + tn * ro[i])
        assert np.array_equal(a, tA[oa: oa + tm * K]), r
        assert np.array_equal(b, tB[ob: ob + K * tn]), r
        assert np.array_equal(pc[cb: cb + tm * tn], tC[oc: oc + tm * tn]), r
        oa, ob, oc = oa + tm * K, ob + K * tn, oc + tm * tn


@pytest.mark.parametrize("rows,cols,mb,nb,P,Q", [(23, 17, 3, 2, 7, 5), (64, 64, 8, 8, 4, 4), (1000, 777, 64, 100, 3, 2),
                                                 (5, 5, 8, 8, 2, 2), (512, 520, 512, 512, 4, 4),
                                                 (2100, 5, 1500, 2, 2, 2), (2, 1100000, 1, 3, 2, 3)])
def test_block_cyclic_pack_unpack(dev, rows, cols, mb, nb, P, Q):
    rng = np.random.default_rng(rows + 7 * cols)
    ldg = rows + 3
    G = rng.standard_normal(ldg * cols)
    ref = oracle.pack_block_cyclic(G, ldg, rows, cols, mb, nb, P, Q)
    assert dgamd.block_cyclic_size(rows, cols, mb, nb, P, Q) == ref.size
    Gd = dev_copy(G, dev)
    Pk = torch.full((ref.size,), float("nan"), dtype=torch.float64, device=dev)
    dgamd.pack_block_cyclic_t(Gd, ldg, rows, cols, mb, nb, P, Q, Pk)
    assert np.array_equal(Pk.cpu().numpy(), ref)
    G2 = torch.full((ldg * cols,), -1.0, dtype=torch.float64, device=dev)
    dgamd.unpack_block_cyclic_t(Pk, mb, nb, P, Q, G2, ldg, rows, cols)
    g2 = G2.cpu().numpy().reshape(cols, ldg)
    assert np.array_equal(g2[:, :rows], G.reshape(cols, ldg)[:, :rows])
    assert np.all(g2[:, rows:] == -1.0)


def test_block_cyclic_full_size_round_trip(dev):
    # the 16384^2, 512x512-block, 4x4-grid layout at full size: unpack(pack(X)) == X
    n = 16384
    X = torch.empty(n * n, dtype=torch.float64, device=dev)
    dgamd.generate_t(99, dgamd.DG_DIST_UNIFORM_PM1, X)
    Pk = torch.empty_like(X)
    dgamd.pack_block_cyclic_t(X, n, n, n, 512, 512, 4, 4, Pk)
    Y = torch.empty_like(X)
    dgamd.unpack_block_cyclic_t(Pk, 512, 512, 4, 4, Y, n, n, n)
    assert torch.equal(X, Y)
    # spot-check a few packed elements against the definition
    Pk_h = Pk.view(-1)
    for (I, J) in [(0, 0), (511, 512), (513, 1), (16383, 16383), (4097, 9000)]:
        pr, pc = (I // 512) % 4, (J // 512) % 4
        li = (I // 512 // 4) * 512 + I % 512
        lj = (J // 512 // 4) * 512 + J % 512
        base = (pr * 4 + pc) * 4096 * 4096
        assert Pk_h[base + li + lj * 4096].item() == X[I + J * n].item()


def test_copy_block(dev):
    rng = np.random.default_rng(3)
    src = rng.standard_normal(101 * 50)
    dst0 = np.full(77 * 50, 9.0)
    ref = dst0.copy()
    oracle.lib().or_copy_block(60, 50, oracle._p(src), 101, oracle._p(ref), 77)
    S, D = dev_copy(src, dev), dev_copy(dst0, dev)
    dgamd.copy_block_t(60, 50, S, 101, D, 77)
    assert np.array_equal(D.cpu().numpy(), ref)


@pytest.mark.parametrize("rows,cols,lds,ldd,soff,doff", [(1, 1100000, 1, 2, 0, 0), (3, 400000, 5, 3, 1, 0),
                                                         (1030, 70, 1031, 1030, 1, 1), (515, 9, 515, 517, 0, 1),
                                                         (2048, 5, 2049, 2048, 1, 1), (17, 3, 17, 17, 1, 0)])
def test_copy_block_offsets_and_wide(dev, rows, cols, lds, ldd, soff, doff):
    """Strided FP64 copies whose buffers start 8 bytes past a 16-byte
    boundary (same or different offsets on the two sides) and whose column
    count exceeds one launch's column groups, against the oracle."""
    rng = np.random.default_rng(rows * 7 + cols)
    src = rng.standard_normal(lds * cols + soff)
    dst0 = np.full(ldd * cols + doff, 9.0)
    ref = dst0.copy()
    oracle.lib().or_copy_block(rows, cols, oracle._p(src[soff:]), lds, oracle._p(ref[doff:]), ldd)
    S, D = dev_copy(src, dev), dev_copy(dst0, dev)
    dgamd.copy_block_t(rows, cols, S[soff:], lds, D[doff:], ldd)
    assert np.array_equal(D.cpu().numpy(), ref)


@pytest.mark.parametrize("rows,cols,ldg,pr,pc,off", [(2857, 6, 2858, 7, 2, 1), (1030, 4, 1031, 3, 2, 1),
                                                     (5, 1100000, 5, 2, 3, 0)])
def test_block2d_offsets_and_wide(dev, rows, cols, ldg, pr, pc, off):
    """2D-block pack and unpack from a global matrix that starts 8 bytes past
    a 16-byte boundary (bands of odd heights, so the two sides' offsets
    differ from column to column) and with more columns than one launch's
    column groups, against the oracle."""
    rng = np.random.default_rng(rows + cols + off)
    G = rng.standard_normal(ldg * cols + off)
    ro, co = oracle.block_splits(rows, pr), oracle.block_splits(cols, pc)
    ref = oracle.pack_block2d(G[off:], ldg, rows, cols, ro, co)
    Gd = dev_copy(G, dev)
    P = torch.full((rows * cols,), float("nan"), dtype=torch.float64, device=dev)
    dgamd.pack_block2d_t(Gd[off:], ldg, rows, cols, ro, co, P)
    assert np.array_equal(P.cpu().numpy(), ref)
    G2 = torch.full((ldg * cols + off,), 3.5, dtype=torch.float64, device=dev)
    dgamd.unpack_block2d_t(P, ro, co, G2[off:], ldg, rows, cols)
    g2 = G2.cpu().numpy()[off:].reshape(cols, ldg)
    assert np.array_equal(g2[:, :rows], G[off:].reshape(cols, ldg)[:, :rows])
    assert np.all(g2[:, rows:] == 3.5) and np.all(G2.cpu().numpy()[:off] == 3.5)


@pytest.mark.parametrize("rows,cols,lds,ldd", [(60, 50, 101, 77), (1, 1, 1, 1), (4099, 3, 4099, 4100), (257, 129, 300, 257)])
def test_copy_block_f32(dev, rows, cols, lds, ldd):
    """FP32 strided copy: every element of the block moved bit for bit (the
    oracle's FP64 copy restated in float32 by index), nothing else written."""
    rng = np.random.default_rng(rows + cols)
    src = rng.standard_normal(lds * cols).astype(np.float32)
    dst0 = np.full(ldd * cols, 9.0, dtype=np.float32)
    ref = dst0.reshape(cols, ldd).copy()
    ref[:, :rows] = src.reshape(cols, lds)[:, :rows]
    S, D = dev_copy(src, dev), dev_copy(dst0, dev)
    dgamd.copy_block_t(rows, cols, S, lds, D, ldd)
    assert np.array_equal(D.cpu().numpy().view(np.uint32), ref.reshape(-1).view(np.uint32))


@pytest.mark.parametrize("n,layers", [(1 << 20, 1), (1000001, 3), (7, 2)])
def test_layer_reduce_bit_exact(dev, n, layers):
    rng = np.random.default_rng(n)
    C = rng.standard_normal(n)
    part = rng.standard_normal(n * layers)
    ref = oracle.layer_reduce(C.copy(), part, n, layers)
    Cd, Pd = dev_copy(C, dev), dev_copy(part, dev)
    dgamd.layer_reduce_t(Cd, Pd, n, layers)
    assert np.array_equal(Cd.cpu().numpy(), ref)
