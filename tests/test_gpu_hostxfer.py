"""Host <-> device column-block copies (dg_copy2d_h2d / dg_copy2d_d2h): the
bytes that land on each side against a numpy view of the same columns, for
pageable (staged through the pinned ring) and pinned host memory, strided
pitches on either side, columns wider than one staging buffer, many pieces
with a partial last one, and the gaps between columns left untouched."""
import numpy as np
import pytest

import dgamd

pytestmark = pytest.mark.gpu

SLOT = 32 << 20  # bytes per staging buffer of the library


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def cols(buf, pitch, width, height):
    """Column j of a byte buffer as row j of a (height, width) view."""
    return np.lib.stride_tricks.as_strided(buf, shape=(height, width), strides=(pitch, 1))


CASES = [
    # (width, height, host pitch, device pitch)
    (100, 1, 100, 100),                      # one small column, the calling thread alone
    (8192, 4096, 8192, 8192),                # exactly one staging buffer
    (12000, 10000, 12008, 12288),            # strided on both sides, a partial last piece
    (131072, 700, 131072, 131072),           # 16384-row FP64 columns, contiguous
    (SLOT + 24, 2, SLOT + 32, SLOT + 1024),  # columns wider than a staging buffer
    (3, 200000, 5, 7),                       # many tiny odd columns
]


@pytest.mark.parametrize("pinned", [False, True])
@pytest.mark.parametrize("width,height,hpitch,dpitch", CASES)
def test_round_trip(torch_dev, width, height, hpitch, dpitch, pinned):
    torch = torch_dev
    print("copy2d case", width, height, hpitch, dpitch, "pinned" if pinned else "pageable", flush=True)
    rng = np.random.default_rng(width * 7 + height)
    hbytes = hpitch * (height - 1) + width
    dbytes = dpitch * (height - 1) + width
    src = rng.integers(0, 256, hbytes, dtype=np.uint8)
    if pinned:
        host_t = torch.from_numpy(src.copy()).pin_memory()
        host = host_t
    else:
        host = src
    dev = torch.full((dbytes,), 0xAB, dtype=torch.uint8, device="cuda")
    dgamd.copy2d_h2d(dev, dpitch, host, hpitch, width, height)
    got = dev.cpu().numpy()
    want = np.full(dbytes, 0xAB, dtype=np.uint8)
    cols(want, dpitch, width, height)[:] = cols(src, hpitch, width, height)
    assert np.array_equal(got, want)

    # back into a buffer of another pitch; its gaps keep their fill
    opitch = hpitch + 3 if height > 1 else hpitch
    obytes = opitch * (height - 1) + width
    if pinned:
        out_t = torch.full((obytes,), 0xCD, dtype=torch.uint8).pin_memory()
        out = out_t
    else:
        out = np.full(obytes, 0xCD, dtype=np.uint8)
    dgamd.copy2d_d2h(out, opitch, dev, dpitch, width, height)
    torch.cuda.synchronize()
    out_np = out.numpy() if pinned else out
    want_out = np.full(obytes, 0xCD, dtype=np.uint8)
    cols(want_out, opitch, width, height)[:] = cols(src, hpitch, width, height)
    assert np.array_equal(out_np, want_out)


def test_follows_stream_order_and_repeats(torch_dev):
    """A device-to-host copy waits for the kernel that writes its source on
    the same stream; back-to-back copies reuse the staging ring."""
    torch = torch_dev
    n = 16384 * 1024
    for seed in (1, 2, 3):
        x = torch.empty(n, dtype=torch.float64, device="cuda")
        dgamd.generate_t(seed, dgamd.DG_DIST_UNIFORM_PM1, x)   # enqueued, not waited for
        out = np.empty(n, dtype=np.float64)
        dgamd.copy2d_d2h(out, 8 * 16384, x, 8 * 16384, 8 * 16384, 1024)
        assert np.array_equal(out, x.cpu().numpy())
        y = torch.zeros(n, dtype=torch.float64, device="cuda")
        dgamd.copy2d_h2d(y, 8 * 16384, out, 8 * 16384, 8 * 16384, 1024)
        assert torch.equal(x, y)


def test_memory_kinds_are_checked(torch_dev):
    torch = torch_dev
    dev = torch.zeros(64, dtype=torch.uint8, device="cuda")
    host = np.zeros(64, dtype=np.uint8)
    L = dgamd.lib()
    import ctypes
    st = L.dg_copy2d_h2d(ctypes.c_void_p(host.ctypes.data), 8, ctypes.c_void_p(dev.data_ptr()), 8, 8, 8, None)
    assert st == dgamd.DG_ERR_INVALID and "device" in dgamd.last_error()
    st = L.dg_copy2d_d2h(ctypes.c_void_p(dev.data_ptr()), 8, ctypes.c_void_p(dev.data_ptr()), 8, 8, 8, None)
    assert st == dgamd.DG_ERR_INVALID and "device memory" in dgamd.last_error()
