"""The direct-to-LDS FP64 GEMM's load addresses (gemm_f64.hip, operand_cursor
and issue_operand, edge instantiation) on the CPU: for every tile, wave, lane
and k-step of ragged shapes, each 16-byte load is 16-byte aligned and holds
at least one element of the operand, so the loads that also take a
neighbouring element (the row before a base 8 bytes past a 16-byte boundary,
the row after an odd row count, clamped rows past the end) never leave a
16-byte chunk of the caller's matrix and cannot cross a page.  The operand
shapes are the ones the routing in prepare_gemm sends to that kernel: even
leading dimensions; a k-contiguous operand 16-byte aligned with an even k
count; a row-contiguous one 0 or 8 bytes past a 16-byte boundary."""
import numpy as np
import pytest

DBM, DBK = 128, 8


def swz(r):
    return (r >> 1) & 3


def loads_kc(rows, ld, klen):
    """Element offsets (from the base) of every 16-byte load of a k-contiguous
    operand with `rows` lines of `klen` elements, ld apart."""
    out = []
    lane = np.arange(64)
    row = lane >> 2
    chunk = (lane & 3) ^ swz(row)
    k2 = 2 * chunk
    steps = (klen + DBK - 1) // DBK
    tail = klen % DBK
    for r0 in range(0, rows, DBM):
        for wave in range(4):
            ra = np.minimum(r0 + 32 * wave + row, rows - 1)
            rb = np.minimum(r0 + 32 * wave + row + 16, rows - 1)
            p = ra * ld + 2 * chunk
            second = (rb - ra) * ld
            for s in range(steps):
                q0, q1 = p + s * DBK, p + second + s * DBK
                live = np.ones(64, bool) if not (s == steps - 1 and tail) else k2 < tail
                out += [q0[live], q1[live]]
    return np.concatenate(out)


def loads_rc(rows, ld, klen, shift):
    """Element offsets of every 16-byte load of a row-contiguous operand
    (element (row, k) at row + k*ld), its tile grid starting `shift` rows
    before the base."""
    out = []
    lane = np.arange(64)
    steps = (klen + DBK - 1) // DBK
    tail = klen % DBK
    last = (rows + shift - 1) & ~1
    for r0 in range(0, rows + shift, DBM):
        r = np.minimum(r0 + 2 * lane, last)
        for wave in range(4):
            p = -shift + r + 2 * wave * ld
            for s in range(steps):
                q0, q1 = p + s * DBK * ld, p + ld + s * DBK * ld
                if s == steps - 1 and tail:
                    if 2 * wave < tail:
                        out.append(q0)
                    if 2 * wave + 1 < tail:
                        out.append(q1)
                else:
                    out += [q0, q1]
    return np.concatenate(out) if out else np.zeros(0, np.int64)


def valid_kc(e, rows, ld, klen):
    return (e >= 0) & (e // ld < rows) & (e % ld < klen)


def valid_rc(e, rows, ld, klen):
    return (e >= 0) & (e % ld < rows) & (e // ld < klen)


@pytest.mark.parametrize("rows", [1, 2, 3, 127, 128, 129, 200, 257])
@pytest.mark.parametrize("klen", [2, 6, 8, 10, 24, 34])
@pytest.mark.parametrize("pad", [0, 2, 6])
def test_kc_loads_inside_chunks(rows, klen, pad):
    ld = klen + pad + (klen + pad) % 2  # even, >= klen
    e = loads_kc(rows, ld, klen)
    assert np.all(e % 2 == 0)  # 16-byte aligned: base aligned, ld and 2*chunk even
    ok = valid_kc(e, rows, ld, klen) | valid_kc(e + 1, rows, ld, klen)
    assert ok.all(), e[~ok][:8]


@pytest.mark.parametrize("rows", [1, 2, 3, 127, 128, 129, 200, 257])
@pytest.mark.parametrize("klen", [1, 3, 8, 9, 17, 24])
@pytest.mark.parametrize("shift", [0, 1])
@pytest.mark.parametrize("pad", [0, 2, 5])
def test_rc_loads_inside_chunks(rows, klen, shift, pad):
    ld = rows + pad + (rows + pad) % 2  # even, >= rows
    e = loads_rc(rows, ld, klen, shift)
    # base 8*shift bytes past a 16-byte boundary: the load address 8*(shift + e) is 16-byte aligned
    assert np.all((shift + e) % 2 == 0)
    ok = valid_rc(e, rows, ld, klen) | valid_rc(e + 1, rows, ld, klen)
    assert ok.all(), e[~ok][:8]
    # the neighbour elements read: the row before the base (shift 1) and one past the last row
    outside = ~valid_rc(e, rows, ld, klen) | ~valid_rc(e + 1, rows, ld, klen)
    if shift == 1:
        assert outside.any()
