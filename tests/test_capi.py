"""The C ABI without a GPU: the library loads, exports every entry point
include/dgamd.h declares, its host-side helpers agree with the oracle, and
argument errors come back as status codes before any device work."""
import os
import re

import pytest

import dgamd
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for hdr in os.listdir(os.path.join(ROOT, "include")):
        if hdr.endswith(".h"):
            text = open(os.path.join(ROOT, "include", hdr)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(dg_[a-z0-9_]+)\s*\(", text))
    return names


def test_every_declared_symbol_is_exported():
    L = dgamd.lib()
    names = declared_symbols()
    assert "dg_dgemm" in names and "dg_cannon25d" in names
    for n in sorted(names):
        assert hasattr(L, n), n
    assert names <= set(dgamd.SIGNATURES), sorted(names - set(dgamd.SIGNATURES))


def test_version_and_no_device_here():
    assert dgamd.lib().dg_version() >= 100


def test_numroc_matches_oracle():
    for n, nb, P in [(16384, 512, 4), (23, 3, 7), (17, 2, 5), (0, 4, 3), (100, 7, 3)]:
        for p in range(P):
            for s in range(P):
                assert dgamd.numroc(n, nb, p, s, P) == oracle.numroc(n, nb, p, s, P)


def test_virtual_grid_matches_oracle():
    for size in range(1, 300):
        assert dgamd.virtual_grid(size) == oracle.virtual_grid(size)


def test_block_splits_and_cyclic_size():
    for n, parts in [(20000, 7), (2048, 2), (5, 5), (7, 3)]:
        assert dgamd.block_splits(n, parts) == oracle.block_splits(n, parts)
    for args in [(23, 17, 3, 2, 7, 5), (16384, 16384, 512, 512, 4, 4), (10, 7, 16, 16, 3, 2)]:
        assert dgamd.block_cyclic_size(*args) == oracle.block_cyclic_size(*args)


@pytest.mark.parametrize("p,c", [(2, 1), (4, 2), (4, 1), (8, 2), (8, 4), (16, 4), (6, 3), (2, 2)])
def test_cannon_m_index_matches_simulation(p, c):
    mA, _ = oracle.cannon_simulate(p, c)
    for k in range(c):
        for i in range(p):
            for j in range(p):
                for r in range(p // c):
                    assert dgamd.cannon_m_index(p, c, dgamd.DG_SCHED_REFERENCE, i, j, k, r) == mA[k, i, j, r]


@pytest.mark.parametrize("p,c", [(3, 1), (5, 1), (6, 2), (7, 1), (6, 1)])
def test_standard_schedule_covers_where_reference_does_not(p, c):
    assert dgamd.cannon_m_index(p, c, dgamd.DG_SCHED_REFERENCE, 0, 0, 0, 0) == -1
    for i in range(p):
        for j in range(p):
            ms = sorted(dgamd.cannon_m_index(p, c, dgamd.DG_SCHED_STANDARD, i, j, k, r)
                        for k in range(c) for r in range(p // c))
            assert ms == list(range(p))
            ms_auto = sorted(dgamd.cannon_m_index(p, c, dgamd.DG_SCHED_AUTO, i, j, k, r)
                             for k in range(c) for r in range(p // c))
            assert ms_auto == list(range(p))


def test_workspace_size():
    assert dgamd.cannon25d_workspace(16384, 8192, 4, 2) == 16384 * 8192
    assert dgamd.cannon25d_workspace(2048, 2048, 2, 1) == 0


def test_argument_errors_need_no_device():
    L = dgamd.lib()
    assert L.dg_dgemm(b"Q", b"N", 1, 1, 1, 1.0, None, 1, None, 1, 0.0, None, 1, None) == dgamd.DG_ERR_INVALID
    assert "opA" in dgamd.last_error()
    assert L.dg_dgemm(b"N", b"N", -1, 1, 1, 1.0, None, 1, None, 1, 0.0, None, 1, None) == dgamd.DG_ERR_INVALID
    assert L.dg_dgemm(b"N", b"N", 5, 5, 5, 1.0, None, 4, None, 5, 0.0, None, 5, None) == dgamd.DG_ERR_INVALID
    # empty problems are a no-op
    assert L.dg_dgemm(b"N", b"N", 0, 5, 5, 1.0, None, 1, None, 5, 0.0, None, 1, None) == dgamd.DG_OK
    st = L.dg_cannon25d(b"N", b"N", 9, 9, 9, 1.0, None, None, 0.0, None, 3, 1, dgamd.DG_SCHED_REFERENCE, None, 0, None)
    assert st == dgamd.DG_ERR_SCHEDULE
    st = L.dg_cannon25d(b"N", b"N", 8, 8, 8, 1.0, None, None, 0.0, None, 4, 3, dgamd.DG_SCHED_AUTO, None, 0, None)
    assert st == dgamd.DG_ERR_INVALID  # c must divide p
    st = L.dg_cannon25d(b"N", b"N", 8, 8, 8, 1.0, None, None, 0.0, None, 4, 2, dgamd.DG_SCHED_AUTO, None, 0, None)
    assert st == dgamd.DG_ERR_WORKSPACE
    st = L.dg_summa(b"N", b"T", 8, 8, 8, 1.0, None, None, 0.0, None, 2, 2, 2, 2, 2, None)
    assert st == dgamd.DG_ERR_INVALID   # transposed operands are supported; the operands are null
    st = L.dg_summa(b"N", b"Z", 8, 8, 8, 1.0, None, None, 0.0, None, 2, 2, 2, 2, 2, None)
    assert st == dgamd.DG_ERR_INVALID and "op" in dgamd.last_error()
    st = L.dg_generate(1, 7, None, 10, None)
    assert st == dgamd.DG_ERR_INVALID
    import ctypes
    h = ctypes.c_void_p()
    st = L.dg_grouped_plan_create_form(None, 0, dgamd.DG_PLAN_AUTOTUNE, 7, 0, ctypes.byref(h))
    assert st == dgamd.DG_ERR_INVALID and "form" in dgamd.last_error()
    st = L.dg_grouped_plan_create_form(None, 0, dgamd.DG_PLAN_AUTOTUNE, dgamd.DG_FORM_OZAKI_INT8, 1, ctypes.byref(h))
    assert st == dgamd.DG_ERR_INVALID and "moduli" in dgamd.last_error()
    mod = ctypes.c_int(5)
    assert L.dg_grouped_plan_problem_form(None, 0, ctypes.byref(mod)) == -1 and mod.value == 0
    # host <-> device copies: sizes, pitches and pointers are checked first
    buf = (ctypes.c_char * 64)()
    hb = ctypes.cast(buf, ctypes.c_void_p)
    for fn in (L.dg_copy2d_h2d, L.dg_copy2d_d2h):
        assert fn(hb, 8, hb, 8, -1, 2, None) == dgamd.DG_ERR_INVALID and "negative" in dgamd.last_error()
        assert fn(hb, 8, hb, 16, 16, 2, None) == dgamd.DG_ERR_INVALID and "pitch" in dgamd.last_error()
        assert fn(hb, 16, hb, 8, 16, 2, None) == dgamd.DG_ERR_INVALID and "pitch" in dgamd.last_error()
        assert fn(None, 8, hb, 8, 8, 2, None) == dgamd.DG_ERR_INVALID and "null" in dgamd.last_error()
        assert fn(hb, 1 << 62, hb, 1 << 62, 8, 3, None) == dgamd.DG_ERR_INVALID and "64 bits" in dgamd.last_error()
        assert fn(None, 0, None, 0, 0, 5, None) == dgamd.DG_OK   # empty copies are a no-op
        assert fn(None, 8, None, 8, 8, 0, None) == dgamd.DG_OK


def test_device_calls_fail_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    L = dgamd.lib()
    assert L.dg_device_ok() == 0
    import ctypes
    buf = (ctypes.c_double * 4)()
    st = L.dg_generate(1, 0, ctypes.cast(buf, ctypes.c_void_p), 4, None)
    assert st == dgamd.DG_ERR_NO_DEVICE
    hb = ctypes.cast(buf, ctypes.c_void_p)
    assert L.dg_copy2d_h2d(hb, 8, hb, 8, 8, 4, None) == dgamd.DG_ERR_NO_DEVICE
    assert L.dg_copy2d_d2h(hb, 8, hb, 8, 8, 4, None) == dgamd.DG_ERR_NO_DEVICE
