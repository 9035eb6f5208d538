"""C-ABI checks that need no GPU: the library loads, exports every function
declared in include/qsft_amd.h, the ctypes descriptor layouts match, and the
host-side planning/validation entry points behave."""
import ctypes
import os
import re

import pytest

from qsft_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "qsft_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qsft_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported_and_bound():
    names = header_functions()
    assert len(names) >= 20
    for nm in names:
        assert hasattr(L.lib, nm), nm
        assert nm in L.SIGNATURES, f"{nm} has no ctypes signature"


def test_struct_layouts_match():
    for which, cls in L.STRUCTS.items():
        assert L.lib.qsft_sizeof(which) == ctypes.sizeof(cls)
    assert L.lib.qsft_abi_version() == 2


@pytest.mark.parametrize("q,b,prec,passes", [(4, 6, L.QSFT_F32, 1), (4, 6, L.QSFT_F64, 2), (4, 7, L.QSFT_F32, 2),
                                             (20, 3, L.QSFT_F32, 1), (20, 4, L.QSFT_F32, 2), (3, 9, L.QSFT_F64, 2),
                                             (2, 16, L.QSFT_F32, 2), (7, 2, L.QSFT_F64, 1)])
def test_dft_plan_covers_every_axis_once(q, b, prec, passes):
    d = L.DftDesc(q=q, b=b, rows=5, in_stride=1, in_row_stride=q ** b)
    L.check(L.lib.qsft_dft_plan(ctypes.byref(d), prec))
    assert d.n_pass == passes
    axes = []
    elem = 8 if prec == L.QSFT_F32 else 16
    for p in range(d.n_pass):
        d0, d1, ti, ta = d.pass_d0[p], d.pass_d1[p], d.pass_ti[p], d.pass_ta[p]
        axes += list(range(d0, d1))
        inner = q ** (b - d1)
        assert inner % ti == 0 and (ta == 1 or ti == inner) and q ** d0 % ta == 0
        assert q ** (d1 - d0) * ti * ta * elem <= 65536 - 64 * 16
        assert d.pass_blocks[p] == 5 * (q ** d0 // ta) * (inner // ti)
        assert 0 < L.lib.qsft_dft_lds_bytes(ctypes.byref(d), p, prec) <= 65536
    assert sorted(axes) == list(range(b))


def test_bad_arguments_are_rejected_with_message():
    d = L.DftDesc(q=1, b=3, rows=1, in_stride=1, in_row_stride=1)
    assert L.lib.qsft_dft_plan(ctypes.byref(d), L.QSFT_F32) == L.QSFT_ERR_ARG
    assert b"bad descriptor" in L.lib.qsft_last_error()
    with pytest.raises(L.QsftLibraryError):
        L.check(L.lib.qsft_dft_plan(ctypes.byref(d), L.QSFT_F32), "plan")
    e = L.EvalDesc(q=4, n=50, b=6, C=3, P=102, S=1000, n_pad=60)     # n_pad not a multiple of 64
    assert L.lib.qsft_eval_blocks(ctypes.byref(e)) == -1
    e.n_pad = 64
    assert L.lib.qsft_eval_blocks(ctypes.byref(e)) == 3 * 102 * 64
    # launches with no descriptors are refused before touching the GPU
    assert L.lib.qsft_eval_cosets(None, None, 0, 0, 0, None) == L.QSFT_ERR_ARG


def test_peel_workspace_and_bind():
    nbytes = L.lib.qsft_peel_workspace_bytes(4, 50, 6, 3, 102, 15, 0)
    assert 5e6 < nbytes < 5e7
    buf = ctypes.create_string_buffer(nbytes)
    d = L.PeelDesc()
    base = ctypes.addressof(buf)
    rc = L.lib.qsft_peel_bind(ctypes.byref(d), 4, 50, 6, 3, 102, 51, L.QSFT_DETECT_NSO1, 15, 1.0, 0,
                              None, None, None, base, nbytes, 64)
    assert rc == L.QSFT_OK
    assert d.Q == 4096 and d.hcap_log2 == 15 and d.out_cap == 1 << 15
    ptrs = [d.kbin, d.rho_bin, d.slot_bin, d.tag, d.owner, d.last, d.head, d.count, d.key, d.rec_time,
            d.rec_rho, d.rec_next, d.ctr, d.out_k, d.out_val, d.out_order]
    assert all(base <= p < base + nbytes and (p - base) % 256 == 0 for p in ptrs)
    assert len(set(ptrs)) == len(ptrs)
    # noiseless detection needs P == n + 1; NSO needs p1 == n + 1
    assert L.lib.qsft_peel_bind(ctypes.byref(d), 4, 50, 6, 3, 102, 51, L.QSFT_DETECT_NOISELESS, 15, 1.0, 0,
                                None, None, None, base, nbytes, 64) == L.QSFT_ERR_ARG
    assert L.lib.qsft_peel_bind(ctypes.byref(d), 4, 50, 6, 3, 102, 51, L.QSFT_DETECT_NSO1, 15, 1.0, 0,
                                None, None, None, base, 1000, 64) == L.QSFT_ERR_ARG


def test_batch_runtime_sizes_on_the_host():
    """The batch runtime plans on the host: a batch's two arenas hold exactly
    the per-request footprints (the admission unit), and a bad request is
    refused with a message."""
    reqs = [L.Request(q=4, n=50, b=6, C=3, R=2, S=1000, method=L.QSFT_DETECT_NSO1, cutoff=1.3, sigma=0.9,
                      amax=10.0, seed=5),
            L.Request(q=20, n=16, b=3, C=3, R=2, S=2000, method=L.QSFT_DETECT_NSO1, cutoff=1.0, sigma=1.0,
                      amax=9.0, seed=6),
            L.Request(q=3, n=12, b=4, C=3, R=1, S=40, method=L.QSFT_DETECT_NOISELESS, cutoff=1e-9, sigma=0.0,
                      amax=3.0, seed=7)]
    fps = [L.lib.qsft_request_footprint(ctypes.byref(r), 15) for r in reqs]
    assert 40 << 20 < fps[0] < 48 << 20                          # K1: 24 B per sample + peel workspace
    arr = (L.Request * 3)(*reqs)
    sz = L.BatchSizes()
    h = L.lib.qsft_batch_create(ctypes.cast(arr, ctypes.c_void_p), 3, 15, 3, ctypes.byref(sz))
    assert h
    try:
        assert sz.input_bytes + sz.work_bytes == sum(fps)
        assert sz.n_req == 3 and sz.row_stride == 64 + 24 and sz.cap_rows == sum(4 * r.S + 1024 for r in reqs)
        assert 0 < sz.plan_bytes < 1 << 16
        offs = (ctypes.c_size_t * 4)()
        assert L.lib.qsft_batch_input_offset(h, 1, offs) == L.QSFT_OK
        assert list(offs) == sorted(offs) and all(o % 256 == 0 for o in offs)
    finally:
        L.lib.qsft_batch_free(h)
    bad = (L.Request * 1)(L.Request(q=4, n=10, b=4, C=3, R=2, S=10, method=L.QSFT_DETECT_NOISELESS, cutoff=1.0,
                                    sigma=0.0, amax=1.0, seed=0))              # noiseless detection needs R = 1
    assert not L.lib.qsft_batch_create(ctypes.cast(bad, ctypes.c_void_p), 1, 15, 3, ctypes.byref(sz))
    assert b"bad request shape" in L.lib.qsft_last_error()
