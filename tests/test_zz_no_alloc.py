"""The no-allocation peel entry point (qsft_peel_transform_noalloc) against
the allocating one (qsft_peel_bind + qsft_peel_run + qsft_peel_pack through
PeelBatch) on the same inputs: the same recovered rows, bit for bit, in the
same detection order, and the same U afterwards. Two calls in a row share the
library's static scratch; the second is a smaller shape, so it reuses memory
the first left behind. A shape whose workspace exceeds the scratch is
refused with a message."""
import numpy as np
import pytest
import torch

from oracle import qsft_oracle as O

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    from qsft_amd import _lib as L
    from qsft_amd import device as dv
    from qsft_amd import kernels as K


def _instance(q, n, S, b, C, R, noise_sd, seed, method="nso"):
    rng = np.random.default_rng(seed)
    src = np.vstack([np.zeros(n), np.eye(n)]).astype(int)
    if method == "nso":
        D = np.vstack([(rng.integers(0, q, n) - src) % q for _ in range(R)])
    else:
        D = src
    Ms = [rng.integers(0, q, size=(n, b)) for _ in range(C)]
    locq = rng.integers(0, q, size=(n, S))
    a = rng.uniform(1, 10, S) * np.exp(2j * np.pi * rng.uniform(size=S))
    Q = q ** b
    Us = np.array([O.closed_form_U(M, D, locq, a, q) for M in Ms])
    Us = Us + noise_sd / np.sqrt(2 * Q) * (rng.standard_normal(Us.shape) + 1j * rng.standard_normal(Us.shape))
    cutoff = 1e-9 + 1.5 * noise_sd ** 2 / Q
    return Ms, D, np.ascontiguousarray(Us.transpose(0, 2, 1)), cutoff, method


def _item(inst, q, n, b, C):
    Ms, D, Ubm, cutoff, method = inst
    dev = dv.current_device()
    n_pad = dv.n_pad_of(n)
    return dict(q=q, n=n, b=b, C=C, P=D.shape[0], p1=n + 1,
                method=L.QSFT_DETECT_NSO1 if method == "nso" else L.QSFT_DETECT_NOISELESS, cutoff=cutoff,
                Mt=dv.Mt_tensor(Ms, b, n_pad, dev), D=dv.D_tensor([D] * C, n_pad, dev),
                U=torch.from_numpy(Ubm.copy()).to(dev))


def _noalloc(item):
    d = L.PeelDesc()
    n_pad = item["Mt"].shape[-1]
    nbytes = L.lib.qsft_peel_workspace_bytes(item["q"], item["n"], item["b"], item["C"], item["P"], 15, 0)
    ws = torch.empty(nbytes, dtype=torch.uint8)                  # host bytes, only to size the output
    L.check(L.lib.qsft_peel_bind(d, item["q"], item["n"], item["b"], item["C"], item["P"], item["p1"],
                                 item["method"], 15, float(item["cutoff"]), 0, None, None, None,
                                 ws.data_ptr(), nbytes, n_pad))
    cap_rows, row_stride = 1 << d.hcap_log2, n_pad + 24
    hdr_bytes = L.lib.qsft_peel_pack_header_bytes(1)
    dst = torch.empty(hdr_bytes + cap_rows * row_stride, dtype=torch.uint8, device=item["U"].device)
    K.peel_noalloc(item, dst, cap_rows, row_stride)
    host = dst.cpu().numpy()
    hdr = host[:hdr_bytes].view(np.int64)
    assert hdr[3] == 0, f"device error bits {hdr[3]:#x}"
    rows = host[hdr_bytes:hdr_bytes + int(hdr[2]) * row_stride].reshape(-1, row_stride)
    return K.unpack_rows(hdr, rows, 1, [item["n"]], [n_pad])[0]


SHAPES = [
    # the headline peel shape first, then a smaller one reusing the scratch
    (4, 50, 1000, 6, 3, 2, 8.0, 1, "nso"),
    (3, 12, 40, 3, 3, 2, 0.3, 2, "nso"),
]


def test_noalloc_peel_matches_allocating_entry_point_twice():
    for q, n, S, b, C, R, sd, seed, method in SHAPES:
        print(f"noalloc peel q={q} n={n} S={S} b={b} C={C} R={R} sd={sd}", flush=True)
        inst = _instance(q, n, S, b, C, R, sd, seed, method)
        ref_item = _item(inst, q, n, b, C)
        batch = K.PeelBatch([ref_item])
        batch.run()
        k0, v0, o0, _ = batch.results()[0]
        item = _item(inst, q, n, b, C)
        k1, v1, o1 = _noalloc(item)
        torch.cuda.synchronize()
        assert len(k0) > 0.5 * S
        np.testing.assert_array_equal(k1, k0)
        np.testing.assert_array_equal(o1, o0)
        assert np.array_equal(v1.view(np.float64), v0.view(np.float64))          # same bits
        assert torch.equal(item["U"], ref_item["U"])                          # U left as the reference leaves it
        print(f"  {len(k1)} rows identical", flush=True)


def test_noalloc_refuses_a_shape_beyond_the_scratch():
    assert L.lib.qsft_noalloc_scratch_bytes() == 64 << 20
    # q = 4, n = 100, b = 7, P = 202: a 74 MiB workspace
    dev = dv.current_device()
    dummy = torch.zeros(1024, dtype=torch.uint8, device=dev)
    rc = L.lib.qsft_peel_transform_noalloc(4, 100, 7, 3, 202, 101, L.QSFT_DETECT_NSO1, 15, 1.0, 0,
                                           dummy.data_ptr(), dummy.data_ptr(), dummy.data_ptr(), 128,
                                           dummy.data_ptr(), 16, 152, None)
    assert rc == L.QSFT_ERR_ARG
    assert b"exceeds the static scratch" in L.lib.qsft_last_error()
