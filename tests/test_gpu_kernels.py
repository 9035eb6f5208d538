"""Kernel-level parity on the GPU, through the C ABI, against the CPU oracle.

Tolerances: integer/index outputs bit-exact; FP64 samples and transforms
within 1e-12 relative (max-abs over max |value|); FP32 within 2e-6 relative
(FP32 accumulation of up to S terms plus the w-table rounding).
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import qsft_oracle as O

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    from qsft_amd import _lib as L
    from qsft_amd import device as dv
    from qsft_amd import kernels as K


def rel_err(got, ref):
    ref = np.asarray(ref)
    return float(np.max(np.abs(np.asarray(got) - ref)) / max(np.max(np.abs(ref)), 1e-300))


def _setup(q, n, b, S, C, R, seed, source="identity"):
    rng = np.random.default_rng(seed)
    Ms = [rng.integers(0, q, size=(n, b)) for _ in range(C)]
    if source == "identity":
        src = np.vstack([np.zeros(n), np.eye(n)]).astype(int)
        Dblocks = [(rng.integers(0, q, n) - src) % q for _ in range(R)]
    else:
        Dblocks = [rng.integers(0, q, size=(n + 1, n)) for _ in range(R)]
    D = np.vstack(Dblocks)
    locq = rng.integers(0, q, size=(n, S))
    a = rng.uniform(1, 10, S) * np.exp(2j * np.pi * rng.uniform(size=S))
    return Ms, D, locq, a


CASES = [(4, 50, 6, 1000, 3, 2), (4, 10, 4, 100, 3, 1), (20, 16, 3, 200, 3, 2), (3, 7, 4, 37, 2, 3),
         (2, 70, 5, 120, 3, 1), (4, 100, 3, 300, 2, 2), (5, 130, 2, 50, 2, 1), (20, 200, 2, 64, 1, 1)]


@pytest.mark.parametrize("q,n,b,S,C,R", CASES)
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_eval_cosets_matches_oracle(q, n, b, S, C, R, precision):
    print(f"eval_cosets q={q} n={n} b={b} S={S} C={C} R={R} {precision}", flush=True)
    Ms, D, locq, a = _setup(q, n, b, S, C, R, seed=q * 1000 + n)
    dev = dv.current_device()
    n_pad = dv.n_pad_of(n)
    Mt = dv.Mt_tensor(Ms, b, n_pad, dev)
    Dd = dv.D_tensor([D] * C, n_pad, dev)
    Kd = dv.digits_tensor(locq.T, n_pad, dev)
    ad = dv.strengths_tensor(a, precision, dev)
    x = dv.evaluate_cosets(q, n, b, Mt, Dd, Kd, ad, precision).cpu().numpy()
    # the oracle (like the reference) exponentiates the unreduced phase
    # sum_i m_i 2 pi k_i / q, up to 2 pi n (q-1)^2 / q rad, so its own error grows with it
    tol = max(1e-12, 4e-16 * 2 * np.pi * n * (q - 1) ** 2 / q) if precision == "fp64" else 2e-6
    # check a subset of rows against the oracle's per-query evaluation
    for c in range(C):
        for p in [0, D.shape[0] // 2, D.shape[0] - 1]:
            dig = O.coset_digits(Ms[c], D[p:p + 1], q, b)[0]
            ref = O.eval_synthetic(dig, locq, a, q)
            assert rel_err(x[c, p], ref) < tol, (c, p)


# q in {2, 4} with n <= 64 (the two-waves-per-SIMD kernel): P = 72 is nine
# 8-row tiles (two workgroups of 5 tiles, the last one partly empty), P = 82
# eleven (6 + 5); q = 2 exercises the one-bit W selector
MFMA_CASES = CASES + [(8, 30, 3, 77, 2, 2), (2, 250, 4, 33, 3, 1), (4, 64, 5, 129, 3, 2), (4, 35, 4, 150, 2, 2),
                      (2, 40, 5, 77, 2, 2)]


@pytest.mark.parametrize("q,n,b,S,C,R", MFMA_CASES)
def test_eval_mfma_matches_oracle(q, n, b, S, C, R):
    """Matrix-core evaluation (fp16x2 limbs, FP32 accumulation) against the
    oracle's per-query FP64 evaluation, every row."""
    print(f"eval_mfma q={q} n={n} b={b} S={S} C={C} R={R}", flush=True)
    Ms, D, locq, a = _setup(q, n, b, S, C, R, seed=q * 1000 + n + 7)
    dev = dv.current_device()
    n_pad = dv.n_pad_of(n)
    Mt = dv.Mt_tensor(Ms, b, n_pad, dev)
    Dd = dv.D_tensor([D] * C, n_pad, dev)
    Kd = dv.digits_tensor(locq.T, n_pad, dev)
    a64 = dv.strengths_tensor(a, "fp64", dev)
    x = dv.evaluate_cosets_mfma(q, n, b, Mt, Dd, Kd, a64).cpu().numpy()
    P = D.shape[0]
    rows = range(P) if P * q ** b * S <= 2e7 else sorted({0, 1, 7, 8, P // 2, P - 2, P - 1})
    for c in range(C):
        for p in rows:
            dig = O.coset_digits(Ms[c], D[p:p + 1], q, b)[0]
            ref = O.eval_synthetic(dig, locq, a, q)
            assert rel_err(x[c, p], ref) < 2e-6, (c, p)


@pytest.mark.parametrize("q,scale", [(4, 1e-7), (4, 3e6), (3, 1e-5), (4, 0.0)])
def test_eval_mfma_strength_scale(q, scale):
    """The fp16 limbs are scaled by 2^bexp from max|a|: tiny and huge
    strengths keep the FP32-level error; scale 0 also covers S = 0."""
    print(f"eval_mfma scale q={q} scale={scale}", flush=True)
    n, b, C, R = 12, 3, 2, 2
    S = 0 if scale == 0.0 else 70
    Ms, D, locq, a = _setup(q, n, b, max(S, 1), C, R, seed=77)
    locq, a = locq[:, :S], a[:S] * scale
    dev = dv.current_device()
    n_pad = dv.n_pad_of(n)
    x = dv.evaluate_cosets_mfma(q, n, b, dv.Mt_tensor(Ms, b, n_pad, dev), dv.D_tensor([D] * C, n_pad, dev),
                                dv.digits_tensor(locq.T, n_pad, dev), dv.strengths_tensor(a, "fp64", dev))
    x = x.cpu().numpy()
    if S == 0:
        assert not np.any(x)
        return
    for c in range(C):
        dig = O.coset_digits(Ms[c], D, q, b)
        ref = O.eval_synthetic(dig.reshape(-1, n), locq, a, q).reshape(D.shape[0], q ** b)
        assert rel_err(x[c], ref) < 2e-6, c


# support counts around the 32-support blocks of the sparse form (an odd
# number of 16-support blocks pads a zero half-block), one to four digit
# blocks, p-tile splits 5..8 and the generic-ptb kernels
SPARSE_CASES = [(4, 10, 3, 1, 2, 1), (4, 10, 3, 16, 2, 2), (4, 20, 4, 17, 3, 2), (4, 30, 4, 32, 2, 3),
                (4, 50, 5, 33, 3, 2), (2, 40, 6, 95, 2, 2), (4, 64, 4, 200, 2, 2), (4, 100, 4, 70, 2, 2),
                (2, 200, 5, 50, 2, 1), (4, 12, 4, 64, 3, 11)]


@pytest.mark.parametrize("q,n,b,S,C,R", SPARSE_CASES)
def test_eval_mfma_sparse_matches_dense_and_oracle(q, n, b, S, C, R):
    """q in {2, 4}: the sparse-MFMA form (compressed +-1 W, 2-bit row
    indices) against the dense form of the same fp16 products and against
    the oracle's FP64 evaluation, every row."""
    print(f"eval_mfma sparse q={q} n={n} b={b} S={S} C={C} R={R}", flush=True)
    Ms, D, locq, a = _setup(q, n, b, S, C, R, seed=q * 31 + n + S)
    dev = dv.current_device()
    n_pad = dv.n_pad_of(n)
    args = (q, n, b, dv.Mt_tensor(Ms, b, n_pad, dev), dv.D_tensor([D] * C, n_pad, dev),
            dv.digits_tensor(locq.T, n_pad, dev), dv.strengths_tensor(a, "fp64", dev))
    xs = dv.evaluate_cosets_mfma(*args).cpu().numpy()
    xd = dv.evaluate_cosets_mfma(*args, dense=True).cpu().numpy()
    assert rel_err(xs, xd) < 1e-6
    P = D.shape[0]
    for c in range(C):
        dig = O.coset_digits(Ms[c], D, q, b)
        ref = O.eval_synthetic(dig.reshape(-1, n), locq, a, q).reshape(P, q ** b)
        assert rel_err(xs[c], ref) < 2e-6, c
        assert rel_err(xd[c], ref) < 2e-6, c


@pytest.mark.parametrize("q", [4, 3])
def test_eval_mfma_distinct_delays_per_group(q):
    """Groups with their own delay rows (B computed per group)."""
    n, b, S, C, P = 20, 3, 90, 3, 21
    rng = np.random.default_rng(41)
    Ms = [rng.integers(0, q, size=(n, b)) for _ in range(C)]
    Ds = [rng.integers(0, q, size=(P, n)) for _ in range(C)]
    locq = rng.integers(0, q, size=(n, S))
    a = rng.uniform(1, 10, S) * np.exp(2j * np.pi * rng.uniform(size=S))
    dev = dv.current_device()
    n_pad = dv.n_pad_of(n)
    x = dv.evaluate_cosets_mfma(q, n, b, dv.Mt_tensor(Ms, b, n_pad, dev), dv.D_tensor(Ds, n_pad, dev),
                                dv.digits_tensor(locq.T, n_pad, dev), dv.strengths_tensor(a, "fp64", dev))
    x = x.cpu().numpy()
    for c in range(C):
        dig = O.coset_digits(Ms[c], Ds[c], q, b)
        ref = O.eval_synthetic(dig.reshape(-1, n), locq, a, q).reshape(P, q ** b)
        assert rel_err(x[c], ref) < 2e-6, c


@pytest.mark.parametrize("q,n,b,C,P", [(4, 50, 6, 3, 102), (20, 16, 3, 2, 17), (3, 7, 4, 2, 8), (5, 130, 2, 1, 3)])
def test_query_digits_bit_exact(q, n, b, C, P):
    print(f"query_digits q={q} n={n} b={b}", flush=True)
    rng = np.random.default_rng(7)
    Ms = [rng.integers(0, q, size=(n, b)) for _ in range(C)]
    D = rng.integers(0, q, size=(P, n))
    dev = dv.current_device()
    n_pad = dv.n_pad_of(n)
    m = dv.query_digits(q, n, b, dv.Mt_tensor(Ms, b, n_pad, dev), dv.D_tensor([D] * C, n_pad, dev)).cpu().numpy()
    for c in range(C):
        ref = O.coset_digits(Ms[c], D, q, b)
        np.testing.assert_array_equal(m[c, :, :, :n], ref)
        assert not m[c, :, :, n:].any()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_eval_points_matches_oracle(precision):
    q, n, S, N = 4, 50, 700, 3001
    rng = np.random.default_rng(8)
    locq = rng.integers(0, q, size=(n, S))
    a = rng.uniform(1, 10, S) * np.exp(2j * np.pi * rng.uniform(size=S))
    pts = rng.integers(0, q, size=(N, n))
    dev = dv.current_device()
    n_pad = dv.n_pad_of(n)
    x = dv.evaluate_points(q, n, dv.digits_tensor(pts, n_pad, dev), dv.digits_tensor(locq.T, n_pad, dev),
                           dv.strengths_tensor(a, precision, dev), precision).cpu().numpy()
    ref = O.eval_synthetic(pts, locq, a, q)
    assert rel_err(x, ref) < (1e-12 if precision == "fp64" else 2e-6)


DFT_CASES = [(4, 6, 7), (4, 7, 3), (20, 3, 5), (20, 4, 2), (3, 5, 4), (2, 13, 3), (7, 3, 2), (5, 5, 3), (4, 1, 9),
             (2, 16, 1)]


@pytest.mark.parametrize("q,b,rows", DFT_CASES)
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_dft_matches_fftn(q, b, rows, precision):
    print(f"dft q={q} b={b} rows={rows} {precision}", flush=True)
    rng = np.random.default_rng(q * 100 + b)
    x = rng.normal(size=(rows, q ** b)) + 1j * rng.normal(size=(rows, q ** b))
    dt = torch.complex128 if precision == "fp64" else torch.complex64
    xt = torch.from_numpy(x).to(dt).to(dv.current_device())
    got = dv.subtransform(xt, q, b, precision=precision).cpu().numpy()
    ref = np.stack([O.gwht(r, q, b) for r in x])
    assert rel_err(got, ref) < (1e-13 if precision == "fp64" else 3e-6)


@pytest.mark.parametrize("b", [4, 6])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_dft_whole_row_kernel_matches_pass_kernel(b, precision):
    """q = 4 rows through the whole-row kernel and through the per-pass kernel
    agree with each other and with the DFT definition."""
    print(f"dft rows4 b={b} {precision}", flush=True)
    rng = np.random.default_rng(13)
    rows, q = 7, 4
    x = rng.normal(size=(rows, q ** b)) + 1j * rng.normal(size=(rows, q ** b))
    dt = torch.complex128 if precision == "fp64" else torch.complex64
    code = dv.PRECISIONS[precision][0]
    dev = dv.current_device()
    outs = []
    for rows4 in (True, False):
        xt = torch.from_numpy(x).to(dt).to(dev)
        yt = torch.empty_like(xt)
        K.Dft([dict(q=q, b=b, rows=rows, inp=xt, out=yt)], code, rows4=rows4).launch()
        outs.append(yt.cpu().numpy())
    ref = np.stack([O.gwht(r, q, b) for r in x])
    tol = 1e-13 if precision == "fp64" else 3e-6
    assert rel_err(outs[0], ref) < tol and rel_err(outs[1], ref) < tol
    # both apply the four-point butterflies innermost digit first: the same bits
    fl = np.float64 if precision == "fp64" else np.float32
    assert np.array_equal(outs[0].view(fl), outs[1].view(fl))


@pytest.mark.parametrize("q,b", [(20, 4), (20, 3), (4, 7), (4, 5), (3, 8), (2, 12), (5, 5), (7, 3)])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_dft_every_schedule_gives_the_same_bits(q, b, precision):
    """The planner's candidate schedules (different pass splits, tile shapes,
    LDS budgets) apply the same axis transforms in the same order, so they
    give the same bits -- and match the DFT definition. The q = 20 transform
    is the prime-factor 4 x 5 one; 3 and 5 are the symmetric-pair forms."""
    rows = 3
    print(f"dft schedules q={q} b={b} {precision}", flush=True)
    rng = np.random.default_rng(q * 31 + b)
    x = rng.normal(size=(rows, q ** b)) + 1j * rng.normal(size=(rows, q ** b))
    dt = torch.complex128 if precision == "fp64" else torch.complex64
    code = dv.PRECISIONS[precision][0]
    dev = dv.current_device()
    outs = []
    for cand in range(6):
        d = L.DftDesc(q=q, b=b, rows=rows, in_stride=1, in_row_stride=q ** b)
        if L.lib.qsft_dft_plan_candidate(ctypes.byref(d), code, cand) != L.QSFT_OK:
            break
        xt = torch.from_numpy(x).to(dt).to(dev)
        yt = torch.empty_like(xt)
        K.Dft([dict(q=q, b=b, rows=rows, inp=xt, out=yt)], code, rows4=False, candidate=cand).launch()
        outs.append(yt.cpu().numpy())
    assert len(outs) >= 2
    ref = np.stack([O.gwht(r, q, b) for r in x])
    assert rel_err(outs[0], ref) < (1e-13 if precision == "fp64" else 3e-6)
    for o in outs[1:]:
        assert np.array_equal(o.view(np.float64 if precision == "fp64" else np.float32),
                              outs[0].view(np.float64 if precision == "fp64" else np.float32))


def test_dft_autotune_records_the_fastest_schedule():
    """qsft_dft_autotune times the candidates on the device and records the
    winner, which later plans of the shape use (same bits as candidate 0)."""
    q, b, rows = 20, 4, 4
    before, tuned0 = K.dft_schedule(q, b, L.QSFT_F32)
    ms = K.dft_autotune(q, b, L.QSFT_F32, rows)
    after, tuned1 = K.dft_schedule(q, b, L.QSFT_F32)
    print(f"q=20 b=4: model {before} -> tuned {after} ({ms:.3f} ms for {rows} rows)", flush=True)
    assert not tuned0 and tuned1 and ms > 0
    rng = np.random.default_rng(3)
    x = rng.normal(size=(rows, q ** b)) + 1j * rng.normal(size=(rows, q ** b))
    xt = torch.from_numpy(x).to(torch.complex64).to(dv.current_device())
    got = dv.subtransform(xt, q, b, precision="fp32").cpu().numpy()
    y0 = torch.empty_like(xt)
    K.Dft([dict(q=q, b=b, rows=rows, inp=xt, out=y0)], L.QSFT_F32, rows4=False, candidate=0).launch()
    assert np.array_equal(got.view(np.float32), y0.cpu().numpy().view(np.float32))


@pytest.mark.parametrize("q,b,bb", [(4, 6, 4), (3, 5, 2), (20, 3, 2)])
def test_dft_strided_sub_b(q, b, bb):
    """all_bs: the b' < b transform of row[::q^(b-b')]."""
    rng = np.random.default_rng(11)
    x = rng.normal(size=(5, q ** b)) + 1j * rng.normal(size=(5, q ** b))
    xt = torch.from_numpy(x).to(dv.current_device())
    got = dv.subtransform(xt, q, bb, in_stride=q ** (b - bb), precision="fp64").cpu().numpy()
    ref = np.stack([O.gwht(r[::q ** (b - bb)], q, bb) for r in x])
    assert rel_err(got, ref) < 1e-13


@pytest.mark.parametrize("p1,Q", [(5, 64), (43, 27), (65, 20)])
def test_assemble_gather_and_noise_array(p1, Q):
    """Exact gather + FP64 noise add; (43, 27): 86 of 129 source rows, odd Q;
    (65, 20): 130 rows, so the bin-major tiles have a 2-row second chunk."""
    print(f"assemble array p1={p1} Q={Q}", flush=True)
    rng = np.random.default_rng(12)
    Csrc, Rsrc = 3, 2
    src = rng.normal(size=(Csrc, Rsrc * p1, Q)) + 1j * rng.normal(size=(Csrc, Rsrc * p1, Q))
    noise = rng.normal(size=(2, 2 * p1, Q)) + 1j * rng.normal(size=(2, 2 * p1, Q))
    dev = dv.current_device()
    gsel, rsel = [2, 0], [1, 0]
    for prec, dt in (("fp64", torch.complex128), ("fp32", torch.complex64)):
        s = torch.from_numpy(src).to(dt).to(dev)
        sref = s.to(torch.complex128).cpu().numpy()
        ref = np.stack([np.vstack([sref[g, r * p1:(r + 1) * p1] for r in rsel]) for g in gsel]) + noise
        for bm in (False, True):
            shape = (2, Q, 2 * p1) if bm else (2, 2 * p1, Q)
            U = torch.empty(shape, dtype=torch.complex128, device=dev)
            K.assemble([dict(Q=Q, C=2, R=2, p1=p1, src=s, src_rows=Rsrc * p1, gsel=gsel, rsel=rsel, U=U,
                             noise=torch.from_numpy(noise).to(dev), bin_major=bm)])
            got = U.cpu().numpy()
            np.testing.assert_array_equal(got.transpose(0, 2, 1) if bm else got, ref)


def test_assemble_philox_noise_statistics():
    dev = dv.current_device()
    Q, C, R, p1 = 4096, 3, 2, 51
    src = torch.zeros((C, R * p1, Q), dtype=torch.complex64, device=dev)
    U = torch.empty((C, R * p1, Q), dtype=torch.complex128, device=dev)
    sigma = 0.37
    K.assemble([dict(Q=Q, C=C, R=R, p1=p1, src=src, src_rows=R * p1, gsel=[0, 1, 2], rsel=[0, 1], U=U,
                     sigma=sigma, seed=1234)])
    z = U.cpu().numpy().ravel()
    assert abs(np.mean(np.abs(z) ** 2) / sigma ** 2 - 1) < 0.01
    assert abs(np.mean(z.real)) < 0.01 * sigma and abs(np.mean(z.imag)) < 0.01 * sigma
    assert abs(np.var(z.real) / np.var(z.imag) - 1) < 0.02
    assert abs(np.mean(z.real * z.imag)) < 0.01 * sigma ** 2
    U2 = torch.empty_like(U)
    K.assemble([dict(Q=Q, C=C, R=R, p1=p1, src=src, src_rows=R * p1, gsel=[0, 1, 2], rsel=[0, 1], U=U2,
                     sigma=sigma, seed=1234)])
    assert torch.equal(U, U2)                      # counter-based: same seed, same draws


@pytest.mark.parametrize("R,p1,Q", [(2, 51, 4096), (1, 100, 33)])
def test_assemble_f32_tile_matches_f64_tile(R, p1, Q):
    """An FP32 source goes through the FP32 LDS tile; forcing the FP64 tile
    gives the same U and energies bit for bit (the FP32 -> FP64 promotion is
    exact wherever it happens)."""
    print(f"assemble tiles R={R} p1={p1} Q={Q}", flush=True)
    dev = dv.current_device()
    C = 3
    rng = np.random.default_rng(8)
    src = torch.from_numpy(rng.normal(size=(C, R * p1, Q)) + 1j * rng.normal(size=(C, R * p1, Q)))
    src = src.to(torch.complex64).to(dev)
    out = []
    for force64 in (False, True):
        U = torch.empty((C, Q, R * p1), dtype=torch.complex128, device=dev)
        e = torch.empty((C, Q), dtype=torch.float64, device=dev)
        a = K.Assemble([dict(Q=Q, C=C, R=R, p1=p1, src=src, src_rows=R * p1, gsel=[2, 0, 1], rsel=list(range(R)),
                             sigma=0.5, seed=7, U=U, energy=e, bin_major=True)])
        assert a.word == 1                           # QSFT_ASM_F32_TILE
        if force64:
            a.word = 0
        a.launch()
        torch.cuda.synchronize()
        out.append((U, e))
    assert torch.equal(out[0][0], out[1][0])
    assert torch.equal(out[0][1], out[1][1])


@pytest.mark.parametrize("R,p1,Q", [(2, 51, 4096), (3, 43, 27), (1, 129, 20)])
def test_assemble_philox_layout_independent(R, p1, Q):
    """The bin-major store draws exactly the row-major noise (the noise is a
    function of the logical element), including odd Q and a 1-row chunk."""
    print(f"assemble philox layouts R={R} p1={p1} Q={Q}", flush=True)
    dev = dv.current_device()
    C = 2
    rng = np.random.default_rng(5)
    src = torch.from_numpy(rng.normal(size=(C, R * p1, Q)) + 1j * rng.normal(size=(C, R * p1, Q)))
    src = src.to(torch.complex64).to(dev)
    args = dict(Q=Q, C=C, R=R, p1=p1, src=src, src_rows=R * p1, gsel=[1, 0], rsel=list(range(R))[::-1],
                sigma=0.8, seed=99)
    Ur = torch.empty((C, R * p1, Q), dtype=torch.complex128, device=dev)
    Ub = torch.empty((C, Q, R * p1), dtype=torch.complex128, device=dev)
    K.assemble([dict(args, U=Ur, bin_major=False)])
    K.assemble([dict(args, U=Ub, bin_major=True)])
    assert torch.equal(Ub.transpose(1, 2), Ur)
