"""End-to-end parity on the GPU: the drop-in API (qsft_amd) against the CPU
oracle on the same seed, plus the oracle's hand-built decision cases run
through the device peeler.

Bars: recovered support (set of k) bit-exact; coefficients within 1e-9
relative of the oracle in FP64 mode. Full-size configurations, which the
oracle cannot run in seconds, are checked against the reference's golden
vectors (tests/test_gpu_golden.py) and through size-independent properties:
noiseless support == planted support; at 10 dB an NMSE at the reference
decoder's own level on the same inputs.
"""
import numpy as np
import pytest
import torch

from oracle import qsft_oracle as O

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    from qsft_amd import _lib as L
    from qsft_amd import device as dv
    from qsft_amd import kernels as K
    from qsft_amd.qsft import QSFT
    from qsft_amd.synthetic_signal import get_random_subsampled_signal


def device_peel(Ms, Ds, Us, q, n, cutoff, method="nso", host_loop=False, unit_d=None):
    """Run the device peeler on numpy inputs; returns (gwht dict, info, final U).
    unit_d=0 forces the full dot product per delay row."""
    dev = dv.current_device()
    C, P, Q = np.asarray(Us).shape
    b = Ms[0].shape[1]
    n_pad = dv.n_pad_of(n)
    Ubm = np.ascontiguousarray(np.asarray(Us, dtype=np.complex128).transpose(0, 2, 1))     # bin-major
    U = torch.from_numpy(Ubm).to(dev)
    item = dict(q=q, n=n, b=b, C=C, P=P, p1=n + 1,
                method=L.QSFT_DETECT_NSO1 if method == "nso" else L.QSFT_DETECT_NOISELESS,
                cutoff=cutoff, Mt=dv.Mt_tensor(Ms, b, n_pad, dev), D=dv.D_tensor(Ds, n_pad, dev), U=U)
    if unit_d is not None:
        item["unit_d"] = unit_d
    batch = K.PeelBatch([item])
    batch.run(host_loop=host_loop)
    k, v, _, info = batch.results()[0]
    return {tuple(r): x for r, x in zip(k, v)}, info, U.cpu().numpy().transpose(0, 2, 1)


def assert_same_result(got, ref, rtol=1e-9):
    assert set(got) == set(ref), (len(set(got) - set(ref)), len(set(ref) - set(got)))
    for k, v in ref.items():
        assert abs(got[k] - v) <= rtol * max(abs(v), 1e-300), k
    assert list(got) == list(ref)          # insertion order = first detection order


# ------------------------------------------------------- decision-rule cases


def _nso_D(q, n, R, rng):
    src = np.vstack([np.zeros(n), np.eye(n)]).astype(int)
    return np.vstack([(rng.integers(0, q, n) - src) % q for _ in range(R)])


def _group_U(M, D, ks, amps, q, Q):
    b = M.shape[1]
    U = np.zeros((D.shape[0], Q), complex)
    w = q ** np.arange(b - 1, -1, -1)
    for k, a in zip(ks, amps):
        U[:, int(w @ ((M.T @ k) % q))] += a * np.exp(2j * np.pi * ((D @ k) % q) / q)
    return U


def _find(q, n, b, C, nballs, pred, rng):
    w = q ** np.arange(b - 1, -1, -1)
    for _ in range(20000):
        Ms = [rng.integers(0, q, size=(n, b)) for _ in range(C)]
        ks = [rng.integers(0, q, n) for _ in range(nballs)]
        bins = [[int(w @ ((M.T @ k) % q)) for k in ks] for M in Ms]
        if pred(bins):
            return Ms, ks, bins
    raise RuntimeError


@pytest.mark.parametrize("host_loop", [False, True])
def test_jacobi_case(host_loop):
    q, n, b, C = 4, 8, 2, 3
    rng = np.random.default_rng(16)
    Ms, ks, _ = _find(q, n, b, C, 2, lambda bn: bn[0][0] != bn[0][1] and bn[1][0] == bn[1][1]
                      and bn[2][0] != bn[2][1], rng)
    D = _nso_D(q, n, 2, rng)
    Us = np.array([_group_U(M, D, ks, [2.0 + 1j, -1.5 + 0.5j], q, q ** b) for M in Ms])
    tr = []
    _, ref, it = O.peel_decode(Ms, [D] * C, Us.copy(), q, n, 1e-9, trace=tr)
    got, info, _ = device_peel(Ms, [D] * C, Us, q, n, 1e-9, host_loop=host_loop)
    assert_same_result(got, ref)
    assert info["iterations"] == it and info["trace"] == tr


def test_last_writer_case():
    q, n, b, C = 4, 6, 2, 2
    rng = np.random.default_rng(17)
    Ms, ks, bins = _find(q, n, b, C, 1, lambda bn: True, rng)
    D = _nso_D(q, n, 2, rng)
    Us = np.array([_group_U(M, D, ks, [2.0 - 1j], q, q ** b) for M in Ms])
    Us[0, :, bins[0][0]] += 0.01
    Us[1, :, bins[1][0]] -= 0.02j
    Uo = Us.copy()
    _, ref, _ = O.peel_decode(Ms, [D] * C, Uo, q, n, 1e-2)
    got, info, Ud = device_peel(Ms, [D] * C, Us, q, n, 1e-2)
    assert_same_result(got, ref, rtol=1e-14)
    # peeled with the later group's rho; the device signature comes from an exact
    # w-table, numpy's from repeated multiplication (a few ulp of |rho| apart)
    np.testing.assert_allclose(Ud, Uo, rtol=0, atol=1e-14)


def test_iteration_cap_case():
    q, n, b = 4, 6, 2
    rng = np.random.default_rng(20)
    w = q ** np.arange(b - 1, -1, -1)
    M0, M1 = rng.integers(0, q, size=(n, b)), rng.integers(0, q, size=(n, b))
    D = _nso_D(q, n, 2, rng)
    k = rng.integers(0, q, n)
    A, B = int(w @ ((M0.T @ k) % q)), int(w @ ((M1.T @ k) % q))
    sig = np.exp(2j * np.pi * ((D @ k) % q) / q)
    U = np.zeros((2, D.shape[0], q ** b), complex)
    U[0, :, A] = 2.0 * sig
    U[1, :, B] = 1.0 * sig
    k2 = rng.integers(0, q, n)
    X = (int(w @ ((M0.T @ k2) % q)) + 1) % q ** b
    if X == A:
        X = (X + 1) % q ** b
    U[0, :, X] = 1.5 * np.exp(2j * np.pi * ((D @ k2) % q) / q)
    tr = []
    _, ref, it = O.peel_decode([M0, M1], [D, D], U.copy(), q, n, 1e-3, trace=tr)
    got, info, _ = device_peel([M0, M1], [D, D], U, q, n, 1e-3)
    assert it == 15 and info["iterations"] == 15 and info["trace"] == tr
    assert_same_result(got, ref, rtol=1e-14)


def test_bin_mismatch_and_residual_cases():
    q, n, b = 4, 7, 2
    rng = np.random.default_rng(19)
    Ms, ks, _ = _find(q, n, b, 1, 2, lambda bn: bn[0][0] == bn[0][1], rng)
    D = _nso_D(q, n, 2, rng)
    U = np.array([_group_U(Ms[0], D, ks, [1.0, 0.9j], q, q ** b)])
    k3 = rng.integers(0, q, n)
    w = q ** np.arange(b - 1, -1, -1)
    wrong = (int(w @ ((Ms[0].T @ k3) % q)) + 3) % q ** b
    U[0, :, wrong] += 1.5 * np.exp(2j * np.pi * ((D @ k3) % q) / q)
    tr = []
    _, ref, it = O.peel_decode(Ms, [D], U.copy(), q, n, 1e-9, trace=tr)
    got, info, _ = device_peel(Ms, [D], U, q, n, 1e-9)
    assert info["trace"] == tr and got == ref == {}


def test_noiseless_detection_case():
    q, n, b, C = 3, 9, 2, 3
    rng = np.random.default_rng(23)
    Ms = [rng.integers(0, q, size=(n, b)) for _ in range(C)]
    D = np.vstack([np.zeros(n), np.eye(n)]).astype(int)
    ks = [rng.integers(0, q, n) for _ in range(4)]
    amps = [1.0, 2j, -3.0, 0.5 + 0.5j]
    Us = np.array([_group_U(M, D, ks, amps, q, q ** b) for M in Ms])
    tr = []
    _, ref, _ = O.peel_decode(Ms, [D] * C, Us.copy(), q, n, 1e-9, method_channel="identity", trace=tr)
    got, info, _ = device_peel(Ms, [D] * C, Us, q, n, 1e-9, method="identity")
    assert info["trace"] == tr
    assert_same_result(got, ref, rtol=1e-12)


# ------------------------------------------------------------ end to end


E2E = [
    # (q, n, S, b, C, R, channel, noise_sd, seed)
    (4, 10, 100, 4, 3, 1, "identity", 0.0, 0),      # the reference's CPU-runnable configuration
    (4, 20, 60, 4, 3, 2, "nso", 0.5, 1),
    (4, 20, 200, 4, 3, 2, "nso", 10.0, 2),          # heavy noise: false and missed detections
    (20, 8, 150, 2, 3, 2, "nso", 0.5, 3),
    (3, 12, 40, 3, 3, 2, "nso", 0.2, 4),
    (2, 30, 50, 6, 3, 3, "nso", 0.3, 5),
    (3, 9, 30, 3, 3, 1, "identity", 0.0, 6),
]


def _make(mod, q, n, S, b, C, R, channel, noise_sd, seed, **kw):
    qa = dict(query_method="complex", num_subsample=C, b=b, delays_method_source="identity",
              delays_method_channel=channel, num_repeat=R, subsampling_method="qsft")
    np.random.seed(seed)
    if mod == "oracle":
        sig = O.get_random_subsampled_signal(n, q, noise_sd, S, 1, 10, qa)
    else:
        sig = get_random_subsampled_signal(n=n, q=q, noise_sd=noise_sd, sparsity=S, a_min=1, a_max=10,
                                           query_args=qa, verbose=False, **kw)
    return sig


@pytest.mark.parametrize("q,n,S,b,C,R,channel,noise_sd,seed", E2E)
def test_transform_matches_oracle_same_seed(q, n, S, b, C, R, channel, noise_sd, seed):
    print(f"e2e q={q} n={n} S={S} b={b} C={C} R={R} {channel} sd={noise_sd} seed={seed}", flush=True)
    ref_sig = _make("oracle", q, n, S, b, C, R, channel, noise_sd, seed)
    ref = O.qsft_transform(ref_sig, C, R, b, "identity", channel, report=True, sort=True)
    st_ref = np.random.get_state()
    sig = _make("dev", q, n, S, b, C, R, channel, noise_sd, seed)
    for i in range(C):
        for j in range(R):
            np.testing.assert_array_equal(sig.Ds[i][j], ref_sig.Ds[i][j])
        np.testing.assert_array_equal(sig.Ms[i], ref_sig.Ms[i])
    sft = QSFT(num_subsample=C, num_repeat=R, reconstruct_method_source="identity",
               reconstruct_method_channel=channel, b=b, noise_sd=noise_sd)
    res = sft.transform(sig, report=True, sort=True)
    # the same global RNG consumption as the reference
    assert np.random.get_state()[1].tolist() == st_ref[1].tolist()
    assert_same_result(res["gwht"], ref["gwht"])
    np.testing.assert_array_equal(res["locations"], ref["locations"])
    assert res["n_samples"] == ref["n_samples"]
    assert res["avg_hamming_weight"] == ref["avg_hamming_weight"]
    assert res["max_hamming_weight"] == ref["max_hamming_weight"]


def test_stored_transforms_match_oracle_and_get_MDU():
    q, n, S, b, C, R = 4, 12, 30, 3, 3, 2
    ref_sig = _make("oracle", q, n, S, b, C, R, "nso", 0.1, 9)
    sig = _make("dev", q, n, S, b, C, R, "nso", 0.1, 9)
    for i in range(C):
        for j in range(R):
            np.testing.assert_allclose(np.array(sig.Us[i][j][b]), np.array(ref_sig.Us[i][j][b]), rtol=0,
                                       atol=1e-12)
    np.random.seed(3)
    Ms, Ds, Us = sig.get_MDU(C, R, b)
    np.random.seed(3)
    Mr, Dr, Ur = ref_sig.get_MDU(C, R, b)
    for i in range(C):
        np.testing.assert_array_equal(Ms[i], Mr[i])
        for j in range(R):
            np.testing.assert_allclose(Us[i][j], Ur[i][j], rtol=0, atol=1e-12)


def test_generic_subsample_callback_path():
    """A user subclass with only subsample(): indices from the stage-(a) kernel,
    samples from the callback, sub-transforms on the device."""
    from qsft_amd.input_signal_subsampled import SubsampledSignal

    q, n, S, b = 3, 8, 12, 3
    rng = np.random.default_rng(31)
    locq = rng.integers(0, q, size=(n, S))
    a = rng.normal(size=S) + 1j * rng.normal(size=S)
    calls = []

    class Plugin(SubsampledSignal):
        def subsample(self, query_indices):
            calls.append(len(query_indices))
            dig = O.dec_to_qary_vec(query_indices, q, n).T
            return O.eval_synthetic(dig, locq, a, q)

    qa = dict(query_method="complex", num_subsample=3, b=b, delays_method_source="identity",
              delays_method_channel="identity", subsampling_method="qsft")
    np.random.seed(4)
    sig = Plugin(n=n, q=q, query_args=dict(qa))
    assert calls == [(n + 1) * q ** b] * 3
    for i in range(3):
        cf = O.closed_form_U(sig.Ms[i], sig.Ds[i][0], locq, a, q)
        np.testing.assert_allclose(np.array(sig.Us[i][0][b]), cf, rtol=0, atol=1e-12)
    np.random.seed(7)
    gw = QSFT(num_subsample=3, num_repeat=1, reconstruct_method_source="identity",
              reconstruct_method_channel="identity", b=b).transform(sig)
    # the oracle decodes the same stored transforms after the same group and
    # repeat permutation (the base get_MDU adds no noise)
    np.random.seed(7)
    sub = np.random.choice(3, 3, replace=False)
    rep = np.random.choice(1, 1, replace=False)
    Ms = [sig.Ms[i][:, :b] for i in sub]
    Ds = [np.vstack([sig.Ds[i][j] for j in rep]) for i in sub]
    Us = np.array([np.vstack([O.closed_form_U(sig.Ms[i], sig.Ds[i][j], locq, a, q) for j in rep]) for i in sub])
    _, ref, _ = O.peel_decode(Ms, Ds, Us, q, n, 1e-9, method_channel="identity")
    assert_same_result(gw, ref)
    truth = {}
    for s in range(S):
        truth[tuple(locq[:, s])] = truth.get(tuple(locq[:, s]), 0) + a[s]
    assert set(gw) <= set(truth) and len(gw) >= len(truth) // 2, (len(gw), len(truth))
    for k, v in gw.items():
        assert abs(v - truth[k]) <= 1e-9 * abs(truth[k])


# ------------------------------------------------- full-size configurations


def test_k1_noiseless_fp32_support_exact():
    """q=4, n=50, S=1000, b=6, C=3, NSO R=2: noiseless FP32 samples/DFT,
    FP64 peel -> support equals the planted support exactly."""
    q, n, S, b = 4, 50, 1000, 6
    qa = dict(query_method="complex", num_subsample=3, b=b, delays_method_source="identity",
              delays_method_channel="nso", num_repeat=2, subsampling_method="qsft")
    np.random.seed(101)
    sig = get_random_subsampled_signal(n=n, q=q, noise_sd=0.0, sparsity=S, a_min=1, a_max=10, query_args=qa,
                                       verbose=False, precision="fp32")
    gw = QSFT(num_subsample=3, num_repeat=2, reconstruct_method_source="identity",
              reconstruct_method_channel="nso", b=b).transform(sig)
    assert set(gw) == set(sig.signal_w)
    err = max(abs(gw[k] - v) / abs(v) for k, v in sig.signal_w.items())
    assert err < 1e-5


def test_k1_noisy_fp32_nmse_at_the_references_level():
    """q=4, n=50, S=1000, b=6, NSO R=2 at 10 dB, on the reference's own inputs
    (tests/golden/k1_10db.npz): the served FP32 path with device noise, over 8
    noise draws, has a median NMSE within 1.3x of the reference decoder's
    NMSE on those inputs (4.9e-3: nso1 with R = 2 misses small coefficients
    at this SNR, so 1e-3 is not reached by the reference either), and the
    drop-in FP32 route with device noise lands in the same range."""
    import golden_data as G
    from qsft_amd import engine as E
    fx = G.load("k1_10db")
    q, n, b, C, R = fx["q"], fx["n"], fx["b"], fx["C"], fx["R"]
    truth = {tuple(int(x) for x in k): v for k, v in zip(fx["signal_w_keys"], fx["signal_w_vals"])}
    est = {tuple(int(x) for x in k): v for k, v in zip(fx["gwht_keys"], fx["gwht_vals"])}
    ref_nmse = O.nmse(est, truth)
    specs = [E.TransformSpec(q=q, n=n, b=b, C=C, R=R, locq=fx["locq"].astype(np.int64),
                             strengths=fx["signal_w_vals"], Ms=[m.astype(np.int64) for m in fx["Ms"]],
                             D=np.vstack(fx["Ds"][0]).astype(np.int64), noise_sd=fx["noise_sd"], seed=s)
             for s in range(8)]
    bt = E.Batch(specs, precision="fp32")
    bt.launch(seed_offset=1)
    nm = [E.nmse(k, v, sp) for (k, v, _, _), sp in zip(bt.results(), specs)]
    print(f"reference NMSE {ref_nmse:.3e}; served FP32 median {np.median(nm):.3e} "
          f"(min {min(nm):.3e}, max {max(nm):.3e})", flush=True)
    assert 3e-3 < ref_nmse < 8e-3
    assert np.median(nm) <= 1.3 * ref_nmse
    qa = dict(query_method="complex", num_subsample=3, b=b, delays_method_source="identity",
              delays_method_channel="nso", num_repeat=2, subsampling_method="qsft")
    np.random.seed(102)
    sig = get_random_subsampled_signal(n=n, q=q, noise_sd=0.0, sparsity=1000, a_min=1, a_max=10, query_args=qa,
                                       verbose=False, precision="fp32", noise_mode="device")
    sig.noise_sd = float(np.sqrt(np.sum(np.abs(sig.strengths) ** 2) / 10 ** (10 / 10)))   # 10 dB
    gw = QSFT(num_subsample=3, num_repeat=2, reconstruct_method_source="identity",
              reconstruct_method_channel="nso", b=b).transform(sig)
    nm_dropin = O.nmse(gw, sig.signal_w)
    print(f"drop-in FP32 device noise: NMSE {nm_dropin:.3e}, recovered {len(gw)}", flush=True)
    assert nm_dropin <= 2.0 * ref_nmse


@pytest.mark.parametrize("channel,R", [("identity", 1), ("nso", 2)])
def test_large_amplitudes_noiseless_match_oracle(channel, R):
    """Strengths of 5e3..2e4 against the noiseless cutoff (1e-9 per row): a
    residual formed as |col|^2 - P|rho|^2 would carry a rounding error above
    the cutoff and call true singletons multitons. Same seed as the oracle:
    the same support (the planted one), order and values."""
    q, n, S, b, C = 4, 16, 120, 4, 3
    qa = dict(query_method="complex", num_subsample=C, b=b, delays_method_source="identity",
              delays_method_channel=channel, num_repeat=R, subsampling_method="qsft")
    print(f"large amplitudes, {channel}", flush=True)
    np.random.seed(31)
    ref_sig = O.get_random_subsampled_signal(n, q, 0.0, S, 5e3, 2e4, dict(qa))
    ref = O.qsft_transform(ref_sig, C, R, b, "identity", channel)
    np.random.seed(31)
    sig = get_random_subsampled_signal(n=n, q=q, noise_sd=0.0, sparsity=S, a_min=5e3, a_max=2e4,
                                       query_args=dict(qa), verbose=False)
    got = QSFT(num_subsample=C, num_repeat=R, reconstruct_method_source="identity",
               reconstruct_method_channel=channel, b=b).transform(sig)
    assert set(ref) == set(sig.signal_w)
    assert_same_result(got, ref)


@pytest.mark.parametrize("channel", ["nso", "identity", "random"])
def test_unit_delay_signatures_match_row_products(channel):
    """The structured-delay signature (R offset products + one digit per row)
    gives the same decode as a full dot product per row; a random delay matrix
    is detected as unstructured."""
    q, n, b, C, S = 4, 12, 4, 3, 60
    rng = np.random.default_rng(31)
    Ms = [rng.integers(0, q, size=(n, b)) for _ in range(C)]
    if channel == "nso":
        D, R = _nso_D(q, n, 2, rng), 2
    elif channel == "identity":
        D, R = np.vstack([np.zeros(n), np.eye(n)]).astype(int), 1
    else:
        D, R = rng.integers(0, q, size=(2 * (n + 1), n)), 2
    locq = rng.integers(0, q, size=(n, S))
    a = rng.uniform(1, 10, S) * np.exp(2j * np.pi * rng.uniform(size=S))
    Us = np.array([O.closed_form_U(M, D, locq, a, q) for M in Ms])
    method = "identity" if channel == "identity" else "nso"
    print(f"unit delays {channel}", flush=True)
    n_pad = dv.n_pad_of(n)
    sign = K.unit_delay_sign(dv.D_tensor([D] * C, n_pad, dv.current_device()), q, n, n + 1)
    assert sign == {"nso": -1, "identity": 1, "random": 0}[channel]
    got, info, U1 = device_peel(Ms, [D] * C, Us.copy(), q, n, 1e-9, method=method)
    ref, info0, U0 = device_peel(Ms, [D] * C, Us.copy(), q, n, 1e-9, method=method, unit_d=0)
    assert_same_result(got, ref, rtol=0.0)
    assert info["trace"] == info0["trace"]
    np.testing.assert_array_equal(U1, U0)


def test_synthetic_helper_heldout_nmse(tmp_path):
    """TestHelper / SyntheticHelper (qsft/test_helper.py, synt_exp/synt_src/
    synthetic_helper.py): decode the training signal, score the spectrum on
    noiseless uniform held-out queries with the GPU evaluator. Equal to the
    oracle's _test_qary on the same queries: ~0 without noise, 1e-9 relative
    with noise; doubled coefficients score 1, an empty spectrum 1."""
    from qsft_amd.synthetic_helper import SyntheticHelper
    from qsft_amd.synthetic_signal import generate_signal_w

    q, n, S, b = 4, 20, 40, 4
    np.random.seed(11)
    signal_w, locq, strengths = generate_signal_w(n, q, S, 1, 10, full=False)
    signal_args = dict(n=n, q=q, locq=locq, strengths=strengths, noise_sd=0.0, signal_w=signal_w)
    sub_args = dict(num_subsample=3, num_repeat=2, b=b)
    helper = SyntheticHelper(signal_args, ["qsft"], sub_args, dict(n_samples=4000), tmp_path)
    assert (tmp_path / "config.json").is_file() and (tmp_path / "test").is_dir()
    test_t = helper.test_signal.signal_t
    idx, samples = list(test_t.keys()), list(test_t.values())
    assert len(idx) == 4000
    for noise_sd, tol in ((0.0, None), (3.0, 1e-9)):
        out = helper.compute_model("qsft", dict(sub_args, noise_sd=noise_sd), report=True)
        beta = out["gwht"]
        got = helper.test_model("qsft", beta=beta)
        ref = O.heldout_nmse(idx, samples, beta, q, n)
        print(f"noise_sd {noise_sd}: found {len(beta)}, held-out NMSE {got:.3e} (oracle {ref:.3e})", flush=True)
        if tol is None:
            assert got < 1e-20 and ref < 1e-20
            exact = beta
        else:
            assert 1e-8 < ref < 1 and abs(got - ref) <= tol * ref
    assert helper.test_model("qsft", beta={k: 2 * v for k, v in exact.items()}) == pytest.approx(1.0, rel=1e-9)
    assert helper.test_model("qsft", beta={}) == 1
    with pytest.raises(NotImplementedError):
        helper.compute_model("lasso", {})


@pytest.mark.gpu
def test_folder_cache_of_delays_samples_and_transforms(tmp_path):
    """The `folder` cache (qsft/input_signal_subsampled.py:92-155): the first
    construction writes M/D, samples and every b's sub-DFTs and takes the same
    RNG draws as no cache; a second one loads M/D (no M/D draws) and the
    transforms (bit-identical, same decode); with only samples cached the
    sub-DFTs are recomputed from them, bit-identical again."""
    import shutil

    from qsft_amd.qsft import QSFT
    from qsft_amd.synthetic_signal import get_random_subsampled_signal

    q, n, S, b, C, R, sd = 4, 10, 30, 4, 3, 2, 0.5
    qa = dict(query_method="complex", num_subsample=C, b=b, all_bs=[b - 1, b], delays_method_source="identity",
              delays_method_channel="nso", num_repeat=R, subsampling_method="qsft")

    def make(folder):
        np.random.seed(7)
        sig = get_random_subsampled_signal(n=n, q=q, noise_sd=sd, sparsity=S, a_min=1, a_max=10,
                                           query_args=dict(qa), verbose=False, folder=folder)
        return sig, np.random.rand()

    def decode(sig):
        np.random.seed(99)
        return QSFT(num_subsample=C, num_repeat=R, reconstruct_method_source="identity",
                    reconstruct_method_channel="nso", b=b).transform(sig)

    def same_us(a, b_):
        for bb in qa["all_bs"]:
            for i in range(C):
                for j in range(R):
                    assert np.array_equal(np.asarray(a.Us[i][j][bb]), np.asarray(b_.Us[i][j][bb]))

    plain, r_plain = make(None)
    first, r_first = make(str(tmp_path))
    assert r_first == r_plain                                          # writing the cache changes no draw
    assert (tmp_path / "Ms_and_Ds.npz").is_file()
    assert all((tmp_path / "samples" / f"M{i}_D{j}.npz").is_file() for i in range(C) for j in range(R))
    assert all((tmp_path / "transforms" / f"U{i}_{j}.npz").is_file() for i in range(C) for j in range(R))
    loaded, r_loaded = make(str(tmp_path))
    assert r_loaded != r_plain                                         # M/D came from the cache
    assert all(np.array_equal(a, c) for a, c in zip(loaded.Ms, plain.Ms))
    same_us(loaded, plain)
    want = decode(plain)
    got = decode(loaded)
    assert list(got) == list(want) and all(got[k] == want[k] for k in want)
    shutil.rmtree(tmp_path / "transforms")
    resampled, _ = make(str(tmp_path))                                 # samples cached, transforms recomputed
    same_us(resampled, plain)
    assert all((tmp_path / "transforms" / f"U{i}_{j}.npz").is_file() for i in range(C) for j in range(R))
