"""Coded (Reed-Solomon) source delays on the GPU: the device syndrome decoder
inside the peeling kernels against the oracle.

  * The drop-in API with delays_method_source='coded' on the same seed as
    the oracle: quick_example's own configuration with the coded source
    (q=3, n=40, t=4, supports of weight <= 4), noiseless and with its
    noise, an NSO channel, and q=5: same support, order, values within 1e-9.
  * The device peeler on hand-built bins (closed-form sub-transforms) with
    planted keys of weight <= t (decoded) and of weight > t (a failed or
    wrong decode makes a multiton): the same pass trace and result as the
    oracle's Peterson-Gorenstein-Zierler decoder.
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
    from qsft.qsft import QSFT
    from qsft.query import get_reed_solomon_dec
    from synt_exp.synt_src.synthetic_signal import get_random_subsampled_signal


CASES = [
    # (q, n, S, b, t, max_weight, channel, R, noise_sd, a_min, a_max, seed)
    (3, 40, 100, 4, 4, 4, "identity", 1, 0.0, 1, 1, 20),         # quick_example, noiseless
    (3, 40, 100, 4, 4, 4, "identity", 1, 1.0, 1, 1, 20),         # quick_example as shipped (noise_sd = 1)
    (3, 40, 100, 4, 4, 4, "nso", 2, 0.5, 1, 10, 21),
    (5, 30, 60, 3, 3, 3, "identity", 1, 0.0, 1, 10, 22),
    (2, 24, 50, 5, 2, 2, "nso", 2, 0.2, 1, 10, 23),
]


@pytest.mark.parametrize("q,n,S,b,t,mw,channel,R,sd,amin,amax,seed", CASES)
def test_coded_transform_matches_oracle_same_seed(q, n, S, b, t, mw, channel, R, sd, amin, amax, seed):
    print(f"coded q={q} n={n} S={S} b={b} t={t} {channel} R={R} sd={sd}", flush=True)
    qa = dict(query_method="complex", num_subsample=3, b=b, delays_method_source="coded",
              delays_method_channel=channel, num_repeat=R, subsampling_method="qsft", t=t)
    np.random.seed(seed)
    ref_sig = O.get_random_subsampled_signal(n, q, sd, S, amin, amax, dict(qa), max_weight=mw)
    ref = O.qsft_transform(ref_sig, 3, R, b, "coded", channel, report=True, sort=True)
    st_ref = np.random.get_state()
    np.random.seed(seed)
    dec = get_reed_solomon_dec(n, t, q)
    sig = get_random_subsampled_signal(n=n, q=q, noise_sd=sd, sparsity=S, a_min=amin, a_max=amax,
                                       query_args=dict(qa), max_weight=mw, verbose=False)
    for i in range(3):
        for j in range(R):
            np.testing.assert_array_equal(sig.Ds[i][j], ref_sig.Ds[i][j])
    res = QSFT(num_subsample=3, num_repeat=R, reconstruct_method_source="coded", reconstruct_method_channel=channel,
               b=b, noise_sd=sd, source_decoder=dec).transform(sig, report=True, sort=True)
    assert np.random.get_state()[1].tolist() == st_ref[1].tolist()
    got, want = res["gwht"], ref["gwht"]
    assert set(got) == set(want), (len(set(got) - set(want)), len(set(want) - set(got)))
    assert list(got) == list(want)
    for k, v in want.items():
        assert abs(got[k] - v) <= 1e-9 * abs(v), k
    np.testing.assert_array_equal(res["locations"], ref["locations"])
    if sd == 0:
        assert set(got) == set(sig.signal_w)                     # noiseless: the planted support
    print(f"  {len(got)} keys of {len(sig.signal_w)}", flush=True)


def _group_U(M, D, keys, amps, q):
    locq = np.array(keys).T
    return O.closed_form_U(M, D, locq, np.asarray(amps, dtype=complex), q)


def test_device_decoder_on_hand_built_bins_matches_oracle():
    """Planted keys of weight <= t decode; heavier keys (t + 1, t + 3) fail or
    mis-decode and stay multitons: trace and result equal the oracle's."""
    q, n, t, b, C = 3, 40, 4, 3, 3
    rng = np.random.default_rng(5)
    D = O.rs_delay_matrix(n, t, q)
    Ms = [rng.integers(0, q, size=(n, b)) for _ in range(C)]
    keys, amps = [], []
    for w in [1, 2, 3, 4, 4, 3, 2, 5, 7, 4, 1, 0]:
        k = np.zeros(n, dtype=np.int64)
        k[rng.choice(n, w, replace=False)] = rng.integers(1, q, w)
        keys.append(k)
        amps.append(rng.uniform(1, 3) * np.exp(2j * np.pi * rng.uniform()))
    Us = np.array([_group_U(M, D, keys, amps, q) for M in Ms])
    tr = []
    _, ref, _ = O.peel_decode(Ms, [D] * C, Us.copy(), q, n, 1e-9, method_source="coded",
                              method_channel="identity", source_parity=D.shape[0], trace=tr, t=t)
    dev = dv.current_device()
    n_pad = dv.n_pad_of(n)
    from qsft_amd.reed_solomon import ReedSolomon
    rs = ReedSolomon(n, t, q)
    item = dict(q=q, n=n, b=b, C=C, P=D.shape[0], p1=D.shape[0],
                method=L.QSFT_DETECT_NOISELESS | L.QSFT_DETECT_CODED, cutoff=1e-9,
                Mt=dv.Mt_tensor(Ms, b, n_pad, dev), D=dv.D_tensor([D] * C, n_pad, dev),
                U=torch.from_numpy(np.ascontiguousarray(Us.transpose(0, 2, 1))).to(dev),
                coded=dict(t=t, s=rs.s, tables=torch.from_numpy(rs.device_tables()).to(dev)))
    batch = K.PeelBatch([item])
    batch.run()
    k, v, _, info = batch.results()[0]
    got = {tuple(int(x) for x in r): val for r, val in zip(k, v)}
    print("trace", info["trace"], "oracle", tr, flush=True)
    assert info["trace"] == tr
    assert set(got) == set(ref) and list(got) == list(ref)
    for key, val in ref.items():
        assert abs(got[key] - val) <= 1e-12 * abs(val)
    light = {tuple(int(x) for x in kk) for kk in keys if np.count_nonzero(kk) <= t}
    assert light <= set(got)
