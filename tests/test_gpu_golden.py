"""The device path against golden vectors of the reference (tests/golden/*.npz).

Each fixture is the reference's own run of quick_example's call sequence on a
fixed seed (`synt_exp/quick_example.py:7-58`: seed, get_random_subsampled_signal,
QSFT(**qsft_args).transform(report=True, sort=True)). Here the same calls go
through the reference's import paths (`qsft.qsft`, `synt_exp.synt_src.
synthetic_signal`, served by qsft_amd) on the GPU, in FP64:

  * the inputs the global RNG draws (support, strengths, M_c, delay rows) and
    the RNG state after the signal and after the transform are the reference's
    bit for bit;
  * the stored sub-DFTs agree (projections within 1e-11 relative: the device
    evaluates w^(<m,k> mod q) from an exact table, the reference exponentiates
    the unreduced phase);
  * the decoded support is the reference's bit for bit, in the same insertion
    order, coefficients within 1e-9 relative, and the report fields are equal.

Shapes: K0 (q=4 n=10 S=100 b=4, identity), a heavy-noise NSO shape, quick_example
itself (q=3 n=40 S=100 b=4, max_weight=4, noise_sd=1, identity channel), and the
headline workloads: q=4 n=50 S=1000 b=6 NSO R=2 noiseless and at 10 dB, and
q=20 n=16 S=2000 b=3 NSO R=2 at 20 dB. The served FP32 engine is held to the
reference's noiseless support at the q=4 n=50 shape.
"""
import numpy as np
import pytest
import torch

import golden_data as G

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    from qsft.qsft import QSFT
    from qsft.query import get_reed_solomon_dec
    from qsft_amd import engine as E
    from synt_exp.synt_src.synthetic_signal import get_random_subsampled_signal


@pytest.mark.parametrize("name", G.SMALL + G.LARGE)
def test_dropin_matches_reference_fixture(name):
    fx = G.load(name)
    print(f"golden {name}: q={fx['q']} n={fx['n']} S={fx['S']} b={fx['b']} R={fx['R']} "
          f"{fx['channel']} sd={fx['noise_sd']:.3g} seed={fx['seed']}", flush=True)
    np.random.seed(fx["seed"])
    if name == "quick_example":
        get_reed_solomon_dec(fx["n"], 4, fx["q"])          # built as quick_example builds it, before the draws
    kw = {} if G.max_weight(fx) is None else {"max_weight": G.max_weight(fx)}
    sig = get_random_subsampled_signal(n=fx["n"], q=fx["q"], sparsity=fx["S"], a_min=fx["a_min"],
                                       a_max=fx["a_max"], noise_sd=fx["noise_sd"], query_args=G.query_args(fx),
                                       verbose=False, **kw)
    G.assert_rng_state(fx, "signal")
    G.assert_signal(fx, sig.locq, sig.signal_w, sig.Ms, sig.Ds)
    G.assert_projections(fx, sig.Us, rtol=1e-11)
    res = QSFT(**G.qsft_args(fx)).transform(sig, report=True, sort=True)
    G.assert_rng_state(fx, "transform")
    err = G.assert_result(fx, res, rtol=1e-9)
    print(f"  {len(res['gwht'])} keys, same order, max rel diff {err:.2e}", flush=True)


def test_served_fp32_noiseless_support_is_the_references():
    """The served engine (matrix-core FP32 evaluation, FP32 DFT, FP64 peel) on
    the reference's own noiseless q=4 n=50 S=1000 b=6 inputs recovers exactly
    the reference's support, with values within FP32 evaluation accuracy."""
    fx = G.load("k1_noiseless")
    q, n, b, C, R = fx["q"], fx["n"], fx["b"], fx["C"], fx["R"]
    spec = E.TransformSpec(q=q, n=n, b=b, C=C, R=R, locq=fx["locq"].astype(np.int64),
                           strengths=fx["signal_w_vals"], Ms=[m.astype(np.int64) for m in fx["Ms"]],
                           D=np.vstack(fx["Ds"][0]).astype(np.int64), noise_sd=0.0, method="nso")
    bt = E.Batch([spec, spec], precision="fp32")
    bt.launch()
    want = {tuple(int(x) for x in r): v for r, v in zip(fx["gwht_keys"], fx["gwht_vals"])}
    for k, v, _, _ in bt.results():
        got = {tuple(int(x) for x in r): val for r, val in zip(k, v)}
        assert set(got) == set(want)
        rel = max(abs(got[key] - want[key]) / abs(want[key]) for key in want)
        print(f"served FP32: {len(got)} keys, max rel diff {rel:.2e}", flush=True)
        assert rel < 1e-5
