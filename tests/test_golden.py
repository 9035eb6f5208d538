"""The CPU oracle against golden vectors of the reference (tests/golden/*.npz).

Same seed, same calls as the reference run that wrote each fixture: the
oracle must draw the same support, strengths, M_c and delay rows, leave the
global legacy RNG in the same state after the signal and after the
transform (which pins the order and count of the get_MDU permutations and
noise draws), store the same sub-DFTs, and decode the same dict in the same
order with the same report fields.

The large shapes (q=4 n=50 S=1000 b=6; q=20 n=16 S=2000 b=3) take the
oracle minutes, so they are checked here only on the inputs the RNG draws
first; the device path is checked against their full outputs in
tests/test_gpu_golden.py.
"""
import numpy as np
import pytest

import golden_data as G
from oracle import qsft_oracle as O


def _oracle_signal(fx):
    np.random.seed(fx["seed"])
    return O.get_random_subsampled_signal(fx["n"], fx["q"], fx["noise_sd"], fx["S"], fx["a_min"], fx["a_max"],
                                          G.query_args(fx), max_weight=G.max_weight(fx))


@pytest.mark.parametrize("name", G.SMALL)
def test_oracle_matches_reference_fixture(name):
    fx = G.load(name)
    sig = _oracle_signal(fx)
    G.assert_signal(fx, sig.locq, sig.signal_w, sig.Ms, sig.Ds)
    G.assert_rng_state(fx, "signal")
    G.assert_projections(fx, sig.Us, rtol=1e-13)
    res = O.qsft_transform(sig, fx["C"], fx["R"], fx["b"], "identity", fx["channel"], report=True, sort=True)
    G.assert_rng_state(fx, "transform")
    G.assert_result(fx, res, rtol=1e-12)


@pytest.mark.parametrize("name", G.LARGE)
def test_oracle_draws_match_large_fixtures(name):
    """Support, strengths, M_c and delay rows of the large shapes, drawn in the
    reference's order (the sampling itself is left to the device tests)."""
    fx = G.load(name)
    q, n, S, b, C = fx["q"], fx["n"], fx["S"], fx["b"], fx["C"]
    np.random.seed(fx["seed"])
    signal_w, locq, _ = O.generate_signal_w(n, q, S, fx["a_min"], fx["a_max"], G.max_weight(fx))
    qa = G.query_args(fx)
    Ms, Ds = O.get_Ms_and_Ds(n, q, **qa)
    G.assert_signal(fx, locq, signal_w, Ms, Ds)


def test_fixtures_are_the_documented_shapes():
    """The fixtures hold the workloads they are named for (and the decoded
    supports are non-trivial: a decode that found nothing would pass nothing)."""
    shapes = {"k0": (4, 10, 100, 4, 3, 1), "nso_heavy_noise": (4, 20, 200, 4, 3, 2),
              "quick_example": (3, 40, 100, 4, 3, 1), "k1_noiseless": (4, 50, 1000, 6, 3, 2),
              "k1_10db": (4, 50, 1000, 6, 3, 2), "k2_20db": (20, 16, 2000, 3, 3, 2)}
    for name, shape in shapes.items():
        fx = G.load(name)
        assert (fx["q"], fx["n"], fx["S"], fx["b"], fx["C"], fx["R"]) == shape
        assert len(fx["gwht_keys"]) >= 0.8 * fx["S"], name
        assert fx["n_samples"] == fx["C"] * fx["R"] * (fx["n"] + 1) * fx["q"] ** fx["b"]
    clean = G.load("k1_noiseless")
    truth = {tuple(int(x) for x in c) for c in clean["locq"].T}
    assert {tuple(int(x) for x in r) for r in clean["gwht_keys"]} == truth
