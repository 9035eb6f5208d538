"""Golden vectors of the reference q-SFT path (tests/golden/*.npz).

The fixtures were written by tests/golden/make_ref_fixtures.py, which runs the
reference's own `get_random_subsampled_signal` and `QSFT.transform` on a fixed
seed (quick_example's call sequence, `synt_exp/quick_example.py:7-58`). Each
holds the RNG-drawn inputs, the legacy MT19937 state after the signal and after
the transform, projections of the stored sub-DFTs, and the decoded dict in
insertion order with the report fields. Loaded with allow_pickle=False.
"""
import os

import numpy as np

DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SMALL = ["k0", "nso_heavy_noise", "quick_example"]
LARGE = ["k1_noiseless", "k1_10db", "k2_20db"]


def load(name):
    with np.load(os.path.join(DIR, f"{name}.npz"), allow_pickle=False) as z:
        fx = {k: z[k] for k in z.files}
    for k in ("q", "n", "S", "b", "C", "R", "seed", "max_weight", "n_samples", "max_hamming_weight",
              "pos_after_signal", "pos_after_transform", "proj_seed"):
        fx[k] = int(fx[k])
    for k in ("noise_sd", "a_min", "a_max", "avg_hamming_weight"):
        fx[k] = float(fx[k])
    fx["channel"] = str(fx["channel"])
    fx["name"] = name
    return fx


def query_args(fx):
    qa = {"query_method": "complex", "num_subsample": fx["C"], "delays_method_source": "identity",
          "subsampling_method": "qsft", "delays_method_channel": fx["channel"], "num_repeat": fx["R"],
          "b": fx["b"]}
    if fx["name"] == "quick_example":
        qa["t"] = 4
    return qa


def qsft_args(fx):
    return {"num_subsample": fx["C"], "num_repeat": fx["R"], "reconstruct_method_source": "identity",
            "reconstruct_method_channel": fx["channel"], "b": fx["b"], "noise_sd": fx["noise_sd"],
            "source_decoder": None}


def max_weight(fx):
    return None if fx["max_weight"] < 0 else fx["max_weight"]


def projection_vector(fx):
    """The fixed complex weights the stored sub-DFT rows were projected on
    (PCG64 stream of the stored seed: the same on every numpy version)."""
    g = np.random.default_rng(fx["proj_seed"])
    Q = fx["q"] ** fx["b"]
    return g.standard_normal(Q) + 1j * g.standard_normal(Q)


def assert_rng_state(fx, which):
    st = np.random.get_state()
    assert st[2] == fx[f"pos_after_{which}"], which
    assert np.array_equal(np.asarray(st[1], dtype=np.uint32), fx[f"mt_after_{which}"]), which


def assert_signal(fx, locq, signal_w, Ms, Ds):
    """RNG-drawn inputs equal the reference's, bit for bit."""
    np.testing.assert_array_equal(np.asarray(locq), fx["locq"])
    keys = np.array([list(k) for k in signal_w], dtype=np.int64).reshape(-1, fx["n"])
    np.testing.assert_array_equal(keys, fx["signal_w_keys"])
    np.testing.assert_array_equal(np.array(list(signal_w.values()), dtype=np.complex128), fx["signal_w_vals"])
    for i in range(fx["C"]):
        np.testing.assert_array_equal(np.asarray(Ms[i]), fx["Ms"][i])
        for j in range(fx["R"]):
            np.testing.assert_array_equal(np.asarray(Ds[i][j]), fx["Ds"][i][j])


def assert_projections(fx, Us, rtol):
    """Every stored sub-DFT U_{c,r} (rows x bins), projected on the fixed
    weights, within rtol of the reference's projection (relative to its norm)."""
    w = projection_vector(fx)
    b = fx["b"]
    for i in range(fx["C"]):
        for j in range(fx["R"]):
            got = np.asarray(Us[i][j][b]) @ w
            want = fx["U_proj"][i][j]
            assert np.linalg.norm(got - want) <= rtol * np.linalg.norm(want), (i, j)


def assert_result(fx, res, rtol):
    """The report of transform(report=True, sort=True): the same keys in the
    same insertion order, values within rtol, the same report fields."""
    gw = res["gwht"]
    keys = [tuple(int(x) for x in k) for k in gw]
    want = [tuple(int(x) for x in r) for r in fx["gwht_keys"]]
    assert set(keys) == set(want), (len(set(keys) - set(want)), len(set(want) - set(keys)))
    assert keys == want                                   # first-detection order
    vals = np.array(list(gw.values()), dtype=np.complex128)
    err = np.abs(vals - fx["gwht_vals"]) / np.maximum(np.abs(fx["gwht_vals"]), 1e-300)
    assert err.max(initial=0.0) <= rtol, float(err.max())
    np.testing.assert_array_equal(np.asarray(res["locations"]).reshape(-1, fx["n"]), fx["locations"])
    assert int(res["n_samples"]) == fx["n_samples"]
    assert float(res["avg_hamming_weight"]) == fx["avg_hamming_weight"]
    assert int(res["max_hamming_weight"]) == fx["max_hamming_weight"]
    return float(err.max(initial=0.0))
