#!/usr/bin/env python3
"""Write golden vectors of the reference q-SFT path (run in the build container only).

    python tests/golden/make_ref_fixtures.py [case ...]

Imports the reference implementation from /root/reference (pure Python) and
runs its own public calls, exactly as `synt_exp/quick_example.py:7-58` does:

    np.random.seed(seed)
    signal = get_random_subsampled_signal(n, q, noise_sd, sparsity, a_min, a_max, query_args[, max_weight])
    result = QSFT(**qsft_args).transform(signal, report=True, sort=True)

and stores, per case, in `tests/golden/<case>.npz` (data only, no pickles):

  * the inputs the global RNG produced, in the reference's draw order:
    locq and the signal_w dict of planted strengths
    (`synt_exp/synt_src/synthetic_signal.py:10-38`), M_c and the delay rows
    (`qsft/query.py:26-143,182-203`);
  * the full legacy MT19937 state after the signal is built and after the
    transform (pins how many draws the sub-sampling, get_MDU's permutations
    `qsft/input_signal_subsampled.py:246-247` and the noise blocks
    `synt_exp/synt_src/synthetic_signal.py:125-129` consumed);
  * a projection of every stored noiseless sub-DFT U_{c,r} (rows x bins) on a
    fixed complex vector w (regenerated from the seed stored beside it), so
    sample values and DFT are pinned without storing q^b bins per row;
  * the decoded output: gwht keys in insertion order and their values, the
    sorted locations, n_samples and the Hamming-weight fields of the report
    (`qsft/qsft.py:247-281`).

The reference needs three packages this image lacks only off the path being
recorded: `group_lasso` (`qsft/utils.py:6`, LASSO baseline) and `galois`
(`qsft/ReedSolomon.py:1-2`, coded delays). They are replaced by stub modules
that raise as soon as anything in them is used, so a case that reached them
would fail rather than record stub output. numpy 2 removed the `np.complex`
and `np.int` aliases the reference uses (`qsft/input_signal_subsampled.py:136`,
`synt_exp/synt_src/synthetic_signal.py:129`); they are restored as the Python
builtins they aliased. NEP 19 keeps the legacy RandomState streams unchanged,
so the draws are those of the reference's numpy 1.23.

quick_example builds a Reed-Solomon decoder with galois before drawing the
signal (`synt_exp/quick_example.py:24`); it is passed to QSFT only as
`source_decoder`, which the identity source never calls, and building it draws
nothing from np.random, so the case here passes source_decoder=None.
"""
import hashlib
import os
import sys
import time
import types

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

# (name, q, n, S, b, C, R, channel, noise_sd, a_min, a_max, max_weight, seed)
# noise_sd for the 10 / 20 dB shapes: noise_sd^2 = E[sum |a|^2] / 10^(snr/10)
# with E|a|^2 = (10^3 - 1) / (3 * 9) = 37 for |a| ~ U[1, 10].
CASES = {
    "k0": (4, 10, 100, 4, 3, 1, "identity", 0.0, 1, 10, None, 0),
    "nso_heavy_noise": (4, 20, 200, 4, 3, 2, "nso", 10.0, 1, 10, None, 2),
    "quick_example": (3, 40, 100, 4, 3, 1, "identity", 1.0, 1, 1, 4, 20),
    "k1_noiseless": (4, 50, 1000, 6, 3, 2, "nso", 0.0, 1, 10, None, 11),
    "k1_10db": (4, 50, 1000, 6, 3, 2, "nso", float(np.sqrt(37000.0 / 10.0)), 1, 10, None, 12),
    "k2_20db": (20, 16, 2000, 3, 3, 2, "nso", float(np.sqrt(74000.0 / 100.0)), 1, 10, None, 13),
}
PROJ_SEED = 20240601


def projection_vector(Q):
    """Fixed complex weights for the U projections (PCG64 stream, stable across numpy versions)."""
    g = np.random.default_rng(PROJ_SEED)
    return g.standard_normal(Q) + 1j * g.standard_normal(Q)


def _stub_module(name, **attrs):
    mod = types.ModuleType(name)
    mod.__dict__.update(attrs)

    def __getattr__(attr):
        raise AttributeError(f"stub module {name}: '{attr}' is not available (package not installed)")
    mod.__getattr__ = __getattr__
    sys.modules[name] = mod
    return mod


def _refuse(*_a, **_k):
    raise RuntimeError("stub of a package that is not installed was called")


class _Refuse:
    def __init__(self, *a, **k):
        _refuse()


def import_reference():
    _stub_module("group_lasso", GroupLasso=_Refuse)
    galois = _stub_module("galois", ReedSolomon=_Refuse, GF=_refuse)
    codes = _stub_module("galois._codes")
    rs = _stub_module("galois._codes._reed_solomon", decode_jit=_refuse)
    galois._codes = codes
    codes._reed_solomon = rs
    np.complex = complex
    np.int = int
    repo_root = os.path.dirname(os.path.dirname(HERE))
    sys.path[:] = [p for p in sys.path if os.path.abspath(p or ".") not in (repo_root, HERE)]
    sys.path.insert(0, REF)
    from qsft.qsft import QSFT
    from synt_exp.synt_src.synthetic_signal import get_random_subsampled_signal
    return QSFT, get_random_subsampled_signal


def rng_state():
    st = np.random.get_state()
    return np.asarray(st[1], dtype=np.uint32), int(st[2])


def run_case(name, QSFT, get_sig):
    q, n, S, b, C, R, channel, noise_sd, a_min, a_max, max_weight, seed = CASES[name]
    query_args = {"query_method": "complex", "num_subsample": C, "delays_method_source": "identity",
                  "subsampling_method": "qsft", "delays_method_channel": channel, "num_repeat": R, "b": b}
    qsft_args = {"num_subsample": C, "num_repeat": R, "reconstruct_method_source": "identity",
                 "reconstruct_method_channel": channel, "b": b, "noise_sd": noise_sd, "source_decoder": None}
    if name == "quick_example":
        query_args["t"] = 4
    t0 = time.time()
    np.random.seed(seed)
    kw = {} if max_weight is None else {"max_weight": max_weight}
    sig = get_sig(n=n, q=q, noise_sd=noise_sd, sparsity=S, a_min=a_min, a_max=a_max, query_args=query_args, **kw)
    mt_sig, pos_sig = rng_state()
    t1 = time.time()
    res = QSFT(**qsft_args).transform(sig, report=True, sort=True)
    t2 = time.time()
    mt_tr, pos_tr = rng_state()
    w = projection_vector(q ** b)
    proj = np.array([[np.asarray(sig.Us[i][j][b]) @ w for j in range(R)] for i in range(C)])
    gw = res["gwht"]
    keys = np.array([list(k) for k in gw], dtype=np.int8).reshape(len(gw), n)
    vals = np.array([complex(v) for v in gw.values()], dtype=np.complex128)
    out = dict(
        q=q, n=n, S=S, b=b, C=C, R=R, channel=channel, noise_sd=noise_sd, a_min=a_min, a_max=a_max,
        max_weight=-1 if max_weight is None else max_weight, seed=seed,
        locq=np.asarray(sig.locq, dtype=np.int8),
        signal_w_keys=np.array([list(k) for k in sig.signal_w], dtype=np.int8).reshape(-1, n),
        signal_w_vals=np.array([complex(v) for v in sig.signal_w.values()], dtype=np.complex128),
        Ms=np.array([np.asarray(M) for M in sig.Ms], dtype=np.int8),
        Ds=np.array([[np.asarray(D) for D in Dc] for Dc in sig.Ds], dtype=np.int8),
        mt_after_signal=mt_sig, pos_after_signal=pos_sig, mt_after_transform=mt_tr, pos_after_transform=pos_tr,
        proj_seed=PROJ_SEED, U_proj=proj,
        gwht_keys=keys, gwht_vals=vals,
        locations=np.asarray(res["locations"], dtype=np.int8).reshape(-1, n),
        n_samples=int(res["n_samples"]),
        avg_hamming_weight=float(res["avg_hamming_weight"]),
        max_hamming_weight=int(res["max_hamming_weight"]),
    )
    path = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(path, **out)
    digest = hashlib.sha256(open(path, "rb").read()).hexdigest()[:16]
    print(f"{name}: {len(gw)} keys, signal {t1 - t0:.1f} s, transform {t2 - t1:.1f} s -> {path} ({digest})",
          flush=True)


def main():
    if not os.path.isdir(REF):
        sys.exit(f"{REF} is not here: the fixtures are written in the build container only")
    QSFT, get_sig = import_reference()
    names = sys.argv[1:] or list(CASES)
    for name in names:
        run_case(name, QSFT, get_sig)


if __name__ == "__main__":
    main()
