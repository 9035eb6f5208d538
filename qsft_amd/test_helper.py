"""TestHelper: the reference's experiment helper (qsft/test_helper.py:9-296)
with the held-out evaluator on the GPU.

A helper loads a training signal (q-SFT queries) and a held-out test signal
(uniform random queries, noiseless), decodes the training signal with
`compute_model` and scores a recovered spectrum with `test_model`: the NMSE of
its prediction sum_k beta_k w^<m, k> at every held-out query m
(qsft/test_helper.py:235-260). The prediction is a dense (queries x recovered
support) contraction and runs in the qsft_eval_points kernel, FP64, with an
exact w table (the reference exponentiates the unreduced integer phase; the
two agree to ~1e-13 relative).

Only the q-SFT method with identity source delays is accelerated; "gwht"
(dense transform of the training data), "qsft_coded", "qsft_binary" and
"lasso" raise NotImplementedError, as the LASSO baseline, coded delays and the
binary expansion are outside this engine.
"""
import json

import numpy as np
import torch

from . import device as dv
from .qsft import QSFT
from .utils import NpEncoder, dec_to_qary_vec


def heldout_prediction(sample_idx_dec, beta, q, n):
    """y_hat[i] = sum_k beta[k] w^<m_i, k> for the decimal query indices
    (qsft/test_helper.py:245-255), evaluated on the GPU in FP64."""
    keys = np.array(list(beta.keys()), dtype=np.int64).reshape(len(beta), n)
    vals = np.array(list(beta.values()), dtype=np.complex128)
    dev = dv.current_device()
    n_pad = dv.n_pad_of(n)
    Kd = dv.digits_tensor(keys, n_pad, dev)
    a = dv.strengths_tensor(vals, "fp64", dev)
    out = []
    idx = list(sample_idx_dec)
    chunk = 1 << 20                                   # queries per launch (digit staging on the host)
    for i in range(0, len(idx), chunk):
        m = dv.digits_tensor(dec_to_qary_vec(idx[i:i + chunk], q, n).T, n_pad, dev)
        out.append(dv.evaluate_points(q, n, m, Kd, a, "fp64").cpu().numpy())
    return np.concatenate(out) if out else np.zeros(0, np.complex128)


class TestHelper:
    __test__ = False                                  # not a pytest test class

    def __init__(self, signal_args, methods, subsampling_args, test_args, exp_dir, subsampling=True):
        self.n = signal_args["n"]
        self.q = signal_args["q"]
        self.exp_dir = exp_dir
        self.subsampling = subsampling
        config_path = self.exp_dir / "config.json"
        if not config_path.is_file():
            with open(config_path, "w") as f:
                json.dump({"query_args": subsampling_args}, f, cls=NpEncoder)
        self.signal_args = signal_args
        self.subsampling_args = subsampling_args
        self.test_args = test_args
        if self.subsampling:
            if "qsft" in methods:
                self.train_signal = self.load_train_data()
            if "qsft_binary" in methods:
                self.train_signal_binary = self.load_train_data_binary()
            if "lasso" in methods:
                self.train_signal_uniform = self.load_train_data_uniform()
            if "qsft_coded" in methods:
                self.train_signal_coded = self.load_train_data_coded()
            self.test_signal = self.load_test_data()
        else:
            self.train_signal = self.load_full_data()
            self.test_signal = self.train_signal
            if any(m.startswith("binary") for m in methods):
                raise NotImplementedError

    def generate_signal(self, signal_args):
        raise NotImplementedError

    def load_train_data(self):
        """q-SFT queries, identity source delays, NSO channel (qsft/test_helper.py:59-70)."""
        signal_args = self.signal_args.copy()
        query_args = self.subsampling_args.copy()
        query_args.update({"subsampling_method": "qsft", "query_method": "complex",
                           "delays_method_source": "identity", "delays_method_channel": "nso"})
        signal_args["folder"] = self.exp_dir / "train"
        signal_args["query_args"] = query_args
        return self.generate_signal(signal_args)

    def load_train_data_coded(self):
        raise NotImplementedError("coded (Reed-Solomon) source delays are outside this engine")

    def load_train_data_binary(self):
        return None

    def load_train_data_uniform(self):
        """n_samples = C q^b R (n + 1) uniform queries (qsft/test_helper.py:98-106)."""
        signal_args = self.signal_args.copy()
        qa = self.subsampling_args
        n_samples = qa["num_subsample"] * (signal_args["q"] ** qa["b"]) * qa["num_repeat"] * (signal_args["n"] + 1)
        signal_args["folder"] = self.exp_dir / "train_uniform"
        signal_args["query_args"] = {"subsampling_method": "uniform", "n_samples": n_samples}
        return self.generate_signal(signal_args)

    def load_test_data(self):
        """Noiseless uniform held-out queries (qsft/test_helper.py:108-114)."""
        signal_args = self.signal_args.copy()
        (self.exp_dir / "test").mkdir(exist_ok=True)
        signal_args["query_args"] = {"subsampling_method": "uniform", "n_samples": self.test_args.get("n_samples")}
        signal_args["folder"] = self.exp_dir / "test"
        signal_args["noise_sd"] = 0
        return self.generate_signal(signal_args)

    def load_full_data(self):
        return None

    def compute_model(self, method, model_kwargs, report=False, verbosity=0):
        if method == "qsft":
            return self._calculate_qsft(model_kwargs, report, verbosity)
        if method in ("gwht", "qsft_binary", "qsft_coded", "lasso"):
            raise NotImplementedError(f"method {method!r} is outside this engine")
        raise NotImplementedError()

    def test_model(self, method, **kwargs):
        if method in ("qsft", "qsft_coded", "lasso"):
            return self._test_qary(**kwargs)
        raise NotImplementedError()

    def _calculate_qsft(self, model_kwargs, report=False, verbosity=0):
        """QSFT with identity source and NSO channel (qsft/test_helper.py:147-166)."""
        qsft = QSFT(reconstruct_method_source="identity", reconstruct_method_channel="nso",
                    num_subsample=model_kwargs["num_subsample"], num_repeat=model_kwargs["num_repeat"],
                    b=model_kwargs["b"])
        self.train_signal.noise_sd = model_kwargs["noise_sd"]
        return qsft.transform(self.train_signal, verbosity=verbosity, timing_verbose=(verbosity >= 1),
                              report=report)

    def _test_qary(self, beta):
        """Held-out NMSE ||y_hat - y||^2 / ||y||^2 (qsft/test_helper.py:235-260); 1 for an empty beta."""
        if len(beta.keys()) == 0:
            return 1
        test_signal = self.test_signal.signal_t
        idx, samples = list(test_signal.keys()), np.asarray(list(test_signal.values()))
        y_hat = heldout_prediction(idx, beta, self.q, self.n)
        return np.linalg.norm(y_hat - samples) ** 2 / np.linalg.norm(samples) ** 2
