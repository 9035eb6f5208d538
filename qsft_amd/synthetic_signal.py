"""Synthetic sparse-spectrum signals (synt_exp/synt_src/synthetic_signal.py).

`get_random_subsampled_signal` and `SyntheticSubsampledSignal` keep the
reference's names, arguments and global-RNG consumption order (support, then
strengths, then M/D inside the constructor, then per transform the group and
repeat permutation and the noise draws). Sampling runs on the device: the
coset queries are generated and evaluated in one kernel (int8 MFMA phase
contraction + w-table accumulation) instead of the reference's per-query
complex exponentials on a process pool.

Noise (`noise_mode`): "host" (default) draws the transform-domain noise with
numpy's legacy RNG in the reference's order, so a seed gives the reference's
noise exactly; "device" draws it with a Philox4x32-10 counter RNG in the
assemble kernel (same distribution, no host work, not the reference's
stream).
"""
import time

import numpy as np
import torch

from . import device as dv
from .input_signal import Signal
from .input_signal_subsampled import SubsampledSignal
from .utils import dec_to_qary_vec, qary_vec_to_dec, random_signal_strength_model, sort_qary_vecs


def generate_signal_w(n, q, sparsity, a_min, a_max, noise_sd=0, full=True, max_weight=None):
    """Sparse spectrum: support locq (n, S) lexsorted, strengths U[a_min, a_max] e^{i U[0, 2 pi)}
    (synthetic_signal.py:10-38). Support is not made distinct (as in the reference)."""
    max_weight = n if max_weight is None else max_weight
    N = q ** n
    if max_weight == n:
        locq = sort_qary_vecs(np.random.randint(q, size=(n, sparsity)).T).T
    else:
        vals = np.random.randint(q - 1, size=(max_weight, sparsity)) + 1
        pos = np.random.choice(a=n, size=(sparsity, max_weight))
        locq = np.zeros((n, sparsity), dtype=int)
        for s in range(sparsity):
            locq[pos[s, :], s] = vals[:, s]
        locq = sort_qary_vecs(locq.T).T
    loc = qary_vec_to_dec(locq, q)
    strengths = random_signal_strength_model(sparsity, a_min, a_max)
    if full:
        w = np.zeros((N,), dtype=complex)
        for idx, s in zip(loc, strengths):
            w[idx] = s
        w = w + np.random.normal(0, noise_sd, size=(N, 2)).view(complex).reshape(N)
        return np.reshape(w, [q] * n), locq, strengths
    return dict(zip(list(map(tuple, locq.T)), strengths)), locq, strengths


def get_random_subsampled_signal(n, q, noise_sd, sparsity, a_min, a_max, query_args, max_weight=None, **kwargs):
    """synthetic_signal.py:70-85. Extra kwargs (precision, noise_mode) pass to the signal."""
    start = time.time()
    signal_w, locq, strengths = generate_signal_w(n, q, sparsity, a_min, a_max, noise_sd, full=False,
                                                  max_weight=max_weight)
    if kwargs.pop("verbose", True):
        print(f"Generation Time:{time.time() - start}", flush=True)
    return SyntheticSubsampledSignal(signal_w=signal_w, locq=locq, strengths=strengths, noise_sd=noise_sd,
                                     n=n, q=q, query_args=query_args, **kwargs)


class SyntheticSignal(Signal):
    """Dense synthetic signal holder (synthetic_signal.py:59-67)."""

    def __init__(self, locq, strengths, **kwargs):
        super().__init__(**kwargs)
        self.locq = locq
        self.strengths = strengths


class SyntheticSubsampledSignal(SubsampledSignal):
    """SubsampledSignal whose samples are sum_s a_s w^{<m, k_s>} (synthetic_signal.py:88-130)."""

    def __init__(self, **kwargs):
        self.q = kwargs["q"]
        self.n = kwargs["n"]
        self.locq = np.asarray(kwargs["locq"])
        self.noise_sd = kwargs["noise_sd"]
        self.strengths = np.asarray(kwargs["strengths"])
        self.noise_mode = kwargs.get("noise_mode", "host")
        if self.noise_mode not in ("host", "device"):
            raise ValueError("noise_mode must be 'host' or 'device'")
        self._noise_seed = int(kwargs.get("noise_seed", 0))
        self._noise_calls = 0
        super().__init__(**kwargs)

    def _support_tensors(self):
        if getattr(self, "_K", None) is None:
            dev = dv.current_device()
            n_pad = dv.n_pad_of(self.n)
            self._K = dv.digits_tensor(self.locq.T, n_pad, dev)
            self._a = dv.strengths_tensor(self.strengths, self.precision, dev)
        return self._K, self._a

    def _device_samples(self):
        Kd, a = self._support_tensors()
        return dv.evaluate_cosets(self.q, self.n, self.b, self._Mt, self._Dall, Kd, a, self.precision)

    def subsample(self, query_indices):
        """Evaluate at decimal indices (the generic hook, synthetic_signal.py:108-118), on the device."""
        Kd, a = self._support_tensors()
        n_pad = dv.n_pad_of(self.n)
        dig = dec_to_qary_vec(query_indices, self.q, self.n).T
        m = dv.digits_tensor(dig, n_pad, dv.current_device())
        return dv.evaluate_points(self.q, self.n, m, Kd, a, self.precision).to(torch.complex128).cpu().numpy()

    def get_MDU(self, ret_num_subsample, ret_num_repeat, b, trans_times=False):
        """Base selection, then fresh noise on every (i, j) block (synthetic_signal.py:120-130)."""
        mdu = super().get_MDU(ret_num_subsample, ret_num_repeat, b, trans_times)
        nu = self.noise_sd / np.sqrt(2 * self.q ** b)
        for i in range(len(mdu[2])):
            for j in range(len(mdu[2][i])):
                shape = np.array(mdu[2][i][j]).shape
                mdu[2][i][j] = np.array(mdu[2][i][j]) + \
                    np.random.normal(0, nu, size=shape + (2,)).view(complex).reshape(shape)
        return mdu

    def _transform_noise(self, C, R, p1, b):
        Q = self.q ** b
        if self.noise_mode == "device":
            self._noise_calls += 1
            if self.noise_sd == 0:
                return None
            seed = (self._noise_seed * 0x9E3779B97F4A7C15 + self._noise_calls) & 0xFFFFFFFFFFFFFFFF
            return dict(sigma=float(self.noise_sd / np.sqrt(Q)), seed=seed)
        nu = self.noise_sd / np.sqrt(2 * Q)
        blocks = [np.random.normal(0, nu, size=(p1, Q, 2)).view(complex).reshape(p1, Q)
                  for _ in range(C) for _ in range(R)]
        if nu == 0:
            return None
        host = np.stack(blocks).reshape(C, R * p1, Q)
        return torch.from_numpy(host).to(self._dev)
