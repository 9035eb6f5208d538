"""SubsampledSignal: the reference's sampling front-end, with the numeric
stages on the GPU (qsft/input_signal_subsampled.py:12-269).

Constructor semantics are the reference's: parameters from kwargs, then M/D
generation (global legacy RNG, reference order), then sampling and the
q^b-point sub-transforms for every b in all_bs -- all in __init__.

Sampling has two routes:
  * a subclass that provides `_device_samples()` (the synthetic signal)
    evaluates every coset query on the device (fused stages (a)+(b));
  * any other subclass implements the reference's plugin hook
    `subsample(query_indices)`: the query digits come from the stage-(a)
    kernel, are turned into Python-int decimal indices on the host (the
    hook's contract) and the returned samples are uploaded.
Either way the sub-transforms run in the device DFT kernel and stay resident
in HBM; `Us` materialises them as the reference's nested lists on demand.

Extra keyword arguments (not in the reference): `precision` ("fp64" default,
or "fp32") selects the sample/sub-transform precision.
"""
import random
import time
from math import floor
from pathlib import Path

import numpy as np
import torch

from . import device as dv
from .input_signal import Signal
from .query import get_Ms_and_Ds
from .utils import qary_ints, qary_vec_to_dec


class SubsampledSignal(Signal):
    def _set_params(self, **kwargs):
        self.n = kwargs.get("n")
        self.q = kwargs.get("q")
        self.N = self.q ** self.n
        self.signal_w = kwargs.get("signal_w")
        self.query_args = kwargs.get("query_args")
        self.b = self.query_args.get("b")
        self.all_bs = self.query_args.get("all_bs", [self.b])
        self.num_subsample = self.query_args.get("num_subsample")
        if "num_repeat" not in self.query_args:
            self.query_args["num_repeat"] = 1
        self.num_repeat = self.query_args.get("num_repeat")
        self.subsampling_method = self.query_args.get("subsampling_method")
        self.delays_method_source = self.query_args.get("delays_method_source")
        self.delays_method_channel = self.query_args.get("delays_method_channel")
        self.L = None
        self.foldername = kwargs.get("folder")
        if not hasattr(self, "noise_sd"):
            self.noise_sd = kwargs.get("noise_sd", 0)
        self.precision = kwargs.get("precision", getattr(self, "precision", "fp64"))
        if self.precision not in dv.PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(dv.PRECISIONS)}")
        self._Us_host = None

    def _init_signal(self):
        if self.subsampling_method == "uniform":
            self._subsample_uniform()
        else:
            self._set_Ms_and_Ds_qsft()
            self._subsample_qsft()

    def _set_Ms_and_Ds_qsft(self):
        """M and D from query_args, or from the folder cache when it holds them
        (qsft/input_signal_subsampled.py:92-105: then no M/D draws are taken
        from the global RNG). The cache is this package's own format: numpy
        archives read with allow_pickle=False (the reference pickles; its
        cache files are not read)."""
        path = None
        if self.foldername:
            Path(self.foldername).mkdir(parents=True, exist_ok=True)
            path = Path(self.foldername) / "Ms_and_Ds.npz"
            if path.is_file():
                with np.load(path, allow_pickle=False) as z:
                    Ms, Ds = z["Ms"], z["Ds"]
                # reuse only a cache that covers this request: enough groups,
                # columns and repeats, and digits of this q
                fits = (Ms.ndim == 3 and Ds.ndim == 4 and Ms.shape[1] == self.n and Ds.shape[3] == self.n
                        and Ms.shape[0] >= self.num_subsample and Ms.shape[2] >= self.b
                        and Ds.shape[0] == Ms.shape[0] and Ds.shape[1] >= self.num_repeat
                        and Ms.size > 0 and Ds.size > 0
                        and 0 <= Ms.min() and Ms.max() < self.q and 0 <= Ds.min() and Ds.max() < self.q)
                if fits:
                    self.Ms = [Ms[i] for i in range(Ms.shape[0])]
                    self.Ds = [[Ds[i, j] for j in range(Ds.shape[1])] for i in range(Ds.shape[0])]
                    return
        self.Ms, self.Ds = get_Ms_and_Ds(self.n, self.q, **self.query_args)
        if path is not None:
            np.savez_compressed(path, Ms=np.stack([np.asarray(m) for m in self.Ms]),
                                Ds=np.stack([np.stack([np.asarray(d) for d in row]) for row in self.Ds]))

    # ------------------------------------------------------------ sampling

    def _device_samples(self):
        """Samples [C][R*p1][q^b] computed on the device, or None to use subsample()."""
        return None

    def _layout(self):
        C = len(self.Ms)
        R = len(self.Ds[0])
        p1 = self.Ds[0][0].shape[0]
        return C, R, p1

    def _upload_queries(self):
        dev = dv.current_device()
        n_pad = dv.n_pad_of(self.n)
        C, R, p1 = self._layout()
        self._dev = dev
        self._n_pad = n_pad
        self._Mt = dv.Mt_tensor(self.Ms, self.b, n_pad, dev)
        self._Dall = dv.D_tensor([np.vstack(self.Ds[i]) for i in range(C)], n_pad, dev)

    def _subsample_qsft(self):
        """input_signal_subsampled.py:107-155: samples, then sub-DFTs per b in all_bs.
        With a folder, the transforms (every (group, repeat) and every b in
        all_bs) are loaded when all are cached, else the samples when all are
        cached; whatever is computed is written back (samples/M{i}_D{j}.npz,
        transforms/U{i}_{j}.npz)."""
        self._upload_queries()
        q, b = self.q, self.b
        C, R, p1 = self._layout()
        folder = Path(self.foldername) if self.foldername else None
        if folder is not None:
            (folder / "samples").mkdir(parents=True, exist_ok=True)
            (folder / "transforms").mkdir(parents=True, exist_ok=True)
            if self._load_cached_transforms(folder):
                return
        x = self._load_cached_samples(folder) if folder is not None else None
        if x is None:
            x = self._device_samples()
            if x is None:
                x = self._callback_samples()
            if folder is not None:
                h = x.cpu().numpy()
                for i in range(C):
                    for j in range(R):
                        np.savez_compressed(folder / "samples" / f"M{i}_D{j}.npz",
                                            samples=h[i, j * p1:(j + 1) * p1])
        rows = C * R * p1
        xr = x.reshape(rows, q ** b)
        self._U = {}
        self.transformTimes = [[{} for _ in range(R)] for _ in range(C)]
        for bb in self.all_bs:
            torch.cuda.synchronize()
            t0 = time.time()
            U = dv.subtransform(xr, q, bb, in_stride=q ** (b - bb), precision=self.precision)
            torch.cuda.synchronize()
            per = (time.time() - t0) / (C * R)
            self._U[bb] = U.reshape(C, R * p1, q ** bb)
            for i in range(C):
                for j in range(R):
                    self.transformTimes[i][j][bb] = per
        self._Us_host = None
        if folder is not None:
            hs = {bb: self._U[bb].cpu().numpy() for bb in self.all_bs}
            for i in range(C):
                for j in range(R):
                    arrs = {f"U{bb}": hs[bb][i, j * p1:(j + 1) * p1] for bb in self.all_bs}
                    arrs.update({f"t{bb}": np.float64(self.transformTimes[i][j][bb]) for bb in self.all_bs})
                    np.savez_compressed(folder / "transforms" / f"U{i}_{j}.npz", **arrs)

    def _load_cached_samples(self, folder):
        """Samples [C][R*p1][q^b] from the folder when every (group, repeat) is there, else None."""
        C, R, p1 = self._layout()
        files = [[folder / "samples" / f"M{i}_D{j}.npz" for j in range(R)] for i in range(C)]
        if not all(f.is_file() for row in files for f in row):
            return None
        h = np.zeros((C, R * p1, self.q ** self.b), dtype=np.complex128)
        for i in range(C):
            for j in range(R):
                with np.load(files[i][j], allow_pickle=False) as z:
                    if "samples" not in z.files or z["samples"].shape != (p1, self.q ** self.b):
                        return None                                # written for other sizes: recompute
                    h[i, j * p1:(j + 1) * p1] = z["samples"]
        return torch.from_numpy(h).to(self._dev).to(dv.PRECISIONS[self.precision][1])

    def _load_cached_transforms(self, folder):
        """Sub-DFTs of every b in all_bs and their times from the folder; False unless all are there."""
        C, R, p1 = self._layout()
        files = [[folder / "transforms" / f"U{i}_{j}.npz" for j in range(R)] for i in range(C)]
        if not all(f.is_file() for row in files for f in row):
            return False
        hs = {bb: np.zeros((C, R * p1, self.q ** bb), dtype=np.complex128) for bb in self.all_bs}
        times = [[{} for _ in range(R)] for _ in range(C)]
        for i in range(C):
            for j in range(R):
                with np.load(files[i][j], allow_pickle=False) as z:
                    if not all(f"U{bb}" in z.files and z[f"U{bb}"].shape == (p1, self.q ** bb)
                               for bb in self.all_bs):
                        return False                               # other b or sizes: recompute
                    for bb in self.all_bs:
                        hs[bb][i, j * p1:(j + 1) * p1] = z[f"U{bb}"]
                        times[i][j][bb] = float(z[f"t{bb}"])
        dt = dv.PRECISIONS[self.precision][1]
        self._U = {bb: torch.from_numpy(h).to(self._dev).to(dt) for bb, h in hs.items()}
        self.transformTimes = times
        self._Us_host = None
        return True

    def _callback_samples(self):
        """Generic plugin route: decimal query indices -> self.subsample()."""
        q, b, n = self.q, self.b, self.n
        C, R, p1 = self._layout()
        Q = q ** b
        dt = dv.PRECISIONS[self.precision][1]
        digits = dv.query_digits(q, n, b, self._Mt, self._Dall).cpu().numpy()[..., :n]
        samples = np.zeros((C, R * p1, Q), dtype=np.complex128)
        for i in range(C):
            for j in range(R):
                idx = [qary_vec_to_dec(digits[i, j * p1 + r].T, q) for r in range(p1)]
                blk = samples[i, j * p1:(j + 1) * p1]
                if Q > 10000:
                    for r in range(p1):
                        blk[r] = np.asarray(self.subsample(idx[r]))
                else:
                    allv = np.asarray(self.subsample(np.concatenate(idx)))
                    for r in range(p1):
                        blk[r] = allv[r * Q:(r + 1) * Q]
        return torch.from_numpy(samples).to(self._dev).to(dt)

    def _subsample_uniform(self):
        """n_samples uniform random queries (input_signal_subsampled.py:157-173, 208-223)."""
        idx = self._get_random_query_indices(self.query_args["n_samples"])
        vals = self.subsample(idx)
        self.signal_t = dict(zip(idx, vals))

    def get_all_qary_vectors(self):
        if self.L is None:
            self.L = np.array(qary_ints(self.b, self.q))
        return self.L

    def subsample(self, query_indices):
        raise NotImplementedError

    def _get_qsft_query_indices(self, M, D_sub):
        """Decimal indices of {M l + d_i} for each delay row d_i (input_signal_subsampled.py:183-206)."""
        n_pad = dv.n_pad_of(self.n)
        dev = dv.current_device()
        b = M.shape[1]
        Mt = dv.Mt_tensor([M], b, n_pad, dev)
        D = dv.D_tensor([np.asarray(D_sub)], n_pad, dev)
        dig = dv.query_digits(self.q, self.n, b, Mt, D).cpu().numpy()[0, :, :, :self.n]
        return [qary_vec_to_dec(dig[r].T, self.q) for r in range(dig.shape[0])]

    def _get_random_query_indices(self, n_samples):
        return [floor(random.uniform(0, 1) * self.N) for _ in range(n_samples)]

    # ------------------------------------------------------ decoder access

    @property
    def Us(self):
        """Sub-transforms as the reference's nested lists Us[i][j][b] = list of rows."""
        if self._Us_host is None:
            C, R, p1 = self._layout()
            self._Us_host = [[{} for _ in range(R)] for _ in range(C)]
            for bb, U in self._U.items():
                h = U.to(torch.complex128).cpu().numpy()
                for i in range(C):
                    for j in range(R):
                        self._Us_host[i][j][bb] = list(h[i, j * p1:(j + 1) * p1])
        return self._Us_host

    def _select(self, ret_num_subsample, ret_num_repeat, b):
        if not (ret_num_subsample <= self.num_subsample and ret_num_repeat <= self.num_repeat and b <= self.b):
            raise ValueError("There are not enough Ms or Ds.")
        gsel = np.random.choice(self.num_subsample, ret_num_subsample, replace=False)
        rsel = np.random.choice(self.num_repeat, ret_num_repeat, replace=False)
        return gsel, rsel

    def get_MDU(self, ret_num_subsample, ret_num_repeat, b, trans_times=False):
        """(Ms, Ds, Us[, Ts]) as numpy lists (input_signal_subsampled.py:225-262)."""
        gsel, rsel = self._select(ret_num_subsample, ret_num_repeat, b)
        Us_all = self.Us
        Ms = [self.Ms[i][:, :b] for i in gsel]
        Ds = [[self.Ds[i][j] for j in rsel] for i in gsel]
        Us = [[Us_all[i][j][b] for j in rsel] for i in gsel]
        Ts = [[self.transformTimes[i][j][b] for j in rsel] for i in gsel]
        return (Ms, Ds, Us, Ts) if trans_times else (Ms, Ds, Us)

    def _transform_noise(self, C, R, p1, b):
        """Noise for the working transform: None here (no noise model)."""
        return None

    def _transform_input(self, ret_num_subsample, ret_num_repeat, b):
        """Device counterpart of get_MDU used by QSFT.transform: (Ms, Ds, U, Ts) with U
        the working transform, complex128 bin-major [C][Q][P] on the device. The
        stored sub-transforms are not modified."""
        gsel, rsel = self._select(ret_num_subsample, ret_num_repeat, b)
        C, R = len(gsel), len(rsel)
        p1 = self.Ds[0][0].shape[0]
        Q = self.q ** b
        Ms = [self.Ms[i][:, :b] for i in gsel]
        Ds = [[self.Ds[i][j] for j in rsel] for i in gsel]
        Ts = [[self.transformTimes[i][j][b] for j in rsel] for i in gsel]
        U = torch.empty((C, Q, R * p1), dtype=torch.complex128, device=self._dev)
        noise = self._transform_noise(C, R, p1, b)
        item = dict(Q=Q, C=C, R=R, p1=p1, src=self._U[b], src_rows=self._U[b].shape[1], gsel=gsel, rsel=rsel,
                    U=U, bin_major=True)
        if isinstance(noise, torch.Tensor):
            item["noise"] = noise
        elif isinstance(noise, dict):
            item.update(noise)
        from . import kernels
        kernels.assemble([item])
        return Ms, Ds, U, Ts

    def get_source_parity(self):
        return self.Ds[0][0].shape[0]
