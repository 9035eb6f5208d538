"""QSFT: the peeling decoder with the reference's interface (qsft/qsft.py:13-281).

`QSFT(**kwargs).transform(signal, verbosity=0, report=False,
timing_verbose=False, **kwargs)` returns the {tuple(k): coefficient} dict
(or the report dict) of the reference. The working transform is assembled on
the device (selected groups/repeats + noise, FP64), and detection + peeling
run entirely on the device (no host round trip per iteration); only the
recovered (k, value) arrays come back to build the dict.

Keyword arguments beyond the reference: `host_loop=True` on the constructor
reads the stop flag after every pass (the plain single-instance path).
"""
import time

import numpy as np
import torch

from . import _lib as L
from . import device as dv
from . import kernels
from .input_signal_subsampled import SubsampledSignal
from .reed_solomon import ReedSolomon
from .utils import calc_hamming_weight, sort_qary_vecs

MAX_ITER = 15


class QSFT:
    def __init__(self, **kwargs):
        self.reconstruct_method_source = kwargs.get("reconstruct_method_source")
        self.reconstruct_method_channel = kwargs.get("reconstruct_method_channel")
        self.num_subsample = kwargs.get("num_subsample")
        self.num_repeat = kwargs.get("num_repeat")
        self.b = kwargs.get("b")
        self.source_decoder = kwargs.get("source_decoder", None)
        self.host_loop = bool(kwargs.get("host_loop", False))

    def _method(self):
        src = self.reconstruct_method_source
        if src not in (None, "identity", "coded"):
            raise NotImplementedError(f"reconstruct_method_source={src!r} (only 'identity' and 'coded')")
        ch = self.reconstruct_method_channel
        if ch == "nso":
            m = L.QSFT_DETECT_NSO1
        elif ch == "identity":
            m = L.QSFT_DETECT_NOISELESS
        else:
            raise NotImplementedError(f"reconstruct_method_channel={ch!r} (only 'nso' and 'identity')")
        return m | (L.QSFT_DETECT_CODED if src == "coded" else 0)

    def _code(self, n, q):
        """The Reed-Solomon code behind source_decoder (a get_reed_solomon_dec
        decoder, qsft/query.py:231-243) for coded source delays."""
        dec = self.source_decoder
        rs = getattr(dec, "__self__", None)
        if not isinstance(rs, ReedSolomon):
            raise ValueError("reconstruct_method_source='coded' needs source_decoder=get_reed_solomon_dec(n, t, q)")
        if rs.ns != n or rs.q != q:
            raise ValueError(f"source_decoder is for n={rs.ns}, q={rs.q}; the signal has n={n}, q={q}")
        return rs

    def _working_input(self, signal):
        """(Ms, Ds, U device complex128 bin-major [C][Q][P], Ts) for this call."""
        base_get = SubsampledSignal.get_MDU
        own = type(signal).get_MDU
        from .synthetic_signal import SyntheticSubsampledSignal
        if own is base_get or own is SyntheticSubsampledSignal.get_MDU:
            return signal._transform_input(self.num_subsample, self.num_repeat, self.b)
        # a user subclass overrides get_MDU with its own semantics: honour it
        Ms, Ds, Us, Ts = signal.get_MDU(self.num_subsample, self.num_repeat, self.b, trans_times=True)
        U = np.array([np.vstack(u) for u in Us]).astype(np.complex128)          # [C][P][Q]
        U = np.ascontiguousarray(U.transpose(0, 2, 1))                        # -> [C][Q][P]
        return Ms, Ds, torch.from_numpy(U).to(dv.current_device()), Ts

    def transform(self, signal, verbosity=0, report=False, timing_verbose=False, **kwargs):
        if not isinstance(signal, SubsampledSignal):
            raise NotImplementedError("QSFT currently only supports signals that inherit from SubsampledSignal")
        method = self._method()
        q, n, b = signal.q, signal.n, self.b
        Ms, Ds, U, Ts = self._working_input(signal)
        Ds = [np.vstack(D) for D in Ds]
        transform_time = float(np.sum(Ts))
        if timing_verbose:
            print(f"Transform Time:{transform_time}", flush=True)
        cutoff = 1e-9 + (1 + 0.5) * (signal.noise_sd ** 2) / (q ** b)
        cutoff = kwargs.get("cutoff", cutoff)
        if verbosity >= 2:
            print("cutoff = ", cutoff, flush=True)
        C, Q, P = U.shape
        dev = U.device
        n_pad = dv.n_pad_of(n)
        Mt = dv.Mt_tensor(Ms, b, n_pad, dev)
        Dd = dv.D_tensor(Ds, n_pad, dev)
        torch.cuda.synchronize()
        t0 = time.time()
        item = dict(q=q, n=n, b=b, C=C, P=P, p1=signal.get_source_parity(), method=method, cutoff=cutoff, Mt=Mt,
                    D=Dd, U=U)
        if method & L.QSFT_DETECT_CODED:
            rs = self._code(n, q)
            item["coded"] = dict(t=rs.t, s=rs.s, tables=torch.from_numpy(rs.device_tables()).to(dev))
        batch = kernels.PeelBatch([item], max_iter=MAX_ITER)
        # passes 3.. as one launch (kernels.PeelBatch.run): the same results, no launch per pass
        batch.run(host_loop=self.host_loop, fuse_from=None if self.host_loop else 3)
        k, vals, _, info = batch.results()[0]
        peeling_time = time.time() - t0
        if timing_verbose:
            print(f"Peeling Time:{peeling_time}", flush=True)
        if verbosity >= 2:
            for it, (ns, nm, nb) in enumerate(info["trace"]):
                print(f"iter {it + 1}: singletons {ns} multitons {nm} balls {nb}", flush=True)
        gwht = {tuple(row): v for row, v in zip(k, vals)}
        self.last_info = info
        if not report:
            return gwht
        loc = list(gwht.keys())
        if loc:
            if kwargs.get("sort", False):
                loc = sort_qary_vecs(loc)
            hw = calc_hamming_weight(loc)
            avg_hw, max_hw = np.mean(hw), np.max(hw)
        else:
            loc, avg_hw, max_hw = [], 0, 0
        return {"gwht": gwht, "runtime": transform_time + peeling_time, "n_samples": C * P * Q,
                "locations": loc, "avg_hamming_weight": avg_hw, "max_hamming_weight": max_hw}
