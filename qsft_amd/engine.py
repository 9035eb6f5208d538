"""Batched q-SFT engine: many synthetic transforms per launch, state resident in HBM.

A `Batch` uploads its instances once (M_c, D, support digits, strengths),
allocates samples, sub-transforms, the FP64 working transform and the peel
workspace per instance, builds every stage's descriptors once, and then
`launch()` enqueues the whole pipeline with no host work between stages:

    stage (a)+(b)  eval      x[c][p][l] = sum_s a_s w^{<M_c l + d_p, k_s>}   (FP32 or FP64)
    stage (c)      DFT       in place, x -> q^-b fftn over [q]*b
    get_MDU step   assemble  FP64 working copy + transform-domain noise (Philox)
    stage (d)      peel      device loop (no host sync), finalize, pack

`results()` is the only synchronisation: one small header copy and one copy
of all recovered rows. Uneven instances (different q, n, b, S, C, R) share
the same launches.
"""
import collections
import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib as L
from . import device as dv
from . import kernels as K



def _addr(x):
    """Address of a C-contiguous array's data (ctypes.from_buffer is the
    cheaper route for the writable arrays wire() builds)."""
    try:
        return ctypes.addressof(ctypes.c_char.from_buffer(x))
    except (TypeError, ValueError):
        return int(x.ctypes.data)

@dataclass
class TransformSpec:
    """One transform request on a synthetic sparse signal."""
    q: int
    n: int
    b: int
    C: int
    R: int
    locq: np.ndarray                  # (n, S) support digits (int8 from random_spec)
    strengths: np.ndarray             # (S,) complex
    Ms: list                          # C arrays (n, b)
    D: np.ndarray                     # (R * (n + 1), n) delay rows shared by all groups
    noise_sd: float = 0.0
    method: str = "nso"               # "nso" (nso1) or "identity" (noiseless)
    seed: int = 0                     # Philox noise seed
    cutoff: float = None              # default 1e-9 + 1.5 noise_sd^2 / q^b, as the reference
    tag: object = field(default=None)

    @property
    def P(self):
        return self.D.shape[0]

    @property
    def Q(self):
        return self.q ** self.b

    @property
    def S(self):
        return self.locq.shape[1]

    def threshold(self):
        if self.cutoff is not None:
            return self.cutoff
        return 1e-9 + 1.5 * self.noise_sd ** 2 / self.q ** self.b

    def wire(self):
        """The request's payload as the serving runtime reads it, built once:
        M int8 [C][n][b], D int8 [P][n], locq int8 [n][S], a complex128 [S]."""
        w = getattr(self, "_wire", None)
        if w is None:
            M = np.ascontiguousarray(np.stack([np.asarray(m)[:, :self.b] for m in self.Ms]), dtype=np.int8)
            w = (M, np.ascontiguousarray(self.D, dtype=np.int8), np.ascontiguousarray(self.locq, dtype=np.int8),
                 np.ascontiguousarray(self.strengths, dtype=np.complex128))
            self._wire = w
        return w

    def wire_ptrs(self):
        """Addresses of the four wire() arrays (kept alive by the spec)."""
        p = getattr(self, "_wire_ptrs", None)
        if p is None:
            p = tuple(_addr(x) for x in self.wire())
            self._wire_ptrs = p
        return p

    def footprint_bytes(self, precision="fp32"):
        """HBM bytes this transform holds while resident (samples/transform, FP64
        working copy, peel workspace, inputs)."""
        el = 8 if precision == "fp32" else 16
        cells = self.C * self.P * self.Q
        ws = L.lib.qsft_peel_workspace_bytes(self.q, self.n, self.b, self.C, self.P, 15, 0)
        n_pad = dv.n_pad_of(self.n)
        return cells * (el + 16) + ws + (self.S + self.C * (self.b + self.P)) * n_pad + self.S * el + self.C * self.Q * 8


def random_spec(q, n, S, b, C=3, R=2, snr_db=10.0, rng=None, seed=0, distinct=True, method="nso"):
    """A synthetic request: uniform support (distinct unless told otherwise),
    |a| ~ U[1, 10], phase ~ U[0, 2 pi), uniform M_c, NSO delays with R random
    offsets, noise_sd^2 = sum |a|^2 / 10^(snr/10) (None -> noiseless)."""
    rng = np.random.default_rng(seed) if rng is None else rng
    locq = rng.integers(0, q, size=(n, S)).astype(np.int8)        # int8 digits, the same stream as int64 draws
    if distinct:
        while True:
            keys = {tuple(c) for c in locq.T}
            if len(keys) == S:
                break
            _, first = np.unique(locq.T, axis=0, return_index=True)
            dup = np.setdiff1d(np.arange(S), first)
            locq[:, dup] = rng.integers(0, q, size=(n, len(dup)))
    a = rng.uniform(1, 10, S) * np.exp(2j * np.pi * rng.uniform(size=S))
    Ms = [rng.integers(0, q, size=(n, b)) for _ in range(C)]
    src = np.vstack([np.zeros(n), np.eye(n)]).astype(int)
    if method == "nso":
        D = np.vstack([(rng.integers(0, q, n) - src) % q for _ in range(R)])
    else:
        R = 1
        D = src
    noise_sd = 0.0 if snr_db is None else float(np.sqrt(np.sum(np.abs(a) ** 2) / 10 ** (snr_db / 10)))
    return TransformSpec(q=q, n=n, b=b, C=C, R=R, locq=locq, strengths=a, Ms=Ms, D=D, noise_sd=noise_sd,
                         method=method, seed=int(rng.integers(0, 2 ** 62)))


class Batch:
    """eval_mode: "mfma" (matrix-core GEMM form, FP32 only; the default for
    FP32) or "direct" (per-query int8 MFMA phase + VALU w-table accumulation,
    FP32 or FP64; the plain path). fused_screen: the assembler also writes each
    bin's first-pass energy (transforms with P <= 128), so the peeler's
    zeroton screen does not read U again. sparse_u (with those energies and
    FP32 samples): the assembler writes only the columns above the zeroton
    cutoff; the peeler rebuilds another column from the samples and the same
    noise when a ball first lands on it (about a fifth of U is written at the
    K1 shape). one_bin_waves: the first detection pass one bin per wave (the
    default classifies four bins per wave where it can; same results)."""

    def __init__(self, specs, precision="fp32", max_iter=15, device=None, eval_mode=None, fused_screen=True,
                 host_rows=True, tail_stream=True, own_stream=False, sparse_u=True, fuse_from=3,
                 one_bin_waves=False):
        self.specs = list(specs)
        self.precision = precision
        if eval_mode is None:
            eval_mode = "mfma" if precision == "fp32" else "direct"
        if eval_mode == "mfma" and precision != "fp32":
            raise ValueError("the matrix-core evaluator produces FP32 samples")
        self.eval_mode = eval_mode
        code, dt = dv.PRECISIONS[precision]
        dev = device if device is not None else dv.current_device()
        self.device = dev
        ev, df, asm, pl = [], [], [], []
        self.bufs = []
        for sp in self.specs:
            n_pad = dv.n_pad_of(sp.n)
            C, P, Q = sp.C, sp.P, sp.Q
            Mt = dv.Mt_tensor(sp.Ms, sp.b, n_pad, dev)
            Dd = dv.D_tensor([sp.D] * C, n_pad, dev)
            Kd = dv.digits_tensor(sp.locq.T, n_pad, dev)
            a = dv.strengths_tensor(sp.strengths, "fp64" if eval_mode == "mfma" else precision, dev)
            x = torch.empty((C, P, Q), dtype=dt, device=dev)
            U = torch.empty((C, Q, P), dtype=torch.complex128, device=dev)      # bin-major for the peeler
            # first-pass column energies come from the assembler when a bin's rows fit one tile
            energy = torch.empty((C, Q), dtype=torch.float64, device=dev) if (fused_screen and P <= 128) else None
            self.bufs.append(dict(Mt=Mt, D=Dd, K=Kd, a=a, x=x, U=U, energy=energy))
            ev.append(dict(q=sp.q, n=sp.n, b=sp.b, C=C, P=P, S=sp.S, Mt=Mt, D=Dd, K=Kd, a=a, a64=a, x=x,
                           shared_d=True, amax=float(np.max(np.abs(sp.strengths))) if sp.S else 0.0))
            df.append(dict(q=sp.q, b=sp.b, rows=C * P, inp=x, out=x))
            item = dict(Q=Q, C=C, R=P // (sp.n + 1), p1=sp.n + 1, src=x, src_rows=P,
                        gsel=list(range(C)), rsel=list(range(P // (sp.n + 1))), U=U, bin_major=True,
                        energy=energy)
            sigma = float(sp.noise_sd / np.sqrt(Q)) if sp.noise_sd > 0 else 0.0
            if sigma > 0:
                item.update(sigma=sigma, seed=sp.seed)
            asm.append(item)
            pl.append(dict(q=sp.q, n=sp.n, b=sp.b, C=C, P=P, p1=sp.n + 1,
                           method=L.QSFT_DETECT_NSO1 if sp.method == "nso" else L.QSFT_DETECT_NOISELESS,
                           cutoff=sp.threshold(), Mt=Mt, D=Dd, U=U, shared_d=True, energy=energy,
                           one_bin_waves=one_bin_waves))
            if sparse_u and energy is not None and precision == "fp32":
                pl[-1]["sparse"] = dict(x=x, sigma=sigma)
        self.eval = K.EvalMfma(ev) if eval_mode == "mfma" else K.EvalCosets(ev, code)
        self.dft = K.Dft(df, code)
        # packed rows: at most a few per planted support entry (noise adds some), kept
        # in pinned host memory that the pack kernels write directly
        cap = sum(4 * sp.S + 1024 for sp in self.specs) if host_rows else None
        self.peel = K.PeelBatch(pl, max_iter=max_iter, host_rows=host_rows, cap_rows=cap, final_u=False)
        for i, it in enumerate(pl):
            if "sparse" in it:
                written, key = self.peel.sparse_targets(i)
                asm[i].update(written=written, cutoff=it["cutoff"], seed_out=key)
        self.asm_items = asm
        self.asm = K.Assemble(asm)
        # the peeling tail (last queued balls, output table, pack into host
        # memory) runs on a second stream, overlapping the next batch's work;
        # the next launch of this batch waits for it
        # With own_stream the whole pipeline runs on a stream of this batch, so
        # the work of two resident batches interleaves on the GPU (the tail
        # then stays on that stream).
        self.stream = torch.cuda.Stream(device=dev) if own_stream else None
        self.tail = torch.cuda.Stream(device=dev) if (host_rows and tail_stream and not own_stream) else None
        self.tail_event = torch.cuda.Event() if self.tail is not None else None
        self.done = None
        # Passes launched per batch: one more than most recent launches ran (the
        # 80th percentile of the last 16). Passes after the loop stops only cost
        # launches; a launch that falls short (an instance whose loop runs to
        # the pass cap) is finished in fetch(), as the packed header says.
        self.max_iter = max_iter
        # Passes from fuse_from on run as one kernel (one workgroup per
        # instance, looping over the passes): a loop that stops costs nothing
        # after that, and one that runs to the cap pays no launch per pass.
        # None: one launch pair per pass, the count adapted as below.
        self.fuse_from = fuse_from
        self.adaptive = self.tail is not None or self.stream is not None
        self.passes = max_iter
        self.launched = max_iter
        self._ran = collections.deque(maxlen=16)

    def launch(self, seed_offset=0, host_loop=False, events=None, marks=None):
        """Enqueue the whole pipeline on the batch's stream (own_stream) or the
        current one. With `events` (a dict), new HIP events bracket each stage
        for timing; `marks` maps stage names ("start", "eval_prep", "eval",
        "dft", "assemble", "peel") to pre-created events to record instead (no
        event creation per launch)."""
        if self.stream is not None:
            with torch.cuda.stream(self.stream):
                self._enqueue(seed_offset, host_loop, events, marks)
        else:
            self._enqueue(seed_offset, host_loop, events, marks)

    def _enqueue(self, seed_offset, host_loop, events, marks):
        def mark(name):
            if marks is not None and name in marks:
                marks[name].record()
            if events is not None:
                e = torch.cuda.Event(enable_timing=True)
                e.record()
                events.setdefault(name, []).append(e)
        if self.done is not None and self.tail is not None:
            torch.cuda.current_stream().wait_event(self.done)      # the previous tail still reads the tables
        mark("start")
        if (events is not None or marks is not None) and hasattr(self.eval, "launch_split"):
            self.eval.launch_split(lambda: mark("eval_prep"))
        else:
            self.eval.launch()
        mark("eval")
        self.dft.launch()
        mark("dft")
        self.asm.launch(seed_offset)
        mark("assemble")
        if self.fuse_from is not None and not host_loop:
            self.launched = self.max_iter
            self.peel.run(tail_stream=self.tail, tail_event=self.tail_event, fuse_from=self.fuse_from)
        else:
            self.launched = self.passes if (self.adaptive and not host_loop) else self.max_iter
            self.peel.run(host_loop=host_loop, tail_stream=self.tail, tail_event=self.tail_event, passes=self.launched)
        # one completion event per batch, re-recorded every launch (a copy that
        # waits on it captured the previous record when it was enqueued)
        if self.done is None:
            self.done = torch.cuda.Event()
        if self.tail is not None:
            with torch.cuda.stream(self.tail):
                mark("peel")
                self.done.record()
        else:
            mark("peel")
            self.done.record()

    def results(self, with_info=False):
        return self.peel.results(with_info=with_info, packed=self.fetch())

    def fetch(self, copy_stream=None):
        """Packed rows of the last launch once it is done: (counts, offsets, rows).
        Work queued after it keeps running. With host_rows the rows are already
        in pinned host memory; otherwise they are copied on `copy_stream`."""
        packed = self.peel.fetch_packed(copy_stream=copy_stream, ready=self.done)
        if self.peel.last_unfinished and self.launched < self.max_iter:
            # a loop wanted more passes than were launched: run them, then
            # finalize and pack again
            stream = self.stream if self.stream is not None else torch.cuda.current_stream()
            with torch.cuda.stream(stream):
                self.peel.resume(self.launched, tail_stream=self.tail, tail_event=self.tail_event)
                if self.tail is not None:
                    with torch.cuda.stream(self.tail):
                        self.done.record()
                else:
                    self.done.record()
            self.launched = self.max_iter
            packed = self.peel.fetch_packed(copy_stream=copy_stream, ready=self.done)
        if self.adaptive:
            self._ran.append(self.peel.last_iters)
            self.passes = min(self.max_iter, int(np.percentile(list(self._ran), 80)) + 1)
        return packed

    def unpack(self, packed, with_info=False):
        return self.peel.results(with_info=with_info, packed=packed)


def nmse(k, vals, spec):
    """NMSE of recovered (k, vals) against the planted spectrum (the FP64 oracle
    when q^n is too large for a dense transform)."""
    truth = {tuple(c): v for c, v in zip(spec.locq.T, spec.strengths)}
    diff = dict(truth)
    for row, v in zip(k, vals):
        t = tuple(row)
        diff[t] = diff.get(t, 0) - v
    num = sum(abs(v) ** 2 for v in diff.values())
    den = sum(abs(v) ** 2 for v in truth.values())
    return float(num / den)
