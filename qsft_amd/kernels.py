"""Batched launchers over the C ABI (include/qsft_amd.h).

Each function takes a list of per-instance argument dicts holding torch CUDA
tensors, builds the descriptor array and block-prefix array(s) on the host,
uploads them in one buffer and launches on torch's current stream. Instances
of one call may have different shapes; they are split into as few launches
as the kernels' compile-time variants require (q's power-of-two class and
digit-block count for the evaluator, q for the DFT radix).
"""
import ctypes
from collections import defaultdict

import numpy as np
import torch

from . import _lib as L

lib = L.lib


def stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def pack(descs, prefix_lists, device):
    """One device buffer: the descriptor array, then each int32 block prefix.

    Returns (buffer tensor, descriptor pointer, [prefix pointers], [totals]).
    The buffer must outlive the launches that read it; stream-ordered reuse
    by torch's caching allocator makes dropping it after enqueueing safe.
    """
    arr_t = type(descs[0]) * len(descs)
    raw = bytes(arr_t(*descs))
    off = (len(raw) + 255) // 256 * 256
    pres, totals = [], []
    for blocks in prefix_lists:
        pre = np.zeros(len(descs) + 1, np.int32)
        pre[1:] = np.cumsum(np.asarray(blocks, dtype=np.int64))
        if pre[-1] >= 2 ** 31 - 1:
            raise ValueError("launch too large")
        pres.append(pre)
        totals.append(int(pre[-1]))
    size = off + sum((p.nbytes + 255) // 256 * 256 for p in pres)
    host = np.zeros(size, np.uint8)
    host[:len(raw)] = np.frombuffer(raw, np.uint8)
    ptr_offs = []
    o = off
    for p in pres:
        host[o:o + p.nbytes] = p.view(np.uint8)
        ptr_offs.append(o)
        o += (p.nbytes + 255) // 256 * 256
    dev = torch.from_numpy(host).to(device, non_blocking=False)
    base = dev.data_ptr()
    return dev, base, [base + x for x in ptr_offs], totals


def n_pad_of(n):
    return (n + 63) // 64 * 64


# ---------------------------------------------------------------- stage (a)+(b)


def eval_cosets(items, precision):
    """items: dicts q, n, b, C, P, S, Mt, D, K, a, x (tensors on one device)."""
    groups = defaultdict(list)
    for it in items:
        n_pad = it["Mt"].shape[-1]
        nkb = n_pad // 64
        pow2 = (it["q"] & (it["q"] - 1)) == 0
        groups[(nkb, pow2)].append(it)
    keep = []
    for (nkb, pow2), its in groups.items():
        descs, blocks = [], []
        for it in its:
            d = L.EvalDesc(q=it["q"], n=it["n"], b=it["b"], C=it["C"], P=it["P"], S=it["S"],
                           n_pad=it["Mt"].shape[-1], Mt=_ptr(it["Mt"]), D=_ptr(it["D"]), K=_ptr(it["K"]),
                           a=_ptr(it["a"]), x=_ptr(it["x"]))
            nb = lib.qsft_eval_blocks(ctypes.byref(d))
            if nb < 0:
                raise ValueError(f"bad eval shape {dict((k, it[k]) for k in ('q', 'n', 'b', 'C', 'P', 'S'))}")
            descs.append(d)
            blocks.append(nb)
        buf, dptr, (pre,), (tot,) = pack(descs, [blocks], its[0]["x"].device)
        word = precision | (nkb << 4) | (int(pow2) << 8)
        L.check(lib.qsft_eval_cosets(ctypes.c_void_p(dptr), ctypes.c_void_p(pre), len(descs), tot, word,
                                     stream_handle()), "qsft_eval_cosets")
        keep.append(buf)
    return keep


def query_digits(items):
    """items: dicts q, n, b, C, P, Mt, D, m (int8 [C][P][q^b][n_pad] out)."""
    descs, blocks = [], []
    for it in items:
        d = L.EvalDesc(q=it["q"], n=it["n"], b=it["b"], C=it["C"], P=it["P"], S=0,
                       n_pad=it["Mt"].shape[-1], Mt=_ptr(it["Mt"]), D=_ptr(it["D"]), K=None, a=None,
                       x=_ptr(it["m"]))
        nb = lib.qsft_query_digits_blocks(ctypes.byref(d))
        if nb < 0:
            raise ValueError("bad query shape")
        descs.append(d)
        blocks.append(nb)
    buf, dptr, (pre,), (tot,) = pack(descs, [blocks], items[0]["m"].device)
    L.check(lib.qsft_query_digits(ctypes.c_void_p(dptr), ctypes.c_void_p(pre), len(descs), tot,
                                  stream_handle()), "qsft_query_digits")
    return [buf]


def eval_points(items, precision):
    """items: dicts q, n, S, m (int8 [N][n_pad]), K, a, x."""
    groups = defaultdict(list)
    for it in items:
        n_pad = it["m"].shape[-1]
        groups[(n_pad // 64, (it["q"] & (it["q"] - 1)) == 0)].append(it)
    keep = []
    for (nkb, pow2), its in groups.items():
        descs, blocks = [], []
        for it in its:
            d = L.EvalPointsDesc(q=it["q"], n=it["n"], S=it["S"], n_pad=it["m"].shape[-1],
                                 N=it["m"].shape[0], m=_ptr(it["m"]), K=_ptr(it["K"]), a=_ptr(it["a"]),
                                 x=_ptr(it["x"]))
            nb = lib.qsft_eval_points_blocks(ctypes.byref(d))
            if nb < 0:
                raise ValueError("bad eval_points shape")
            descs.append(d)
            blocks.append(nb)
        buf, dptr, (pre,), (tot,) = pack(descs, [blocks], its[0]["x"].device)
        word = precision | (nkb << 4) | (int(pow2) << 8)
        L.check(lib.qsft_eval_points(ctypes.c_void_p(dptr), ctypes.c_void_p(pre), len(descs), tot, word,
                                     stream_handle()), "qsft_eval_points")
        keep.append(buf)
    return keep


# -------------------------------------------------------------------- stage (c)


def dft(items, precision):
    """items: dicts q, b, rows, inp (tensor), in_stride, in_row_stride, out (tensor [rows][q^b])."""
    groups = defaultdict(list)
    for it in items:
        groups[it["q"]].append(it)
    keep = []
    for q, its in groups.items():
        descs = []
        for it in its:
            d = L.DftDesc(q=q, b=it["b"], rows=it["rows"], in_stride=it.get("in_stride", 1),
                          in_row_stride=it.get("in_row_stride", q ** it["b"]), in_=_ptr(it["inp"]),
                          out=_ptr(it["out"]))
            L.check(lib.qsft_dft_plan(ctypes.byref(d), precision), "qsft_dft_plan")
            descs.append(d)
        npass = max(d.n_pass for d in descs)
        pl = [[d.pass_blocks[p] if p < d.n_pass else 0 for d in descs] for p in range(npass)]
        buf, dptr, pres, tots = pack(descs, pl, its[0]["out"].device)
        word = precision | (q << 8)
        for p in range(npass):
            L.check(lib.qsft_dft_pass(ctypes.c_void_p(dptr), ctypes.c_void_p(pres[p]), len(descs), tots[p], p,
                                      word, stream_handle()), "qsft_dft_pass")
        keep.append(buf)
    return keep


# ---------------------------------------------------------------- get_MDU step


def assemble(items):
    """items: dicts Q, C, R, p1, src (tensor), src_rows, gsel, rsel, U (out), and
    noise (tensor, complex128 [C][R*p1][Q]) or sigma+seed (Philox) or neither."""
    descs, blocks = [], []
    for it in items:
        d = L.AssembleDesc(Q=it["Q"], C=it["C"], R=it["R"], p1=it["p1"], src_rows=it["src_rows"])
        src = it["src"]
        d.src_prec = L.QSFT_F32 if src.dtype == torch.complex64 else L.QSFT_F64
        d.src = _ptr(src)
        for i, g in enumerate(it["gsel"]):
            d.gsel[i] = int(g)
        for i, r in enumerate(it["rsel"]):
            d.rsel[i] = int(r)
        if it.get("noise") is not None:
            d.noise_mode = L.QSFT_NOISE_ARRAY
            d.noise = _ptr(it["noise"])
        elif it.get("sigma", 0.0) > 0.0:
            d.noise_mode = L.QSFT_NOISE_PHILOX
            d.sigma = float(it["sigma"])
            d.seed = int(it["seed"]) & 0xFFFFFFFFFFFFFFFF
        else:
            d.noise_mode = L.QSFT_NOISE_NONE
        d.U = _ptr(it["U"])
        nb = lib.qsft_assemble_blocks(ctypes.byref(d))
        if nb < 0:
            raise ValueError("bad assemble shape")
        descs.append(d)
        blocks.append(nb)
    buf, dptr, (pre,), (tot,) = pack(descs, [blocks], items[0]["U"].device)
    L.check(lib.qsft_assemble(ctypes.c_void_p(dptr), ctypes.c_void_p(pre), len(descs), tot, stream_handle()),
            "qsft_assemble")
    return [buf]


# -------------------------------------------------------------------- stage (d)


class PeelBatch:
    """Bound peel descriptors + workspaces for a list of instances."""

    def __init__(self, items, max_iter=15):
        self.items = items
        self.max_iter = max_iter
        self.descs = []
        self.ws = []
        dev = items[0]["U"].device
        for it in items:
            q, n, b, C, P = it["q"], it["n"], it["b"], it["C"], it["P"]
            n_pad = it["Mt"].shape[-1]
            nbytes = lib.qsft_peel_workspace_bytes(q, n, b, C, P, max_iter, 0)
            if nbytes == 0:
                raise ValueError("bad peel shape")
            ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
            d = L.PeelDesc()
            L.check(lib.qsft_peel_bind(ctypes.byref(d), q, n, b, C, P, it["p1"], it["method"], max_iter,
                                       float(it["cutoff"]), 0, _ptr(it["Mt"]), _ptr(it["D"]), _ptr(it["U"]),
                                       _ptr(ws), nbytes, n_pad), "qsft_peel_bind")
            self.descs.append(d)
            self.ws.append(ws)
        self.bins = [lib.qsft_peel_bin_blocks(ctypes.byref(d)) for d in self.descs]
        self.slots = [lib.qsft_peel_slot_blocks(ctypes.byref(d)) for d in self.descs]
        self.buf, self.dptr, (self.pre_b, self.pre_s), (self.tot_b, self.tot_s) = pack(
            self.descs, [self.bins, self.slots], dev)

    def run(self, host_loop=False):
        st = stream_handle()
        n = len(self.descs)
        d, pb, ps = ctypes.c_void_p(self.dptr), ctypes.c_void_p(self.pre_b), ctypes.c_void_p(self.pre_s)
        if not host_loop:
            L.check(lib.qsft_peel_run(d, pb, ps, n, self.tot_b, self.tot_s, self.max_iter, st), "qsft_peel_run")
            return
        # host-driven loop: one sync per pass to test the stop flag (the plain path)
        L.check(lib.qsft_peel_init(d, ps, n, self.tot_s, st), "qsft_peel_init")
        for it in range(self.max_iter):
            L.check(lib.qsft_peel_iteration(d, pb, n, self.tot_b, it, st), "qsft_peel_iteration")
            flags = [self._ctr_word(i, (it + 1) * L.QSFT_CTR_ITER_STRIDE + 3) for i in range(n)] \
                if it + 1 < 16 else [0]
            if not any(flags):
                break
        L.check(lib.qsft_peel_finalize(d, ps, n, self.tot_s, st), "qsft_peel_finalize")

    def _ctr_word(self, i, w):
        base = self.descs[i].ctr - self.ws[i].data_ptr()
        return int(self.ws[i][base + 4 * w: base + 4 * w + 4].view(torch.int32).item())

    def results(self):
        """Per instance: (k int8 [m][n], values complex128 [m], order uint64 [m], info dict),
        rows sorted by first detection (the reference's dict insertion order)."""
        out = []
        for i, (d, ws) in enumerate(zip(self.descs, self.ws)):
            base = ws.data_ptr()
            ctr = ws[d.ctr - base: d.ctr - base + 4 * L.QSFT_CTR_WORDS].view(torch.int32).cpu().numpy()
            ctr = ctr.view(np.uint32)
            err = int(ctr[L.QSFT_CTR_ERROR])
            if err:
                raise L.QsftLibraryError(f"peel device error bits {err:#x} (1 table full, 2 tag collision, "
                                         "4 output full)")
            m = int(ctr[L.QSFT_CTR_OUT])
            n_pad = d.n_pad
            ok = ws[d.out_k - base: d.out_k - base + m * n_pad].view(m, n_pad)[:, :d.n]
            ov = ws[d.out_val - base: d.out_val - base + 16 * m].view(torch.float64).view(m, 2)
            oo = ws[d.out_order - base: d.out_order - base + 8 * m].view(torch.int64)
            k = ok.cpu().numpy().astype(np.int64)
            v = ov.cpu().numpy()
            vals = v[:, 0] + 1j * v[:, 1]
            order = oo.cpu().numpy().view(np.uint64)
            srt = np.argsort(order, kind="stable")
            trace = [(int(ctr[4 * t]), int(ctr[4 * t + 1]), int(ctr[4 * t + 2]))
                     for t in range(int(ctr[L.QSFT_CTR_ITERS]))]
            info = {"iterations": int(ctr[L.QSFT_CTR_ITERS]), "records": int(ctr[L.QSFT_CTR_LOG]),
                    "peeled": int(ctr[L.QSFT_CTR_PEELED]), "trace": trace}
            out.append((k[srt], vals[srt], order[srt], info))
        return out
