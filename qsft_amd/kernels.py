"""Batched launchers over the C ABI (include/qsft_amd.h).

Each stage has a prepared-launch class: construction builds the descriptor
array and the block-prefix array(s) for a list of per-instance argument
dicts (torch CUDA tensors), uploads them once, and `launch()` enqueues the
kernels on torch's current stream with no further host work -- so a batch
can be relaunched every step (and captured in a graph). Instances of one
batch may differ in shape; they are split into as few launches as the
kernels' compile-time variants need (q's power-of-two class and the digit
block count for the evaluator, q for the DFT radix). The lower-case
functions are one-shot conveniences (prepare + launch).
"""
import ctypes
import math
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
    The buffer must outlive the launches that read it (prepared launches keep
    it; one-shot callers may drop it after enqueueing, since torch's caching
    allocator reuses memory in stream order).
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


def _vp(x):
    return ctypes.c_void_p(x)


def is_pow2(q):
    return (q & (q - 1)) == 0


# ---------------------------------------------------------------- stage (a)+(b)


class EvalCosets:
    """items: dicts q, n, b, C, P, S, Mt, D, K, a, x."""

    def __init__(self, items, precision):
        groups = defaultdict(list)
        for it in items:
            groups[(it["Mt"].shape[-1] // 64, is_pow2(it["q"]))].append(it)
        self.launches = []
        for (nkb, pow2), its in groups.items():
            descs, blocks = [], []
            for it in its:
                d = L.EvalDesc(q=it["q"], n=it["n"], b=it["b"], C=it["C"], P=it["P"], S=it["S"],
                               n_pad=it["Mt"].shape[-1], Mt=_ptr(it["Mt"]), D=_ptr(it["D"]), K=_ptr(it["K"]),
                               a=_ptr(it["a"]), x=_ptr(it["x"]))
                nb = lib.qsft_eval_blocks(ctypes.byref(d))
                if nb < 0:
                    raise ValueError(f"bad eval shape {[it[k] for k in ('q', 'n', 'b', 'C', 'P', 'S')]}")
                descs.append(d)
                blocks.append(nb)
            buf, dptr, (pre,), (tot,) = pack(descs, [blocks], its[0]["x"].device)
            word = precision | (nkb << 4) | (int(pow2) << 8)
            self.launches.append((buf, dptr, pre, len(descs), tot, word))

    def launch(self):
        for _, dptr, pre, n, tot, word in self.launches:
            L.check(lib.qsft_eval_cosets(_vp(dptr), _vp(pre), n, tot, word, stream_handle()), "qsft_eval_cosets")


def eval_cosets(items, precision):
    p = EvalCosets(items, precision)
    p.launch()
    return p


def mfma_mode(q):
    """0: q in {2, 4} (w^e exact in fp16); 1: other powers of two; 2: any other q."""
    if q in (2, 4):
        return 0
    return 1 if is_pow2(q) else 2


def mfma_wide(q, b, n_pad):
    """Eight-wave workgroups (512 base queries) of the matrix-core evaluator:
    q in {2, 4}, one digit block and q^b >= 512 (bit 11 of the launch word)."""
    return q in (2, 4) and n_pad == 64 and q ** b >= 512


def mfma_ptb(mode, nkb, npt):
    """p-tiles per workgroup of the matrix-core evaluator: the tiles split as
    evenly as possible over ceil(npt / 8) workgroups (5..8) for q in {2, 4}
    with n <= 64, else 8."""
    if mode != 0 or nkb != 1:
        return 8
    return max(5, -(-npt // -(-npt // 8)))


def mfma_bexp(amax, a64):
    """Scale exponent of the fp16 limbs: 14 - e with max|a| = m 2^e, m in [0.5, 1)."""
    if amax is None:
        amax = float(a64.abs().max().item()) if a64.numel() else 0.0
    if not amax > 0.0:
        return 0
    return int(min(100, max(-100, 14 - math.frexp(amax)[1])))


class EvalMfma:
    """Matrix-core coset evaluation (FP32 out).

    items: dicts q, n, b, C, P, S, Mt, D, K, a64 (complex128 [S]), x (complex64 [C][P][q^b]),
    optional amax (max |a_s|, else read from a64) and shared_d.
    The B-fragment workspace is allocated here, one per instance.
    q in {2, 4} runs the sparse-MFMA form; dense=True selects the dense form
    of the same products instead (for comparisons).
    """

    def __init__(self, items, dense=False):
        groups = defaultdict(list)
        self.empty = [it["x"] for it in items if it["S"] == 0]     # no support: x = 0
        for it in items:
            if it["S"] > 0:
                nkb, mode = it["Mt"].shape[-1] // 64, mfma_mode(it["q"])
                wide = mfma_wide(it["q"], it["b"], it["Mt"].shape[-1])
                groups[(nkb, mode, mfma_ptb(mode, nkb, (it["P"] + 7) // 8), wide)].append(it)
        self.launches = []
        self.ws = []
        for (nkb, mode, ptb, wide), its in groups.items():
            descs, prep, main = [], [], []
            for it in its:
                if it["x"].dtype != torch.complex64 or it["a64"].dtype != torch.complex128:
                    raise TypeError("EvalMfma needs complex64 x and complex128 strengths")
                C, P, S = it["C"], it["P"], it["S"]
                shared = it.get("shared_d")
                if shared is None:
                    Dg = it["D"]
                    shared = all(torch.equal(Dg[0], Dg[c]) for c in range(1, C))
                ws = torch.empty(lib.qsft_eval_mfma_workspace_bytes(1 if shared else C, P, S), dtype=torch.uint8,
                                 device=it["x"].device)
                self.ws.append(ws)
                d = L.EvalMfmaDesc(q=it["q"], n=it["n"], b=it["b"], C=C, P=P, S=S, n_pad=it["Mt"].shape[-1],
                                   nsb=(S + 15) // 16, npt=(P + 7) // 8, shared_d=int(bool(shared)),
                                   bexp=mfma_bexp(it.get("amax", None), it["a64"]), ptb=ptb,
                                   Mt=_ptr(it["Mt"]), D=_ptr(it["D"]), K=_ptr(it["K"]), a=_ptr(it["a64"]),
                                   Bf=_ptr(ws), x=_ptr(it["x"]))
                pb, mb = ctypes.c_int(), ctypes.c_int()
                if lib.qsft_eval_mfma_blocks(ctypes.byref(d), ctypes.byref(pb), ctypes.byref(mb)) != 0:
                    raise ValueError(f"bad mfma eval shape {[it[k] for k in ('q', 'n', 'b', 'C', 'P', 'S')]}")
                descs.append(d)
                prep.append(pb.value)
                main.append(mb.value)
            buf, dptr, (pp, pm), (tp, tm) = pack(descs, [prep, main], its[0]["x"].device)
            word = (nkb << 4) | (mode << 8) | (ptb << 12) | (int(bool(dense) and mode == 0) << 10) | (int(wide) << 11)
            self.launches.append((buf, dptr, pp, pm, len(descs), tp, tm, word))

    def launch(self):
        for x in self.empty:
            x.zero_()
        for _, dptr, pp, pm, n, tp, tm, word in self.launches:
            L.check(lib.qsft_eval_mfma(_vp(dptr), _vp(pp), _vp(pm), n, tp, tm, word, stream_handle()),
                    "qsft_eval_mfma")

    def launch_split(self, between):
        """The B-fragment prep launches, then `between()`, then the GEMM
        launches (so events can time the GEMM kernel on its own)."""
        for x in self.empty:
            x.zero_()
        st = stream_handle()
        for _, dptr, pp, pm, n, tp, tm, word in self.launches:
            L.check(lib.qsft_eval_mfma(_vp(dptr), _vp(pp), _vp(pm), n, tp, 0, word, st), "qsft_eval_mfma(prep)")
        between()
        for _, dptr, pp, pm, n, tp, tm, word in self.launches:
            L.check(lib.qsft_eval_mfma(_vp(dptr), _vp(pp), _vp(pm), n, 0, tm, word, st), "qsft_eval_mfma(main)")


class QueryDigits:
    """items: dicts q, n, b, C, P, Mt, D, m (int8 [C][P][q^b][n_pad] out)."""

    def __init__(self, items):
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
        self.buf, self.dptr, (self.pre,), (self.tot,) = pack(descs, [blocks], items[0]["m"].device)
        self.n = len(descs)

    def launch(self):
        L.check(lib.qsft_query_digits(_vp(self.dptr), _vp(self.pre), self.n, self.tot, stream_handle()),
                "qsft_query_digits")


def query_digits(items):
    p = QueryDigits(items)
    p.launch()
    return p


class EvalPoints:
    """items: dicts q, n, S, m (int8 [N][n_pad]), K, a, x."""

    def __init__(self, items, precision):
        groups = defaultdict(list)
        for it in items:
            groups[(it["m"].shape[-1] // 64, is_pow2(it["q"]))].append(it)
        self.launches = []
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
            self.launches.append((buf, dptr, pre, len(descs), tot, precision | (nkb << 4) | (int(pow2) << 8)))

    def launch(self):
        for _, dptr, pre, n, tot, word in self.launches:
            L.check(lib.qsft_eval_points(_vp(dptr), _vp(pre), n, tot, word, stream_handle()), "qsft_eval_points")


def eval_points(items, precision):
    p = EvalPoints(items, precision)
    p.launch()
    return p


# -------------------------------------------------------------------- stage (c)


class Dft:
    """items: dicts q, b, rows, inp (tensor), in_stride, in_row_stride, out (tensor [rows][q^b]).
    rows4: q = 4, b in {4, 6} contiguous rows go to the whole-row kernel."""

    def __init__(self, items, precision, rows4=True, candidate=None):
        """candidate: use the planner's schedule number `candidate` for the
        pass kernels instead of the tuned / model-best one."""
        groups = defaultdict(list)
        fast = defaultdict(list)
        for it in items:
            if rows4 and it["q"] == 4 and it["b"] in (4, 6) and it.get("in_stride", 1) == 1:
                fast[it["b"]].append(it)                           # whole rows in one launch
            else:
                groups[it["q"]].append(it)
        self.launches = []
        self.rows4 = []
        for b, its in fast.items():
            descs = [L.DftDesc(q=4, b=b, rows=it["rows"], in_stride=1,
                               in_row_stride=it.get("in_row_stride", 4 ** b), in_=_ptr(it["inp"]),
                               out=_ptr(it["out"])) for it in its]
            buf, dptr, (pre,), (tot,) = pack(descs, [[it["rows"] for it in its]], its[0]["out"].device)
            self.rows4.append((buf, dptr, pre, tot, len(descs), precision | (b << 8)))
        for q, its in groups.items():
            descs = []
            for it in its:
                d = L.DftDesc(q=q, b=it["b"], rows=it["rows"], in_stride=it.get("in_stride", 1),
                              in_row_stride=it.get("in_row_stride", q ** it["b"]), in_=_ptr(it["inp"]),
                              out=_ptr(it["out"]))
                if candidate is None:
                    L.check(lib.qsft_dft_plan(ctypes.byref(d), precision), "qsft_dft_plan")
                else:
                    L.check(lib.qsft_dft_plan_candidate(ctypes.byref(d), precision, int(candidate)),
                            "qsft_dft_plan_candidate")
                descs.append(d)
            npass = max(d.n_pass for d in descs)
            pl = [[d.pass_blocks[p] if p < d.n_pass else 0 for d in descs] for p in range(npass)]
            lds = [max(lib.qsft_dft_lds_bytes(ctypes.byref(d), p, precision) for d in descs if p < d.n_pass)
                   for p in range(npass)]
            if min(lds) < 0:
                raise ValueError("DFT tile does not fit the LDS budget")
            buf, dptr, pres, tots = pack(descs, pl, its[0]["out"].device)
            self.launches.append((buf, dptr, pres, tots, len(descs), precision | (q << 8), lds))

    def launch(self):
        for _, dptr, pre, tot, n, word in self.rows4:
            L.check(lib.qsft_dft4_rows(_vp(dptr), _vp(pre), n, tot, word, stream_handle()), "qsft_dft4_rows")
        for _, dptr, pres, tots, n, word, lds in self.launches:
            for p in range(len(pres)):
                L.check(lib.qsft_dft_pass(_vp(dptr), _vp(pres[p]), n, tots[p], p, word, lds[p], stream_handle()),
                        "qsft_dft_pass")


def dft(items, precision, candidate=None, rows4=True):
    p = Dft(items, precision, rows4=rows4, candidate=candidate)
    p.launch()
    return p


_tuned = {}


def dft_autotune(q, b, precision, rows, max_candidates=6, device=None):
    """Time the planner's best-ranked DFT schedules for (q, b, precision) on
    `rows` scratch rows (once per process and shape) and record the fastest;
    returns its time in ms. Later plans of that shape use it."""
    key = (q, b, precision)
    if key in _tuned:
        return _tuned[key]
    elem = 8 if precision == L.QSFT_F32 else 16
    rows = max(1, min(int(rows), (256 << 20) // (elem * q ** b) or 1))
    nbytes = (rows * q ** b * elem + 255) // 256 * 256 + 4096
    dev = device if device is not None else torch.device("cuda", torch.cuda.current_device())
    buf = torch.zeros(nbytes, dtype=torch.uint8, device=dev)
    best = ctypes.c_float()
    L.check(lib.qsft_dft_autotune(q, b, precision, rows, _vp(buf.data_ptr()), nbytes, int(max_candidates),
                                  stream_handle(), ctypes.byref(best)), "qsft_dft_autotune")
    _tuned[key] = float(best.value)
    return _tuned[key]


def dft_schedule(q, b, precision):
    """(passes [(d0, d1, ti, ta)], tuned) the planner uses for (q, b, precision)."""
    n = ctypes.c_int()
    arrs = [(ctypes.c_int * 4)() for _ in range(4)]
    tuned = ctypes.c_int()
    L.check(lib.qsft_dft_schedule(q, b, precision, ctypes.byref(n), *arrs, ctypes.byref(tuned)), "qsft_dft_schedule")
    return [tuple(a[i] for a in arrs) for i in range(n.value)], bool(tuned.value)


# ---------------------------------------------------------------- get_MDU step


class Assemble:
    """items: dicts Q, C, R, p1, src (tensor), src_rows, gsel, rsel, U (out), and
    noise (tensor, complex128 [C][R*p1][Q]) or sigma+seed (Philox) or neither;
    optional energy (float64 [C][Q] out: per-bin |column|^2, bin-major only);
    optional written + cutoff + seed_out (device addresses from a sparse
    PeelBatch: PeelBatch.sparse_targets): only the columns above the zeroton
    cutoff are written."""

    def __init__(self, items):
        descs, blocks = [], []
        for it in items:
            d = L.AssembleDesc(Q=it["Q"], C=it["C"], R=it["R"], p1=it["p1"], src_rows=it["src_rows"],
                               bin_major=int(bool(it.get("bin_major", False))))
            src = it["src"]
            d.src_prec = L.QSFT_F32 if src.dtype == torch.complex64 else L.QSFT_F64
            d.src = _ptr(src)
            if len(it["gsel"]) != it["C"] or len(it["rsel"]) != it["R"]:
                raise ValueError("gsel/rsel lengths must equal C/R")
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
            if it.get("energy") is not None:
                d.energy = _ptr(it["energy"])
            if it.get("written"):                                  # sparse U for the peeler
                d.written = int(it["written"])
                d.cutoff = float(it["cutoff"])
                d.seed_out = int(it["seed_out"])
            nb = lib.qsft_assemble_blocks(ctypes.byref(d))
            if nb < 0:
                raise ValueError("bad assemble shape (energies need bin-major U with R*p1 <= 128)")
            descs.append(d)
            blocks.append(nb)
        self.word = -1
        for d in descs:                                            # launch word: AND over the instances
            self.word &= lib.qsft_assemble_word(ctypes.byref(d))
        self.buf, self.dptr, (self.pre,), (self.tot,) = pack(descs, [blocks], items[0]["U"].device)
        self.n = len(descs)

    def launch(self, seed_offset=0):
        L.check(lib.qsft_assemble(_vp(self.dptr), _vp(self.pre), self.n, self.tot,
                                  int(seed_offset) & 0xFFFFFFFFFFFFFFFF, self.word, stream_handle()),
                "qsft_assemble")


def assemble(items):
    p = Assemble(items)
    p.launch()
    return p


# -------------------------------------------------------------------- stage (d)


def unit_delay_sign(D, q, n, p1):
    """+1 or -1 if in every group each repeat block of delay rows is its row 0
    plus s e_(i-1) (mod q) in row i (the identity and NSO delays of the
    reference's get_D), with zero padding past n; 0 otherwise (the peeler then
    takes a full dot product per row). D: int8 [C][P][n_pad] on the device."""
    C, P, n_pad = D.shape
    if p1 != n + 1 or P % p1 or P // p1 > 64 or bool(D[:, :, n:].any()):
        return 0
    R = P // p1
    Dn = D[:, :, :n].to(torch.int64).reshape(C, R, p1, n)
    E = torch.zeros(p1, n, dtype=torch.int64, device=D.device)
    E[1:] = torch.eye(n, dtype=torch.int64, device=D.device)
    for sgn in (-1, 1):
        if torch.equal((Dn[:, :, :1, :] + sgn * E) % q, Dn):
            return sgn
    return 0


def peel_flags(it):
    """QSFT_PEEL_SHARED_D / UNIT_D / UNIT_D_NEG (/ ONE_BIN_WAVES) bits of a peel item (see PeelBatch)."""
    C, q, n = it["C"], it["q"], it["n"]
    shared = it.get("shared_d")
    if shared is None:
        Dg = it["D"]
        shared = all(torch.equal(Dg[0], Dg[c]) for c in range(1, C))
    flags = L.QSFT_PEEL_SHARED_D if shared else 0
    unit = it.get("unit_d")
    if unit is None:
        unit = unit_delay_sign(it["D"][:1] if shared else it["D"], q, n, it["p1"])
    if unit:
        flags |= L.QSFT_PEEL_UNIT_D | (L.QSFT_PEEL_UNIT_D_NEG if unit < 0 else 0)
    if it.get("one_bin_waves"):
        flags |= L.QSFT_PEEL_ONE_BIN_WAVES
    return flags


def peel_noalloc(it, dst, cap_rows, row_stride, max_iter=15):
    """One instance's whole peeling loop through qsft_peel_transform_noalloc:
    the descriptor and workspace live in the library's static scratch, so
    nothing is allocated; `dst` (device or pinned host memory) receives the
    packed header and rows exactly as PeelBatch's pack does for one instance.
    `it` is a PeelBatch item without energy / sparse entries."""
    if it.get("energy") is not None or it.get("sparse") is not None:
        raise ValueError("the no-allocation peel reads U itself (no assembler energies, no sparse U)")
    L.check(lib.qsft_peel_transform_noalloc(
        it["q"], it["n"], it["b"], it["C"], it["P"], it["p1"], it["method"], max_iter, float(it["cutoff"]),
        peel_flags(it), _ptr(it["Mt"]), _ptr(it["D"]), _ptr(it["U"]), it["Mt"].shape[-1], _vp(dst.data_ptr()),
        int(cap_rows), int(row_stride), stream_handle()), "qsft_peel_transform_noalloc")


def unpack_rows(hdr, rows, n_inst, ns, n_pads):
    """Per instance (k int64 [m][n], values complex128 [m], order uint64 [m]),
    sorted by first detection, from a packed header (int64) and rows (uint8
    [total][row_stride])."""
    counts, offs = hdr[:n_inst], hdr[n_inst:2 * n_inst]
    out = []
    for i in range(n_inst):
        r = rows[offs[i]:offs[i] + counts[i]]
        k = r[:, :ns[i]].astype(np.int64)
        v = r[:, n_pads[i]:n_pads[i] + 16].copy().view(np.float64).reshape(-1, 2)
        order = r[:, n_pads[i] + 16:n_pads[i] + 24].copy().view(np.uint64).reshape(-1)
        srt = np.argsort(order, kind="stable")
        out.append((k[srt], (v[:, 0] + 1j * v[:, 1])[srt], order[srt]))
    return out


class PeelBatch:
    """Bound peel descriptors + workspaces for a list of instances.

    items: dicts q, n, b, C, P, p1, method, cutoff, Mt, D, U (complex128,
    bin-major [C][Q][P]), optional shared_d (all groups use the delay rows of
    group 0; compared on the device when absent), unit_d (+1 / -1: delay rows
    are block offsets + s e_i, 0: arbitrary rows; detected when absent),
    coded (dict t, s, tables: the Reed-Solomon code of coded source delays,
    reed_solomon.ReedSolomon.device_tables() on the device; method carries
    QSFT_DETECT_CODED), energy (float64 [C][Q] first-pass column energies written by Assemble into
    the same U) and sparse (dict x = the assembler's FP32 samples [C][P][Q],
    sigma = its Philox sigma or 0): U then holds only the columns the
    assembler wrote (see sparse_targets).
    """

    def __init__(self, items, max_iter=15, host_rows=False, cap_rows=None, final_u=True):
        """host_rows: the pack kernels write the header and rows straight into
        pinned host memory (read after the launch's completion event; no copy
        and no second stream). cap_rows: capacity of the packed rows (default:
        every hash slot of every instance); overflow is reported as an error.
        final_u: subtract the balls the last pass queued, so U ends as the
        reference's loop leaves it (the recovered rows do not depend on it)."""
        self.items = items
        self.fin_word = L.QSFT_FIN_FLUSH if final_u else 0
        self.max_iter = max_iter
        self.host_rows = host_rows
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
            d.flags = peel_flags(it)
            cd = it.get("coded")
            if cd is not None:
                L.check(lib.qsft_peel_set_coded(ctypes.byref(d), int(cd["t"]), int(cd["s"]), _ptr(cd["tables"])),
                        "qsft_peel_set_coded")
            if it.get("energy") is not None:
                d.energy = _ptr(it["energy"])
            sp = it.get("sparse")
            if sp is not None:
                x = sp["x"]
                if (it.get("energy") is None or x.dtype != torch.complex64
                        or tuple(x.shape) != (C, P, it["U"].shape[1])):
                    raise ValueError("a sparse U needs the assembler's energies and FP32 samples [C][P][Q]")
                d.xsrc = _ptr(x)
                d.noise_sigma = float(sp.get("sigma", 0.0))
                d.flags |= L.QSFT_PEEL_SPARSE_U
            self.descs.append(d)
            self.ws.append(ws)
        self.bins = [lib.qsft_peel_bin_blocks(ctypes.byref(d)) for d in self.descs]
        self.slots = [lib.qsft_peel_slot_blocks(ctypes.byref(d)) for d in self.descs]
        self.buf, self.dptr, (self.pre_b, self.pre_s), (self.tot_b, self.tot_s) = pack(
            self.descs, [self.bins, self.slots], dev)
        self.n_pad_max = max(d.n_pad for d in self.descs)
        self.max_p = max(d.P for d in self.descs)
        self.row_stride = self.n_pad_max + 24
        self.cap_rows = sum(1 << d.hcap_log2 for d in self.descs) if cap_rows is None else int(cap_rows)
        self.hdr_bytes = lib.qsft_peel_pack_header_bytes(len(self.descs))
        size = self.hdr_bytes + self.cap_rows * self.row_stride
        if host_rows:
            self.packed = torch.empty(size, dtype=torch.uint8).pin_memory()
        else:
            self.packed = torch.empty(size, dtype=torch.uint8, device=dev)
            self.hdr_host = torch.empty(self.hdr_bytes, dtype=torch.uint8).pin_memory()

    def sparse_targets(self, i):
        """(written, noise_key) device addresses of instance i, for its assembler."""
        return self.descs[i].written, self.descs[i].noise_key

    def run(self, host_loop=False, tail_stream=None, tail_event=None, passes=None, fuse_from=None):
        """The peeling passes on the current stream, then the tail (last
        pass's queued balls, output table, pack). With `tail_stream` the tail
        runs there, after `tail_event` (recorded here on the current stream),
        so the current stream can go on with other work meanwhile. `passes`
        (< max_iter) launches only that many passes; the packed header then
        says whether a loop wanted more (resume() runs the rest). `fuse_from`
        (without host_loop) launches passes fuse_from .. max_iter - 1 as one
        kernel (qsft_peel_passes) after the per-pass launches before it."""
        st = stream_handle()
        n = len(self.descs)
        d, pb, ps = _vp(self.dptr), _vp(self.pre_b), _vp(self.pre_s)
        passes = self.max_iter if passes is None else min(int(passes), self.max_iter)
        if fuse_from is not None and not host_loop:
            fuse_from = max(0, min(int(fuse_from), self.max_iter))
            L.check(lib.qsft_peel_init(d, pb, ps, n, self.tot_b, self.tot_s, st), "qsft_peel_init")
            for it in range(fuse_from):
                L.check(lib.qsft_peel_iteration(d, pb, n, self.tot_b, self.max_p, it, st), "qsft_peel_iteration")
            L.check(lib.qsft_peel_passes(d, n, self.max_p, fuse_from, self.max_iter, st), "qsft_peel_passes")
            self._tail(tail_stream, tail_event)
            return
        if (not host_loop and tail_stream is None and self.fin_word == L.QSFT_FIN_FLUSH
                and passes == self.max_iter):
            L.check(lib.qsft_peel_run(d, pb, ps, n, self.tot_b, self.tot_s, self.max_p, self.max_iter, st),
                    "qsft_peel_run")
            L.check(lib.qsft_peel_pack(d, n, _vp(self.packed.data_ptr()), self.cap_rows, self.row_stride, st),
                    "qsft_peel_pack")
            return
        L.check(lib.qsft_peel_init(d, pb, ps, n, self.tot_b, self.tot_s, st), "qsft_peel_init")
        for it in range(passes):
            L.check(lib.qsft_peel_iteration(d, pb, n, self.tot_b, self.max_p, it, st), "qsft_peel_iteration")
            if not host_loop or it + 1 >= self.max_iter:
                continue
            # host-driven loop: one sync per pass to read the stop flag (the plain path)
            if not any(self._ctr_word(i, (it + 1) * L.QSFT_CTR_ITER_STRIDE + 3) for i in range(n)):
                break
        self._tail(tail_stream, tail_event)

    def resume(self, start, tail_stream=None, tail_event=None):
        """Passes start .. max_iter - 1 (each stops at once if its loop already
        has), then the tail again: for a run() that launched too few passes."""
        st = stream_handle()
        n = len(self.descs)
        for it in range(start, self.max_iter):
            L.check(lib.qsft_peel_iteration(_vp(self.dptr), _vp(self.pre_b), n, self.tot_b, self.max_p, it, st),
                    "qsft_peel_iteration")
        self._tail(tail_stream, tail_event)

    def _tail(self, tail_stream, tail_event):
        n = len(self.descs)
        d, ps = _vp(self.dptr), _vp(self.pre_s)
        st = stream_handle()
        if tail_stream is not None:
            tail_event.record()
            tail_stream.wait_event(tail_event)
            st = stream_handle(tail_stream)
        L.check(lib.qsft_peel_finalize(d, ps, n, self.tot_s, self.fin_word, st), "qsft_peel_finalize")
        L.check(lib.qsft_peel_pack(d, n, _vp(self.packed.data_ptr()), self.cap_rows, self.row_stride, st),
                "qsft_peel_pack")

    def _ctr_word(self, i, w):
        base = self.descs[i].ctr - self.ws[i].data_ptr()
        return int(self.ws[i][base + 4 * w: base + 4 * w + 4].view(torch.int32).item())

    def fetch_packed(self, copy_stream=None, ready=None):
        """The packed rows on the host: (counts, offsets, rows uint8 [total][row_stride]).

        host_rows: wait for `ready` (an event recorded after run(); else the
        current stream) and read the pinned buffer. Otherwise copy: with
        `copy_stream` and `ready` the copies wait only for that event on the
        copy stream, so later work queued on the compute stream keeps running.
        """
        n = len(self.descs)
        if self.host_rows:
            if ready is not None:
                ready.synchronize()
            else:
                torch.cuda.current_stream().synchronize()
            hdr = self.packed[:self.hdr_bytes].numpy().view(np.int64)
            self._check_header(hdr, n)
            nbytes = int(hdr[2 * n]) * self.row_stride
            rows = self.packed[self.hdr_bytes:self.hdr_bytes + nbytes].numpy().copy()
            return hdr[:n].copy(), hdr[n:2 * n].copy(), rows.reshape(-1, self.row_stride)
        if copy_stream is None:
            self.hdr_host.copy_(self.packed[:self.hdr_bytes])          # synchronising copy
        else:
            copy_stream.wait_event(ready)
            with torch.cuda.stream(copy_stream):
                self.hdr_host.copy_(self.packed[:self.hdr_bytes], non_blocking=True)
            copy_stream.synchronize()
        hdr = self.hdr_host.numpy().view(np.int64)
        self._check_header(hdr, n)
        total = int(hdr[2 * n])
        nbytes = total * self.row_stride
        if getattr(self, "rows_host", None) is None or self.rows_host.numel() < nbytes:
            self.rows_host = torch.empty(max(nbytes, 1 << 20), dtype=torch.uint8).pin_memory()
        src = self.packed[self.hdr_bytes:self.hdr_bytes + nbytes]
        if copy_stream is None:
            self.rows_host[:nbytes].copy_(src)
        else:
            with torch.cuda.stream(copy_stream):
                self.rows_host[:nbytes].copy_(src, non_blocking=True)
            copy_stream.synchronize()
        rows = self.rows_host[:nbytes].numpy().copy()
        return hdr[:n].copy(), hdr[n:2 * n].copy(), rows.reshape(total, self.row_stride)

    def _check_header(self, hdr, n):
        self.last_iters = int(hdr[2 * n + 2])                      # most passes any instance ran
        self.last_unfinished = int(hdr[2 * n + 3])                  # instances that wanted more passes
        err = int(hdr[2 * n + 1])
        if err:
            raise L.QsftLibraryError(f"peel device error bits {err:#x} (1 table full, 2 tag collision, "
                                     "4 output full)")

    def results(self, with_info=True, packed=None):
        """Per instance: (k int64 [m][n], values complex128 [m], order uint64 [m], info dict),
        rows sorted by first detection (the reference's dict insertion order)."""
        counts, offs, rows = self.fetch_packed() if packed is None else packed
        out = []
        for i, d in enumerate(self.descs):
            r = rows[offs[i]:offs[i] + counts[i]]
            k = r[:, :d.n].astype(np.int64)
            v = r[:, d.n_pad:d.n_pad + 16].copy().view(np.float64).reshape(-1, 2)
            order = r[:, d.n_pad + 16:d.n_pad + 24].copy().view(np.uint64).reshape(-1)
            srt = np.argsort(order, kind="stable")
            info = self.info(i) if with_info else {}
            out.append((k[srt], (v[:, 0] + 1j * v[:, 1])[srt], order[srt], info))
        return out

    def info(self, i):
        d, ws = self.descs[i], self.ws[i]
        base = d.ctr - ws.data_ptr()
        ctr = ws[base: base + 4 * L.QSFT_CTR_WORDS].view(torch.int32).cpu().numpy().view(np.uint32)
        iters = int(ctr[L.QSFT_CTR_ITERS])
        s = L.QSFT_CTR_ITER_STRIDE
        return {"iterations": iters, "records": int(ctr[L.QSFT_CTR_LOG]), "peeled": int(ctr[L.QSFT_CTR_PEELED]),
                "trace": [(int(ctr[s * t]), int(ctr[s * t + 1]), int(ctr[s * t + 2])) for t in range(iters)],
                "scanned": [int(ctr[s * t + 4]) for t in range(iters)]}
