"""Serving loop: transforms of mixed shapes admitted as they arrive, batched
continuously onto the GPU, results delivered per request.

The reference serves one request per `QSFT.transform` call
(`qsft/qsft.py:56-281`) and scales only through process pools
(`qsft/parallel_tests.py:91-97`). Here a request is a `TransformSpec`
(q, n, b, S, C, R, noise), and a `Server`:

  * admits it by its HBM footprint (`qsft_request_footprint`: inputs,
    samples, FP64 working transform, B fragments, peel workspace) against a
    byte budget; a request that could never fit is rejected at submit
    (`AdmissionError`), one that does not fit now waits in arrival order;
  * keeps up to `max_inflight` batches on the GPU at once, each on its own
    stream with its finalize/pack on a tail stream, so a batch formed from
    requests that arrived while others run overlaps their work on the device;
    a batch that completes frees its slot and bytes at once and the next
    batch is formed from everything that fits (continuous batching);
  * builds each batch with the native batch runtime (`qsft_batch_*`: one
    staging arena, one plan, one upload, every stage enqueued by one call);
  * delivers each request's recovered (k, value) rows as views of the packed
    output in pinned host memory, already in the reference's dict order
    (first detection); `Result.to_dict()` builds the reference's dict on
    demand.

Latency is measured per request from its arrival (submit time) to the moment
its batch's results are in host memory and handed to it.

While the loop serves (`drain`, `run_open_loop`, or any loop inside
`gc_paused()`), Python's automatic garbage collection is off and only the
youngest generation is collected between polls: a full collection walks the
whole heap (every request and result object), which here stalled the host for
50-60 ms, and an open-loop queue carries such a stall as latency for hundreds
of milliseconds after it.
"""
import collections
import contextlib
import ctypes
import gc
import os
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib as L
from .engine import TransformSpec

lib = L.lib
# host threads that pack a batch's request inputs into its staging arena
PACK_THREADS = max(1, min(8, (os.cpu_count() or 1) // 2))


class AdmissionError(ValueError):
    """The request's HBM footprint exceeds the server's whole budget."""


# objects allocated since the last young collection above which the serving
# loop collects generation 0 (Python's own default threshold is 700)
GC_YOUNG = 2000


@contextlib.contextmanager
def gc_paused():
    """Automatic garbage collection off for the duration (restored after);
    `collect_young()` between polls keeps the youngest generation small."""
    was = gc.isenabled()
    gc.disable()
    try:
        yield
    finally:
        if was:
            gc.enable()


def collect_young():
    if gc.get_count()[0] > GC_YOUNG:
        gc.collect(0)


def request_struct(sp, max_iter=15):
    """The qsft_request of a TransformSpec."""
    if sp.method == "nso":
        method = L.QSFT_DETECT_NSO1
    elif sp.method == "identity":
        method = L.QSFT_DETECT_NOISELESS
    else:
        raise ValueError(f"detection method {sp.method!r}")
    Q = sp.q ** sp.b
    # amax 0: the runtime takes max |a_s| from the strengths while it packs them
    return L.Request(sp.q, sp.n, sp.b, sp.C, sp.R, sp.strengths.size, method, 0, float(sp.threshold()),
                     float(sp.noise_sd) / Q ** 0.5 if sp.noise_sd > 0 else 0.0, 0.0,
                     int(sp.seed) & 0xFFFFFFFFFFFFFFFF)


def footprint(sp, max_iter=15):
    """Device bytes the request holds while its batch runs (0: not servable)."""
    r = request_struct(sp, max_iter)
    return int(lib.qsft_request_footprint(ctypes.byref(r), max_iter))


class Result:
    """One request's recovered spectrum: rows in first-detection order.

    `k` (int8 [m][n]) and `values` (complex128 [m]) are views of the batch's
    packed output in pinned host memory (kept alive by this object)."""

    __slots__ = ("k", "values", "order", "iterations")

    def __init__(self, k, values, order, iterations):
        self.k, self.values, self.order, self.iterations = k, values, order, iterations

    def to_dict(self):
        """The reference's {tuple(k): coefficient} dict (qsft/qsft.py:247-272)."""
        return {tuple(int(x) for x in row): complex(v) for row, v in zip(self.k, self.values)}


def _arena_take(arena, name, nbytes, device, pinned=False):
    """A uint8 buffer of at least nbytes from `arena` (a dict kept per server
    slot), grown by a quarter beyond the request when it is too small; None:
    a fresh allocation."""
    nbytes = max(int(nbytes), 256)
    if arena is not None:
        buf = arena.get(name)
        if buf is not None and buf.numel() >= nbytes:
            return buf[:nbytes]
        nbytes_alloc = nbytes + nbytes // 4
    else:
        nbytes_alloc = nbytes
    if pinned:
        buf = torch.empty(nbytes_alloc, dtype=torch.uint8, pin_memory=True)
    else:
        buf = torch.empty(nbytes_alloc, dtype=torch.uint8, device=device)
    if arena is not None:
        arena[name] = buf
    return buf[:nbytes]


class ServedBatch:
    """One batch of requests on the native runtime: staged, bound, launched.

    arena: a dict of staging and device buffers reused across batches (the
    server keeps one per slot; a slot runs one batch at a time and takes the
    next only once the GPU is done with the last), so a steady serving loop
    allocates only each batch's output. None: buffers of this batch alone."""

    def __init__(self, specs, device, stream=None, tail_stream=None, max_iter=15, fuse_from=3, structs=None,
                 arena=None):
        self.specs = list(specs)
        n = len(self.specs)
        if structs is None:
            structs = [request_struct(sp, max_iter) for sp in self.specs]
        self.reqs = (L.Request * n)(*structs)
        sz = L.BatchSizes()
        h = lib.qsft_batch_create(ctypes.cast(self.reqs, ctypes.c_void_p), n, max_iter, fuse_from, ctypes.byref(sz))
        if not h:
            raise L.QsftLibraryError("qsft_batch_create: " + lib.qsft_last_error().decode(errors="replace"))
        self.h = ctypes.c_void_p(h)
        self.sizes = sz
        self.bytes = int(sz.input_bytes + sz.work_bytes + sz.plan_bytes)
        self.in_host = _arena_take(arena, "in_host", sz.input_bytes, device, pinned=True)
        self.plan_host = _arena_take(arena, "plan_host", sz.plan_bytes, device, pinned=True)
        # the output is handed to the requests (views): never from the arena
        self.out = torch.empty(int(sz.out_bytes), dtype=torch.uint8, pin_memory=True)
        self.in_dev = _arena_take(arena, "in_dev", sz.input_bytes, device)
        self.work_dev = _arena_take(arena, "work_dev", sz.work_bytes, device)
        self.plan_dev = _arena_take(arena, "plan_dev", sz.plan_bytes, device)
        base = self.in_host.data_ptr()
        # [4][n] addresses of every request's M, D, locq and strengths
        ptrs = np.ascontiguousarray(np.array([sp.wire_ptrs() for sp in self.specs], dtype=np.uint64).T)
        p0 = ptrs.ctypes.data
        L.check(lib.qsft_batch_pack_inputs(self.h, p0, p0 + 8 * n, p0 + 16 * n, p0 + 24 * n, base,
                                           PACK_THREADS if n >= 8 else 1), "qsft_batch_pack_inputs")
        L.check(lib.qsft_batch_bind(self.h, self.in_dev.data_ptr(), self.work_dev.data_ptr(),
                                    self.plan_host.data_ptr(), self.plan_dev.data_ptr(), self.out.data_ptr()),
                "qsft_batch_bind")
        self.stream = stream
        self.tail_stream = tail_stream
        self.uploaded = False

    def launch(self, seed_offset=0):
        """Enqueue the whole pipeline (the first launch also uploads inputs and plan)."""
        st = (self.stream or torch.cuda.current_stream()).cuda_stream
        ts = self.tail_stream.cuda_stream if self.tail_stream is not None else None
        upload = 0 if self.uploaded else 3
        L.check(lib.qsft_batch_launch(self.h, upload, self.in_host.data_ptr(), int(seed_offset) & (2 ** 64 - 1),
                                      st, ts), "qsft_batch_launch")
        self.uploaded = True

    def capture(self):
        """Capture the batch's relaunch (the whole pipeline, tail stream
        included) as a hipGraph; needs a dedicated stream and one launch
        first. Replay k draws the noise of launch(seed_offset=k)."""
        if self.stream is None:
            raise ValueError("graph capture needs the batch's own stream (not the default stream)")
        ts = self.tail_stream.cuda_stream if self.tail_stream is not None else None
        L.check(lib.qsft_batch_capture(self.h, self.stream.cuda_stream, ts), "qsft_batch_capture")
        self.replays = 0

    def replay(self):
        L.check(lib.qsft_batch_replay(self.h, self.stream.cuda_stream), "qsft_batch_replay")
        self.replays += 1

    def done(self):
        rc = lib.qsft_batch_done(self.h)
        if rc < 0:
            raise L.QsftLibraryError("qsft_batch_done: " + lib.qsft_last_error().decode(errors="replace"))
        return rc == 1

    def wait(self):
        L.check(lib.qsft_batch_wait(self.h), "qsft_batch_wait")

    def results(self):
        """Per request a Result (views of the pinned output) once the batch is done."""
        n = len(self.specs)
        hb = int(lib.qsft_peel_pack_header_bytes(n))
        buf = self.out.numpy()
        hdr = buf[:hb].view(np.int64)
        err = int(hdr[2 * n + 1])
        if err:
            raise L.QsftLibraryError(f"peel device error bits {err:#x} (1 table full, 2 tag collision, "
                                     "4 output full)")
        stride = int(self.sizes.row_stride)
        total = int(hdr[2 * n])
        rows = buf[hb:hb + total * stride]
        iters = int(hdr[2 * n + 2])
        out = []
        for i, sp in enumerate(self.specs):
            cnt, off = int(hdr[i]), int(hdr[n + i])
            n_pad = (sp.n + 63) // 64 * 64
            blk = rows[off * stride:(off + cnt) * stride]
            k = np.ndarray((cnt, sp.n), dtype=np.int8, buffer=blk, strides=(stride, 1)) if cnt else \
                np.zeros((0, sp.n), np.int8)
            vals = np.ndarray((cnt,), dtype=np.complex128, buffer=blk, offset=n_pad, strides=(stride,)) if cnt \
                else np.zeros(0, np.complex128)
            order = np.ndarray((cnt,), dtype=np.uint64, buffer=blk, offset=n_pad + 16, strides=(stride,)) if cnt \
                else np.zeros(0, np.uint64)
            out.append(Result(k, vals, order, iters))
        return out

    def free(self):
        if self.h:
            lib.qsft_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


@dataclass
class Request:
    spec: TransformSpec
    footprint: int
    t_submit: float
    struct: object = None
    t_start: float = None
    t_done: float = None
    result: Result = None
    batch_size: int = 0

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_submit


class Server:
    """Continuous-batching server on one GPU.

    hbm_budget: bytes that requests in flight may hold together (admission);
    max_batch: requests per batch; max_inflight: batches on the GPU at once
    (each on its own stream; finalize and pack on a tail stream per slot).
    A free slot takes whatever waits at once (holding requests back until a
    batch fills measured no better at 80-90 % load: p99 2.6-3.5 ms either way).

    Each slot keeps its batches' staging and device buffers (grown to the
    largest batch it has run, outside the budget's count), so a steady loop
    allocates only each batch's pinned output."""

    def __init__(self, hbm_budget, max_batch=64, max_inflight=2, max_iter=15, fuse_from=3, device=None):
        self.device = device if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.budget = int(hbm_budget)
        self.max_batch = int(max_batch)
        self.max_inflight = int(max_inflight)
        self.max_iter, self.fuse_from = max_iter, fuse_from
        self.waiting = collections.deque()
        self.inflight = []                    # (ServedBatch, [Request], slot)
        self.used = 0
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(self.max_inflight)]
        self.tails = [torch.cuda.Stream(device=self.device) for _ in range(self.max_inflight)]
        self.free_slots = list(range(self.max_inflight))
        self.arenas = [dict() for _ in range(self.max_inflight)]    # per slot: staging and device buffers
        self.completed = []
        self.batches = 0
        self.host_s = {"submit": 0.0, "build": 0.0, "launch": 0.0, "retire": 0.0}     # host seconds per phase

    def tune(self, specs, max_candidates=6):
        """Autotune the sub-DFT schedule of every (q, b) among `specs` not
        tuned yet (FP32, C * P rows of the first request of each shape), before
        serving them: tuning times kernels synchronously."""
        from . import kernels as K
        seen = {}
        for sp in specs:
            if not (sp.q == 4 and sp.b in (4, 6)):                # those use the whole-row kernel
                seen.setdefault((sp.q, sp.b), sp.C * sp.P)
        for (q, b), rows in seen.items():
            K.dft_autotune(q, b, L.QSFT_F32, rows, max_candidates=max_candidates, device=self.device)

    def submit(self, spec, t_submit=None):
        t0 = time.perf_counter()
        rs = request_struct(spec, self.max_iter)
        fp = int(lib.qsft_request_footprint(ctypes.byref(rs), self.max_iter))
        if fp <= 0:
            raise AdmissionError(f"request shape not servable (q={spec.q}, n={spec.n}, b={spec.b}, R={spec.R})")
        if fp > self.budget:
            raise AdmissionError(f"request needs {fp} bytes of HBM, more than the whole budget {self.budget}")
        req = Request(spec=spec, footprint=fp, t_submit=time.perf_counter() if t_submit is None else t_submit,
                      struct=rs)
        self.host_s["submit"] += time.perf_counter() - t0
        self.waiting.append(req)
        return req

    def _retire(self):
        still = []
        for bt, reqs, slot in self.inflight:
            if bt.done():
                t0 = time.perf_counter()
                res = bt.results()
                now = time.perf_counter()
                for r, x in zip(reqs, res):
                    r.result = x
                    r.t_done = now
                self.completed.extend(reqs)
                self.used -= bt.bytes
                self.free_slots.append(slot)
                bt.free()
                self.host_s["retire"] += time.perf_counter() - t0
            else:
                still.append((bt, reqs, slot))
        self.inflight = still

    def _admit(self):
        while self.waiting and self.free_slots:
            take, used = [], self.used
            while self.waiting and len(take) < self.max_batch:
                r = self.waiting[0]
                if used + r.footprint > self.budget:
                    break                                          # arrival order: wait for bytes to free up
                take.append(self.waiting.popleft())
                used += r.footprint
            if not take:
                return
            slot = self.free_slots.pop()
            t0 = time.perf_counter()
            bt = ServedBatch([r.spec for r in take], self.device, stream=self.streams[slot],
                             tail_stream=self.tails[slot], max_iter=self.max_iter, fuse_from=self.fuse_from,
                             structs=[r.struct for r in take], arena=self.arenas[slot])
            now = time.perf_counter()
            self.host_s["build"] += now - t0
            for r in take:
                r.t_start = now
                r.batch_size = len(take)
            bt.launch()
            self.host_s["launch"] += time.perf_counter() - now
            self.used += bt.bytes
            self.inflight.append((bt, take, slot))
            self.batches += 1

    def poll(self):
        """Retire finished batches, then start new ones from the waiting requests."""
        self._retire()
        self._admit()

    def idle(self):
        return not self.waiting and not self.inflight

    def drain(self):
        with gc_paused():
            while not self.idle():
                self.poll()
                collect_young()
                if self.inflight and not self.free_slots:
                    time.sleep(20e-6)

    def run_open_loop(self, arrivals, specs):
        """Replay arrivals (seconds from now, ascending) of `specs`: submit each
        at its time, keep polling, until every request has its result.
        Returns the Requests in arrival order."""
        with gc_paused():
            t0 = time.perf_counter()
            reqs = []
            i, n = 0, len(specs)
            while i < n or not self.idle():
                now = time.perf_counter()
                while i < n and t0 + arrivals[i] <= now:
                    reqs.append(self.submit(specs[i], t_submit=t0 + arrivals[i]))
                    i += 1
                self.poll()
                collect_young()
                if i < n and not self.inflight and not self.waiting:
                    wait = t0 + arrivals[i] - time.perf_counter()
                    if wait > 50e-6:
                        time.sleep(wait - 30e-6)
        return reqs


def latency_summary(reqs):
    lat = np.array([r.latency for r in reqs], dtype=np.float64)
    return {"p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99)),
            "mean_ms": 1e3 * float(lat.mean()), "n": int(lat.size)}
