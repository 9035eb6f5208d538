#!/usr/bin/env python3
"""q-SFT throughput on one MI355X: served FP32 engine vs plain FP64 path.

Workload (the configuration the headline metric is quoted on): q=4, n=50,
S=1000 uniform distinct support, |a| ~ U[1,10], random phase, 10 dB SNR
(noise_sd^2 = sum|a|^2 / 10), C=3 uniform M_c, b=6 (4096 bins), NSO delays
with R=2 repeats (P=102 rows per group), nso1 detection, 15-pass cap.

A step is one pass of the whole hot path over a batch of `--batch`
independent transforms whose inputs (M_c, D, support, strengths) are already
resident in HBM: query generation + signal evaluation, sub-DFTs, working
transform + noise, device peeling, and the recovered (k, value) rows written
by the pack kernels into pinned host memory. `value` = transforms per second
over the timed steps.

Reported beside it:
  * p50/p99 latency of a step (submit -> rows on the host);
  * the plain single-instance FP64 path (same kernels in FP64, batch 1,
    host-synchronised peel loop) and the speedup over it;
  * NMSE of every transform of the last step against the planted spectrum;
  * per-stage kernel time from HIP events, and the roofline of the dominant
    kernel (signal evaluation);
  * the CPU oracle (numpy restatement of the reference path, process pool
    for the evaluation as the reference does) on one transform.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "transforms/s + p99 latency @ NMSE≤1e-3: served FP32 vs plain FP64 GPU"
K1 = dict(q=4, n=50, S=1000, b=6, C=3, R=2, snr_db=10.0)
PEAK_INT8_TOPS = 5000.0          # MI355X dense int8 MFMA, TOP/s
PEAK_F16_TFLOPS = 2500.0         # MI355X dense fp16 / bf16 MFMA, TFLOP/s


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--plain", type=int, default=3, help="plain FP64 transforms to time")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-workers", type=int, default=16)
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--streams", choices=("shared", "per-batch"), default="shared",
                    help="shared: both resident batches on one stream (the peeling tail on a second); "
                         "per-batch: each batch on its own stream, so their kernels interleave")
    ap.add_argument("--fuse-from", default="3",
                    help="first peeling pass of the fused tail kernel (one workgroup per instance looping over "
                         "the remaining passes); 'none': one detect/apply launch pair per pass")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a known-size reference for "
                         "hardware byte counters)")
    return ap.parse_args()


def log(msg, quiet=False):
    if not quiet:
        print(msg, file=sys.stderr, flush=True)


def measured_traffic(batch):
    """HBM bytes per k_eval_mfma launch at this batch size, from the committed
    rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE per dispatch, each
    scaled by what the counter reads for a copy of known size); None if absent."""
    path = os.path.join(ROOT, "profiles", "k_eval_mfma_k1_traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("bytes_per_transform"):
                return float(line.split("=")[1]) * batch
    return None


def cpu_baseline(seed, workers):
    """One K1 transform through the CPU oracle: decimal query indices as the
    reference builds them, evaluation on a process pool, scipy fftn, the
    Python peeling loop."""
    from multiprocessing import Pool

    from oracle import qsft_oracle as O
    from qsft_amd.engine import random_spec
    sp = random_spec(K1["q"], K1["n"], K1["S"], K1["b"], K1["C"], K1["R"], K1["snr_db"], seed=seed)
    qa = dict(query_method="complex", num_subsample=sp.C, b=sp.b, delays_method_source="identity",
              delays_method_channel="nso", num_repeat=sp.R, subsampling_method="qsft")
    cores = max(1, min(workers, os.cpu_count() or 1))
    np.random.seed(seed)
    t0 = time.perf_counter()
    with Pool(cores) as pool:
        sig = O.SyntheticSignal(sp.n, sp.q, sp.locq, sp.strengths, sp.noise_sd, qa, pool=pool,
                                faithful_indices=True)
    t1 = time.perf_counter()
    gw = O.qsft_transform(sig, sp.C, sp.R, sp.b, "identity", "nso")
    t2 = time.perf_counter()
    truth = {tuple(c): v for c, v in zip(sp.locq.T, sp.strengths)}
    return {"value": 1.0 / (t2 - t0), "unit": "transforms/s", "cores": cores, "kind": "port",
            "sample": f"1 transform of the same workload (construct {t1 - t0:.1f} s with a {cores}-process pool "
                      f"for evaluation, decode {t2 - t1:.1f} s single-threaded)",
            "nmse": O.nmse(gw, truth)}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (transforms are independent: "
              "scale-out is one engine per GPU with no collective).", file=sys.stderr)
        sys.exit(2)
    import torch

    from qsft_amd import engine as E

    if args.copy_calibration_mib > 0:
        nb = args.copy_calibration_mib << 20
        src = torch.ones(nb, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {nb} bytes read, {nb} bytes written", args.quiet)
        del src, dst

    q, n, S, b, C, R, snr = (K1[k] for k in ("q", "n", "S", "b", "C", "R", "snr_db"))
    rng = np.random.default_rng(args.seed)
    log(f"generating 2 x {args.batch} synthetic transforms", args.quiet)
    spec_sets = [[E.random_spec(q, n, S, b, C, R, snr, rng=rng) for _ in range(args.batch)] for _ in range(2)]
    # two resident batches alternate: while the host reads back one (its pack
    # kernels write the rows straight into pinned host memory), the GPU runs the other
    fuse = None if args.fuse_from == "none" else int(args.fuse_from)
    engines = [E.Batch(sp, precision="fp32", own_stream=args.streams == "per-batch", fuse_from=fuse)
               for sp in spec_sets]
    eng, specs = engines[0], spec_sets[0]

    log("warmup", args.quiet)
    for w in range(args.warmup):
        for e in engines:
            e.launch(seed_offset=w)
        for e in engines:
            e.unpack(e.fetch())
    torch.cuda.synchronize()

    lat = []
    launched = {}
    last = {}
    # HIP events around the evaluation GEMM of every timed step, created up front
    # (creating events inside the loop stalls the queue now and then)
    k0 = "eval_prep" if eng.eval_mode == "mfma" else "start"          # GEMM only: after the B-fragment prep
    gemm_marks = [{k0: torch.cuda.Event(enable_timing=True), "eval": torch.cuda.Event(enable_timing=True)}
                  for _ in range(args.steps)]

    def launch(s):
        engines[s % 2].launch(seed_offset=1000 + s, marks=gemm_marks[s])
        launched[s] = time.perf_counter()

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(min(2, args.steps)):
        launch(s)
    for s in range(args.steps):
        e = engines[s % 2]
        packed = e.fetch()
        lat.append(time.perf_counter() - launched[s])
        last[s % 2] = packed
        if s + 2 < args.steps:
            launch(s + 2)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = args.batch * args.steps / elapsed
    log(f"served: {value:.1f} transforms/s ({1e3 * elapsed / args.steps:.2f} ms/step)", args.quiet)
    nm = []
    for slot, packed in last.items():
        res = engines[slot].unpack(packed)
        nm += [E.nmse(k, v, sp) for (k, v, _, _), sp in zip(res, spec_sets[slot])]

    # the dominant kernel's launch time inside the timed steps
    gemm_ms = float(np.mean([m[k0].elapsed_time(m["eval"]) for m in gemm_marks]))
    # per-stage kernel time: three more launches with events between all stages
    stage_ms = {}
    for rep in range(3):
        ev = {}
        eng.launch(seed_offset=99 + rep, events=ev)
        eng.results(with_info=False)
        torch.cuda.synchronize()
        prev = ev["start"][0]
        for x in ("eval_prep", "eval", "dft", "assemble", "peel"):
            if x in ev:
                stage_ms[x] = stage_ms.get(x, 0.0) + prev.elapsed_time(ev[x][0]) / 3
                prev = ev[x][0]
    eval_s = gemm_ms / 1e3
    samples = C * (R * (n + 1)) * q ** b                                # B = C * P * Q per transform
    if eng.eval_mode == "mfma":
        # algorithmic work: B*S complex multiply-adds = 8 B S real flops per transform
        flops = 8.0 * samples * S * args.batch
        achieved = flops / eval_s / 1e12
        # executed on the matrix cores: 2 fp16 limbs of the (Q x 2S) x (2S x 2P) real GEMM
        # (padded to 16-support blocks and 8-row tiles, ptb tiles per workgroup) plus the int8 E0 products
        from qsft_amd.kernels import mfma_mode, mfma_ptb
        nsb, npt = (S + 15) // 16, (R * (n + 1) + 7) // 8
        ptb = mfma_ptb(mfma_mode(q), 1, npt)
        npt_run = ptb * ((npt + ptb - 1) // ptb)
        mfma_f16 = 2 * 2.0 * C * q ** b * (32 * nsb) * (16 * npt_run) * args.batch
        roof = {"kernel": "k_eval_mfma (int8 MFMA phase + fp16x2 MFMA complex GEMM, FP32 accumulation)",
                "bound": "mfma", "achieved": achieved, "peak": PEAK_F16_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_F16_TFLOPS, "traffic": measured_traffic(args.batch),
                "algorithmic_flops_per_launch": flops, "launch_ms": gemm_ms,
                "mfma_f16_executed_tflops": mfma_f16 / eval_s / 1e12}
    else:
        ops = 2.0 * samples * S * n * args.batch                         # int8 MACs x 2 per launch
        achieved = ops / eval_s / 1e12
        roof = {"kernel": "k_eval_cosets (int8 MFMA phase + FP32 w-table accumulation)", "bound": "mfma",
                "achieved": achieved, "peak": PEAK_INT8_TOPS, "unit": "TOP/s (int8)",
                "frac": achieved / PEAK_INT8_TOPS, "traffic": None, "launch_ms": gemm_ms}
    log(f"stages (ms per batch): {stage_ms}", args.quiet)

    # plain FP64 single-instance path
    plain = None
    if args.plain > 0:
        pspecs = specs[:args.plain]
        batches = [E.Batch([sp], precision="fp64") for sp in pspecs]
        batches[0].launch(host_loop=True)
        batches[0].results()
        torch.cuda.synchronize()
        tp = time.perf_counter()
        pn = []
        for i, bt in enumerate(batches):
            bt.launch(seed_offset=2000 + i, host_loop=True)
            (k, v, _, _), = bt.results()
            pn.append((k, v))
        torch.cuda.synchronize()
        tplain = (time.perf_counter() - tp) / len(batches)
        plain = {"value": 1.0 / tplain, "ms_per_transform": 1e3 * tplain,
                 "nmse_mean": float(np.mean([E.nmse(k, v, sp) for (k, v), sp in zip(pn, pspecs)]))}
        log(f"plain fp64: {plain['value']:.1f} transforms/s", args.quiet)

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (one transform through the oracle)", args.quiet)
        cpu = cpu_baseline(args.seed, args.cpu_workers)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "transforms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 eval/dft, fp64 peel",
        "data": "synthetic",
        "config": {"workload": "q=4 n=50 S=1000 b=6 C=3 nso R=2 10dB", "batch": args.batch, "streams": args.streams,
                   "eval": eng.eval_mode,
                   "parallelism": "batched instances, 1 GPU"},
        "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)), "p99": 1e3 * float(np.percentile(lat, 99)),
                       "of": "a batch from launch to rows on the host, two batches in flight"},
        "plain_fp64": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        "nmse": {"mean": float(np.mean(nm)), "max": float(np.max(nm)), "n": len(nm)},
        "stage_ms_per_batch": stage_ms,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
