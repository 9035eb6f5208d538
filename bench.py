#!/usr/bin/env python3
"""q-SFT throughput on one MI355X: served FP32 engine vs plain FP64 path.

Headline workload: q=4, n=50, S=1000 uniform distinct support, |a| ~ U[1,10],
random phase, 10 dB SNR (noise_sd^2 = sum|a|^2 / 10), C=3 uniform M_c, b=6
(4096 bins), NSO delays with R=2 repeats (P=102 rows per group), nso1
detection, 15-pass cap.

`value` is the serving loop's throughput on that workload: `--steps x
--batch` distinct requests (each with its own support, strengths, M_c,
delays and noise seed; the requests carry their digits as int8 arrays) are
submitted back to back to a `qsft_amd.serving.Server` (admission by HBM
footprint, continuous batching of up to --batch requests over
--inflight streams, per-request delivery of the recovered rows into pinned
host memory, in the reference's dict order), the server polling after every
--batch submissions so that batches start while the burst is still being
submitted; timed from the first submit to the last result. A "step" is one batch of --batch requests. Latency p50/p99
come from a second, open-loop run with Poisson arrivals at 80 % of that rate
(per request, arrival -> result).

Reported beside it:
  * the resident relaunch engine (two resident batches relaunched with fresh
    noise; per-stage kernel times from HIP events; the roofline of the
    dominant kernel, the evaluation GEMM, timed with HIP events on its stream);
  * the plain single-instance FP64 path (same kernels in FP64, batch 1,
    host-synchronised peel loop) and the served/plain ratio;
  * NMSE against the planted spectrum, the reference decoder's own level at
    10 dB, and a served run at 16 dB where the NMSE <= 1e-3 gate holds;
  * the serving mix (configs[4] reading: the K3 uneven mix plus 5 % large
    q=4 n=100 S=5000 b=7 jobs): saturation and a Poisson rate sweep;
  * 256 uneven K3 instances in one batch, eager vs hipGraph replay;
  * the q=20 n=16 S=2000 b=3 (20 dB) stage table and its DFT bandwidth;
  * the CPU oracle (numpy restatement of the reference path, evaluation on a
    process pool as the reference does) over --cpu-seeds seeds.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "transforms/s + p99 latency @ NMSE≤1e-3: served FP32 vs plain FP64 GPU"
K1 = dict(q=4, n=50, S=1000, b=6, C=3, R=2, snr_db=10.0)
PEAK_F16_TFLOPS = 2500.0         # MI355X dense fp16 / bf16 MFMA, TFLOP/s
PEAK_HBM_GBS = 8000.0            # MI355X HBM3E, GB/s
# NMSE the reference decoder itself reaches on this workload (median over 8
# seeds of the CPU oracle, which matches the reference's golden vectors)
REFERENCE_NMSE_10DB = 4.6e-3


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40, help="batches of --batch requests in the timed serving run")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--inflight", type=int, default=4, help="batches on the GPU at once (one stream each)")
    ap.add_argument("--plain", type=int, default=3, help="plain FP64 transforms to time")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seeds", type=int, default=5)
    ap.add_argument("--cpu-workers", type=int, default=16,
                    help="process pool of the CPU oracle's evaluation (capped at the host's cores)")
    ap.add_argument("--no-extras", action="store_true", help="only the headline and the resident relaunch engine")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--fuse-from", default="3",
                    help="first peeling pass of the fused tail kernel; 'none': one detect/apply launch pair per pass "
                         "(resident relaunch engine only)")
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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(seed, workers, seeds):
    """K1 transforms through the CPU oracle, one per seed: decimal query
    indices as the reference builds them, evaluation on a process pool (as
    the reference's Pool), scipy fftn, the Python peeling loop. Median."""
    from multiprocessing import Pool

    from oracle import qsft_oracle as O
    from qsft_amd.engine import random_spec
    cores = max(1, min(workers, os.cpu_count() or 1))
    per = []
    nm = []
    with Pool(cores) as pool:
        for k in range(seeds):
            sp = random_spec(K1["q"], K1["n"], K1["S"], K1["b"], K1["C"], K1["R"], K1["snr_db"], seed=seed + k)
            qa = dict(query_method="complex", num_subsample=sp.C, b=sp.b, delays_method_source="identity",
                      delays_method_channel="nso", num_repeat=sp.R, subsampling_method="qsft")
            np.random.seed(seed + k)
            t0 = time.perf_counter()
            sig = O.SyntheticSignal(sp.n, sp.q, sp.locq.astype(np.int64), sp.strengths, sp.noise_sd, qa, pool=pool,
                                    faithful_indices=True)
            t1 = time.perf_counter()
            gw = O.qsft_transform(sig, sp.C, sp.R, sp.b, "identity", "nso")
            t2 = time.perf_counter()
            per.append((t1 - t0, t2 - t1))
            nm.append(O.nmse(gw, {tuple(c): v for c, v in zip(sp.locq.T, sp.strengths)}))
    tot = sorted(a + b for a, b in per)
    med = tot[len(tot) // 2]
    return {"value": 1.0 / med, "unit": "transforms/s", "cores": cores, "kind": "port",
            "sample": f"{seeds} transforms of the same workload (seeds {seed}..{seed + seeds - 1}), median; "
                      f"evaluation on a {cores}-process pool, decode single-threaded",
            "per_seed_s": [round(a + b, 3) for a, b in per],
            "construct_s_median": float(np.median([a for a, _ in per])),
            "decode_s_median": float(np.median([b for _, b in per])),
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(), "nmse_median": float(np.median(nm))}


def k1_specs(count, rng, snr=K1["snr_db"]):
    from qsft_amd.engine import random_spec
    q, n, S, b, C, R = (K1[k] for k in ("q", "n", "S", "b", "C", "R"))
    specs = [random_spec(q, n, S, b, C, R, snr, rng=rng) for _ in range(count)]
    for sp in specs:
        sp.wire()                                   # the request payload: int8 digits, complex128 strengths
    return specs


def mix_specs(count, rng, large=0.05):
    """The K3 uneven mix (q in {2, 3, 4, 20}, n in [8, 64], S log-uniform in
    [10, 2000], b = ceil(log_q S) + 1, C = 3, NSO R = 2, SNR U[5, 30] dB) with
    a `large` share of q=4 n=100 S=5000 b=7 jobs."""
    from qsft_amd.engine import random_spec
    specs = []
    for _ in range(count):
        if rng.uniform() < large:
            specs.append(random_spec(4, 100, 5000, 7, 3, 2, float(rng.uniform(5, 30)), rng=rng))
            continue
        q = int(rng.choice([2, 3, 4, 20]))
        n = int(rng.integers(8, 65))
        S = int(np.exp(rng.uniform(np.log(10), np.log(2000))))
        S = max(2, min(S, q ** n // 8))
        b = int(np.ceil(np.log(S) / np.log(q))) + 1
        specs.append(random_spec(q, n, S, b, 3, 2, float(rng.uniform(5, 30)), rng=rng))
    for sp in specs:
        sp.wire()
    return specs


def serve_closed(specs, batch, inflight, budget):
    """A burst of requests submitted back to back, the server polling after
    every `batch` submissions (host submission overlaps the GPU work of the
    batches already formed); transforms/s from first submit to last result."""
    import torch

    from qsft_amd import serving as SV
    srv = SV.Server(hbm_budget=budget, max_batch=batch, max_inflight=inflight)
    torch.cuda.synchronize()
    with SV.gc_paused():
        t0 = time.perf_counter()
        reqs = []
        for i, sp in enumerate(specs):
            reqs.append(srv.submit(sp))
            if (i + 1) % batch == 0:
                srv.poll()
                SV.collect_young()
        srv.drain()
        el = time.perf_counter() - t0
    return reqs, el, srv


def serve_open(specs, rate, batch, inflight, budget, rng, warm=256):
    """Poisson arrivals at `rate` requests/s replayed against the clock, after
    `warm` of the same requests served the same way on the same server (its
    streams and the allocators' pinned and device blocks are then warm)."""
    from qsft_amd import serving as SV
    srv = SV.Server(hbm_budget=budget, max_batch=batch, max_inflight=inflight)
    wspecs = specs[:min(warm, len(specs))]
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(specs)))
    with SV.gc_paused():                                           # (the clock read below allocates)
        srv.run_open_loop(np.cumsum(rng.exponential(1.0 / rate, size=len(wspecs))), wspecs)
        srv.batches = 0
        t0 = time.perf_counter()
        reqs = srv.run_open_loop(arrivals, specs)
        el = time.perf_counter() - t0
    lat = SV.latency_summary(reqs)
    return {"offered_per_s": rate, "achieved_per_s": len(specs) / el, "requests": len(specs),
            "p50_ms": lat["p50_ms"], "p99_ms": lat["p99_ms"], "batches": srv.batches,
            "mean_batch": len(specs) / max(srv.batches, 1)}


def resident_engine(args, E, torch):
    """Two resident batches relaunched with fresh noise (inputs stay in HBM):
    stage times, the GEMM roofline, per-batch latency."""
    q, n, S, b, C, R, snr = (K1[k] for k in ("q", "n", "S", "b", "C", "R", "snr_db"))
    rng = np.random.default_rng(args.seed)
    spec_sets = [[E.random_spec(q, n, S, b, C, R, snr, rng=rng) for _ in range(args.batch)] for _ in range(2)]
    fuse = None if args.fuse_from == "none" else int(args.fuse_from)
    engines = [E.Batch(sp, precision="fp32", fuse_from=fuse) for sp in spec_sets]
    eng = engines[0]
    for w in range(args.warmup):
        for e in engines:
            e.launch(seed_offset=w)
        for e in engines:
            e.unpack(e.fetch())
    torch.cuda.synchronize()
    steps = args.steps
    lat, launched, last = [], {}, {}
    k0 = "eval_prep"
    marks = [{k0: torch.cuda.Event(enable_timing=True), "eval": torch.cuda.Event(enable_timing=True)}
             for _ in range(steps)]

    def launch(s):
        engines[s % 2].launch(seed_offset=1000 + s, marks=marks[s])
        launched[s] = time.perf_counter()

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(min(2, steps)):
        launch(s)
    for s in range(steps):
        packed = engines[s % 2].fetch()
        lat.append(time.perf_counter() - launched[s])
        last[s % 2] = packed
        if s + 2 < steps:
            launch(s + 2)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    nm = []
    for slot, packed in last.items():
        res = engines[slot].unpack(packed)
        nm += [E.nmse(k, v, sp) for (k, v, _, _), sp in zip(res, spec_sets[slot])]
    gemm_ms = float(np.mean([m[k0].elapsed_time(m["eval"]) for m in marks]))
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
    samples = C * (R * (n + 1)) * q ** b                           # B = C * P * Q per transform
    flops = 8.0 * samples * S * args.batch                         # B*S complex multiply-adds per launch
    achieved = flops / (gemm_ms / 1e3) / 1e12
    from qsft_amd.kernels import mfma_mode, mfma_ptb
    nsp, npt = (S + 31) // 32, (R * (n + 1) + 7) // 8
    ptb = mfma_ptb(mfma_mode(q), 1, npt)
    npt_run = ptb * ((npt + ptb - 1) // ptb)
    # q in {2, 4}: v_smfmac_f32_16x16x64_f16, 32 supports x 16 queries x 16 columns per
    # instruction (one nonzero of each (Re W, Im W) k-row pair), two fp16 limbs of B
    mfma_f16 = 2 * 2.0 * C * q ** b * (32 * nsp) * (16 * npt_run) * args.batch
    roof = {"kernel": "k_eval_smfmac_occ2 (int8 MFMA phase + fp16x2 sparse-MFMA complex GEMM, FP32 accumulation)",
            "bound": "mfma", "achieved": achieved, "peak": PEAK_F16_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_F16_TFLOPS, "traffic": measured_traffic(args.batch),
            "algorithmic_flops_per_launch": flops, "launch_ms": gemm_ms,
            "mfma_f16_nonzero_executed_tflops": mfma_f16 / (gemm_ms / 1e3) / 1e12}
    out = {"value": args.batch * steps / elapsed, "ms_per_step": 1e3 * elapsed / steps,
           "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)), "p99": 1e3 * float(np.percentile(lat, 99)),
                          "of": "a batch from launch to rows on the host, two batches in flight"},
           "stage_ms_per_batch": stage_ms, "nmse": {"mean": float(np.mean(nm)), "max": float(np.max(nm))}}
    return out, roof, spec_sets[0]


def plain_fp64(specs, E, torch):
    batches = [E.Batch([sp], precision="fp64") for sp in specs]
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
    return {"value": 1.0 / tplain, "ms_per_transform": 1e3 * tplain,
            "nmse_mean": float(np.mean([E.nmse(k, v, sp) for (k, v), sp in zip(pn, specs)]))}


def k3_graph(torch, SV, quiet):
    """256 K3-mix instances in one batch: eager launch vs hipGraph replay."""
    rng = np.random.default_rng(77)
    specs = mix_specs(256, rng, large=0.0)
    dev = torch.device("cuda", torch.cuda.current_device())
    SV.Server(hbm_budget=1 << 40, device=dev).tune(specs)
    s, t = torch.cuda.Stream(device=dev), torch.cuda.Stream(device=dev)
    sb = SV.ServedBatch(specs, dev, stream=s, tail_stream=t)
    sb.launch(seed_offset=0)
    sb.wait()
    reps = 10
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in range(reps):
        sb.launch(seed_offset=1 + r)
    sb.wait()
    eager = (time.perf_counter() - t0) / reps
    sb.capture()
    sb.replay()
    sb.wait()
    t0 = time.perf_counter()
    for r in range(reps):
        sb.replay()
    sb.wait()
    graph = (time.perf_counter() - t0) / reps
    samples = sum(sp.C * sp.P * sp.Q for sp in specs)
    log(f"K3 x256: eager {1e3 * eager:.2f} ms, graph {1e3 * graph:.2f} ms", quiet)
    return {"instances": len(specs), "samples": int(samples), "eager_ms": 1e3 * eager, "graph_ms": 1e3 * graph,
            "transforms_per_s_graph": len(specs) / graph, "transforms_per_s_eager": len(specs) / eager,
            "shapes_qb": sorted({f"{sp.q}^{sp.b}" for sp in specs})}


def k2_stages(E, torch, quiet):
    """q=20 n=16 S=2000 b=3, NSO R=2, 20 dB: 16 transforms per batch."""
    rng = np.random.default_rng(20)
    specs = [E.random_spec(20, 16, 2000, 3, 3, 2, 20.0, rng=rng) for _ in range(16)]
    from qsft_amd import kernels as K
    from qsft_amd import _lib as L
    K.dft_autotune(20, 3, L.QSFT_F32, specs[0].C * specs[0].P)
    bt = E.Batch(specs, precision="fp32")
    bt.launch(seed_offset=1)
    bt.results()
    stage = {}
    reps = 5
    for rep in range(reps):
        ev = {}
        bt.launch(seed_offset=10 + rep, events=ev)
        res = bt.results(with_info=False)
        torch.cuda.synchronize()
        prev = ev["start"][0]
        for x in ("eval_prep", "eval", "dft", "assemble", "peel"):
            if x in ev:
                stage[x] = stage.get(x, 0.0) + prev.elapsed_time(ev[x][0]) / reps
                prev = ev[x][0]
    samples = sum(sp.C * sp.P * sp.Q for sp in specs)
    flops = 8.0 * sum(sp.C * sp.P * sp.Q * sp.S for sp in specs)
    nm = [E.nmse(k, v, sp) for (k, v, _, _), sp in zip(res, specs)]
    out = {"batch": len(specs), "stage_ms_per_batch": stage,
           "dft_GBps": 16.0 * samples / (stage["dft"] / 1e3) / 1e9,
           "eval_gemm_TFLOPs": flops / (stage["eval"] / 1e3) / 1e12,
           "transforms_per_s_stage_sum": len(specs) / (sum(stage.values()) / 1e3),
           "nmse_mean": float(np.mean(nm)), "dft_schedule": K.dft_schedule(20, 3, L.QSFT_F32)[0]}
    log(f"K2 stages: {stage}", quiet)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (transforms are independent: "
              "scale-out is one engine per GPU with no collective).", file=sys.stderr)
        sys.exit(2)
    import torch

    from qsft_amd import engine as E
    from qsft_amd import serving as SV

    if args.copy_calibration_mib > 0:
        nb = args.copy_calibration_mib << 20
        src = torch.ones(nb, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {nb} bytes read, {nb} bytes written", args.quiet)
        del src, dst

    budget = 200 << 30                                              # HBM the served requests may hold
    rng = np.random.default_rng(args.seed)
    nreq = args.steps * args.batch
    log(f"generating {nreq} + {2 * nreq} K1 requests", args.quiet)
    warm = k1_specs(args.warmup * args.batch, rng)
    sat_specs = k1_specs(nreq, rng)
    open_specs = k1_specs(2 * nreq, rng)                          # twice as long: a steadier p99

    # ---- headline: the serving loop at saturation, then open loop at 80 %
    serve_closed(warm, args.batch, args.inflight, budget)
    reqs, el, srv = serve_closed(sat_specs, args.batch, args.inflight, budget)
    value = nreq / el
    sat_lat = SV.latency_summary(reqs)
    nm = [E.nmse(r.result.k, r.result.values, sp) for r, sp in zip(reqs, sat_specs)]
    host_us = {k: 1e6 * v / nreq for k, v in srv.host_s.items()}
    log(f"served: {value:.1f} transforms/s ({srv.batches} batches); host us/request {host_us}", args.quiet)
    openl = serve_open(open_specs, 0.8 * value, args.batch, args.inflight, budget, rng)
    log(f"open loop at {0.8 * value:.0f}/s: p50 {openl['p50_ms']:.2f} ms p99 {openl['p99_ms']:.2f} ms", args.quiet)

    # ---- the resident relaunch engine (stages, roofline of the GEMM)
    engine, roof, pspecs = resident_engine(args, E, torch)
    log(f"resident relaunch: {engine['value']:.1f} transforms/s; stages {engine['stage_ms_per_batch']}", args.quiet)

    plain = plain_fp64(pspecs[:args.plain], E, torch) if args.plain > 0 else None
    if plain:
        log(f"plain fp64: {plain['value']:.1f} transforms/s", args.quiet)

    extras = {}
    if not args.no_extras:
        # NMSE gate: the same workload at 16 dB, where the reference decoder reaches <= 1e-3
        g_specs = k1_specs(16 * args.batch, np.random.default_rng(args.seed + 16), snr=16.0)
        greqs, gel, _ = serve_closed(g_specs, args.batch, args.inflight, budget)
        gnm = [E.nmse(r.result.k, r.result.values, sp) for r, sp in zip(greqs, g_specs)]
        extras["nmse_gate_16db"] = {"transforms_per_s": len(g_specs) / gel, "nmse_mean": float(np.mean(gnm)),
                                    "nmse_max": float(np.max(gnm)), "gate_met": bool(np.mean(gnm) <= 1e-3)}
        log(f"16 dB: {len(g_specs) / gel:.0f}/s, NMSE {np.mean(gnm):.2e}", args.quiet)
        # serving mix: K3 mix + 5 % large jobs
        mrng = np.random.default_rng(args.seed + 4)
        mix_all = mix_specs(1024, mrng)
        SV.Server(hbm_budget=budget).tune(mix_all)
        serve_closed(mix_all[:64], args.batch, args.inflight, budget)
        mreqs, mel, msrv = serve_closed(mix_all[64:576], args.batch, args.inflight, budget)
        msat = 512 / mel
        sweep = []
        pool = mix_all[576:]
        for frac in (0.25, 0.5, 0.75, 0.9):
            sub = pool[:max(64, int(frac * msat * 0.25))] if len(pool) else []
            sweep.append(dict(serve_open(sub, frac * msat, args.batch, args.inflight, budget, mrng), frac=frac))
        mlat = SV.latency_summary(mreqs)
        extras["serving_mix"] = {"saturation_per_s": msat, "saturation_latency_ms": mlat, "batches": msrv.batches,
                                 "sweep": sweep, "large_share": 0.05}
        log(f"mix: saturation {msat:.0f}/s; sweep {[(s['frac'], round(s['p99_ms'], 2)) for s in sweep]}", args.quiet)
        extras["k3_batch_256"] = k3_graph(torch, SV, args.quiet)
        extras["k2"] = k2_stages(E, torch, args.quiet)

    cpu = None
    if not args.no_cpu_baseline:
        log(f"cpu baseline ({args.cpu_seeds} transforms through the oracle)", args.quiet)
        cpu = cpu_baseline(args.seed, args.cpu_workers, args.cpu_seeds)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "transforms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 eval/dft, fp64 peel",
        "data": "synthetic",
        "config": {"workload": "q=4 n=50 S=1000 b=6 C=3 nso R=2 10dB, distinct requests through the serving loop",
                   "requests": nreq, "max_batch": args.batch, "inflight": args.inflight,
                   "parallelism": "continuous batching on 1 GPU"},
        "latency_ms": {"p50": openl["p50_ms"], "p99": openl["p99_ms"],
                       "of": "per request, arrival -> result in host memory, open-loop Poisson arrivals at 80% of "
                             "the saturation rate, after 256 warm-up requests at that rate; twice the saturation run's "
                             "request count",
                       "open_loop": openl, "saturation_p50": sat_lat["p50_ms"], "saturation_p99": sat_lat["p99_ms"]},
        "serving_host_us_per_request": host_us,
        "plain_fp64": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        "nmse": {"mean": float(np.mean(nm)), "max": float(np.max(nm)), "n": len(nm),
                 "reference_decoder_10db": REFERENCE_NMSE_10DB,
                 "gate_1e-3_met": bool(np.mean(nm) <= 1e-3),
                 "note": "at 10 dB the reference decoder itself reaches ~4.6e-3 (nso1, R=2); the gate holds from "
                         "~14 dB, see nmse_gate_16db"},
        "resident_relaunch": engine,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    out.update(extras)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
