"""Benchmark of the batched American FD pricer on one MI355X.

One step = one pass of the pricer over the whole batch with inputs resident in HBM:
32768 seeded American options (K U[50,150], T U[1/12,3], sigma U[0.1,0.6],
r U[0,0.1], q U[0,0.08], puts and calls, spot 100) on a 1024 x-point x 1024 t-point
grid, FP64, through kwgpu_fd1d_price_device (one kernel launch per step).

Prints one JSON line. `value` is options/s of the whole batch; `roofline` prices the
kernel's algorithmic FP64 work (10 flops per point and step) against the FP64 peak
using the kernel duration measured with HIP events on the launch stream;
`cpu_baseline` times the reference pricer itself (its sources compiled into
oracle/_ref, its thread pool on every hardware thread) on a bounded sample of the same
workload, with the oracle's restatement of it timed beside (or alone, where the
reference library was not built).

    python bench.py [--gpus 1] [--steps 10] [--warmup 2]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "kwinto-cuda_amd"))

N_OPTIONS = 32768
X_DIM = 1024
T_DIM = 1024
SEED = 1
FP64_PEAK_TFLOPS = 78.6          # MI355X FP64 (vector and dense matrix)
HBM_PEAK_GBS = 8000.0
ERR_BAR = 1e-4                   # max abs price error vs the FP64 FD price for a form to count
# published FP64 GPU throughput of the reference's own GPU solver at this shape
# (GTX 1070, batch 32768, 1024 x 1024), options/s
REFERENCE_GPU_FP64 = 949.0


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--layout", default="wave", choices=["wave", "lane", "auto"])
    ap.add_argument("--cpu-sample", type=int, default=8192,
                    help="options of the workload the CPU baseline prices (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--uneven", type=int, default=65536,
                    help="options of the uneven-grid batch measured beside the headline "
                         "(0 = skip)")
    ap.add_argument("--surrogate", type=int, default=1,
                    help="also train the premium ensemble on FD labels and time "
                         "surrogate-first serving (0 = skip)")
    ap.add_argument("--sur-train", type=int, default=2_000_000,
                    help="training options, each labelled at its two bracketing mesh nodes "
                         "(FD FP64 1024x1024); 5 %% of them are held out")
    ap.add_argument("--sur-serve", type=int, default=1 << 20)
    ap.add_argument("--sur-epochs", type=int, default=320)
    ap.add_argument("--sur-save", default="",
                    help="save the fitted 1024 x 1024 surrogate to this file")
    ap.add_argument("--sur-load", default="",
                    help="load the 1024 x 1024 surrogate from this file (saved by --sur-save "
                         "from this configuration) instead of fitting it; it is evaluated and "
                         "its gate calibrated as after a fit")
    ap.add_argument("--sur-po-calibrate", type=int, default=16 << 20,
                    help="calibration options of the per-option-grid gate")
    ap.add_argument("--sur-per-option", type=int, default=1,
                    help="also fit a surrogate to the per-option-grid FD pricer and serve "
                         "the batch on per-option grids")
    ap.add_argument("--next-rows", type=int, default=1,
                    help="also time FD1D-BS, a compressed chain book and Greeks (host entry "
                         "points, copies included)")
    ap.add_argument("--peak-probe", type=int, default=1,
                    help="run bin/kwgpu_peak after the timed work (measured FP64 FMA rate "
                         "and copy bandwidth)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the timed region, run one device copy of this size "
                         "(calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    return ap.parse_args()


def read_traffic():
    """HBM bytes per launch measured by rocprofv3 --pmc (profiles/fd1d_wave_traffic.md)."""
    path = os.path.join(ROOT, "profiles", "fd1d_wave_traffic.md")
    try:
        txt = open(path).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", txt)
    return float(m.group(1)) if m else None


def error_by_class(opts, d):
    """Max abs error per (put | call) x (T < 1 | T >= 1)."""
    out = {}
    for w, wn in ((-1, "put"), (1, "call")):
        for lo, hi, tn in ((0.0, 1.0, "T<1"), (1.0, np.inf, "T>=1")):
            m = (opts["w"] == w) & (opts["t"] >= lo) & (opts["t"] < hi)
            out[f"{wn} {tn}"] = float(d[m].max()) if m.any() else None
    return out


def device_ceilings(achieved_tflops):
    """Measured FP64 FMA throughput and copy bandwidth of this device (bin/kwgpu_peak,
    run as a child process after all timed work), beside the published figures."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kwinto-cuda_amd", "bin",
                       "kwgpu_peak")
    if not os.path.exists(exe):
        return {"error": "kwgpu_peak not built"}
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        return {"error": r.stderr.strip()[-300:]}
    m = json.loads(r.stdout.strip().splitlines()[-1])
    m["spec_fp64_tflops"] = FP64_PEAK_TFLOPS
    m["spec_hbm_gb_per_s"] = 8000.0
    m["fd_fp64_frac_of_measured_fma"] = achieved_tflops / m["fp64_fma_tflops"]
    return m


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(opts, sample, threads, gpu_prices):
    """The reference pricer itself (its sources compiled into oracle/_ref, its
    ThreadPool on every hardware thread) where that library was built, else the
    oracle's restatement of it on `threads` threads; the other one is timed beside it.

    A leg's time is accepted only with its prices checked: the reference's
    ThreadPool::wait can return before the last job has run
    (src/kwThreadPool.cpp:59-72), which would show as unset prices and a faster CPU.
    So the reference leg must equal the port bit for bit on every run, and both must
    match the GPU FP64 prices of the same options to 1e-8."""
    import oracle
    sub = opts[:sample]
    gpu = np.asarray(gpu_prices[:sample])

    def timed(price, reps):
        price(sub[:min(16, sample)])  # warm-up
        runs, outs = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            p, err = price(sub)
            runs.append(time.perf_counter() - t0)
            if err:
                raise RuntimeError(err)
            outs.append(p.copy())
        return float(np.median(runs)), runs, outs

    port = lambda s: oracle.fd1d_price(s, X_DIM, T_DIM, nthreads=threads)
    legs = [("port", port, threads, "the oracle's restatement, one job per PDE", 1)]
    if oracle.ref_available():
        ref = lambda s: oracle.ref_price("FD1D", s, X_DIM, T_DIM)
        legs.insert(0, ("reference", ref, oracle.ref_threads(),
                        "the reference's Fd1d_Pricer and ThreadPool, compiled from its sources",
                        3))
    out, first_prices = None, None
    for kind, price, cores, what, reps in legs:
        dt, runs, outs = timed(price, reps)
        for p in outs:
            d_gpu = float(np.max(np.abs(p - gpu)))
            if not d_gpu <= 1e-8:
                raise RuntimeError(f"cpu_baseline {kind} leg: max |cpu - gpu| = {d_gpu:.3e}")
            if first_prices is not None and not np.array_equal(p, first_prices):
                raise RuntimeError(f"cpu_baseline {kind} leg: prices differ from the "
                                   f"first leg's")
        if first_prices is None:
            first_prices = outs[0]
        rate = sample / dt
        leg = {"value": rate, "unit": "options/s", "cores": cores, "kind": kind,
               "sample": f"first {sample} of the {len(opts)} options of the same workload "
                         f"({100.0 * sample / len(opts):.0f} %), {X_DIM}x{T_DIM} FP64, "
                         f"{what}, {cores} threads, "
                         + (f"median of {reps} runs " if reps > 1 else "one run ")
                         + f"({', '.join(f'{r:.2f}' for r in runs)} s); every run's prices "
                         f"checked (equal to the other leg's bit for bit, within 1e-8 of the "
                         f"GPU's); the whole batch at this rate: "
                         f"{len(opts) / rate:.1f} s; host {cpu_model()}, "
                         f"{os.cpu_count()} logical CPUs visible"}
        if out is None:
            out = leg
        else:
            out["beside"] = leg
    return out


def bench_uneven(ctx, torch, stream, n, steps, log):
    """Uneven-grid batch: x-points per option in {256,512,1024,2048}, t-points
    max(32, ceil(365 T)), sigma U[0.05,0.8] with a 2% low-vol/long-maturity slice
    (seed 2); autotuned per bucket, captured in a hipGraph and replayed."""
    import kwgpu
    from kwgpu import workloads as W
    o, xd, td = W.uneven(n, 2)
    dev = torch.device("cuda:0")
    d_o = torch.from_numpy(o.view(np.uint8).copy()).to(dev)
    d_x = torch.from_numpy(xd).to(dev)
    d_t = torch.from_numpy(td).to(dev)
    d_p = torch.empty(n, dtype=torch.float64, device=dev)
    d_s = torch.empty(n, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    flops = float(np.sum(10.0 * xd.astype(np.float64) * (td - 1)))
    out = {"workload": f"{n} options, x-points in {{256,512,1024,2048}}, "
                       f"t-points max(32, ceil(365 T)), seed 2",
           "point_steps": float(np.sum(xd.astype(np.float64) * (td - 1)))}
    prices = {}
    for name, prec in (("fp64", kwgpu.FP64), ("fp32_mixed", kwgpu.FP32_MIXED)):
        prm = kwgpu.params(precision=prec, x_dims=d_x.data_ptr(), t_dims=d_t.data_ptr())
        tune = ctx.autotune(d_o.data_ptr(), n, d_p.data_ptr(), d_s.data_ptr(), prm)
        chosen = {b["bucket"]: b["candidates"][b["chosen"]] for b in tune["buckets"]}

        def label(cd):
            if cd["layout"] == "lane":
                return "lane" + (" (lower bound, pruned)" if cd.get("pruned") else "")
            return (f'{cd["rows_per_lane"]}x{cd["lanes"]}'
                    f'{" persistent" if cd["persistent"] else " static"}')
        # every candidate's time on the bucket's sample (ms), for the record
        trials = {str(b["bucket"]): {"sampled": b["sampled"],
                                     "ms": {label(cd): round(cd["ms"], 4) for cd in b["candidates"]}}
                  for b in tune["buckets"]}
        ctx.fd1d_price_device(d_o.data_ptr(), n, d_p.data_ptr(), d_s.data_ptr(), prm,
                              stream=stream.cuda_stream)  # plans and captures the graph
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            ctx.fd1d_price_device(d_o.data_ptr(), n, d_p.data_ptr(), d_s.data_ptr(), prm,
                                  stream=stream.cuda_stream)
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        if not bool((d_s == 0).all()):
            raise RuntimeError("uneven batch: failed options")
        prices[name] = d_p.cpu().numpy().copy()
        # FP32-mixed against the non-packed FP32 rate this build issues (packed FP32 is
        # switched off), which equals the FP64 vector rate
        peak = FP64_PEAK_TFLOPS
        out[name] = {"options_per_s": n / (ms * 1e-3), "ms_per_step": ms,
                     "tflops": flops / (ms * 1e-3) / 1e12,
                     "roofline_frac": flops / (ms * 1e-3) / 1e12 / peak,
                     "autotuned": {str(k): v for k, v in chosen.items()},
                     "autotune_trials": trials,
                     "bucket_schedule": tune.get("schedule")}
        log(f"bench: uneven {name} {ms:.3f} ms/step")
    d = np.abs(prices["fp32_mixed"] - prices["fp64"])
    out["fp32_mixed"]["max_abs_err_vs_fp64_fd"] = float(d.max())
    out["fp32_mixed"]["rms_abs_err_vs_fp64_fd"] = float(np.sqrt(np.mean(d * d)))
    return out


def bench_next_rows(ctx, opts, cpu_threads, log):
    """The rows beside the solve, through the host-pointer entry points (timings include
    the host<->device copies): FD1D-BS on the headline workload, a chain-structured
    book whose PDE compression leaves one solve per 10 options, and Greeks."""
    import kwgpu
    import oracle
    from kwgpu import workloads as W

    def timed(fn, reps):
        fn()  # warm-up (first call sizes buffers)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    out = {}
    n = len(opts)
    dt = timed(lambda: ctx.fd1d_bs_price(opts, X_DIM, T_DIM), 3)
    out["fd1d_bs"] = {"workload": f"{n} options (the headline batch), {X_DIM}x{T_DIM}, "
                                  "American + European FD legs and the BS kernel",
                      "ms_per_call": dt * 1e3, "options_per_s": n / dt}
    log(f"bench: FD1D-BS {dt * 1e3:.3f} ms/call")

    base = W.plain(600, 7)
    book = np.repeat(base, 10)
    book["s"] = np.tile(np.linspace(80.0, 125.0, 10), 600)
    book["k"] = np.tile(np.linspace(90.0, 115.0, 10)[::-1], 600)
    dt = timed(lambda: ctx.fd1d_price(book, 512, 512), 5)
    out["chain_book"] = {"workload": "6000 options = 600 PDEs x 10 (spot, strike) pairs, "
                                     "512x512 (the CLI default), seed 7",
                         "pdes": 600, "ms_per_call": dt * 1e3, "options_per_s": 6000 / dt}
    if cpu_threads:  # CPU baseline of this row: the oracle compresses the same way
        t0 = time.perf_counter()
        _, err = oracle.fd1d_price(book, 512, 512, nthreads=cpu_threads)
        cpu = time.perf_counter() - t0
        if err:
            raise RuntimeError(err)
        out["chain_book"].update({"cpu_port_ms": cpu * 1e3, "cpu_threads": cpu_threads,
                                  "speedup_vs_cpu": cpu / dt})
    log(f"bench: chain book {dt * 1e3:.3f} ms/call")

    sub = opts[:4096]
    dt = timed(lambda: ctx.fd1d_greeks(sub, X_DIM, T_DIM), 3)
    out["greeks"] = {"workload": f"first 4096 options of the headline batch, {X_DIM}x{T_DIM}, "
                                 "price + 8 central bumps of 1 %, spot bumps sharing the PDE",
                     "ms_per_call": dt * 1e3, "options_per_s": 4096 / dt}
    log(f"bench: Greeks {dt * 1e3:.3f} ms/call")
    return out


def time_surrogate_first(ctx, sur, d_opts, n, gate, steps, warmup, p64):
    """The surrogate-first form under `gate` (a gate dict of kwgpu.pipeline) on a batch
    already in HBM, timed like the FD forms:
    `warmup` untimed calls, then `steps` calls between two synchronizations (each call
    gates, runs the ensemble on the candidates, prices the rest by FP64 FD and scatters).
    Its prices are checked against the FP64 FD prices `p64` of the same options."""
    import torch
    import kwgpu
    from kwgpu import pipeline
    d_p = torch.empty(n, dtype=torch.float64, device="cuda")
    d_r = torch.empty(n, dtype=torch.int32, device="cuda")
    fd = kwgpu.params(X_DIM, T_DIM)
    g = pipeline.to_gate(gate)

    def call():
        return sur.price_gated(d_opts.data_ptr(), n, g, fd, d_p.data_ptr(), d_r.data_ptr())

    for _ in range(warmup):
        call()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        nfd = call()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    p, route = d_p.cpu().numpy(), d_r.cpu().numpy()
    d = np.abs(p - p64)
    return {"options_per_s": n / (ms * 1e-3), "ms_per_step": ms,
            "max_abs_err_vs_fp64_fd": float(d.max()),
            "rms_abs_err_vs_fp64_fd": float(np.sqrt(np.mean(d * d))),
            "served_by_surrogate": int(np.sum(route == 0)), "routed_to_fd": int(nfd)}


def bench_surrogate(ctx, args, log, head=None):
    """Surrogate training (2M options, FD FP64 labels at 1024 x 1024, 4 members) and
    surrogate-first serving of the 1M-option batch: on the 1024 x 1024 grid the surrogate
    was fitted to, with the FP64 and the FP32-mixed FD fallback; on per-option grids
    (x-points from {256, 512, 1024, 2048}, t-points max(32, ceil(365 T))) with a second
    ensemble fitted to the FD pricer on such grids. Each is checked against FP64 FD on the
    same grids. `head` = (device options, n, FP64 prices, steps, warmup): the headline
    batch, on which the surrogate-first form is timed as well (time_surrogate_first)."""
    import kwgpu
    from kwgpu import pipeline
    from kwgpu import workloads as W
    n_test = args.sur_train // 20  # 5 % of the labelled set held out
    fitted = kwgpu.Surrogate.load(ctx, args.sur_load) if args.sur_load else None
    sur, train = pipeline.train_and_evaluate(ctx, n_train=args.sur_train - n_test, n_test=n_test,
                                             epochs=args.sur_epochs, fitted=fitted, log=log)
    if args.sur_save:
        sur.save(args.sur_save)
    opts = W.serving(args.sur_serve, 4)
    gate = train["gate"]
    headline = None
    if head is not None:
        d_h, n_h, p64_h, steps, warmup = head
        headline = time_surrogate_first(ctx, sur, d_h, n_h, gate, steps, warmup, p64_h)
        log(f"bench: surrogate-first on the headline batch {headline['ms_per_step']:.3f} "
            f"ms/step, {headline['served_by_surrogate']} served, max err "
            f"{headline['max_abs_err_vs_fp64_fd']:.2e}")
    serve = pipeline.serve(ctx, sur, opts, gate, log=log)
    serve["meets_bar"] = serve["max_abs_err_vs_fp64_fd"] <= pipeline.ERR_BAR
    one = train["one_class_gate"]
    serve1 = pipeline.serve(ctx, sur, opts, pipeline.one_class_gate(one["spread_bound"],
                                                                     one["scale_bound"]),
                            repeats=1, log=log)
    serve1["meets_bar"] = serve1["max_abs_err_vs_fp64_fd"] <= pipeline.ERR_BAR
    serve32 = pipeline.serve(ctx, sur, opts, gate, log=log, precision=kwgpu.FP32_MIXED)
    serve32["meets_bar"] = serve32["max_abs_err_vs_fp64_fd"] <= pipeline.ERR_BAR
    # the trade-off at looser bars (not the metric: its bar is 1e-4)
    looser = {}
    for bar, v in train["looser_bars"].items():
        r = pipeline.serve(ctx, sur, opts, v["gate"], repeats=1, log=log)
        r["meets_bar"] = r["max_abs_err_vs_fp64_fd"] <= float(bar)
        looser[bar] = r
    sur.close()
    uneven = {}
    if args.sur_per_option:
        # the per-option gate rests on a 16M calibration set: its errors have a longer
        # tail than the one-grid ensemble's, and a 4M set once let 1.07e-4 through
        sur2, train2 = pipeline.train_and_evaluate(ctx, n_train=args.sur_train - n_test,
                                                   n_test=n_test, epochs=args.sur_epochs,
                                                   per_option_grids=True,
                                                   n_calibrate=args.sur_po_calibrate, log=log)
        xd, td = W.serving_grids(opts, 5)
        uneven["training"] = train2
        for name, prec in (("fp64", kwgpu.FP64), ("fp32_mixed", kwgpu.FP32_MIXED)):
            r = pipeline.serve(ctx, sur2, opts, train2["gate"], log=log, x_dims=xd, t_dims=td,
                               precision=prec)
            r["meets_bar"] = r["max_abs_err_vs_fp64_fd"] <= pipeline.ERR_BAR
            uneven[name + "_fallback"] = r
        sur2.close()
    return {"training": train, "headline_batch": headline, "serving": serve,
            "serving_one_class_gate": serve1, "serving_fp32_mixed_fallback": serve32,
            "serving_per_option_grids": uneven, "serving_at_looser_bars": looser}


def main():
    args = parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this pricer runs on one GPU; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import torch

    import kwgpu
    from kwgpu import workloads as W

    layout = {"wave": kwgpu.LAYOUT_WAVE_REG, "lane": kwgpu.LAYOUT_LANE_SOA,
              "auto": kwgpu.LAYOUT_AUTO}[args.layout]
    opts = W.plain(N_OPTIONS, SEED)
    dev = torch.device("cuda:0")
    d_opts = torch.from_numpy(opts.view(np.uint8).copy()).to(dev)
    d_prices = torch.empty(N_OPTIONS, dtype=torch.float64, device=dev)
    d_status = torch.empty(N_OPTIONS, dtype=torch.int32, device=dev)
    ctx = kwgpu.Context(0)
    # a stream of our own (default priority): kernels launched on it are bracketed by the
    # HIP events recorded on it (the legacy null stream would not order with them)
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.synchronize()

    def run_form(precision, steps, warmup):
        """Times `steps` launches of one form; returns (ms/step wall, kernel ms list, prices)."""
        prm = kwgpu.params(X_DIM, T_DIM, precision=precision, layout=layout)

        def step():
            ctx.fd1d_price_device(d_opts.data_ptr(), N_OPTIONS, d_prices.data_ptr(),
                                  d_status.data_ptr(), prm, stream=stream.cuda_stream)

        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        # HIP events around every launch on the launch stream
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev[0].record(stream)
        for i in range(steps):
            step()
            ev[i + 1].record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kms = [ev[i].elapsed_time(ev[i + 1]) for i in range(steps)]
        status = d_status.cpu().numpy()
        if not np.all(status == 0):
            raise RuntimeError(f"{int(np.sum(status != 0))} options failed")
        return wall * 1e3 / steps, kms, d_prices.cpu().numpy().copy()

    ms_per_step, kernel_ms, p64 = run_form(kwgpu.FP64, args.steps, args.warmup)
    print(f"bench: FP64 timed {args.steps} steps, {ms_per_step:.3f} ms/step", file=sys.stderr,
          flush=True)
    forms = {"fd_fp64": {"options_per_s": N_OPTIONS / (ms_per_step * 1e-3),
                         "ms_per_step": ms_per_step, "max_abs_err_vs_fp64_fd": 0.0}}
    if args.layout != "lane":
        ms_mx, kms_mx, pmx = run_form(kwgpu.FP32_MIXED, args.steps, 1)
        d = np.abs(pmx - p64)
        forms["fd_fp32_mixed"] = {
            "options_per_s": N_OPTIONS / (ms_mx * 1e-3), "ms_per_step": ms_mx,
            "kernel_ms": float(np.mean(kms_mx)),
            "max_abs_err_vs_fp64_fd": float(d.max()),
            "rms_abs_err_vs_fp64_fd": float(np.sqrt(np.mean(d * d))),
            "max_abs_err_by_class": error_by_class(opts, d),
            "speedup_vs_fp64": ms_per_step / ms_mx}
        print(f"bench: FP32-mixed {ms_mx:.3f} ms/step, max err {d.max():.2e}", file=sys.stderr,
              flush=True)

    if args.copy_probe_mib:
        nbytes = args.copy_probe_mib << 20
        a = torch.empty(nbytes // 8, dtype=torch.float64, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()

    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    extra = {}
    if args.surrogate:
        # the third form: the surrogate (fitted first, outside any timed region) in front
        # of FP64 FD, timed on this same batch; its 1M-option serving batch is reported
        # under extra.surrogate
        extra["surrogate"] = bench_surrogate(
            ctx, args, log, head=(d_opts, N_OPTIONS, p64, args.steps, args.warmup))
        hb = extra["surrogate"]["headline_batch"]
        forms["surrogate_first"] = dict(
            hb, speedup_vs_fp64=forms["fd_fp64"]["ms_per_step"] / hb["ms_per_step"],
            fallback="FP64 FD of the routed options (the fd_fp64 kernel)")

    # the metric's value: the fastest form on this workload within the error bar
    chosen = max((k for k in ("fd_fp64", "fd_fp32_mixed", "surrogate_first")
                  if k in forms and forms[k]["max_abs_err_vs_fp64_fd"] <= ERR_BAR),
                 key=lambda k: forms[k]["options_per_s"])
    mixed_chosen = chosen == "fd_fp32_mixed"
    value = forms[chosen]["options_per_s"]
    ms_value = forms[chosen]["ms_per_step"]
    # the dominant kernel: the FD march (for the surrogate-first form, the same FP64
    # kernel on the routed options; its launch over the whole batch is the one timed)
    avg_kernel_s = float(np.mean(kms_mx if mixed_chosen else kernel_ms)) * 1e-3
    flops = W.fd_flops_per_option(X_DIM, T_DIM) * N_OPTIONS
    achieved = flops / avg_kernel_s / 1e12
    traffic = read_traffic()

    if args.uneven > 0 and args.layout != "lane":
        extra["uneven"] = bench_uneven(ctx, torch, stream, args.uneven, max(1, args.steps // 2),
                                       log)
    if args.next_rows and args.layout != "lane":
        extra["next_rows"] = bench_next_rows(
            ctx, opts, min(args.cpu_threads, os.cpu_count() or 1) if args.cpu_sample > 0 else 0,
            log)

    if args.peak_probe:
        extra["device_ceilings"] = device_ceilings(achieved)

    cpu = None
    if args.cpu_sample > 0:
        threads = min(args.cpu_threads, os.cpu_count() or 1)
        print(f"bench: CPU baseline, {args.cpu_sample} options on {threads} threads",
              file=sys.stderr, flush=True)
        cpu = cpu_baseline(opts, args.cpu_sample, threads, p64)
        for f in forms.values():
            f["speedup_vs_cpu"] = f["options_per_s"] / cpu["value"]

    out = {
        "metric": "options/sec at max abs err <=1e-4 vs FP64 FD: plain FP64 vs FP32 vs surrogate",
        "value": value,
        "unit": "options/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_value,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / REFERENCE_GPU_FP64,
        "dtype": {"fd_fp32_mixed": "f32+f64", "surrogate_first": "f64+bf16"}.get(chosen, "f64"),
        "data": "synthetic (seeded SplitMix64 portfolio)",
        "config": {
            "workload": f"plain method: {N_OPTIONS} American options, {X_DIM} x-points x "
                        f"{T_DIM} t-points, FP64, Crank-Nicolson, seed {SEED}",
            "form": chosen,
            "layout": args.layout,
            "max_abs_err_vs_fp64_fd": forms[chosen]["max_abs_err_vs_fp64_fd"],
            "error_bar": ERR_BAR,
            "forms": forms,
        },
        "roofline": {
            "bound": "valu",
            "roof": "FP64 VALU (MI355X FP64 vector peak = FP64 dense peak = 78.6 TFLOP/s"
                    + ("; non-packed FP32 issues at the same rate)" if mixed_chosen else ")"),
            "achieved": achieved,
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": traffic,
            "kernel": ("fd1d_lane_f64" if args.layout == "lane" else
                       "fd1d_wave<M=16,L=64,W=1,PERSIST=false,MIXED=%s,GTH=false>"
                       % ("true" if mixed_chosen else "false")),
            "kernel_ms": avg_kernel_s * 1e3,
            "algorithmic_flops_per_launch": flops,
        },
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
