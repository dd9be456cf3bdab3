#!/usr/bin/env python3
"""Gridding throughput of the MI355X device on one GPU.

Workload (the reference's single-pol gridding configuration): 1024x1024
grid, 1 polarization, 1 channel; 1,000,000 VisData<1> records (N(0,1) values,
weights in (0.5, 1.5], u, v uniform within +-0.45 of the grid extent, w = 0)
gridded through one 7x7-support separable Gaussian CF (sigma 1.5 px,
oversampling 20).  A step is one grid_visibilities call over the whole batch,
with the records already resident in HBM, into the complex<float> grid with
compensated FP32 accumulation (FP64 weights).  The same step on the
complex<double> grid is reported beside it, as are the finishing passes
(normalize_by_weights, FFT, shift) and the relative error of both grids
against the FP64 CPU grid after normalization and after FFT + shift.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s
FP32_PEAK_TFLOPS = 157.3       # MI355X FP32 vector
NUM_VIS = 1_000_000
GRID = 1024
CF_RADIUS, CF_OS = 3, 20


def algorithmic_bytes(num_vis, grid_bytes_per_pixel):
    """Compulsory HBM bytes of one gridding call: the records, the CF in the
    reference's 6-D layout, the grid read and written, the weights read and
    written."""
    rec = num_vis * 88
    support = 2 * CF_RADIUS + 1
    cf = (support + 4) ** 2 * CF_OS ** 2 * 8
    grid = GRID * GRID * grid_bytes_per_pixel
    return rec + cf + 2 * grid + 2 * 8


def algorithmic_flops(num_vis):
    support = 2 * CF_RADIUS + 1
    return num_vis * support * support * 8


def run_gpu(hpg, mode, dev_records, cf, steps, warmup):
    import torch
    gs = hpg.GridderState.create(hpg.Device.Hip, 0, NUM_VIS, cf, (GRID, GRID, 1, 1),
                                 (1.0, 1.0), [[0]], grid_mode=mode)
    gs.set_convolution_function(hpg.Device.Serial, cf)
    for _ in range(warmup):
        gs.grid_visibilities(hpg.Device.Serial, dev_records)
    gs.fence()
    torch.cuda.synchronize()
    gs.event_record(0)
    t0 = time.perf_counter()
    for _ in range(steps):
        gs.grid_visibilities(hpg.Device.Serial, dev_records)
    gs.event_record(1)
    gs.fence()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    dev_ms = gs.event_elapsed_ms(0, 1) / steps
    # finishing passes, each timed on the state's stream (the FFT plan is
    # created by a first, untimed transform)
    gs.apply_grid_fft(1.0, hpg.FFTSign.POSITIVE, True)
    gs.fence()
    gs.event_record(2)
    gs.normalize_by_weights(1.0)
    gs.event_record(3)
    gs.apply_grid_fft(1.0, hpg.FFTSign.POSITIVE, True)
    gs.event_record(4)
    gs.shift_grid(hpg.ShiftDirection.FORWARD)
    gs.event_record(5)
    gs.fence()
    fin = {"normalize_ms": gs.event_elapsed_ms(2, 3), "fft_ms": gs.event_elapsed_ms(3, 4),
           "shift_ms": gs.event_elapsed_ms(4, 5)}
    gs.release()
    return wall * 1e3 / steps, dev_ms, fin


def precision_grids(hpg, mode, dev_records, cf):
    """One call on a fresh grid; returns the normalized grid and the grid after
    FFT + shift (complex128, 1024x1024)."""
    gs = hpg.GridderState.create(hpg.Device.Hip, 0, NUM_VIS, cf, (GRID, GRID, 1, 1),
                                 (1.0, 1.0), [[0]], grid_mode=mode)
    gs.set_convolution_function(hpg.Device.Serial, cf)
    gs.grid_visibilities(hpg.Device.Serial, dev_records)
    gs.normalize_by_weights(1.0)
    g_norm = gs.grid_values()[0, 0]
    gs.apply_grid_fft(1.0, hpg.FFTSign.POSITIVE, True)
    gs.shift_grid(hpg.ShiftDirection.FORWARD)
    g_img = gs.grid_values()[0, 0]
    gs.release()
    return g_norm, g_img


def rel_err(got, ref):
    import numpy as np
    d = np.abs(got - ref)
    return {"e_max": float(d.max() / np.abs(ref).max()),
            "e_rms": float(np.linalg.norm(d) / np.linalg.norm(ref))}


def cpu_baseline_leg(records, cf, gpu_grids):
    """The OpenMP restatement of the reference's host gridding (oracle/,
    per-tap FP64 sin/cos, FP64 atomics), timed on this host's cores on the
    same 1e6-record batch; its FP64 grid is also the reference for the GPU
    grids' relative error."""
    import numpy as np
    from oracle import pyoracle
    threads = pyoracle.num_threads()
    recs = np.ascontiguousarray(records).view(pyoracle.visdata_dtype(1))
    groups = [(cf.extents(0), cf.values(0))]

    def one(openmp):
        o = pyoracle.Oracle((GRID, GRID, 1, 1), (1.0, 1.0), [[0]])
        o.set_cf(CF_OS, groups)
        t = time.perf_counter()
        o.grid_visibilities(recs, True, openmp=openmp)
        return time.perf_counter() - t, o

    one(True)  # warm-up
    times = []
    o = None
    while len(times) < 5 or (sum(times) < 10.0 and len(times) < 50):
        dt, o = one(True)
        times.append(dt)
    med = float(np.median(times))
    o.normalize_by_weights(1.0)
    ref_norm = o.grid_values()[0, 0]
    o.apply_fft(1.0, "POSITIVE")
    o.shift("FORWARD")
    ref_img = o.grid_values()[0, 0]
    prec = {}
    for name, (g_norm, g_img) in gpu_grids.items():
        prec[name] = {"after_normalize": rel_err(g_norm, ref_norm),
                      "after_fft_shift": rel_err(g_img, ref_img)}
    base = {"value": NUM_VIS / med, "unit": "vis/s", "cores": threads, "kind": "port",
            "sample": "the full 1,000,000-record batch, gridding stage, median of "
                      f"{len(times)} runs after 1 warm-up (OpenMP, {threads} threads)"}
    return base, prec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baseline and precision legs")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import hpg_amd as hpg
    from hpg_amd import workloads

    hpg.initialize(0)
    cf = workloads.gaussian_cf(CF_RADIUS, CF_OS, 1.5)
    records = workloads.config1_visibilities(NUM_VIS, GRID)
    dev_records = torch.from_numpy(records.view(np.uint8).copy()).cuda()
    torch.cuda.synchronize()

    results = {}
    for name, mode in [("fp32c", hpg.GridMode.FP32_COMPENSATED), ("fp64", hpg.GridMode.FP64)]:
        results[name] = run_gpu(hpg, mode, dev_records, cf, args.steps, args.warmup)

    wall_ms, dev_ms, fin = results["fp32c"]
    nbytes = algorithmic_bytes(NUM_VIS, 8)
    achieved = nbytes / (dev_ms * 1e-3) / 1e9
    flops = algorithmic_flops(NUM_VIS)
    out = {
        "metric": "gridded vis/s (FP32-compensated grid, 1 GPU)",
        "value": NUM_VIS / (wall_ms * 1e-3),
        "unit": "vis/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "grid_visibilities, 1024x1024x1x1 complex<float> grid, "
                               "1e6 VisData<1>, 7x7 Gaussian CF, oversampling 20",
                   "num_visibilities": NUM_VIS, "grid": [GRID, GRID, 1, 1],
                   "cf_support": 2 * CF_RADIUS + 1, "oversampling": CF_OS,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "grid_visibilities call (k_bin + k_tile4 + k_halo)",
                     "bytes_per_call": nbytes, "device_ms_per_call": dev_ms,
                     "fp32_frac": flops / (dev_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS},
        "finishing_ms": fin,
        "fp64_grid": {"value": NUM_VIS / (results["fp64"][0] * 1e-3),
                      "ms_per_step": results["fp64"][0],
                      "device_ms_per_call": results["fp64"][1],
                      "hbm_frac": algorithmic_bytes(NUM_VIS, 16)
                      / (results["fp64"][1] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                      "finishing_ms": results["fp64"][2]},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        grids = {"fp32c": precision_grids(hpg, hpg.GridMode.FP32_COMPENSATED, dev_records, cf),
                 "fp64": precision_grids(hpg, hpg.GridMode.FP64, dev_records, cf)}
        base, prec = cpu_baseline_leg(records, cf, grids)
        out["cpu_baseline"] = base
        out["precision_vs_fp64_cpu"] = prec
        out["speedup_vs_cpu"] = out["value"] / base["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
