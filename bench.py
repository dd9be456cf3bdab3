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
(normalize_by_weights, FFT, shift), degridding of the same batch from a
complex<double> model, and the relative error of both grids against the
FP64 CPU grid after normalization and after FFT + shift.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints one JSON line.  Traffic (HBM bytes per call) comes from two counter
runs of the same workload, summarized into profiles/:

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d F -- python3 bench.py --pmc-workload
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d W -- python3 bench.py --pmc-workload
    python3 bench.py --pmc-summary F W > profiles/traffic_grid_1024sq_1e6vis.json

`--configs-only` runs only the multi-polarization workloads' legs (the
`configs` figures), for a kernel trace of them.
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
FIN_REPS = 20
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
    # finishing passes, each timed over FIN_REPS back-to-back calls on the
    # state's stream (the FFT plan is created by a first, untimed transform;
    # the values are not used afterwards)
    gs.apply_grid_fft(1.0, hpg.FFTSign.POSITIVE, True)
    gs.fence()
    fin = {}
    passes = [("normalize_ms", lambda: gs.normalize_by_weights(1.0)),
              ("fft_ms", lambda: gs.apply_grid_fft(1.0, hpg.FFTSign.POSITIVE, True)),
              ("shift_ms", lambda: gs.shift_grid(hpg.ShiftDirection.FORWARD))]
    for name, call in passes:
        gs.event_record(2)
        for _ in range(FIN_REPS):
            call()
        gs.event_record(3)
        gs.fence()
        fin[name] = gs.event_elapsed_ms(2, 3) / FIN_REPS
    # the three as one chain (normalize and the shift deferred into one pass
    # before the FFT for even grids; the chain runs when the next one starts
    # or the closing event is recorded)
    gs.event_record(2)
    for _ in range(FIN_REPS):
        gs.normalize_by_weights(1.0)
        gs.apply_grid_fft(1.0, hpg.FFTSign.POSITIVE, True)
        gs.shift_grid(hpg.ShiftDirection.FORWARD)
    gs.event_record(3)
    gs.fence()
    fin["chain_ms"] = gs.event_elapsed_ms(2, 3) / FIN_REPS
    gs.release()
    return wall * 1e3 / steps, dev_ms, fin


def run_grid_sizes(hpg, cf, steps, warmup):
    """The single-pol workload's FP32-compensated gridding call on larger
    grids (1e6 records, u, v uniform within +-0.45 of each extent): device
    ms per call on the state's stream and the HBM fraction of its
    algorithmic bytes, per grid edge."""
    import numpy as np
    import torch
    from hpg_amd import workloads
    out = {}
    for grid in (2048, 4096):
        rec = workloads.config1_visibilities(NUM_VIS, grid, seed=100 + grid)
        dev = torch.from_numpy(rec.view(np.uint8).copy()).cuda()
        torch.cuda.synchronize()
        gs = hpg.GridderState.create(hpg.Device.Hip, 0, NUM_VIS, cf, (grid, grid, 1, 1),
                                     (1.0, 1.0), [[0]],
                                     grid_mode=hpg.GridMode.FP32_COMPENSATED)
        gs.set_convolution_function(hpg.Device.Serial, cf)
        for _ in range(warmup):
            gs.grid_visibilities(hpg.Device.Serial, dev)
        gs.fence()
        gs.event_record(4)
        for _ in range(steps):
            gs.grid_visibilities(hpg.Device.Serial, dev)
        gs.event_record(5)
        gs.fence()
        ms = gs.event_elapsed_ms(4, 5) / steps
        nbytes = NUM_VIS * 88 + 387200 + 2 * grid * grid * 8
        out[f"{grid}x{grid}"] = {"device_ms_per_call": ms, "vis_per_s": NUM_VIS / (ms * 1e-3),
                                 "hbm_frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        gs.release()
        del dev
    return out


def run_degrid(hpg, dev_records, cf, model, steps, warmup):
    """Degridding of the same batch from a complex<double> model, as
    degrid_get_predicted_visibilities: the records come back with their
    predicted values (a device copy of the batch, rewritten in place), timed
    on the state's stream.  Returns device ms per call."""
    import ctypes
    import numpy as np
    import torch
    from hpg_amd import _lib
    gs = hpg.GridderState.create(hpg.Device.Hip, 0, NUM_VIS, cf, (GRID, GRID, 1, 1),
                                 (1.0, 1.0), [[0]], grid_mode=hpg.GridMode.FP32_COMPENSATED)
    gs.set_convolution_function(hpg.Device.Serial, cf)
    gs.set_model(hpg.Device.Serial, model)
    # the C ABI call of degrid_get_predicted_visibilities on device records,
    # into one preallocated device batch, calls queued back to back
    out = torch.empty_like(dev_records)
    torch.cuda.synchronize()
    lib = _lib.lib()
    src = ctypes.c_void_p(dev_records.data_ptr())
    dst = ctypes.c_void_p(out.data_ptr())

    def call():
        rc = lib.hpgx_grid_visibilities(gs._h, src, NUM_VIS, 1, 0, 1, 1, 0, dst, 1)
        if rc != 0:
            raise RuntimeError(f"hpgx_grid_visibilities failed ({rc})")

    for _ in range(warmup):
        call()
    gs.fence()
    gs.event_record(6)
    for _ in range(steps):
        call()
    gs.event_record(7)
    gs.fence()
    # the returned batch holds the predictions (the model is not zero)
    pred = out.cpu().numpy().view(hpg.visdata_dtype(1))["visibilities"]
    if not np.any(pred != 0):
        raise RuntimeError("degridding returned no predicted values")
    ms = gs.event_elapsed_ms(6, 7) / steps
    gs.release()
    return ms


def run_host_batches(hpg, records, cf, tasks, calls):
    """End-to-end gridding from host memory (the minigridder's timed loop,
    src/minigridder.cc:753-782): `calls` grid_visibilities calls back to back
    on the host batch, the upload of each call's records included, on a
    state with `tasks` added tasks (tasks > 0: each call's upload overlaps
    the previous calls' kernels).  Returns wall ms per call."""
    import torch
    gs = hpg.GridderState.create(hpg.Device.Hip, tasks, NUM_VIS, cf, (GRID, GRID, 1, 1),
                                 (1.0, 1.0), [[0]], grid_mode=hpg.GridMode.FP32_COMPENSATED)
    gs.set_convolution_function(hpg.Device.Serial, cf)
    for _ in range(2):
        gs.grid_visibilities(hpg.Device.Serial, records)
    gs.fence()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(calls):
        gs.grid_visibilities(hpg.Device.Serial, records)
    gs.fence()
    wall = (time.perf_counter() - t0) / calls
    gs.release()
    return wall * 1e3


# ---- the multi-polarization workloads, one timed call each -------------------

def _inside_support_sq(torch, rec_fields, radii, grid_size):
    """per record: S^2 of its CF group if its support lies inside the grid
    (grid_scale 1, frequency c: position u + N/2), else 0"""
    grp = rec_fields["cf_index"][:, 1].long()
    r = torch.as_tensor(radii, device=grp.device)[grp]
    s = 2 * r + 1
    ok = torch.ones_like(grp, dtype=torch.bool)
    for d in range(2):
        pos = rec_fields["uvw"][:, d] + grid_size[d] / 2.0
        corner = torch.round(pos).long() - r
        ok &= (corner >= 0) & (corner + s <= grid_size[d])
    return torch.where(ok, s * s, torch.zeros_like(s)).double()


def _cf_bytes(radii, os_, nmueller):
    return sum((2 * r + 1 + 4) ** 2 * os_ ** 2 * nmueller * 8 for r in radii)


def _time_call(gs, fn):
    import torch
    torch.cuda.synchronize()
    gs.fence()
    gs.event_record(8)
    fn()
    gs.event_record(9)
    gs.fence()
    return gs.event_elapsed_ms(8, 9) * 1e-3


def config_legs(hpg):
    """Device time, vis/s, FP32-issue and HBM fractions of the full-pol
    W-projection (4096^2 x 4 x 8, 64 groups, 5e7 VisData<4>), differentiable
    (2048^2 x 4, 16 groups, 2e7) and A/W-projection (8192^2 x 4 x 2, 512
    groups, 2.5e7) workloads, one timed call each after a warm-up call on
    the first 1e6 records.  FLOPs: 8 per (tap, vpol) gridded and 10
    degridded, over records whose support lies inside the grid; bytes: the
    records, the CF in the reference's layout, the grid read and written,
    the weights, the model read when degridding.  gather_tbps: the CF and
    model bytes the kernels gather per second (from L2 or HBM; per tap and
    Mueller term 8 bytes of CF, gridding and degridding each, and 16 bytes
    of model per tap and grid row degridded), the traffic that bounds
    these workloads, as against the compulsory bytes of hbm_frac."""
    import torch
    from hpg_amd import workloads
    out = {}
    M = workloads.MUELLER_FULL

    def fractions(flops, nbytes, sec, gathered):
        return {"fp32_frac": flops / sec / (FP32_PEAK_TFLOPS * 1e12),
                "hbm_frac": nbytes / sec / (HBM_PEAK_GBS * 1e9),
                "tflops": flops / sec / 1e12,
                "gather_tbps": gathered / sec / 1e12}

    # full-pol W-projection
    gsz = (4096, 4096, 4, 8)
    radii, chirps = workloads.config2_geometry()
    n = 50_000_000
    cf = workloads.taper_chirp_cf_device(20, radii, chirps, 16)
    rec = workloads.gaussian_uv_device_records(
        n, 4, gsz, lambda w: torch.clamp((w * 64).floor(), max=63), 0.2, ncube=8)
    fields = _record_fields(torch, rec, 4)
    s2 = float(_inside_support_sq(torch, fields, radii, gsz).sum())
    gs = hpg.GridderState.create(hpg.Device.Hip, 0, n, cf, gsz, (1.0, 1.0), M, M,
                                 grid_mode=hpg.GridMode.FP64)
    gs.set_convolution_function(hpg.Device.Serial, cf)
    del cf
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    shape = (8, 4, 4096, 4096)
    gs.set_model(hpg.Device.Serial, torch.complex(
        torch.randn(shape, generator=g, device="cuda", dtype=torch.float64),
        torch.randn(shape, generator=g, device="cuda", dtype=torch.float64)))
    gs.degrid_grid_visibilities(hpg.Device.Serial, rec[:1_000_000])
    gs.reset_grid()
    sec = _time_call(gs, lambda: gs.degrid_grid_visibilities(hpg.Device.Serial, rec))
    grid_b = 4096 * 4096 * 4 * 8 * 16
    flops = s2 * 16 * (8 + 10)
    nbytes = n * 120 + _cf_bytes(radii, 20, 16) + 3 * grid_b + 2 * 8 * 32
    out["fullpol_wproj_4096sq_5e7"] = dict(
        call="degrid_grid_visibilities (FP64 grid, model set)", records=n,
        device_s=sec, vis_per_s=4 * n / sec,
        **fractions(flops, nbytes, sec, s2 * (16 * 8 * 2 + 4 * 16)))
    gs.release()
    del rec, fields
    torch.cuda.empty_cache()

    # differentiable degridding: forward and adjoint
    gsz = (2048, 2048, 4, 1)
    radii = [4 + (12 * k) // 15 for k in range(16)]
    chirps = [0.002 * k for k in range(16)]
    n = 20_000_000
    Md = [[0, -1, -1, -1], [-1, 1, -1, -1], [-1, -1, 2, -1], [-1, -1, -1, 3]]
    cf = workloads.taper_chirp_cf_device(16, radii, chirps, 4)
    rec = workloads.gaussian_uv_device_records(
        n, 4, gsz, lambda w: torch.clamp((w * 16).floor(), max=15), 0.15, seed=61)
    s2 = float(_inside_support_sq(torch, _record_fields(torch, rec, 4), radii, gsz).sum())
    gs = hpg.GridderState.create(hpg.Device.Hip, 0, n, cf, gsz, (1.0, 1.0), Md, Md)
    gs.set_convolution_function(hpg.Device.Serial, cf)
    del cf
    shape = (1, 4, 2048, 2048)
    gs.set_model(hpg.Device.Serial, torch.complex(
        torch.randn(shape, generator=g, device="cuda", dtype=torch.float64),
        torch.randn(shape, generator=g, device="cuda", dtype=torch.float64)))
    gs.degrid_forward(rec[:1_000_000])
    fwd = _time_call(gs, lambda: gs.degrid_forward(rec))
    resid = torch.randn((n, 4), dtype=torch.complex128, device="cuda")
    gs.degrid_adjoint(rec[:1_000_000], resid[:1_000_000])
    adj = _time_call(gs, lambda: gs.degrid_adjoint(rec, resid))
    grid_b = 2048 * 2048 * 4 * 16
    out["differentiable_2048sq_2e7"] = dict(
        forward=dict(call="degrid_forward (FP64 predictions)", records=n, device_s=fwd,
                     vis_per_s=4 * n / fwd,
                     **fractions(s2 * 4 * 10, n * (120 + 64) + _cf_bytes(radii, 16, 4)
                                 + grid_b, fwd, s2 * (4 * 8 + 4 * 16))),
        adjoint=dict(call="degrid_adjoint (FP64 grid)", records=n, device_s=adj,
                     vis_per_s=4 * n / adj,
                     **fractions(s2 * 4 * 8, n * (120 + 64) + _cf_bytes(radii, 16, 4)
                                 + 2 * grid_b, adj, s2 * 4 * 8)))
    gs.release()
    del rec, resid
    torch.cuda.empty_cache()

    # A/W-projection, one GPU's share (2 of 16 channels)
    gsz = (8192, 8192, 4, 2)
    radii, chirps = workloads.config5_geometry()
    n = 25_000_000
    cf = workloads.taper_chirp_cf_device(10, radii, chirps, 16, phase_screen=True)
    pa = torch.randint(0, 16, (n,), generator=g, device="cuda")
    rec = workloads.gaussian_uv_device_records(
        n, 4, gsz, lambda w: torch.clamp((w * 32).floor(), max=31).long() * 16 + pa, 0.25,
        weight_range=(1e-6, 1.0), ncube=2, seed=52)
    s2 = float(_inside_support_sq(torch, _record_fields(torch, rec, 4), radii, gsz).sum())
    gs = hpg.GridderState.create(hpg.Device.Hip, 0, n, cf, gsz, (1.0, 1.0), M, M,
                                 grid_mode=hpg.GridMode.FP32_COMPENSATED)
    gs.set_convolution_function(hpg.Device.Serial, cf)
    del cf
    torch.cuda.empty_cache()
    shape = (2, 4, 8192, 8192)
    gs.set_model(hpg.Device.Serial, torch.complex(
        torch.randn(shape, generator=g, device="cuda", dtype=torch.float64),
        torch.randn(shape, generator=g, device="cuda", dtype=torch.float64)))
    torch.cuda.empty_cache()
    gs.degrid_grid_visibilities(hpg.Device.Serial, rec[:1_000_000])
    gs.reset_grid()
    sec = _time_call(gs, lambda: gs.degrid_grid_visibilities(hpg.Device.Serial, rec))
    grid_b = 8192 * 8192 * 4 * 2 * 8
    flops = s2 * 16 * (8 + 10)
    nbytes = n * 120 + _cf_bytes(radii, 10, 16) + 2 * grid_b + 8192 * 8192 * 8 * 16
    out["awproj_8192sq_2ch_2.5e7"] = dict(
        call="degrid_grid_visibilities (FP32-compensated grid, FP64 model)", records=n,
        device_s=sec, vis_per_s=4 * n / sec,
        **fractions(flops, nbytes, sec, s2 * (16 * 8 * 2 + 4 * 16)))
    gs.release()
    del rec
    torch.cuda.empty_cache()
    return out


def _record_fields(torch, rec, npol):
    """typed views of the cf_index and uvw fields of device records"""
    dt = __import__("hpg_amd").visdata_dtype(npol)
    oc = dt.fields["cf_index"][1]
    ou = dt.fields["uvw"][1]
    return {"cf_index": rec[:, oc:oc + 8].contiguous().view(torch.int32),
            "uvw": rec[:, ou:ou + 24].contiguous().view(torch.float64)}


def degrid_bytes(num_vis):
    """Compulsory HBM bytes of one degridding call: the records read and
    written back whole (with their predicted values), the CF and the FP64
    model read once."""
    support = 2 * CF_RADIUS + 1
    cf = (support + 4) ** 2 * CF_OS ** 2 * 8
    return num_vis * (88 + 88) + cf + GRID * GRID * 16


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


# HBM traffic per call from rocprofv3 counters (FETCH_SIZE and WRITE_SIZE,
# one run each, calibrated by a plain copy of known size); written by
# --pmc-summary, read by the default run
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_grid_1024sq_1e6vis.json")
CAL_BYTES = 256 << 20


def pmc_workload(hpg, dev_records, cf):
    """The program profiled by the counter runs: a 256 MiB device copy
    (calibration), then 3 gridding calls into each grid type and 3
    degridding calls of the batch."""
    import torch
    a = torch.ones(CAL_BYTES, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    for mode in (hpg.GridMode.FP32_COMPENSATED, hpg.GridMode.FP64):
        gs = hpg.GridderState.create(hpg.Device.Hip, 0, NUM_VIS, cf, (GRID, GRID, 1, 1),
                                     (1.0, 1.0), [[0]], grid_mode=mode)
        gs.set_convolution_function(hpg.Device.Serial, cf)
        for _ in range(3):
            gs.grid_visibilities(hpg.Device.Serial, dev_records)
        gs.fence()
        gs.release()
    # and 3 degridding calls of the batch (records returned), as run_degrid
    import numpy as np
    rng = np.random.default_rng(7)
    model = (rng.standard_normal((1, 1, GRID, GRID))
             + 1j * rng.standard_normal((1, 1, GRID, GRID)))
    run_degrid(hpg, dev_records, cf, model, 3, 0)
    print(json.dumps({"pmc_workload": "done", "calibration_bytes": CAL_BYTES}), flush=True)


def read_counters(directory):
    """{kernel name: [per-dispatch value]} from a rocprofv3 counter CSV."""
    import csv
    import glob
    files = glob.glob(os.path.join(directory, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection.csv under {directory}")
    out = {}
    for f in files:
        for row in csv.DictReader(open(f)):
            name = row.get("Kernel_Name", "")
            out.setdefault(name, []).append(float(row["Counter_Value"]))
    return out


def short_name(kernel):
    import re
    k = kernel.replace("void ", "").replace("hpgx::", "").replace("(anonymous namespace)::", "")
    return re.sub(r"\(.*", "", k).strip()


def pmc_summary(fetch_dir, write_dir):
    """Per-call HBM traffic of each grid type's gridding kernels, in bytes:
    counter KiB x 1024 x (calibration bytes / what the counter reads for
    the calibration copy)."""
    import numpy as np
    res = {"calibration_bytes": CAL_BYTES, "kernels": {}}
    for counter, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        vals = read_counters(d)
        cal = max(v for k, vs in vals.items() if "copyBuffer" in k for v in vs)
        factor = CAL_BYTES / (cal * 1024.0)
        res[counter + "_calibration_kib"] = cal
        res[counter + "_factor"] = factor
        for k, vs in vals.items():
            if "copyBuffer" in k or not short_name(k).startswith("k_"):
                continue
            e = res["kernels"].setdefault(short_name(k), {})
            e[counter + "_bytes"] = float(np.median(vs)) * 1024.0 * factor
            e["dispatches"] = len(vs)
    per_call = {"fp32c": 0.0, "fp64": 0.0}
    for k, e in res["kernels"].items():
        b = e.get("FETCH_SIZE_bytes", 0.0) + e.get("WRITE_SIZE_bytes", 0.0)
        if k.startswith("k_bin"):
            per_call["fp32c"] += b
            per_call["fp64"] += b
        elif (k.startswith("k_sub<") or k.startswith("k_quad<") or k.startswith("k_tile4<1, 1")
              or k.startswith("k_halo<HIP_vector_type<float")):
            per_call["fp32c"] += b
        elif k.startswith("k_tile4<1, 0") or k.startswith("k_halo<HIP_vector_type<double"):
            per_call["fp64"] += b
    res["per_call_bytes"] = per_call
    print(json.dumps(res, indent=1), flush=True)


def traffic_per_call(grid):
    """Measured HBM bytes per call for 'fp32c' or 'fp64', or None."""
    try:
        t = json.load(open(TRAFFIC_FILE))
        return t["per_call_bytes"][grid]
    except (OSError, KeyError, ValueError):
        return None


def copy_bandwidth():
    """Achievable HBM bandwidth: a 2 GiB device-to-device copy, GB/s
    (read + write bytes / time)."""
    import torch
    n = 2 << 30
    a = torch.ones(n, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    del a, b
    torch.cuda.empty_cache()
    return 2 * n / (ms * 1e-3) / 1e9


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
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the multi-polarization workloads' legs")
    ap.add_argument("--pmc-workload", action="store_true",
                    help="run the program profiled by the counter runs and exit")
    ap.add_argument("--configs-only", action="store_true",
                    help="run only the multi-polarization workloads' legs (one JSON "
                         "line of their figures) and exit")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="summarize the two counter runs' CSVs and exit")
    args = ap.parse_args()
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import hpg_amd as hpg
    from hpg_amd import workloads

    hpg.initialize(0)
    if args.configs_only:
        print(json.dumps({"configs": config_legs(hpg)}), flush=True)
        return
    cf = workloads.gaussian_cf(CF_RADIUS, CF_OS, 1.5)
    records = workloads.config1_visibilities(NUM_VIS, GRID)
    dev_records = torch.from_numpy(records.view(np.uint8).copy()).cuda()
    torch.cuda.synchronize()
    if args.pmc_workload:
        pmc_workload(hpg, dev_records, cf)
        return
    copy_gbs = copy_bandwidth()

    results = {}
    for name, mode in [("fp32c", hpg.GridMode.FP32_COMPENSATED), ("fp64", hpg.GridMode.FP64)]:
        results[name] = run_gpu(hpg, mode, dev_records, cf, args.steps, args.warmup)
    # the FP32 grid with FP64 tile sums instead of the two-sum tiles
    # (HPGX_FP32_TILE is read when a state is created)
    prev = os.environ.get("HPGX_FP32_TILE")
    os.environ["HPGX_FP32_TILE"] = "fp64"
    try:
        results["fp32_fp64_tiles"] = run_gpu(hpg, hpg.GridMode.FP32_COMPENSATED, dev_records,
                                             cf, max(1, args.steps // 4),
                                             max(1, args.warmup // 4))
    finally:
        if prev is None:
            os.environ.pop("HPGX_FP32_TILE")
        else:
            os.environ["HPGX_FP32_TILE"] = prev

    e2e = {f"tasks_{t}": run_host_batches(hpg, records, cf, t, 10) for t in (0, 2)}
    sizes = run_grid_sizes(hpg, cf, max(1, args.steps // 4), max(1, args.warmup // 4))

    rng = np.random.default_rng(7)
    model = (rng.standard_normal((1, 1, GRID, GRID))
             + 1j * rng.standard_normal((1, 1, GRID, GRID)))
    degrid_ms = run_degrid(hpg, dev_records, cf, model, max(1, args.steps // 4),
                           max(1, args.warmup // 4))

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
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_per_call("fp32c"),
                     "copy_kernel_gbs": copy_gbs,
                     "kernel": "grid_visibilities call (k_bin + k_sub + k_halo)",
                     "bytes_per_call": nbytes, "device_ms_per_call": dev_ms,
                     "fp32_frac": flops / (dev_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS},
        "finishing_ms": fin,
        "fp64_grid": {"value": NUM_VIS / (results["fp64"][0] * 1e-3),
                      "ms_per_step": results["fp64"][0],
                      "device_ms_per_call": results["fp64"][1],
                      "hbm_frac": algorithmic_bytes(NUM_VIS, 16)
                      / (results["fp64"][1] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                      "traffic": traffic_per_call("fp64"),
                      "finishing_ms": results["fp64"][2]},
        "fp32_grid_fp64_tiles": {
            "value": NUM_VIS / (results["fp32_fp64_tiles"][0] * 1e-3),
            "device_ms_per_call": results["fp32_fp64_tiles"][1],
            "hbm_frac": nbytes / (results["fp32_fp64_tiles"][1] * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "degrid": {"vis_per_s": NUM_VIS / (degrid_ms * 1e-3),
                   "device_ms_per_call": degrid_ms,
                   "bytes_per_call": degrid_bytes(NUM_VIS),
                   "hbm_frac": degrid_bytes(NUM_VIS) / (degrid_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                   "kernel": "k_bin (also writes the returned batch) + k_degrid4"},
        "fp32c_larger_grids": sizes,
        "end_to_end": {
            "what": "grid_visibilities from host memory, records uploaded each call "
                    "(10 calls back to back), FP32-compensated grid",
            **{k: {"ms_per_call": v, "vis_per_s": NUM_VIS / (v * 1e-3)}
               for k, v in e2e.items()}},
        "cpu_baseline": None,
    }
    if not args.no_configs:
        out["configs"] = config_legs(hpg)
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
