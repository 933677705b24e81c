"""Benchmark: 3D SRRF radiality on one MI355X (full-volume workload).

Workload: T=200 volumes of 256x256x96 uint16 (2000 emitters, seed 2, PSF
sigma_xy 1.3 / sigma_z 2.6, 400 photons peak, background 100, Markov
blinking 0.3/0.1, Poisson noise), magnification 4 -> 1024x1024x384 FP32 SR
grid, ring radius 0.5 SR px, 26-direction sample sphere, TRAC order 2
(ITC 0), gradient smoothing and intensity weighting as in the shipped
SRRF.ini.  One step = one pass of the fused kernel over the whole series,
input resident in HBM.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints one JSON line.  value = SR voxels / s; roofline = algorithmic HBM
bytes per launch / measured kernel time against 8 TB/s (the metric asks
for %HBM), with the FP32-VALU roof beside it; cpu_baseline = the
multithreaded oracle on a bounded crop of the same series.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "3D SRRF SR-voxels/s + %HBM roofline; 3B batched logL+grad evals/s vs plain"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload_conf(srrf):
    return srrf.default_conf().replace(
        RingRadius=0.5, Magnification=4, Method=3, ITC=0, TRACOrder=2, GradientSmooth=1, MainRange=1,
        XRange=1, YRange=1, ZRange=1, RemoveCon=0, Renormalize=0, IntensityW=1)


def cpu_baseline(series, conf, smp, crop, threads):
    """Time the oracle on a crop [T, cz, cy, cx] of the same series."""
    from oracle import oracle
    cx, cy, cz = crop
    T, Z, Y, X = series.shape
    z0, y0, x0 = (Z - cz) // 2, (Y - cy) // 2, (X - cx) // 2
    sub = np.ascontiguousarray(series[:, z0:z0 + cz, y0:y0 + cy, x0:x0 + cx]).astype(np.float32)
    t0 = time.perf_counter()
    out = oracle.srrf_stack(sub, conf, smp.xyz, smp.divisor, nthreads=threads)
    dt = time.perf_counter() - t0
    n_sr = out.size
    return {"value": n_sr / dt, "unit": "SR-voxels/s", "cores": threads, "kind": "port",
            "sample": "oracle (FP32 transcription, %d threads) on a %dx%dx%d low-res crop x T=%d of the same series "
                      "(%d SR voxels, %.1f s); rate per SR voxel at full T" % (threads, cx, cy, cz, T, n_sr, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-crop", type=str, default="12,12,8")
    ap.add_argument("--frames", type=int, default=200)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run the plain copy kernel over this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_probe_mib:
        from rfba_amd import srrf
        nbytes = args.copy_probe_mib << 20
        ms = srrf.copy_probe(nbytes)
        print(json.dumps({"copy_probe_bytes": nbytes, "ms": ms, "GB/s": 2 * nbytes / ms / 1e6}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only; multi-GPU runs are not part of this build "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from rfba_amd import srrf, synth

    X, Y, Z, T = 256, 256, 96, args.frames
    n_em, seed = 2000, 2
    t0 = time.perf_counter()
    series = synth.srrf_series((X, Y, Z, T), n_em, seed)
    log("generated %s uint16 series in %.1f s" % (series.shape, time.perf_counter() - t0))
    conf = workload_conf(srrf)
    smp = srrf.directions(26, conf.RingRadius)
    M = conf.Magnification
    n_lr = X * Y * Z
    n_sr = n_lr * M ** 3

    d_in = srrf.DeviceBuffer.from_array(series)
    d_out = srrf.DeviceBuffer(n_sr * 4)
    kms = []
    for i in range(args.warmup):
        ms = srrf.srrf_stack_device(d_in, srrf.RFBA_U16, series.shape, conf, smp, d_out)
        log("warmup %d: kernel %.1f ms" % (i, ms))
    srrf.lib().rfba_device_synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        ms = srrf.srrf_stack_device(d_in, srrf.RFBA_U16, series.shape, conf, smp, d_out)
        kms.append(ms)
        log("step %d: kernel %.1f ms" % (i, ms))
    srrf.lib().rfba_device_synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    kernel_ms = float(np.mean(kms))

    out = d_out.to_array((Z * M, Y * M, X * M), np.float32)
    nan_frac = float(np.isnan(out).mean())

    alg_bytes = 2.0 * n_lr * T + 4.0 * n_sr
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9
    K = smp.n
    alg_flops = float(T) * n_sr * (K * 430 + 128 * conf.IntensityW + 6)
    tflops = alg_flops / (kernel_ms * 1e-3) / 1e12

    rec = {
        "metric": METRIC,
        "value": n_sr / (ms_per_step * 1e-3),
        "unit": "SR-voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded emitters, Markov blinking, Poisson noise; generated on the host)",
        "config": {"workload": "plain 3D SRRF, one stack: %dx%dx%d uint16, T=%d, %d emitters, M=%d -> %dx%dx%d SR, "
                               "R=0.5, 26 directions, TRAC order 2 (ITC 0), smoothing 1, IW 1"
                               % (X, Y, Z, T, n_em, M, X * M, Y * M, Z * M),
                   "sr_voxels": n_sr, "frames": T, "nan_fraction": nan_frac},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "srrf_fused<uint16,16>", "kernel_ms": kernel_ms,
                     "algorithmic_bytes_per_launch": alg_bytes,
                     "valu_fp32": {"achieved": tflops, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                                   "frac": tflops / FP32_PEAK_TFLOPS, "algorithmic_flops_per_launch": alg_flops}},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        crop = tuple(int(v) for v in args.cpu_crop.split(","))
        threads = synth.default_threads()
        rec["cpu_baseline"] = cpu_baseline(series, conf, smp, crop, threads)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
