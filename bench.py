"""Benchmark: RFBA's 3D SRRF radiality stage and 3B likelihood core on one MI355X.

Headline workload (value, roofline, cpu_baseline): plain 3D SRRF over one
full-volume stack -- T=200 volumes of 256x256x96 uint16 (2000 emitters,
seed 2, PSF sigma_xy 1.3 / sigma_z 2.6, 400 photons peak, background 100,
Markov blinking 0.3/0.1, Poisson noise), magnification 4 -> 1024x1024x384
FP32 SR grid, ring radius 0.5 SR px, 26-direction sample sphere, TRAC
order 2 (ITC 0), gradient smoothing and intensity weighting on.  One step =
one pass of the fused kernel over the whole series, input resident in HBM.

Beside it, in "workloads" of the same JSON line:
  srrf_batched  512x512x128 T=120 series (8000 emitters, seed 3) split into
                29^3 blocks with 3-voxel overlap (doSplit.m), plus 64 extra
                stacks (edges 8-29, T 30-200, seed 4), magnification per block
                from {2,3,4} (seed 5): one persistent launch in a hipGraph
                with autotuned tile depth vs one launch per block;
  b3_single     29^3 block, T=100, 40 spots (seed 6), 256 perturbed
                parameter sets (seed 7): logL + adjoint for every spot,
                batched vs one launch per parameter set;
  b3_batched    1200 blocks of the same split (seed 8), K ~ Poisson(40)
                clipped to 5-120, T=120, 32 parameter sets per block,
                checkpoint stride 11 in the model (at T=120 the HMM runs
                one wave per item and stores no messages): batched vs one
                launch per block (first 100) and per (block, parameter
                set) (first 10);
  b3_functor    the CG functor of one spot of that block with 10 background
                samples, created once and called per conjugate-gradient step
                (vs the per-call re-upload path), and its FFBS Hessian with
                1000 inner samples;
  b3_gibbs      Gibbs sampling of the blink states of the same 1200 blocks
                (true spot parameters): 10 samples of one sweep per run,
                all blocks in one launch vs one launch per block (first 20);
  prep          image normalisation of the 512x512x128 T=120 series (per
                z-slice high-pass, mean/variance, scaling), and the TIFF path
                end to end on 8 of its frames.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--only c2,c4,functor,c5,gibbs,c3,prep]

Prints one JSON line.  value = SR voxels / s of the headline workload;
roofline = algorithmic FP32 flops per launch / measured kernel time against
MI355X's FP32 peak (the kernel is VALU-bound), with the HBM roof (algorithmic
bytes against 8 TB/s) beside it; cpu_baseline = the multithreaded oracle on a
bounded crop of the same series.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "3D SRRF SR-voxels/s + %HBM roofline; 3B batched logL+grad evals/s vs plain"
# HBM bytes per launch of srrf_fused on the headline workload (T=200), from
# rocprofv3 FETCH_SIZE and WRITE_SIZE scaled by the copy-kernel calibration
# (profiles/srrf_fused_full_volume_256x256x96_T200.md).  Dominated by the TRAC
# per-voxel series written and read back (8 B per SR voxel-frame).
HEADLINE_TRAFFIC_BYTES = 9.482e11
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload_conf(srrf):
    return srrf.default_conf().replace(
        RingRadius=0.5, Magnification=4, Method=3, ITC=0, TRACOrder=2, GradientSmooth=1, MainRange=1,
        XRange=1, YRange=1, ZRange=1, RemoveCon=0, Renormalize=0, IntensityW=1)


def hbm_roof(alg_bytes, kernel_ms):
    gbs = alg_bytes / (kernel_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}


def srrf_flops_per_voxel_frame(n_samples, intensity_w, trac):
    """Algorithmic FP32 flops of one SR voxel-frame of SRRF: 430 per sample
    direction (the gradient interpolation and the radiality term), 128 for the
    intensity weight's interpolation when it is on, and the temporal
    statistic: 6 for TRAC, 4 for the mean (TRA)."""
    return n_samples * 430 + 128 * intensity_w + (6 if trac else 4)


SRRF_ROOF_NOTE = ("FP32 VALU-bound (no matrix instructions; the FP32 peak is the same on the vector units and "
                  "the matrix cores); algorithmic flops count every SR voxel-frame, with the interpolation shared "
                  "by a voxel's samples once; the kernel skips voxels whose value is NaN by definition (a sample "
                  "out of range), the fraction given beside")


# ---------------------------------------------------------------- headline: one full-volume stack
def host_cpu(threads):
    """What the host offers and what a CPU leg used: nproc (os.cpu_count),
    the affinity mask's size, std::thread::hardware_concurrency and the
    threads used (at most 16: the CPU share one GPU's job is given on the
    measured box)."""
    from oracle import oracle
    try:
        aff = len(os.sched_getaffinity(0))
    except Exception:
        aff = None
    return {"nproc": os.cpu_count(), "affinity": aff,
            "hardware_concurrency": int(oracle.lib().oracle_hardware_threads()), "threads_used": threads}


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
    return {"value": n_sr / dt, "unit": "SR-voxels/s", "cores": threads, "kind": "port", "host": host_cpu(threads),
            "sample": "oracle (FP32 transcription, %d threads) on a %dx%dx%d low-res crop x T=%d of the same series "
                      "(%d SR voxels, %.1f s); rate per SR voxel at full T" % (threads, cx, cy, cz, T, n_sr, dt)}


def srrf_accuracy(series, conf, smp, crop, frames, threads):
    """Accuracy of the GPU kernel on a crop [frames, cz, cy, cx] of the series
    (its first frames), run as a stack of its own: against the FP32 oracle (the bar: bit-identical,
    NaN masks equal) and against the FP64 re-evaluation (max-normalised error
    over finite voxels), and the number of (voxel, frame, sample) triples
    whose converging/diverging test differs between FP32 and FP64."""
    from oracle import oracle
    from rfba_amd import srrf
    cx, cy, cz = crop
    T, Z, Y, X = series.shape
    z0, y0, x0 = (Z - cz) // 2 + 3, (Y - cy) // 2 + 5, (X - cx) // 2 + 7
    sub = np.ascontiguousarray(series[:frames, z0:z0 + cz, y0:y0 + cy, x0:x0 + cx])
    T = sub.shape[0]
    g = srrf.srrf_stack(sub, conf, smp)
    o32 = oracle.srrf_stack(sub.astype(np.float32), conf, smp.xyz, smp.divisor, nthreads=threads)
    o64 = oracle.srrf_stack(sub.astype(np.float64), conf, smp.xyz.astype(np.float64), smp.divisor, f64=True,
                            nthreads=threads)
    flips = oracle.gdotr_flips(sub, conf, smp.xyz, nthreads=threads)
    fin = np.isfinite(o32)
    scale = float(np.abs(o64[fin]).max()) if fin.any() else 0.0
    same_nan = bool(np.array_equal(np.isnan(g), np.isnan(o32)))
    rel32 = float(np.max(np.abs(g[fin].astype(np.float64) - o32[fin]) / np.maximum(np.abs(o32[fin]), 1e-30))) \
        if fin.any() else 0.0
    return {"crop_lowres_xyz": [cx, cy, cz], "frames": T, "sr_voxels": int(g.size), "finite_voxels": int(fin.sum()),
            "nan_masks_equal": same_nan, "bit_identical_to_fp32_oracle": bool(np.array_equal(g, o32, equal_nan=True)),
            "max_rel_err_vs_fp32_oracle": rel32,
            "gpu_err_vs_fp64_max_normalised": float(np.abs(g[fin] - o64[fin]).max() / scale) if scale else 0.0,
            "fp32_oracle_err_vs_fp64_max_normalised": float(np.abs(o32[fin] - o64[fin]).max() / scale)
            if scale else 0.0,
            "gdotr_sign_flips_fp32_vs_fp64": flips, "gdotr_tests": int(g.size) * T * smp.n}


def bench_stack(args, rec, lock):
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
    d_in.close()
    d_out.close()
    srrf.lib().rfba_release_cache()

    alg_bytes = 2.0 * n_lr * T + 4.0 * n_sr
    K = smp.n
    alg_flops = float(T) * n_sr * srrf_flops_per_voxel_frame(K, conf.IntensityW, conf.Method == 3)
    tflops = alg_flops / (kernel_ms * 1e-3) / 1e12
    traffic = args.traffic if args.traffic is not None else (HEADLINE_TRAFFIC_BYTES if T == 200 else None)
    # The kernel is bound by FP32 arithmetic on the vector units (it issues no
    # matrix instructions); MI355X's FP32 peak is the same 157.3 TFLOP/s for
    # the vector units and the dense matrix cores, so that is the roof.  The
    # HBM figure is reported beside it.
    roof = {"bound": "mfma", "achieved": tflops, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": tflops / FP32_PEAK_TFLOPS, "traffic": traffic, "kernel": "srrf_fused (uint16 input)",
            "kernel_ms": kernel_ms, "algorithmic_flops_per_launch": alg_flops,
            "note": SRRF_ROOF_NOTE, "nan_by_definition_fraction": nan_frac,
            "hbm": dict(hbm_roof(alg_bytes, kernel_ms), algorithmic_bytes_per_launch=alg_bytes)}
    upd = {
        "value": n_sr / (ms_per_step * 1e-3),
        "ms_per_step": ms_per_step,
        "config": {"workload": "plain 3D SRRF, one stack: %dx%dx%d uint16, T=%d, %d emitters, M=%d -> %dx%dx%d SR, "
                               "R=0.5, 26 directions, TRAC order 2 (ITC 0), smoothing 1, IW 1"
                               % (X, Y, Z, T, n_em, M, X * M, Y * M, Z * M),
                   "sr_voxels": n_sr, "frames": T, "nan_fraction": nan_frac},
        "roofline": roof,
    }
    with lock:
        rec.update(upd)
    if not args.no_cpu:
        crop = tuple(int(v) for v in args.cpu_crop.split(","))
        cb = cpu_baseline(series, conf, smp, crop, synth.default_threads())
        with lock:
            rec["cpu_baseline"] = cb
        acc = srrf_accuracy(series, conf, smp, (8, 8, 6), 50, synth.default_threads())
        with lock:
            rec["accuracy"] = acc


# ---------------------------------------------------------------- batched SRRF over uneven blocks
class _View:
    """A pointer into a larger device buffer (what srrf.Block reads .ptr of)."""

    def __init__(self, base, offset):
        self.ptr = ctypes.c_void_p(base.ptr.value + offset)


_C3 = {}


def c3_series():
    """The 512x512x128 T=120 uint16 series (8000 emitters, seed 3), generated
    once per process."""
    if not _C3:
        from rfba_amd import synth
        t0 = time.perf_counter()
        _C3["series"] = synth.srrf_series((512, 512, 128, 120), 8000, 3)
        log("c3: generated %s series in %.1f s" % (_C3["series"].shape, time.perf_counter() - t0))
    return _C3["series"]


def bench_srrf_batched(args):
    from rfba_amd import srrf, synth
    from rfba_amd.blocks import volume_blocks
    vol, T, n_em = (512, 512, 128), 120, 8000
    series = c3_series()
    split = volume_blocks(vol, 29, 3)
    rng4 = np.random.default_rng(4)
    extra_dims = rng4.integers(8, 30, (64, 3))
    extra_T = rng4.integers(30, 201, 64)
    mags = np.random.default_rng(5).integers(2, 5, len(split) + 64)
    density = n_em / float(np.prod(vol))
    extras = [synth.srrf_series((int(d[0]), int(d[1]), int(d[2]), int(t)),
                                max(1, int(round(density * float(np.prod(d))))), 400 + i)
              for i, (d, t) in enumerate(zip(extra_dims, extra_T))]
    conf = srrf.default_conf().replace(RingRadius=0.5, Method=1, GradientSmooth=1, MainRange=1, XRange=1, YRange=1,
                                       ZRange=1, RemoveCon=0, Renormalize=0, IntensityW=1)
    smp = srrf.directions(6, conf.RingRadius)

    d_vol = srrf.DeviceBuffer.from_array(series)
    d_extra = [srrf.DeviceBuffer.from_array(e) for e in extras]
    geo = [(d_vol, series.shape, o, d, T) for (o, d) in split]
    geo += [(d_extra[i], e.shape, (0, 0, 0), tuple(int(v) for v in e.shape[:0:-1]), e.shape[0])
            for i, e in enumerate(extras)]
    sizes = [int(np.prod(d)) * int(m) ** 3 for (_, _, _, d, _), m in zip(geo, mags)]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]) * 4
    d_out = srrf.DeviceBuffer(int(sum(sizes)) * 4)
    blocks = [srrf.Block(src, shp, o, d, t, int(m), _View(d_out, int(off)))
              for (src, shp, o, d, t), m, off in zip(geo, mags, offs)]
    n_sr = int(sum(sizes))
    sr_frames = float(sum(s * g[4] for s, g in zip(sizes, geo)))
    alg_bytes = float(sum(2 * np.prod(g[3]) * g[4] for g in geo)) + 4.0 * n_sr
    edges = sorted({int(v) for g in geo for v in g[3]})

    bt = srrf.Batch(blocks, srrf.RFBA_U16, conf, smp)
    t0 = time.perf_counter()
    tz, tune = bt.autotune()
    groups = bt.groups
    log("c3: autotune %.1f s -> groups (M, tile depth) %s; largest group's sample ms %s"
        % (time.perf_counter() - t0, groups, tune))
    for i in range(args.aux_warmup):
        log("c3: warmup batched %.1f ms" % bt.run(use_graph=True))
    srrf.lib().rfba_device_synchronize()
    kms = []
    t0 = time.perf_counter()
    for i in range(args.aux_steps):
        kms.append(bt.run(use_graph=True))
        log("c3: batched step %d: %.1f ms" % (i, kms[-1]))
    srrf.lib().rfba_device_synchronize()
    wall_b = (time.perf_counter() - t0) * 1e3 / args.aux_steps
    probe = [0, len(split) // 2, len(split) - 1, len(blocks) - 1]

    def snap():
        out = []
        for i in probe:
            b = blocks[i]
            n = sizes[i]
            buf = np.empty(n, np.float32)
            srrf._check(srrf.lib().rfba_memcpy_d2h(buf.ctypes.data_as(ctypes.c_void_p), b.out.ptr, n * 4))
            out.append(buf)
        return out

    ref = snap()
    same, ms_pb, wall_pb = None, None, None
    if not args.no_plain:
        t0 = time.perf_counter()
        ms_pb = bt.run_per_block()
        srrf.lib().rfba_device_synchronize()
        wall_pb = (time.perf_counter() - t0) * 1e3
        log("c3: per-block %.1f ms" % ms_pb)
        same = all(np.array_equal(a, b, equal_nan=True) for a, b in zip(ref, snap()))
    # Voxels with a sample out of range on any axis are NaN in every frame by
    # definition, and the kernel skips them: their share of the voxel-frames
    # is reported beside the roofline.
    host = np.empty(n_sr, np.float32)
    srrf._check(srrf.lib().rfba_memcpy_d2h(host.ctypes.data_as(ctypes.c_void_p), d_out.ptr, host.nbytes))
    valid = np.add.reduceat((~np.isnan(host)).view(np.uint8), (offs // 4).astype(np.int64), dtype=np.int64)
    del host
    valid_frames = float(sum(int(v) * g[4] for v, g in zip(valid, geo)))
    bt.close()
    for b in [d_vol, d_out] + d_extra:
        b.close()
    kb = float(np.mean(kms))
    fl = sr_frames * srrf_flops_per_voxel_frame(smp.n, conf.IntensityW, conf.Method == 3)
    tf = fl / (kb * 1e-3) / 1e12
    roof = {"bound": "mfma", "achieved": tf, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": tf / FP32_PEAK_TFLOPS, "kernel": "srrf_fused persistent queue (uint16 input)", "kernel_ms": kb,
            "algorithmic_flops_per_launch": fl, "note": SRRF_ROOF_NOTE,
            "nan_by_definition_fraction": 1.0 - valid_frames / sr_frames,
            "hbm": dict(hbm_roof(alg_bytes, kb), algorithmic_bytes_per_launch=alg_bytes)}
    res = {
        "workload": "batched 3D SRRF: 512x512x128 T=120 uint16 (8000 emitters, seed 3) split by doSplit.m into "
                    "%d blocks of 29^3 with 3-voxel overlap (z edge row 13 deep) + 64 extra stacks (edges 8-29, "
                    "T 30-200, seed 4); M per block from {2,3,4} (seed 5); R=0.5, 6 axis directions, TRA, "
                    "smoothing 1, IW 1" % len(split),
        "blocks": len(blocks), "edge_sizes": edges, "sr_voxels": n_sr, "sr_voxel_frames": sr_frames,
        "groups_mag_tile_depth": groups, "autotune_sample_ms_largest_group": {str(k): v for k, v in tune.items()},
        "batched": {"ms": wall_b, "kernel_ms": kb, "sr_voxels_per_s": n_sr / (wall_b * 1e-3)},
        "roofline": roof,
    }
    if not args.no_plain:
        res.update({"per_block": {"ms": wall_pb, "kernel_ms": ms_pb, "sr_voxels_per_s": n_sr / (wall_pb * 1e-3)},
                    "speedup_batched_vs_per_block": wall_pb / wall_b,
                    "batched_equals_per_block": bool(same)})
    if not args.no_cpu:
        from oracle import oracle
        i = 0
        b = blocks[i]
        o, d = geo[i][2], geo[i][3]
        sub = np.ascontiguousarray(series[:, o[2]:o[2] + d[2], o[1]:o[1] + d[1], o[0]:o[0] + 12]).astype(np.float32)
        c = conf.replace(Magnification=int(mags[i]))
        th = synth.default_threads()
        t0 = time.perf_counter()
        out = oracle.srrf_stack(sub[:, :, :12, :], c, smp.xyz, smp.divisor, nthreads=th)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": out.size / dt, "unit": "SR-voxels/s", "cores": th, "kind": "port",
                               "host": host_cpu(th),
                               "sample": "oracle on a 12x12x29 x T=120 part of block 0 (M=%d), %.1f s"
                                         % (int(mags[i]), dt)}
    return res


# ---------------------------------------------------------------- 3B
def b3_model(b3, ckpt):
    return b3.make_model(b3.blink_matrix(0.2, 0.1, 0.01), (0.5, 0.5, 0.0), 100.0, zfactor=2.0, z_exact=0,
                         checkpoint=ckpt, mu_brightness=7.0, sigma_brightness=1.0, mu_blur=0.25, sigma_blur=0.1)


def b3_alg(blocks):
    """(bytes, flops) per batched launch sequence, as algorithmic counts:
    bytes = the block data plus S background planes, 4 T P (1 + S) per block
    (what a fused evaluation must read); flops = 13 FP32 flops per
    (pixel, frame) of every (q, spot, sample) item plus 30 per pixel of its
    spot table."""
    by, fl = 0.0, 0.0
    for b in blocks:
        T, Z, Y, X = b.data.shape
        Q, K, _ = b.params.shape
        S = b.states.shape[0]
        P = X * Y * Z
        by += 4.0 * T * P * (1 + S)
        fl += float(Q) * K * S * (13.0 * T * P + 30.0 * P)
    return by, fl


def b3_roof(by, fl, kernel_ms):
    """The correlation kernel runs on the FP32 matrix cores: the binding roof
    is their peak; the HBM figure is reported beside it."""
    tf = fl / (kernel_ms * 1e-3) / 1e12
    roof = {"bound": "mfma", "achieved": tf, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": tf / FP32_PEAK_TFLOPS, "kernel": "b3_resid + b3_corr_mfma + b3_hmm (one launch sequence)",
            "kernel_ms": kernel_ms, "algorithmic_flops_per_launch": fl}
    roof["hbm"] = hbm_roof(by, kernel_ms)
    roof["hbm"]["algorithmic_bytes_per_launch"] = by
    return roof


def b3_fd_check(data, phi, states, m, grad, spots=4, zfactor=2.0):
    """Central finite differences (h = 1e-3, relative for I and sigma) of the
    FP64 oracle's logL of spot k over the other spots' sampled images plus
    b0, against the GPU gradient grad[k], for the first `spots` spots with a
    nonzero gradient.  The model's z derivative is the reference's, which
    lacks 1/zF^2 (storm.h:78): it is the exact one times zF^2, so the
    difference quotient's z component is scaled by zF^2 before comparing."""
    from oracle import b3oracle as bo
    T, Z, Y, X = data.shape
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    xyz = np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.int32)
    d2 = data.reshape(T, -1).astype(np.float64)
    K = phi.shape[0]
    imgs = [bo.spot(xyz, phi[j], zfactor)[0] for j in range(K)]
    on = states == 0  # [K, T]
    errs = []
    for k in range(K):
        if len(errs) == spots:
            break
        if not np.linalg.norm(grad[k]) > 0:
            continue
        bg = np.full_like(d2, 100.0)
        for j in range(K):
            if j != k:
                bg += on[j][:, None] * imgs[j][None]
        fd = np.empty(5)
        for n in range(5):
            h = 1e-3 * (abs(phi[k, n]) if n < 2 else 1.0)
            e = np.zeros(5)
            e[n] = h
            fd[n] = (bo.logl_grad(xyz, d2, bg, phi[k] + e, m)[0] - bo.logl_grad(xyz, d2, bg, phi[k] - e, m)[0]) / (2 * h)
        fd[4] *= zfactor * zfactor
        errs.append(float(np.linalg.norm(grad[k] - fd) / np.linalg.norm(fd)))
    return {"grad_max_rel": max(errs) if errs else None, "spots": len(errs), "parameter_set": 0,
            "h": "1e-3 (times |I|, |sigma| for those two)"}


def bench_b3_single(args):
    from rfba_amd import b3
    from rfba_amd.synth import default_threads
    m = b3_model(b3, 0)
    data, params, states = b3.synth_block((29, 29, 29), 100, 40, 6)
    P = b3.perturb(params, 256, 7)
    blk = b3.BlockData(data, P, states[None], 100.0)
    bt = b3.Batch([blk], m)
    for i in range(max(1, args.aux_warmup)):
        log("c4: warmup batched %.2f ms" % bt.run(use_graph=True))
    kms = []
    t0 = time.perf_counter()
    for i in range(max(1, args.aux_steps)):
        kms.append(bt.run(use_graph=True))
    wall_b = (time.perf_counter() - t0) * 1e3 / len(kms)
    L, G = bt.results(0)
    bt.run_per_block(per_param_set=True)
    t0 = time.perf_counter()
    ms_pb = bt.run_per_block(per_param_set=True)
    wall_pb = (time.perf_counter() - t0) * 1e3
    L2, G2 = bt.results(0)
    bt.close()
    log("c4: batched %.2f ms, per-parameter-set %.2f ms" % (wall_b, wall_pb))
    by, fl = b3_alg([blk])
    kb = float(np.mean(kms))
    roof = b3_roof(by, fl, kb)
    res = {
        "workload": "3B logL + adjoint, one 29^3 block, T=100, 40 spots (seed 6), 256 parameter sets (+-0.5 px, "
                    "+-20%% intensity, seed 7), S=1, one-wave-per-item HMM (no stored messages)",
        "block_evals": 256, "spot_evals": 256 * 40,
        "batched": {"ms": wall_b, "kernel_ms": kb, "block_evals_per_s": 256 / (wall_b * 1e-3)},
        "per_param_set": {"ms": wall_pb, "kernel_ms": ms_pb, "block_evals_per_s": 256 / (wall_pb * 1e-3)},
        "speedup_batched_vs_per_param_set": wall_pb / wall_b,
        "plain_equals_batched": bool(np.array_equal(L, L2) and np.array_equal(G, G2)),
        "roofline": roof,
    }
    if not args.no_cpu:
        from oracle import b3oracle as bo
        th = default_threads()
        q = 4
        t0 = time.perf_counter()
        Lo, Go = bo.model_eval(data, 100.0, m, P[:q], states[None], threads=th)
        dt = time.perf_counter() - t0
        el = float(np.max(np.abs(L[:q] - Lo) / np.abs(Lo)))
        ng = np.linalg.norm(Go, axis=-1)
        nz = ng > 0  # spots whose posterior on-probability underflows have a zero gradient
        eg = float(np.max(np.linalg.norm(G[:q] - Go, axis=-1)[nz] / ng[nz]))
        res["error_vs_fp64_oracle"] = {"logL_max_rel": el, "grad_max_rel": eg, "parameter_sets": q,
                                       "zero_gradient_items": int((~nz).sum()),
                                       "zero_gradient_items_exact": bool(np.all(G[:q][~nz] == 0))}
        res["error_vs_central_fd"] = b3_fd_check(data, P[0], states, m, G[0, :, 0])
        res["cpu_baseline"] = {"value": q / dt, "unit": "block-evals/s", "cores": th, "kind": "port",
                               "host": host_cpu(th),
                               "sample": "FP64 oracle, %d parameter sets x 40 spots, %.1f s" % (q, dt)}
    return res


def bench_b3_functor(args):
    """The CG functor and the FFBS Hessian of one spot, as the 3B fit calls
    them: SampledBackgroundData built once (multispot5.cc:538-575) for spot 0
    of the b3_single block (29^3, T=100, 40 spots, seed 6) with S = 10 background
    samples (main.gibbs.samples, multispot5.cfg:28: the other spots at 10
    shifted sampled state paths), then operator() per conjugate-gradient call
    (one candidate vector) and sampled_background_spot_hessian_ffbs with
    1000 inner samples (add_remove.hessian.inner_samples, multispot5.cfg:43)."""
    from rfba_amd import b3
    from oracle import b3oracle as bo
    m = b3_model(b3, 0)
    T, S, bs = 100, 10, 1000
    data, params, states = b3.synth_block((29, 29, 29), T, 40, 6)
    xyz = b3.box_pixels((29, 29, 29))
    P = xyz.shape[0]
    d2 = data.reshape(T, -1).astype(np.float64)
    spots = np.stack([bo.spot(xyz, q, m.zfactor)[0] for q in params])
    bg = np.empty((S, T, P))
    for smp in range(S):
        on = (np.roll(states[1:], smp, axis=1) == 0).astype(np.float64)
        bg[smp] = 100.0 + on.T @ spots[1:]
    phi = params[0] * [1.05, 1.02, 1, 1, 1] + [0, 0, 0.2, -0.1, 0.1]
    t0 = time.perf_counter()
    f = b3.Functor(xyz, d2, bg, m, max_candidates=8)
    create_ms = (time.perf_counter() - t0) * 1e3
    for _ in range(3):
        f(phi)
    n_calls = 50
    t0 = time.perf_counter()
    for _ in range(n_calls):
        g, L = f(phi)
    call_ms = (time.perf_counter() - t0) * 1e3 / n_calls
    b3.neg_grad_sampled(xyz, d2.astype(np.float32), bg.astype(np.float32), phi[None], m)
    t0 = time.perf_counter()
    b3.neg_grad_sampled(xyz, d2.astype(np.float32), bg.astype(np.float32), phi[None], m)
    old_ms = (time.perf_counter() - t0) * 1e3
    u = np.random.default_rng(11).random((S, bs, T))
    f.hessian_ffbs(phi, bs, u)
    t0 = time.perf_counter()
    H, D, paths = f.hessian_ffbs(phi, bs, u)
    hess_ms = (time.perf_counter() - t0) * 1e3
    f.close()
    # the plane is read once per call: S x T x P FP32 residuals
    by = 4.0 * S * T * P
    roof = hbm_roof(by, call_ms)
    roof.update({"kernel": "b3_corr_mfma (pixel-split) + b3_corr_reduce + b3_hmm, per call incl. copies",
                 "algorithmic_bytes_per_call": by,
                 "note": "one candidate per call: latency-bound (launches, copies and the FP64 HMM's serial frames), "
                         "not bandwidth-bound"})
    res = {
        "workload": "CG functor + FFBS Hessian of one spot: 29^3 block (P=24389), T=100, 40 spots (seed 6), "
                    "S=10 background samples, one candidate per CG call; Hessian with 1000 inner samples",
        "create_ms": create_ms,
        "per_cg_call": {"ms": call_ms, "calls_per_s": 1e3 / call_ms},
        "per_cg_call_host_reupload_path": {"ms": old_ms, "what": "rfba3b_neg_grad_sampled: residuals formed and "
                                                                 "copied on every call, per-item kernel"},
        "speedup_resident_vs_reupload": old_ms / call_ms,
        "hessian_ffbs": {"ms": hess_ms, "inner_samples": bs, "backward_samplings": S * bs},
        "roofline": roof,
    }
    if not args.no_cpu:
        t0 = time.perf_counter()
        go, Lo = bo.neg_grad(xyz, d2, bg, phi, m)
        dt_g = time.perf_counter() - t0
        t0 = time.perf_counter()
        Ho, Do, po = bo.hessian_ffbs(xyz, d2, bg, phi, m, bs, u)
        dt_h = time.perf_counter() - t0
        res["error_vs_fp64_oracle"] = {
            "grad_rel": float(np.linalg.norm(g[0] - go) / np.linalg.norm(go)),
            "logL_max_rel": float(np.max(np.abs(L[0] - Lo) / np.abs(Lo))),
            "hessian_max_normalised": float(np.abs(H - Ho).max() / np.abs(Ho).max()),
            "hessian_paths_equal": bool(np.array_equal(paths, po))}
        res["cpu_baseline"] = {"value": 1.0 / dt_g, "unit": "CG calls/s", "cores": 1, "kind": "port",
                               "host": host_cpu(1),
                               "sample": "FP64 oracle functor, one call (%.2f s); its FFBS Hessian with the same "
                                         "draws %.2f s" % (dt_g, dt_h),
                               "hessian_s": dt_h}
    return res


_C5 = {}


def c5_blocks():
    """The 1200 blocks of the 512x512x128 doSplit (seed 8): K ~ Poisson(40)
    clipped to 5-120, T=120; (data [T,Z,Y,X], true params [K,5], states
    [K,T]) per block, generated once per process."""
    if not _C5:
        from rfba_amd import b3
        from rfba_amd.blocks import volume_blocks
        split = volume_blocks((512, 512, 128), 29, 3)
        rng = np.random.default_rng(8)
        idx = np.sort(rng.choice(len(split), 1200, replace=False))
        K = np.clip(rng.poisson(40, 1200), 5, 120)
        T = 120
        dims = [split[i][1] for i in idx]
        t0 = time.perf_counter()
        _C5["gen"] = b3.synth_blocks(dims, T, K, 8000)
        _C5["K"], _C5["T"] = K, T
        log("c5: generated %d blocks in %.1f s" % (len(dims), time.perf_counter() - t0))
    return _C5["gen"], _C5["K"], _C5["T"]


def bench_b3_batched(args):
    from rfba_amd import b3
    from rfba_amd.synth import default_threads
    m = b3_model(b3, 11)
    gen, K, T = c5_blocks()
    Q = 32
    blocks = [b3.BlockData(d, b3.perturb(p, Q, 80000 + i), s[None], 100.0) for i, (d, p, s) in enumerate(gen)]
    bt = b3.Batch(blocks, m)
    log("c5: warmup batched %.1f ms" % bt.run(use_graph=True))
    kms = []
    t0 = time.perf_counter()
    for i in range(max(1, args.aux_steps)):
        kms.append(bt.run(use_graph=True))
        log("c5: batched step %d %.1f ms" % (i, kms[-1]))
    wall_b = (time.perf_counter() - t0) * 1e3 / len(kms)
    L0, G0 = bt.results(0)
    # the same launch with the checkpointed lane-form HMM (forward messages
    # kept every 11 frames, the rest recomputed in the adjoint)
    bt.set_hmm_form(1)
    bt.run(use_graph=True)
    t0 = time.perf_counter()
    ms_lane = bt.run(use_graph=True)
    wall_lane = (time.perf_counter() - t0) * 1e3
    L1, G1 = bt.results(0)
    lane_dl = float(np.max(np.abs(L1 - L0) / np.abs(L0)))
    ng = np.linalg.norm(G0, axis=-1)
    nz = ng > 0
    lane_dg = float(np.max(np.linalg.norm(G1 - G0, axis=-1)[nz] / ng[nz])) if nz.any() else 0.0
    log("c5: batched with the checkpointed lane HMM %.1f ms" % wall_lane)
    bt.close()
    nsub = min(args.c5_plain_blocks, len(blocks))
    sub = b3.Batch(blocks[:nsub], m)
    sub.run(use_graph=True)
    t0 = time.perf_counter()
    ms_sub = sub.run(use_graph=True)
    wall_sub = (time.perf_counter() - t0) * 1e3
    sub.run_per_block()
    t0 = time.perf_counter()
    ms_pb = sub.run_per_block()
    wall_pb = (time.perf_counter() - t0) * 1e3
    sub.close()
    nq = min(10, nsub)  # one launch sequence per (block, parameter set): the CG-call granularity
    sq = b3.Batch(blocks[:nq], m)
    sq.run_per_block(per_param_set=True)
    t0 = time.perf_counter()
    ms_pq = sq.run_per_block(per_param_set=True)
    wall_pq = (time.perf_counter() - t0) * 1e3
    sq.close()
    log("c5: batched %.1f ms (all), subset batched %.1f ms, subset per block %.1f ms, %d blocks per (block, q) %.1f ms"
        % (wall_b, wall_sub, wall_pb, nq, wall_pq))
    by, fl = b3_alg(blocks)
    kb = float(np.mean(kms))
    roof = b3_roof(by, fl, kb)
    n_ev = len(blocks) * Q
    res = {
        "workload": "batched 3B logL + adjoint: 1200 blocks of the 512x512x128 doSplit (29^3, 3-voxel overlap; "
                    "seed 8), K ~ Poisson(40) clipped 5-120, T=120, 32 parameter sets per block, S=1, "
                    "checkpoint stride 11 in the model (T <= 512 runs the one-wave-per-item HMM, whose lanes "
                    "hold their frames' messages in registers and store none, so the stride is not used)",
        "blocks": len(blocks), "block_evals": n_ev, "spot_evals": int(K.sum()) * Q,
        "batched": {"ms": wall_b, "kernel_ms": kb, "block_evals_per_s": n_ev / (wall_b * 1e-3)},
        "batched_checkpointed_lane_hmm": {"ms": wall_lane, "kernel_ms": ms_lane, "checkpoint_stride": 11,
                                          "block_evals_per_s": n_ev / (wall_lane * 1e-3),
                                          "block0_vs_wave_form": {"logL_max_rel": lane_dl,
                                                                  "grad_max_rel": lane_dg}},
        "plain_per_block": {"blocks": nsub, "ms": wall_pb, "kernel_ms": ms_pb,
                            "block_evals_per_s": nsub * Q / (wall_pb * 1e-3),
                            "batched_ms_same_blocks": wall_sub,
                            "batched_block_evals_per_s_same_blocks": nsub * Q / (wall_sub * 1e-3)},
        "plain_per_param_set": {"blocks": nq, "ms": wall_pq, "kernel_ms": ms_pq,
                                "block_evals_per_s": nq * Q / (wall_pq * 1e-3)},
        "speedup_batched_vs_plain": (n_ev / wall_b) / (nsub * Q / wall_pb),
        "speedup_batched_vs_per_param_set": (n_ev / wall_b) / (nq * Q / wall_pq),
        "roofline": roof,
    }
    if not args.no_cpu:
        from oracle import b3oracle as bo
        th = default_threads()
        t0 = time.perf_counter()
        b = blocks[0]
        bo.model_eval(b.data, 100.0, m, b.params[:2], b.states, threads=th)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": 2 / dt, "unit": "block-evals/s", "cores": th, "kind": "port",
                               "host": host_cpu(th),
                               "sample": "FP64 oracle, block 0 (K=%d), 2 parameter sets, %.1f s"
                                         % (b.params.shape[1], dt)}
    return res


def gibbs_changes(prev, samples):
    """Frames whose on/off status changes at each spot update of a run:
    prev [K,T] the sample before the run, samples [S,K,T] after each step
    (one sweep per step)."""
    on = np.concatenate([prev[None], samples]) == 0
    return (on[1:] != on[:-1]).sum(axis=(0, 2))  # per spot, summed over steps


def bench_b3_gibbs(args):
    """Gibbs sampling of the blink states over the 1200 blocks: one run =
    main.gibbs.samples = 10 next() steps of every block (multispot5.cfg:28),
    each one sweep (gibbs.mixing_iterations = 1, :20)."""
    from concurrent.futures import ThreadPoolExecutor
    from rfba_amd import b3
    from rfba_amd.synth import default_threads
    m = b3_model(b3, 0)
    gen, K, T = c5_blocks()
    S = 10
    t0 = time.perf_counter()
    chains = []
    for i, (d, p, _) in enumerate(gen):
        Tz, Z, Y, X = d.shape
        chains.append(b3.ChainData(d.reshape(Tz, -1), b3.box_pixels((X, Y, Z)), p, base=100.0, seed=1000 + i,
                                   slots=S))
    g = b3.Gibbs(chains, m)
    setup_ms = g.setup_ms
    log("gibbs: %d chains set up in %.1f s (statistics kernels %.1f ms)"
        % (len(chains), time.perf_counter() - t0, setup_ms))
    masks = [int(mm["offsets"][-1]) for mm in g.meta]
    log("gibbs: warmup %.1f ms" % g.run(S, 1))
    # first samples of every block, for the oracle's check (a bounded CPU sample: one sweep per block)
    first = [g.samples(i, 1)[0].copy() for i in range(len(chains))] if not args.no_cpu else []
    kms, walls = [], []
    for i in range(max(1, args.aux_steps)):
        prev = [g.sample(c) for c in range(len(chains))]
        t1 = time.perf_counter()
        kms.append(g.run(S, 1))
        walls.append((time.perf_counter() - t1) * 1e3)
        log("gibbs: step %d %.1f ms" % (i, kms[-1]))
    # algorithmic bytes of the last run: per spot update, the spot's C and A
    # rows (16 T), its sample row read and written (2 T), and for every frame
    # whose on/off status changed, O and A of the other spots (24 K); the
    # pixels are read only by the plan's statistics (setup below)
    by = 0.0
    ups = 0
    setup_by = 0.0
    for c, mm in enumerate(g.meta):
        off = mm["offsets"].astype(np.int64)
        msz = off[1:] - off[:-1]
        Kc = mm["K"]
        ch = gibbs_changes(prev[c], g.samples(c, S))
        by += float(S * Kc * 18.0 * T + 24.0 * Kc * ch.sum())
        ups += S * Kc
        # statistics: the frames once (4 T P), the mask indices and spot values
        # at them (12 m), C, O and S2 written (8 K T + 8 K^2 + 8 K)
        setup_by += float(4.0 * T * mm["P"] + 12.0 * msz.sum() + 8.0 * (Kc * T + Kc * Kc + Kc))
    kb = float(np.mean(kms))
    wall = float(np.mean(walls))
    roof = hbm_roof(by, kb)
    roof.update({"kernel": "b3_gibbs (one 64-thread workgroup per chain)", "kernel_ms": kb,
                 "algorithmic_bytes_per_launch": by,
                 "note": "bound by the latency of each spot update, not by memory: the spots of a chain are "
                         "updated in order, each by a 64-lane scan of extended-range FP64 3x3 products and a "
                         "backward sampling scan, one wave per chain; the longest chain sets the time"})
    # plain form: one launch per block (a plan per chain) on the first blocks
    nsub = min(20, len(chains))
    subs = [b3.Gibbs([chains[i]], m) for i in range(nsub)]
    for sg in subs:
        sg.run(S, 1)
    t1 = time.perf_counter()
    ms_pb = sum(sg.run(S, 1) for sg in subs)
    wall_pb = (time.perf_counter() - t1) * 1e3
    ups_sub = S * int(sum(chains[i].spots.shape[0] for i in range(nsub)))
    for sg in subs:
        sg.close()
    g.close()
    res = {
        "workload": "Gibbs sampling of blink states (GibbsSampler2::next): the 1200 blocks of the 512x512x128 "
                    "doSplit (seed 8), K ~ Poisson(40) clipped 5-120, T=120, masks of radius 6 sigma + 1 "
                    "(sigma 1.3), 10 samples of one sweep each per run, MT19937 per block",
        "chains": len(chains), "spot_updates": ups, "mask_pixels_mean": float(np.sum(masks) / max(1, K.sum())),
        "batched": {"ms": wall, "kernel_ms": kb, "spot_updates_per_s": ups / (wall * 1e-3),
                    "block_samples_per_s": len(chains) * S / (wall * 1e-3)},
        "plain_per_block": {"blocks": nsub, "ms": wall_pb, "kernel_ms": ms_pb,
                            "spot_updates_per_s": ups_sub / (wall_pb * 1e-3)},
        "speedup_batched_vs_per_block": (ups / wall) / (ups_sub / wall_pb),
        "setup": {"kernel_ms": setup_ms, "algorithmic_bytes": setup_by,
                  "GB_per_s": setup_by / (setup_ms * 1e-3) / 1e9 if setup_ms > 0 else None,
                  "what": "per plan: spot values at the mask pixels and S2 (b3_gibbs_mask_values), C per "
                          "(pixel range, 8 frames) (b3_gibbs_cpart), mask overlaps per (spot, range) "
                          "(b3_gibbs_opart), partial sums added per spot (b3_gibbs_reduce)"},
        "roofline": roof,
    }
    if not args.no_cpu:
        from oracle import b3oracle as bo
        th = default_threads()
        n = len(first)

        def one(i):
            d, p, _ = gen[i]
            Tz = d.shape[0]
            Z, Y, X = d.shape[1:]
            xyz = b3.box_pixels((X, Y, Z))
            data = d.reshape(Tz, -1).astype(np.float64) - 100.0
            spot = np.stack([bo.spot(xyz, q, zfactor=m.zfactor)[0] for q in p])
            off, idx = bo.spot_pixels(xyz, p)
            u = b3.mt19937_uniforms(1000 + i, len(p) * Tz)
            smp = np.full((len(p), Tz), 2, np.uint8)
            intens = np.zeros(data.shape)
            t = time.perf_counter()
            bo.gibbs_next(data, spot, off, idx, np.array(m.A[:]).reshape(3, 3), np.array(m.pi[:]), m.variance, 1,
                          smp, intens, u)
            return time.perf_counter() - t, smp, len(p)

        t0 = time.perf_counter()
        with ThreadPoolExecutor(th) as ex:
            out = list(ex.map(one, range(n)))
        dt = time.perf_counter() - t0
        neq = sum(bool(np.array_equal(o[1], first[i])) for i, o in enumerate(out))
        res["first_sample_equals_oracle"] = {"blocks": n, "equal_blocks": neq, "equal": neq == n}
        res["cpu_baseline"] = {"value": sum(o[2] for o in out) / dt, "unit": "spot-updates/s", "cores": th,
                               "kind": "port", "host": host_cpu(th),
                               "sample": "oracle GibbsSampler2::next, first sweep of every block (%d blocks), %d "
                                         "threads, %.1f s" % (n, th, dt)}
    return res


def bench_prep(args):
    """load_and_normalize_images on the 512x512x128 T=120 series: per z-slice
    high-pass (sigma 5, 31 taps per axis), mean/variance, scaling; plus the
    TIFF path end to end on 8 of its frames."""
    import tempfile
    from rfba_amd import prep
    from rfba_amd.synth import default_threads
    series = c3_series()
    T, Z, Y, X = series.shape
    c = prep.conf(lpf=5.0)
    d_raw = prep.Device.from_array(series)
    d_out = prep.Device(series.size * 4)
    stats, ms = prep.normalize_device(d_raw, series.dtype, series.shape, c, d_out)
    log("prep: warmup %.1f ms" % ms)
    kms, walls = [], []
    for i in range(max(1, args.aux_steps)):
        t0 = time.perf_counter()
        stats, ms = prep.normalize_device(d_raw, series.dtype, series.shape, c, d_out)
        walls.append((time.perf_counter() - t0) * 1e3)
        kms.append(ms)
        log("prep: step %d %.1f ms" % (i, ms))
    d_raw.close()
    d_out.close()
    n = float(series.size)
    kb = float(np.mean(kms))
    # algorithmic bytes per voxel-frame: the uint16 input read and the float
    # high-pass written (2 + 4), then read and written scaled (4 + 4)
    by = 14.0 * n
    roof = hbm_roof(by, kb)
    k, _ = prep.gaussian_kernel(5.0)
    ks = len(k)
    roof.update({"kernel": "prep_highpass + prep_scale + 2 x prep_reduce", "kernel_ms": kb,
                 "algorithmic_bytes_per_launch": by,
                 "fp32_ops_per_voxel": 2 * (3 * ks + 1) + 1})
    res = {
        "workload": "load_and_normalize_images on the 512x512x128 T=120 uint16 series (8000 emitters, seed 3): "
                    "per z-slice high-pass (sigma 5, half-width 15), mean/variance, division by the std",
        "voxel_frames": int(n),
        "normalize": {"ms": float(np.mean(walls)), "kernel_ms": kb, "voxel_frames_per_s": n / (kb * 1e-3),
                      "stats": [float(v) for v in stats]},
        "roofline": roof,
    }
    # the TIFF path end to end: 8 frames written as stacks, read back
    tmp = tempfile.mkdtemp(prefix="rfba_prep_", dir=os.environ.get("TMPDIR", "/tmp"))
    try:
        paths = []
        for t in range(8):
            p = os.path.join(tmp, "frame%03d.tif" % t)
            prep.tiff_write(p, series[t], big_endian=bool(t & 1))
            paths.append(p)
        prep.load_and_normalize(paths, c)
        t0 = time.perf_counter()
        out8, st8 = prep.load_and_normalize(paths, c)
        dt = time.perf_counter() - t0
        mb = 8 * series[0].nbytes / 1e6
        res["tiff_load_and_normalize"] = {"frames": 8, "MB": mb, "s": dt, "MB_per_s": mb / dt,
                                          "note": "files just written: read from the page cache"}
    finally:
        for p in os.listdir(tmp):
            os.remove(os.path.join(tmp, p))
        os.rmdir(tmp)
    if not args.no_cpu:
        from oracle import preporacle as po
        th = default_threads()
        nf = min(16, T)
        sub = series[:nf]
        t0 = time.perf_counter()
        oo, ost = po.normalize(sub.astype(np.float32), lpf=5.0, threads=th)
        dt = time.perf_counter() - t0
        go, gst = prep.normalize(sub, c)
        res["first_frames_equal_oracle"] = {"frames": nf, "equal": bool(np.array_equal(go, oo) and
                                                                         np.array_equal(gst, ost))}
        del oo, go
        res["cpu_baseline"] = {"value": sub.size / dt, "unit": "voxel-frames/s", "cores": min(th, nf), "kind": "port",
                               "host": host_cpu(min(th, nf)),
                               "sample": "oracle on frames 0-%d of the same series (%d voxel-frames), %.1f s; "
                                         "high-pass one thread per frame, statistics and scaling on one thread"
                                         % (nf - 1, sub.size, dt)}
    return res


# Estimated wall seconds of each auxiliary workload at one warm-up and two
# timed passes, CPU legs included (measured on MI355X with 16 host threads,
# rounded up).  A workload starts only if it fits the remaining budget.
EST_S = {"c4": 10, "functor": 10, "c5": 25, "gibbs": 20, "c3": 100, "prep": 15}
C3_SERIES_S = 35  # generating the 512x512x128 T=120 series shared by c3 and prep


def main():
    t_start = time.perf_counter()
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2, help="timed passes of the headline workload")
    ap.add_argument("--warmup", type=int, default=1, help="untimed passes of the headline workload")
    ap.add_argument("--aux-steps", type=int, default=2, help="timed passes of every other workload")
    ap.add_argument("--aux-warmup", type=int, default=1, help="untimed passes of every other workload")
    ap.add_argument("--budget", type=float, default=470.0,
                    help="seconds from start after which no further workload is started")
    ap.add_argument("--deadline", type=float, default=565.0,
                    help="seconds from start at which the JSON line is printed with what has been measured, "
                         "and the process ends")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the batched SRRF workload's one-launch-per-block comparison leg (for profiling the "
                         "batched dispatches, whose work is unchanged)")
    ap.add_argument("--cpu-crop", type=str, default="12,12,8")
    ap.add_argument("--frames", type=int, default=200)
    ap.add_argument("--only", type=str, default="c2,c4,functor,c5,gibbs,c3,prep",
                    help="workloads to run, in this order: c2 (headline stack), c3 (batched SRRF), c4, c5 (3B), "
                         "functor (3B CG functor + FFBS Hessian per call), gibbs (3B blink-state sampling), "
                         "prep (image normalisation)")
    ap.add_argument("--c5-plain-blocks", type=int, default=100)
    ap.add_argument("--traffic", type=float, default=None,
                    help="measured HBM bytes per launch of the headline kernel (rocprofv3 FETCH_SIZE+WRITE_SIZE)")
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
    order = [w for w in args.only.split(",") if w]
    rec = {
        "metric": METRIC, "value": None, "unit": "SR-voxels/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded emitters, Markov blinking, Poisson / Gaussian noise; generated on the host)",
        "config": None, "roofline": None, "cpu_baseline": None, "workloads": {},
        "passes": {"headline": {"warmup": args.warmup, "timed": args.steps},
                   "other_workloads": {"warmup": args.aux_warmup, "timed": args.aux_steps}},
    }
    lock = threading.Lock()
    printed = []

    def emit(reason=None):
        with lock:
            if printed:
                return
            printed.append(True)
            rec["wall_s"] = time.perf_counter() - t_start
            if reason:
                rec["truncated"] = reason
            print(json.dumps(rec), flush=True)

    def on_deadline():
        emit("deadline of %.0f s reached while a workload ran; what it had not finished is missing" % args.deadline)
        os._exit(0)

    watchdog = threading.Timer(max(1.0, args.deadline - (time.perf_counter() - t_start)), on_deadline)
    watchdog.daemon = True
    watchdog.start()
    runners = {"c3": ("srrf_batched", bench_srrf_batched), "c4": ("b3_single", bench_b3_single),
               "functor": ("b3_functor", bench_b3_functor),
               "c5": ("b3_batched", bench_b3_batched), "gibbs": ("b3_gibbs", bench_b3_gibbs),
               "prep": ("prep", bench_prep)}
    try:
        for w in order:
            elapsed = time.perf_counter() - t_start
            if w == "c2":
                bench_stack(args, rec, lock)
                log("c2 done at %.0f s" % (time.perf_counter() - t_start))
                continue
            name, fn = runners[w]
            need = EST_S[w] + (C3_SERIES_S if w in ("c3", "prep") and not _C3 else 0)
            if elapsed + need > args.budget:
                with lock:
                    rec["workloads"][name] = {"skipped": "budget", "elapsed_s": elapsed, "estimated_s": need,
                                              "budget_s": args.budget}
                log("%s skipped: %.0f s elapsed + %.0f s estimated > budget %.0f s" % (w, elapsed, need, args.budget))
                continue
            t0 = time.perf_counter()
            res = fn(args)
            res["wall_s"] = time.perf_counter() - t0
            with lock:
                rec["workloads"][name] = res
            log("%s done in %.0f s (at %.0f s)" % (w, res["wall_s"], time.perf_counter() - t_start))
    finally:
        watchdog.cancel()
        emit()


if __name__ == "__main__":
    main()
