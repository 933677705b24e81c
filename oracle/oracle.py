"""TEST INFRASTRUCTURE ONLY: ctypes wrappers of oracle/liboracle.so.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker.  See srrf_oracle.cpp / b3_oracle.cpp for what each
function restates (reference file:line).
"""
from __future__ import annotations

import contextlib
import ctypes
import os
import threading
from ctypes import POINTER, c_double, c_float, c_int

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_L = None
# one CDLL per process, whose signatures are set once, also when several
# threads ask for it first
LOCK = threading.RLock()


def _load(path):
    if not os.path.exists(path):
        raise ImportError("oracle not built: run `python -m rfba_amd.build oracle`")
    L = ctypes.CDLL(path)
    fp, dp, ip = POINTER(c_float), POINTER(c_double), POINTER(c_int)
    from rfba_amd.srrf import SRRFConf
    cp = POINTER(SRRFConf)
    L.oracle_build_ring_f32.argtypes = [c_int, c_float, fp]
    L.oracle_build_ring_f64.argtypes = [c_int, c_double, dp]
    L.oracle_cubic_f32.argtypes = [c_float]
    L.oracle_cubic_f32.restype = c_float
    L.oracle_cubic_f64.argtypes = [c_double]
    L.oracle_cubic_f64.restype = c_double
    L.oracle_gradient_f32.argtypes = [fp, ip, cp, fp, fp, fp]
    L.oracle_interp_f32.argtypes = [fp, ip, c_int, c_int, fp, c_int, fp]
    L.oracle_interp_f64.argtypes = [dp, ip, c_int, c_int, dp, c_int, dp]
    L.oracle_srrf_stack_f32.argtypes = [fp, ip, cp, fp, c_int, c_float, fp, fp, c_int]
    L.oracle_srrf_stack_f64.argtypes = [dp, ip, cp, dp, c_int, c_double, dp, dp, c_int]
    L.oracle_srrf_stack_box_f32.argtypes = [fp, ip, cp, fp, c_int, c_float, ip, fp, c_int]
    L.oracle_temporal_f32.argtypes = [fp, fp, ip, cp, fp]
    L.oracle_temporal_f64.argtypes = [dp, dp, ip, cp, dp]
    L.oracle_srrf_tiled_f32.argtypes = [fp, ip, cp, fp, c_int]
    L.oracle_hardware_threads.restype = c_int
    return L


def lib():
    global _L
    if _L is None:
        with LOCK:
            if _L is None:
                _L = _load(os.path.join(HERE, "liboracle.so"))
    return _L


@contextlib.contextmanager
def using(path):
    """Route this module's SRRF calls to another build of the oracle (a
    deliberately broken copy compiled by a test) inside the block."""
    global _L
    with LOCK:
        prev = lib()
        _L = _load(path)
    try:
        yield
    finally:
        with LOCK:
            _L = prev


def _p(a, t):
    return a.ctypes.data_as(POINTER(t))


def _dims(shape_tzyx):
    T, Z, Y, X = shape_tzyx
    return (c_int * 4)(X, Y, Z, T)


def threads():
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def build_ring(axes, radius, f64=False):
    n = 2 * axes
    if f64:
        buf = np.zeros((n * n, 3), np.float64)
        lib().oracle_build_ring_f64(axes, radius, _p(buf, c_double))
    else:
        buf = np.zeros((n * n, 3), np.float32)
        lib().oracle_build_ring_f32(axes, radius, _p(buf, c_float))
    return buf


def cubic(x, f64=False):
    f = lib().oracle_cubic_f64 if f64 else lib().oracle_cubic_f32
    return np.array([f(float(v)) for v in np.ravel(x)]).reshape(np.shape(x))


def gradient(series, conf):
    a = np.ascontiguousarray(series, np.float32)
    gx, gy, gz = (np.empty_like(a) for _ in range(3))
    lib().oracle_gradient_f32(_p(a, c_float), _dims(a.shape), ctypes.byref(conf), _p(gx, c_float),
                              _p(gy, c_float), _p(gz, c_float))
    return gx, gy, gz


def interp(arr, mag, t, pts, f64=False):
    """interpolateG of `arr` [T,Z,Y,X] at SR points pts [n,3] of frame t."""
    dt, ct = (np.float64, c_double) if f64 else (np.float32, c_float)
    a = np.ascontiguousarray(arr, dt)
    p = np.ascontiguousarray(pts, dt)
    out = np.empty(len(p), dt)
    dims = _dims(a.shape)
    f = lib().oracle_interp_f64 if f64 else lib().oracle_interp_f32
    f(_p(a, ct), dims, mag, t, _p(p, ct), len(p), _p(out, ct))
    return out


def srrf_stack(series, conf, xyz, divisor, f64=False, with_rad=False, nthreads=None):
    """Whole stack (kernel.cu:560-657 semantics) with an explicit sample list."""
    dt, ct = (np.float64, c_double) if f64 else (np.float32, c_float)
    a = np.ascontiguousarray(series, dt)
    T, Z, Y, X = a.shape
    M = conf.Magnification
    ring = np.ascontiguousarray(xyz, dt)
    out = np.empty((Z * M, Y * M, X * M), dt)
    rad = np.empty((T, Z * M, Y * M, X * M), dt) if with_rad else None
    f = lib().oracle_srrf_stack_f64 if f64 else lib().oracle_srrf_stack_f32
    rc = f(_p(a, ct), _dims(a.shape), ctypes.byref(conf), _p(ring, ct), len(ring), divisor, _p(out, ct),
           _p(rad, ct) if rad is not None else None, nthreads or threads())
    if rc:
        raise ValueError("oracle rejected the input")
    return (out, rad) if with_rad else out


def srrf_stack_box(series, conf, xyz, divisor, box_xyz0, box_shape_xyz, nthreads=None):
    """FP32 stack oracle evaluated only on an SR box (absolute coordinates);
    returns [nz, ny, nx]."""
    a = np.ascontiguousarray(series, np.float32)
    ring = np.ascontiguousarray(xyz, np.float32)
    nx, ny, nz = box_shape_xyz
    box = (c_int * 6)(*box_xyz0, nx, ny, nz)
    out = np.empty((nz, ny, nx), np.float32)
    rc = lib().oracle_srrf_stack_box_f32(_p(a, c_float), _dims(a.shape), ctypes.byref(conf), _p(ring, c_float),
                                         len(ring), divisor, box, _p(out, c_float), nthreads or threads())
    if rc:
        raise ValueError("oracle rejected the input")
    return out


def gdotr_flips(series, conf, xyz, nthreads=None):
    """Number of (SR voxel, frame, sample) triples whose converging/diverging
    test differs between the FP32 arithmetic and its FP64 re-evaluation."""
    L = lib()
    L.oracle_gdotr_flips.argtypes = [POINTER(c_float), POINTER(c_double), POINTER(c_int), POINTER(type(conf)),
                                     POINTER(c_float), POINTER(c_double), c_int, c_int]
    L.oracle_gdotr_flips.restype = ctypes.c_longlong
    a32 = np.ascontiguousarray(series, np.float32)
    a64 = np.ascontiguousarray(series, np.float64)
    r32 = np.ascontiguousarray(xyz, np.float32)
    r64 = np.ascontiguousarray(xyz, np.float64)
    n = L.oracle_gdotr_flips(_p(a32, c_float), _p(a64, c_double), _dims(a32.shape), ctypes.byref(conf),
                             _p(r32, c_float), _p(r64, c_double), len(r32), nthreads or threads())
    if n < 0:
        raise ValueError("oracle rejected the input")
    return int(n)


def temporal(series, rad, conf, f64=False):
    """Temporal reduction alone on a given radiality series rad [T, ZM, YM, XM]."""
    dt, ct = (np.float64, c_double) if f64 else (np.float32, c_float)
    a = np.ascontiguousarray(series, dt)
    r = np.ascontiguousarray(rad, dt)
    out = np.empty(r.shape[1:], dt)
    f = lib().oracle_temporal_f64 if f64 else lib().oracle_temporal_f32
    f(_p(a, ct), _p(r, ct), _dims(a.shape), ctypes.byref(conf), _p(out, ct))
    return out


def srrf_tiled(series, conf, fill=np.nan, nthreads=None):
    """SRRFGradientCalc tiling (GradientCalc.cpp:47-163); unwritten voxels = fill."""
    a = np.ascontiguousarray(series, np.float32)
    T, Z, Y, X = a.shape
    M = conf.Magnification
    out = np.full((Z * M, Y * M, X * M), fill, np.float32)
    rc = lib().oracle_srrf_tiled_f32(_p(a, c_float), _dims(a.shape), ctypes.byref(conf), _p(out, c_float),
                                     nthreads or threads())
    if rc:
        raise ValueError("oracle rejected the input")
    return out


def place_spots(sr, mag, pixels, num_spots, intensity_mode, blur_mode, seed):
    """Spots [n, 5] placed by the restated place_spots_intensity_sampled."""
    L = lib()
    L.oracle_place_spots.argtypes = [ctypes.POINTER(c_float), ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int, c_double, c_double,
                                     ctypes.c_int, ctypes.POINTER(c_double)]
    L.oracle_place_spots.restype = ctypes.c_int
    sr = np.ascontiguousarray(sr, np.float32)
    Zs, Ys, Xs = sr.shape
    dims = (ctypes.c_int * 3)(Xs // mag, Ys // mag, Zs // mag)
    px = np.ascontiguousarray(pixels, np.int32)
    out = np.zeros((num_spots, 5))
    n = L.oracle_place_spots(_p(sr, c_float), dims, mag, _p(px, ctypes.c_int), len(px), num_spots, intensity_mode,
                             blur_mode, seed, _p(out, c_double))
    return out[:n]
