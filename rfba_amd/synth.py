"""Seeded synthetic fluorescence series (librfba_synth.so)."""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_long, c_uint16, c_uint64, c_uint8

import numpy as np

from . import _lib


def _lib_synth():
    L = _lib.load("rfba_synth")
    L.rfba_synth_srrf_u16.argtypes = [c_int, c_int, c_int, c_int, c_int, c_uint64, c_double, c_double, c_double,
                                      c_double, c_double, c_double, POINTER(c_uint16), c_int]
    L.rfba_synth_srrf_u16.restype = c_int
    L.rfba_synth_emitters.argtypes = [c_int, c_int, c_int, c_int, c_int, c_uint64, c_double, c_double,
                                      POINTER(c_double), POINTER(c_uint8)]
    L.rfba_synth_emitters.restype = c_int
    L.rfba_synth_lambda_f64.argtypes = [c_int, c_int, c_int, c_int, c_int, c_uint64, c_double, c_double, c_double,
                                        c_double, c_double, c_double, POINTER(c_double)]
    L.rfba_synth_lambda_f64.restype = c_int
    L.rfba_synth_poisson.argtypes = [c_double, c_uint64, c_int, POINTER(c_long)]
    L.rfba_synth_poisson.restype = c_int
    return L


def default_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def srrf_series(shape_xyzt, n_emitters, seed, sigma_xy=1.3, sigma_z=2.6, peak=400.0, background=100.0,
                p_on_off=0.3, p_off_on=0.1, threads=None) -> np.ndarray:
    """uint16 [T, Z, Y, X] series: Gaussian PSF emitters with two-state
    Markov blinking over a constant background, Poisson counts."""
    X, Y, Z, T = shape_xyzt
    out = np.empty((T, Z, Y, X), np.uint16)
    rc = _lib_synth().rfba_synth_srrf_u16(X, Y, Z, T, n_emitters, seed, sigma_xy, sigma_z, peak, background,
                                          p_on_off, p_off_on, out.ctypes.data_as(POINTER(c_uint16)),
                                          threads or default_threads())
    if rc:
        raise RuntimeError("synthetic generator failed")
    return out


def emitters(shape_xyzt, n, seed, p_on_off=0.3, p_off_on=0.1):
    X, Y, Z, T = shape_xyzt
    pos = np.empty((n, 3), np.float64)
    on = np.empty((T, n), np.uint8)
    _lib_synth().rfba_synth_emitters(X, Y, Z, T, n, seed, p_on_off, p_off_on, pos.ctypes.data_as(POINTER(c_double)),
                                     on.ctypes.data_as(POINTER(c_uint8)))
    return pos, on


def srrf_lambda(shape_xyzt, n_emitters, seed, sigma_xy=1.3, sigma_z=2.6, peak=400.0, background=100.0,
                p_on_off=0.3, p_off_on=0.1) -> np.ndarray:
    X, Y, Z, T = shape_xyzt
    out = np.empty((T, Z, Y, X), np.float64)
    _lib_synth().rfba_synth_lambda_f64(X, Y, Z, T, n_emitters, seed, sigma_xy, sigma_z, peak, background, p_on_off,
                                       p_off_on, out.ctypes.data_as(POINTER(c_double)))
    return out


def poisson_draws(lam, seed, n) -> np.ndarray:
    out = np.empty(n, np.int64)
    _lib_synth().rfba_synth_poisson(lam, seed, n, out.ctypes.data_as(POINTER(c_long)))
    return out
