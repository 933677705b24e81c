"""Python mirror of the SRRF stage interface (librfba_srrf.so via ctypes).

The names follow the reference DLL: SRRFGradientCalc / SRRFGetMag
(RFBA Main/GradientCalc/GradientCalc.h:4-5), SRRFGradientCalcGPU and the
SRRFConf struct (RFBA Main/GradientCalc/kernel.cuh:3-26).  Arrays are numpy,
x-fastest: series [T, Z, Y, X], SR outputs [Z*M, Y*M, X*M].  Errors raise
SRRFError with the library's message; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_float, c_int, c_void_p, c_size_t, c_char_p
from dataclasses import dataclass, fields

import numpy as np

from . import _lib

RFBA_U16 = 1
RFBA_F32 = 2
MAX_SAMPLES = 1024


class SRRFError(RuntimeError):
    pass


class SRRFConf(ctypes.Structure):
    """ABI-identical to kernel.cuh:3-24."""

    _fields_ = [
        ("RingRadius", c_float),
        ("Magnification", c_int),
        ("AxesinRing", c_int),
        ("MainRange", c_int),
        ("XRange", c_int),
        ("YRange", c_int),
        ("ZRange", c_int),
        ("Method", c_int),
        ("ITC", c_int),
        ("TRACOrder", c_int),
        ("RemoveCon", c_int),
        ("Renormalize", c_int),
        ("GradientSmooth", c_int),
        ("IntensityW", c_int),
        ("blockSizeX", c_int),
        ("blockSizeY", c_int),
        ("blockSizeZ", c_int),
    ]

    def replace(self, **kw) -> "SRRFConf":
        c = SRRFConf()
        ctypes.memmove(ctypes.addressof(c), ctypes.addressof(self), ctypes.sizeof(self))
        for k, v in kw.items():
            if not hasattr(c, k):
                raise AttributeError(k)
            setattr(c, k, v)
        return c

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class _Samples(ctypes.Structure):
    _fields_ = [("n", c_int), ("divisor", c_float), ("xyz", POINTER(c_float))]


class _Block(ctypes.Structure):
    _fields_ = [
        ("d_src", c_void_p),
        ("src_dims", c_int * 4),
        ("origin", c_int * 3),
        ("dims", c_int * 3),
        ("frames", c_int),
        ("mag", c_int),
        ("d_out", c_void_p),
    ]


_sig_done = False


def lib() -> ctypes.CDLL:
    global _sig_done
    L = _lib.load("rfba_srrf")
    if not _sig_done:
        L.SRRFGradientCalc.argtypes = [POINTER(c_float), POINTER(c_int), POINTER(c_float)]
        L.SRRFGradientCalc.restype = c_int
        L.SRRFGetMag.argtypes = []
        L.SRRFGetMag.restype = c_int
        L.SRRFGradientCalcGPU.argtypes = [POINTER(c_float), POINTER(c_int), POINTER(c_float), POINTER(SRRFConf)]
        L.SRRFGradientCalcGPU.restype = c_int
        L.rfba_srrf_load_ini.argtypes = [c_char_p, POINTER(SRRFConf)]
        L.rfba_srrf_load_ini.restype = c_int
        L.rfba_srrf_default_conf.argtypes = [POINTER(SRRFConf)]
        L.rfba_srrf_default_conf.restype = None
        L.rfba_srrf_build_ring.argtypes = [c_int, c_float, POINTER(c_float)]
        L.rfba_srrf_build_ring.restype = c_int
        L.rfba_srrf_build_dirs.argtypes = [c_int, c_float, POINTER(c_float)]
        L.rfba_srrf_build_dirs.restype = c_int
        L.rfba_srrf_stack.argtypes = [c_void_p, c_int, POINTER(c_int), POINTER(SRRFConf), POINTER(_Samples),
                                      POINTER(c_float)]
        L.rfba_srrf_stack.restype = c_int
        L.rfba_srrf_stack_device.argtypes = [c_void_p, c_int, POINTER(c_int), POINTER(SRRFConf),
                                             POINTER(_Samples), c_void_p, c_void_p, POINTER(c_float)]
        L.rfba_srrf_stack_device.restype = c_int
        L.rfba_srrf_stack_device_noalloc.argtypes = [c_void_p, c_int, POINTER(c_int), POINTER(SRRFConf),
                                                     POINTER(_Samples), c_void_p, c_void_p]
        L.rfba_srrf_stack_device_noalloc.restype = c_int
        L.rfba_srrf_batch_create.argtypes = [POINTER(_Block), c_int, c_int, POINTER(SRRFConf), POINTER(_Samples),
                                             POINTER(c_void_p)]
        L.rfba_srrf_batch_create.restype = c_int
        L.rfba_srrf_batch_run.argtypes = [c_void_p, c_void_p, c_int, POINTER(c_float)]
        L.rfba_srrf_batch_run.restype = c_int
        L.rfba_srrf_batch_run_per_block.argtypes = [c_void_p, c_void_p, POINTER(c_float)]
        L.rfba_srrf_batch_run_per_block.restype = c_int
        L.rfba_srrf_batch_status.argtypes = [c_void_p]
        L.rfba_srrf_batch_status.restype = c_int
        L.rfba_srrf_batch_autotune.argtypes = [c_void_p, c_void_p, c_int, POINTER(c_int), POINTER(c_float)]
        L.rfba_srrf_batch_autotune.restype = c_int
        L.rfba_srrf_batch_tile_depth.argtypes = [c_void_p]
        L.rfba_srrf_batch_tile_depth.restype = c_int
        L.rfba_srrf_batch_groups.argtypes = [c_void_p, POINTER(c_int), POINTER(c_int), c_int]
        L.rfba_srrf_batch_groups.restype = c_int
        L.rfba_srrf_batch_destroy.argtypes = [c_void_p]
        L.rfba_srrf_batch_destroy.restype = None
        L.rfba_last_error.argtypes = []
        L.rfba_last_error.restype = c_char_p
        L.rfba_device_count.argtypes = []
        L.rfba_device_count.restype = c_int
        L.rfba_malloc_device.argtypes = [POINTER(c_void_p), c_size_t]
        L.rfba_malloc_device.restype = c_int
        L.rfba_free_device.argtypes = [c_void_p]
        L.rfba_free_device.restype = c_int
        L.rfba_memcpy_h2d.argtypes = [c_void_p, c_void_p, c_size_t]
        L.rfba_memcpy_h2d.restype = c_int
        L.rfba_memcpy_d2h.argtypes = [c_void_p, c_void_p, c_size_t]
        L.rfba_memcpy_d2h.restype = c_int
        L.rfba_memset_device.argtypes = [c_void_p, c_int, c_size_t]
        L.rfba_memset_device.restype = c_int
        L.rfba_device_synchronize.argtypes = []
        L.rfba_device_synchronize.restype = c_int
        L.rfba_copy_probe.argtypes = [c_void_p, c_void_p, c_size_t, POINTER(c_float)]
        L.rfba_copy_probe.restype = c_int
        L.rfba_srrf_postprocess_device.argtypes = [c_void_p, c_size_t, c_void_p, POINTER(c_float)]
        L.rfba_srrf_postprocess_device.restype = c_int
        L.rfba_srrf_write_raw.argtypes = [c_char_p, POINTER(c_float), c_size_t]
        L.rfba_srrf_write_raw.restype = c_int
        L.rfba_srrf_read_raw.argtypes = [c_char_p, POINTER(c_float), c_size_t]
        L.rfba_srrf_read_raw.restype = c_int
        L.rfba_srrf_place_spots.argtypes = [c_void_p, POINTER(c_int), c_int, POINTER(c_int), c_int, c_int, ctypes.c_double,
                                            ctypes.c_double, c_int, POINTER(ctypes.c_double),
                                            POINTER(ctypes.c_ulonglong)]
        L.rfba_srrf_place_spots.restype = c_int
        L.rfba_release_cache.argtypes = []
        L.rfba_release_cache.restype = None
        _sig_done = True
    return L


def _check(rc: int):
    if rc != 0:
        raise SRRFError("rfba error %d: %s" % (rc, lib().rfba_last_error().decode(errors="replace")))


# ---------------------------------------------------------------- config
def default_conf() -> SRRFConf:
    c = SRRFConf()
    lib().rfba_srrf_default_conf(ctypes.byref(c))
    return c


def load_ini(path: str | None = None) -> SRRFConf:
    c = SRRFConf()
    _check(lib().rfba_srrf_load_ini(path.encode() if path else None, ctypes.byref(c)))
    return c


def SRRFGetMag() -> int:
    return lib().SRRFGetMag()


# ---------------------------------------------------------------- samples
@dataclass
class Samples:
    """Ordered sample offsets (SR pixels) and the divisor of their signed sum."""

    xyz: np.ndarray  # float32 [n, 3]
    divisor: float

    @property
    def n(self) -> int:
        return int(self.xyz.shape[0])

    def _c(self):
        self._keep = np.ascontiguousarray(self.xyz, dtype=np.float32)
        return _Samples(self.n, float(self.divisor), self._keep.ctypes.data_as(POINTER(c_float)))


def ring(axes_in_ring: int, radius: float) -> Samples:
    """Reference ring (kernel.cu:23-33): (2*axes)^2 samples, divisor 2*axes."""
    n = 2 * axes_in_ring
    buf = np.zeros((max(n * n, 1), 3), np.float32)
    rc = lib().rfba_srrf_build_ring(axes_in_ring, radius, buf.ctypes.data_as(POINTER(c_float)))
    if rc < 0:
        _check(rc)
    return Samples(buf[:rc], float(n))


def directions(kind: int, radius: float, divisor: float | None = None) -> Samples:
    """6 axis or 26 neighbourhood directions at `radius`; divisor defaults to n."""
    buf = np.zeros((26, 3), np.float32)
    rc = lib().rfba_srrf_build_dirs(kind, radius, buf.ctypes.data_as(POINTER(c_float)))
    if rc < 0:
        _check(rc)
    return Samples(buf[:rc], float(rc if divisor is None else divisor))


# ---------------------------------------------------------------- compute
def _dims(a: np.ndarray):
    T, Z, Y, X = a.shape
    return (c_int * 4)(X, Y, Z, T)


def srrf_stack(series: np.ndarray, conf: SRRFConf, samples: Samples | None = None) -> np.ndarray:
    """One stack through the fused kernel (host arrays in and out).

    series: uint16 or float32 [T, Z, Y, X]."""
    if series.dtype == np.uint16:
        dt = RFBA_U16
    elif series.dtype == np.float32:
        dt = RFBA_F32
    else:
        raise TypeError("series must be uint16 or float32")
    a = np.ascontiguousarray(series)
    T, Z, Y, X = a.shape
    M = conf.Magnification
    out = np.empty((Z * M, Y * M, X * M), np.float32)
    smp = samples._c() if samples is not None else None
    _check(lib().rfba_srrf_stack(a.ctypes.data_as(c_void_p), dt, _dims(a), ctypes.byref(conf),
                                 ctypes.byref(smp) if smp is not None else None,
                                 out.ctypes.data_as(POINTER(c_float))))
    return out


def SRRFGradientCalcGPU(series: np.ndarray, conf: SRRFConf) -> np.ndarray:
    """kernel.cu:560: one float stack, reference ring sampling."""
    a = np.ascontiguousarray(series, dtype=np.float32)
    T, Z, Y, X = a.shape
    M = conf.Magnification
    out = np.empty((Z * M, Y * M, X * M), np.float32)
    _check(lib().SRRFGradientCalcGPU(a.ctypes.data_as(POINTER(c_float)), _dims(a),
                                     out.ctypes.data_as(POINTER(c_float)), ctypes.byref(conf)))
    return out


def SRRFGradientCalc(series: np.ndarray, out: np.ndarray | None = None, fill: float = 0.0) -> np.ndarray:
    """GradientCalc.cpp:169: ./SRRF.ini + 16^3 tiling.  Voxels the reference
    never writes (the low band [0, 2M) of every axis) keep `out`'s values."""
    a = np.ascontiguousarray(series, dtype=np.float32)
    T, Z, Y, X = a.shape
    M = SRRFGetMag()
    if out is None:
        out = np.full((Z * M, Y * M, X * M), fill, np.float32)
    _check(lib().SRRFGradientCalc(a.ctypes.data_as(POINTER(c_float)), _dims(a), out.ctypes.data_as(POINTER(c_float))))
    return out


# ---------------------------------------------------------------- device buffers
class DeviceBuffer:
    """A raw device allocation owned by Python (freed on close/del)."""

    def __init__(self, nbytes: int):
        self.ptr = c_void_p()
        self.nbytes = int(nbytes)
        _check(lib().rfba_malloc_device(ctypes.byref(self.ptr), c_size_t(self.nbytes)))

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        _check(lib().rfba_memcpy_h2d(b.ptr, a.ctypes.data_as(c_void_p), c_size_t(a.nbytes)))
        return b

    def to_array(self, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        _check(lib().rfba_memcpy_d2h(out.ctypes.data_as(c_void_p), self.ptr, c_size_t(out.nbytes)))
        return out

    def zero(self):
        self.fill_byte(0)

    def fill_byte(self, value: int):
        """Set every byte of the allocation to `value` (0-255)."""
        _check(lib().rfba_memset_device(self.ptr, int(value), c_size_t(self.nbytes)))

    def close(self):
        if self.ptr:
            lib().rfba_free_device(self.ptr)
            self.ptr = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def srrf_stack_device(d_in: DeviceBuffer, dtype: int, shape_tzyx, conf: SRRFConf, samples: Samples | None,
                      d_out: DeviceBuffer) -> float:
    """Device-resident stack; returns the kernel time (ms) from HIP events."""
    T, Z, Y, X = shape_tzyx
    dims = (c_int * 4)(X, Y, Z, T)
    ms = c_float(0)
    smp = samples._c() if samples is not None else None
    _check(lib().rfba_srrf_stack_device(d_in.ptr, dtype, dims, ctypes.byref(conf),
                                        ctypes.byref(smp) if smp is not None else None, d_out.ptr, None,
                                        ctypes.byref(ms)))
    return ms.value


def copy_probe(nbytes: int) -> float:
    """Run the plain copy kernel over nbytes (device to device); returns ms."""
    a, b = DeviceBuffer(nbytes), DeviceBuffer(nbytes)
    a.zero()
    ms = c_float(0)
    _check(lib().rfba_copy_probe(b.ptr, a.ptr, c_size_t(nbytes), ctypes.byref(ms)))
    a.close()
    b.close()
    return ms.value


# ---------------------------------------------------------------- batched
@dataclass
class Block:
    src: DeviceBuffer
    src_shape_tzyx: tuple
    origin_xyz: tuple
    dims_xyz: tuple
    frames: int
    mag: int
    out: DeviceBuffer


class Batch:
    """Persistent work-queue plan over uneven blocks (rfba_srrf_batch_*)."""

    def __init__(self, blocks: list[Block], dtype: int, conf: SRRFConf, samples: Samples | None):
        arr = (_Block * len(blocks))()
        for i, b in enumerate(blocks):
            T, Z, Y, X = b.src_shape_tzyx
            arr[i].d_src = b.src.ptr
            arr[i].src_dims[:] = [X, Y, Z, T]
            arr[i].origin[:] = list(b.origin_xyz)
            arr[i].dims[:] = list(b.dims_xyz)
            arr[i].frames = b.frames
            arr[i].mag = b.mag
            arr[i].d_out = b.out.ptr
        self._blocks = blocks
        self.plan = c_void_p()
        smp = samples._c() if samples is not None else None
        _check(lib().rfba_srrf_batch_create(arr, len(blocks), dtype, ctypes.byref(conf),
                                            ctypes.byref(smp) if smp is not None else None, ctypes.byref(self.plan)))

    def run(self, use_graph: bool = True) -> float:
        ms = c_float(0)
        _check(lib().rfba_srrf_batch_run(self.plan, None, int(use_graph), ctypes.byref(ms)))
        return ms.value

    def run_per_block(self) -> float:
        ms = c_float(0)
        _check(lib().rfba_srrf_batch_run_per_block(self.plan, None, ctypes.byref(ms)))
        return ms.value

    def run_async(self, use_graph: bool = True):
        """Enqueue a run and return at once; status() waits for it."""
        _check(lib().rfba_srrf_batch_run(self.plan, None, int(use_graph), None))

    def status(self):
        """Wait for the last run; raises if a kernel flagged an error."""
        _check(lib().rfba_srrf_batch_status(self.plan))

    def autotune(self, sample_stride: int = 0):
        """Per magnification group, time tile depths 4/8/16 on a sample of its
        blocks and keep the fastest.  Returns (depth, {depth: sample ms}) of
        the group with the most work."""
        tz = c_int(0)
        ms = (c_float * 3)()
        _check(lib().rfba_srrf_batch_autotune(self.plan, None, sample_stride, ctypes.byref(tz), ms))
        return tz.value, {d: ms[i] for i, d in enumerate((4, 8, 16)) if ms[i] > 0}

    @property
    def tile_depth(self) -> int:
        return lib().rfba_srrf_batch_tile_depth(self.plan)

    @property
    def groups(self):
        """[(magnification, tile depth)] per group, largest work first."""
        m, t = (c_int * 64)(), (c_int * 64)()
        n = lib().rfba_srrf_batch_groups(self.plan, m, t, 64)
        return [(m[i], t[i]) for i in range(min(n, 64))]

    def close(self):
        if self.plan:
            lib().rfba_srrf_batch_destroy(self.plan)
            self.plan = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- caller post-processing
def postprocess_device(d_out: DeviceBuffer, n: int) -> float:
    """In place on the device: NaN -> 0, negative -> 0; returns max(1, max)
    (multispot5_place_methods.cc:260-269)."""
    fm = c_float(0)
    _check(lib().rfba_srrf_postprocess_device(d_out.ptr, c_size_t(n), None, ctypes.byref(fm)))
    return fm.value


def write_raw(path: str, volume: np.ndarray):
    """SRRF.raw: raw float32, x fastest (multispot5_place_methods.cc:270-274)."""
    a = np.ascontiguousarray(volume, np.float32)
    _check(lib().rfba_srrf_write_raw(path.encode(), a.ctypes.data_as(POINTER(c_float)), c_size_t(a.size)))


def read_raw(path: str, shape_zyx) -> np.ndarray:
    """Read SRRF.raw back (multispot5_place_methods.cc:284-289)."""
    out = np.empty(shape_zyx, np.float32)
    _check(lib().rfba_srrf_read_raw(path.encode(), out.ctypes.data_as(POINTER(c_float)), c_size_t(out.size)))
    return out


def place_spots(d_sr: DeviceBuffer, dims_xyz, mag: int, pixels, num_spots: int, intensity_mode: float,
                blur_mode: float, seed: int):
    """SRRF-guided spot placement (multispot5_place_methods.cc:293-341) from a
    device SR volume.  pixels [n, 3] low-res (x, y, z).  Returns (spots [k, 5]
    as (I, sigma, x, y, z), generator draws used)."""
    px = np.ascontiguousarray(pixels, np.int32)
    dims = (c_int * 3)(*dims_xyz)
    out = np.zeros((max(num_spots, 1), 5), np.float64)
    used = ctypes.c_ulonglong(0)
    rc = lib().rfba_srrf_place_spots(d_sr.ptr, dims, mag, px.ctypes.data_as(POINTER(c_int)), len(px), num_spots,
                                     intensity_mode, blur_mode, seed, out.ctypes.data_as(POINTER(ctypes.c_double)),
                                     ctypes.byref(used))
    if rc < 0:
        _check(rc)
    return out[:rc], used.value
