"""Python mirror of the 3B likelihood interface (librfba_3b.so via ctypes).

Reference: the gradient functor SpotNegProbabilityDiffWithSampledBackground
(RFBA Main/multispot5.cc:579-669) over SpotWithBackground
(spot_with_background.hh:57-152) and forward_algorithm_deriv
(forward_algorithm.h:157-332); forward messages for FFBS as
forward_algorithm_delta (forward_algorithm.h:344-382).  Parameters are
(I, sigma, x, y, z); states 0 = on, 1 = off, 2 = bleached.  Errors raise
B3Error; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_float, c_int, c_uint8, c_uint64, c_void_p, c_char_p
from dataclasses import dataclass

import numpy as np

from . import _lib
from .srrf import DeviceBuffer
from .srrf import _check as _srrf_check
from .srrf import lib as _srrf_lib


class B3Error(RuntimeError):
    pass


class Model(ctypes.Structure):
    """ABI of rfba3b_model."""

    _fields_ = [
        ("A", c_double * 9),
        ("pi", c_double * 3),
        ("variance", c_double),
        ("zfactor", c_double),
        ("z_exact", c_int),
        ("checkpoint", c_int),
        ("mu_brightness", c_double),
        ("sigma_brightness", c_double),
        ("mu_blur", c_double),
        ("sigma_blur", c_double),
    ]


def make_model(A, pi, variance, zfactor=1.0, z_exact=0, checkpoint=0, mu_brightness=0.0, sigma_brightness=1.0,
               mu_blur=0.0, sigma_blur=1.0) -> Model:
    m = Model()
    m.A[:] = [float(v) for v in np.asarray(A, np.float64).ravel()]
    m.pi[:] = [float(v) for v in np.asarray(pi, np.float64).ravel()]
    m.variance = variance
    m.zfactor = zfactor
    m.z_exact = z_exact
    m.checkpoint = checkpoint
    m.mu_brightness = mu_brightness
    m.sigma_brightness = sigma_brightness
    m.mu_blur = mu_blur
    m.sigma_blur = sigma_blur
    return m


def blink_matrix(p_on_off=0.2, p_off_on=0.1, p_bleach=0.01):
    """3-state transition matrix (on, off, bleached), A[i][j] = P(i -> j)."""
    return np.array([[1 - p_on_off - p_bleach, p_on_off, p_bleach],
                     [p_off_on, 1 - p_off_on - p_bleach, p_bleach],
                     [0.0, 0.0, 1.0]])


class _Block(ctypes.Structure):
    _fields_ = [
        ("dims", c_int * 3),
        ("T", c_int), ("K", c_int), ("Q", c_int), ("S", c_int),
        ("b0", c_double),
        ("d_data", c_void_p), ("d_params", c_void_p), ("d_states", c_void_p),
        ("d_logL", c_void_p), ("d_grad", c_void_p),
    ]


_sig = False


class _GibbsChain(ctypes.Structure):
    _fields_ = [("P", c_int), ("T", c_int), ("K", c_int), ("d_data", c_void_p), ("d_spot", c_void_p),
                ("d_mask_off", c_void_p), ("d_mask", c_void_p), ("base", c_double), ("d_sample", c_void_p),
                ("d_samples", c_void_p), ("seed", ctypes.c_uint), ("skip", ctypes.c_ulonglong)]


def lib():
    global _sig
    L = _lib.load("rfba_3b")
    if not _sig:
        ip, fp, dp = POINTER(c_int), POINTER(c_float), POINTER(c_double)
        L.rfba3b_neg_grad_sampled.argtypes = [ip, c_int, c_int, fp, fp, c_int, dp, c_int, POINTER(Model), dp, dp]
        L.rfba3b_neg_grad_sampled.restype = c_int
        L.rfba3b_functor_create.argtypes = [ip, c_int, c_int, dp, dp, c_int, POINTER(Model), c_int,
                                            POINTER(c_void_p)]
        L.rfba3b_functor_create.restype = c_int
        L.rfba3b_functor_eval.argtypes = [c_void_p, dp, c_int, dp, dp, c_void_p]
        L.rfba3b_functor_eval.restype = c_int
        L.rfba3b_functor_hessian_ffbs.argtypes = [c_void_p, dp, c_int, dp, dp, dp, POINTER(c_uint8), c_void_p]
        L.rfba3b_functor_hessian_ffbs.restype = c_int
        L.rfba3b_functor_destroy.argtypes = [c_void_p]
        L.rfba3b_functor_destroy.restype = None
        L.rfba3b_forward_messages.argtypes = [ip, c_int, c_int, fp, fp, dp, POINTER(Model), dp, dp]
        L.rfba3b_forward_messages.restype = c_int
        L.rfba3b_ffbs.argtypes = [ip, c_int, c_int, fp, fp, dp, POINTER(Model), dp, POINTER(c_uint8), dp, dp]
        L.rfba3b_ffbs.restype = c_int
        L.rfba3b_backward_sample_device.argtypes = [c_void_p, c_int, c_int, POINTER(Model), c_void_p, c_void_p,
                                                    c_void_p]
        L.rfba3b_backward_sample_device.restype = c_int
        L.rfba3b_mt19937_uniforms.argtypes = [c_int, ctypes.c_ulonglong, ctypes.c_size_t, dp]
        L.rfba3b_mt19937_uniforms.restype = c_int
        L.rfba3b_batch_create.argtypes = [POINTER(_Block), c_int, POINTER(Model), POINTER(c_void_p)]
        L.rfba3b_batch_create.restype = c_int
        L.rfba3b_batch_run.argtypes = [c_void_p, c_void_p, c_int, POINTER(c_float)]
        L.rfba3b_batch_run.restype = c_int
        L.rfba3b_batch_run_per_block.argtypes = [c_void_p, c_void_p, c_int, POINTER(c_float)]
        L.rfba3b_batch_run_per_block.restype = c_int
        L.rfba3b_batch_chunks.argtypes = [c_void_p]
        L.rfba3b_batch_chunks.restype = c_int
        L.rfba3b_batch_set_hmm_form.argtypes = [c_void_p, c_int]
        L.rfba3b_batch_set_hmm_form.restype = c_int
        L.rfba3b_batch_destroy.argtypes = [c_void_p]
        L.rfba3b_batch_destroy.restype = None
        L.rfba3b_gibbs_create.argtypes = [POINTER(_GibbsChain), c_int, POINTER(Model), POINTER(c_void_p),
                                          POINTER(c_float)]
        L.rfba3b_gibbs_create.restype = c_int
        L.rfba3b_gibbs_reset.argtypes = [c_void_p, c_void_p]
        L.rfba3b_gibbs_reset.restype = c_int
        L.rfba3b_gibbs_run.argtypes = [c_void_p, c_int, c_int, POINTER(c_double), c_void_p, POINTER(c_float)]
        L.rfba3b_gibbs_run.restype = c_int
        L.rfba3b_gibbs_sample_intensities.argtypes = [c_void_p, c_int, c_void_p, c_void_p]
        L.rfba3b_gibbs_sample_intensities.restype = c_int
        L.rfba3b_gibbs_draws.argtypes = [c_void_p, c_int, POINTER(ctypes.c_ulonglong)]
        L.rfba3b_gibbs_draws.restype = c_int
        L.rfba3b_gibbs_destroy.argtypes = [c_void_p]
        L.rfba3b_gibbs_destroy.restype = None
        L.rfba3b_spot_pixels.argtypes = [POINTER(c_int), c_int, POINTER(c_double), c_int, POINTER(c_int),
                                         POINTER(c_int), ctypes.c_size_t]
        L.rfba3b_spot_pixels.restype = c_int
        L.rfba3b_spot_intensities_device.argtypes = [c_void_p, c_int, c_void_p, c_int, c_double, c_void_p, c_void_p]
        L.rfba3b_spot_intensities_device.restype = c_int
        L.rfba3b_last_error.argtypes = []
        L.rfba3b_last_error.restype = c_char_p
        _sig = True
    return L


def _check(rc):
    if rc != 0:
        raise B3Error("rfba3b error %d: %s" % (rc, lib().rfba3b_last_error().decode(errors="replace")))


def _p(a, t):
    return a.ctypes.data_as(POINTER(t))


def neg_grad_sampled(xyz, data, bg, phi, model: Model):
    """The CG functor: -(mean_s grad logL(phi; bg_s) + grad log-prior).

    xyz [P,3] int, data [T,P], bg [S,T,P], phi [n,5] -> (grad [n,5], logL [n,S])."""
    xyz = np.ascontiguousarray(xyz, np.int32)
    data = np.ascontiguousarray(data, np.float32)
    bg = np.ascontiguousarray(bg, np.float32)
    phi = np.ascontiguousarray(np.atleast_2d(phi), np.float64)
    S, T, P = bg.shape
    n = phi.shape[0]
    g = np.empty((n, 5), np.float64)
    L = np.empty((n, S), np.float64)
    _check(lib().rfba3b_neg_grad_sampled(_p(xyz, c_int), P, T, _p(data, c_float), _p(bg, c_float), S,
                                         _p(phi, c_double), n, ctypes.byref(model), _p(g, c_double), _p(L, c_double)))
    return g, L


class Functor:
    """SpotNegProbabilityDiffWithSampledBackground kept on the device
    (rfba3b_functor_*): created once from the pixel list, the data [T, P]
    and the S background samples [S, T, P] (FP64), then evaluated at
    candidate parameter vectors without further allocation."""

    def __init__(self, xyz, data, bg, model: Model, max_candidates=1):
        xyz = np.ascontiguousarray(xyz, np.int32)
        data = np.ascontiguousarray(data, np.float64)
        bg = np.ascontiguousarray(bg, np.float64)
        S, T, P = bg.shape
        assert data.shape == (T, P) and xyz.shape == (P, 3)
        self.S, self.nmax = S, max_candidates
        self.model = model
        self.h = c_void_p()
        _check(lib().rfba3b_functor_create(_p(xyz, c_int), P, T, _p(data, c_double), _p(bg, c_double), S,
                                           ctypes.byref(model), max_candidates, ctypes.byref(self.h)))

    def __call__(self, phi):
        """grad [n, 5] (NaN rows for invalid candidates) and logL [n, S]."""
        phi = np.ascontiguousarray(np.atleast_2d(phi), np.float64)
        n = phi.shape[0]
        g = np.empty((n, 5))
        L = np.empty((n, self.S))
        _check(lib().rfba3b_functor_eval(self.h, _p(phi, c_double), n, _p(g, c_double), _p(L, c_double), None))
        return g, L

    def hessian_ffbs(self, phi, bs_iterations, uniforms):
        """sampled_background_spot_hessian_ffbs at phi with the draws
        uniforms [S, bs, T]: (hess [5, 5], deriv [5], paths [S, bs, T])."""
        phi = np.ascontiguousarray(phi, np.float64)
        u = np.ascontiguousarray(uniforms, np.float64)
        assert u.shape[:2] == (self.S, bs_iterations)
        h, d = np.empty((5, 5)), np.empty(5)
        paths = np.empty(u.shape, np.uint8)
        _check(lib().rfba3b_functor_hessian_ffbs(self.h, _p(phi, c_double), bs_iterations, _p(u, c_double),
                                                 _p(h, c_double), _p(d, c_double), _p(paths, c_uint8), None))
        return h, d, paths

    def close(self):
        if self.h:
            lib().rfba3b_functor_destroy(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def forward_messages(xyz, data, bg, phi, model: Model):
    """delta [T,3] (ln forward messages) and logL of one spot over one background."""
    xyz = np.ascontiguousarray(xyz, np.int32)
    data = np.ascontiguousarray(data, np.float32)
    bg = np.ascontiguousarray(bg, np.float32)
    phi = np.ascontiguousarray(phi, np.float64).ravel()
    T, P = data.shape
    delta = np.empty((T, 3), np.float64)
    L = c_double(0)
    _check(lib().rfba3b_forward_messages(_p(xyz, c_int), P, T, _p(data, c_float), _p(bg, c_float),
                                         _p(phi, c_double), ctypes.byref(model), _p(delta, c_double), ctypes.byref(L)))
    return delta, L.value


def ffbs(xyz, data, bg, phi, model: Model, uniforms):
    """Forward filtering / backward sampling of one spot's states (0 on, 1
    off, 2 bleached).  uniforms [T]: the generator's draws, uniforms[0] for
    the last frame.  Returns (states [T] uint8, delta [T,3], logL)."""
    xyz = np.ascontiguousarray(xyz, np.int32)
    data = np.ascontiguousarray(data, np.float32)
    bg = np.ascontiguousarray(bg, np.float32)
    phi = np.ascontiguousarray(phi, np.float64).ravel()
    u = np.ascontiguousarray(uniforms, np.float64)
    T, P = data.shape
    states = np.empty(T, np.uint8)
    delta = np.empty((T, 3), np.float64)
    L = c_double(0)
    _check(lib().rfba3b_ffbs(_p(xyz, c_int), P, T, _p(data, c_float), _p(bg, c_float), _p(phi, c_double),
                             ctypes.byref(model), _p(u, c_double), _p(states, c_uint8), _p(delta, c_double),
                             ctypes.byref(L)))
    return states, delta, L.value


def backward_sample(delta, uniforms, model: Model):
    """Backward sampling of n chains on the device: delta [n,T,3], uniforms
    [n,T] -> states [n,T] uint8."""
    delta = np.ascontiguousarray(delta, np.float64)
    u = np.ascontiguousarray(uniforms, np.float64)
    n, T, _ = delta.shape
    dd, du, ds = DeviceBuffer.from_array(delta), DeviceBuffer.from_array(u), DeviceBuffer(n * T)
    try:
        _check(lib().rfba3b_backward_sample_device(dd.ptr, n, T, ctypes.byref(model), du.ptr, ds.ptr, None))
        return ds.to_array((n, T), np.uint8)
    finally:
        for b in (dd, du, ds):
            b.close()


def mt19937_uniforms(seed: int, n: int, skip: int = 0) -> np.ndarray:
    """The reference's rng() stream for a seed (MT19937, 37 outputs discarded
    at seeding, draws = output * 2^-32)."""
    out = np.empty(n, np.float64)
    _check(lib().rfba3b_mt19937_uniforms(int(seed), skip, n, _p(out, c_double)))
    return out


@dataclass
class BlockData:
    data: np.ndarray     # float32 [T, Z, Y, X]
    params: np.ndarray   # float64 [Q, K, 5]
    states: np.ndarray   # uint8 [S, K, T]
    b0: float


class Batch:
    """Device plan over blocks: logL [Q,K,S] and grad [Q,K,S,5] per block."""

    def __init__(self, blocks: list[BlockData], model: Model, with_grad=True):
        self.model = model
        self.blocks = blocks
        self._bufs = []
        arr = (_Block * len(blocks))()
        self.out = []
        for i, b in enumerate(blocks):
            T, Z, Y, X = b.data.shape
            Q, K, _ = b.params.shape
            S = b.states.shape[0]
            dd = DeviceBuffer.from_array(np.ascontiguousarray(b.data, np.float32))
            dp = DeviceBuffer.from_array(np.ascontiguousarray(b.params, np.float64))
            ds = DeviceBuffer.from_array(np.ascontiguousarray(b.states, np.uint8))
            dl = DeviceBuffer(Q * K * S * 8)
            dg = DeviceBuffer(Q * K * S * 5 * 8) if with_grad else None
            self._bufs += [dd, dp, ds, dl] + ([dg] if dg else [])
            self.out.append((dl, dg, (Q, K, S)))
            a = arr[i]
            a.dims[:] = [X, Y, Z]
            a.T, a.K, a.Q, a.S = T, K, Q, S
            a.b0 = b.b0
            a.d_data, a.d_params, a.d_states = dd.ptr, dp.ptr, ds.ptr
            a.d_logL = dl.ptr
            a.d_grad = dg.ptr if dg else None
        self.plan = c_void_p()
        _check(lib().rfba3b_batch_create(arr, len(blocks), ctypes.byref(model), ctypes.byref(self.plan)))

    def run(self, use_graph=True) -> float:
        ms = c_float(0)
        _check(lib().rfba3b_batch_run(self.plan, None, int(use_graph), ctypes.byref(ms)))
        return ms.value

    def run_per_block(self, per_param_set=False) -> float:
        """Plain form: one launch sequence per block (or per (block, q))."""
        ms = c_float(0)
        _check(lib().rfba3b_batch_run_per_block(self.plan, None, int(per_param_set), ctypes.byref(ms)))
        return ms.value

    def set_hmm_form(self, form: int) -> int:
        """0: one wave per item for T <= 512 (the default), else one lane per
        item; 1: always the lane form with the model's checkpoint stride.
        Returns the previous form."""
        rc = lib().rfba3b_batch_set_hmm_form(self.plan, int(form))
        if rc < 0:
            _check(rc)
        return rc

    @property
    def chunks(self) -> int:
        """Memory chunks of the plan (RFBA3B_PLANE_BUDGET_MB at creation)."""
        return lib().rfba3b_batch_chunks(self.plan)

    def results(self, i):
        dl, dg, (Q, K, S) = self.out[i]
        L = dl.to_array((Q, K, S), np.float64)
        G = dg.to_array((Q, K, S, 5), np.float64) if dg else None
        return L, G

    def close(self):
        if self.plan:
            lib().rfba3b_batch_destroy(self.plan)
            self.plan = c_void_p()
        for b in self._bufs:
            b.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------ Gibbs sampling
def box_pixels(dims_xyz, origin_xyz=(0, 0, 0)):
    """Pixel list [X*Y*Z, 3] of a dense box in the region order (z, then y,
    then x), coordinates offset by origin."""
    X, Y, Z = dims_xyz
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    xyz = np.stack([x.ravel() + origin_xyz[0], y.ravel() + origin_xyz[1], z.ravel() + origin_xyz[2]], 1)
    return np.ascontiguousarray(xyz, np.int32)


def spot_pixels(xyz, phi):
    """get_spot_pixels (multispot5.cc:1001-1020) of spots phi [K,5] in the
    pixel list xyz [P,3] (sorted by z, y, x): (offsets [K+1], indices)."""
    xyz = np.ascontiguousarray(xyz, np.int32)
    phi = np.ascontiguousarray(np.atleast_2d(phi), np.float64)
    K, P = phi.shape[0], len(xyz)
    off = np.empty(K + 1, np.int32)
    _check(lib().rfba3b_spot_pixels(_p(xyz, c_int), P, _p(phi, c_double), K, _p(off, c_int), None, 0))
    idx = np.empty(max(int(off[-1]), 1), np.int32)
    _check(lib().rfba3b_spot_pixels(_p(xyz, c_int), P, _p(phi, c_double), K, _p(off, c_int), _p(idx, c_int),
                                    idx.size))
    return off, idx[:int(off[-1])]


@dataclass
class ChainData:
    data: np.ndarray         # float32 [T, P] (or [T, Z, Y, X] over box_pixels order)
    xyz: np.ndarray          # int32 [P, 3], sorted by (z, y, x)
    spots: np.ndarray        # float64 [K, 5]
    base: float = 0.0        # intensity of the empty sample
    seed: int = 0
    skip: int = 0
    slots: int = 0           # S: sample copies kept by Gibbs.run (d_samples [S, K, T])


class DeviceView:
    """A byte range of a pooled DeviceBuffer; keeps the pool alive."""

    def __init__(self, pool: DeviceBuffer, offset: int, nbytes: int):
        self.pool = pool
        self.nbytes = int(nbytes)
        self.ptr = c_void_p(pool.ptr.value + int(offset))

    def write(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _srrf_check(_srrf_lib().rfba_memcpy_h2d(self.ptr, a.ctypes.data_as(c_void_p), ctypes.c_size_t(a.nbytes)))

    def to_array(self, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        _srrf_check(_srrf_lib().rfba_memcpy_d2h(out.ctypes.data_as(c_void_p), self.ptr,
                                                ctypes.c_size_t(out.nbytes)))
        return out


class Gibbs:
    """Device Gibbs sampler over chains (GibbsSampler2, one workgroup per
    chain).  Spot tables are computed on the device, masks on the host;
    setup_ms is the device time of the plan's statistics."""

    def __init__(self, chains: list[ChainData], model: Model):
        self.model = model
        self._bufs = []
        self.meta = []
        arr = (_GibbsChain * len(chains))()
        # host-side arrays per chain, then one device allocation per kind with
        # every chain at an aligned offset (thousands of chains would
        # otherwise cost thousands of allocations)
        host = []
        for ch in chains:
            T = ch.data.shape[0]
            data = np.ascontiguousarray(ch.data, np.float32).reshape(T, -1)
            P = data.shape[1]
            phi = np.ascontiguousarray(np.atleast_2d(ch.spots), np.float64).reshape(-1, 5)
            K = phi.shape[0]
            xyz = np.ascontiguousarray(ch.xyz, np.int32)
            assert xyz.shape == (P, 3)
            off, idx = spot_pixels(xyz, phi)
            host.append((data, xyz, phi, off, idx if idx.size else np.zeros(1, np.int32), P, T, K, ch))
        kinds = {
            "data": lambda h: h[0].nbytes, "xyz": lambda h: h[1].nbytes, "phi": lambda h: max(h[2].nbytes, 8),
            "spot": lambda h: max(h[7] * h[5], 1) * 8, "off": lambda h: h[3].nbytes, "idx": lambda h: h[4].nbytes,
            "smp": lambda h: max(h[7] * h[6], 1), "out": lambda h: max(h[8].slots * h[7] * h[6], 1)}
        views = [dict() for _ in host]
        for kind, size in kinds.items():
            offs, total = [], 0
            for h in host:
                offs.append(total)
                total += (size(h) + 255) & ~255
            pool = DeviceBuffer(max(total, 256))
            self._bufs.append(pool)
            for i, h in enumerate(host):
                views[i][kind] = DeviceView(pool, offs[i], size(h))
        for i, (data, xyz, phi, off, idx, P, T, K, ch) in enumerate(host):
            v = views[i]
            v["data"].write(data)
            v["xyz"].write(xyz)
            v["off"].write(off)
            v["idx"].write(idx)
            if K:
                v["phi"].write(phi)
                _check(lib().rfba3b_spot_intensities_device(v["xyz"].ptr, P, v["phi"].ptr, K, model.zfactor,
                                                            v["spot"].ptr, None))
            out = v["out"] if ch.slots else None
            self.meta.append(dict(P=P, T=T, K=K, slots=ch.slots, spot=v["spot"], sample=v["smp"], out=out,
                                  offsets=off, idx=idx))
            a = arr[i]
            a.P, a.T, a.K = P, T, K
            a.d_data, a.d_spot, a.d_mask_off, a.d_mask = v["data"].ptr, v["spot"].ptr, v["off"].ptr, v["idx"].ptr
            a.base = float(ch.base)
            a.d_sample = v["smp"].ptr
            a.d_samples = out.ptr if out else None
            a.seed = int(ch.seed) & 0xFFFFFFFF
            a.skip = int(ch.skip)
        self.plan = c_void_p()
        ms = c_float(0)
        _check(lib().rfba3b_gibbs_create(arr, len(chains), ctypes.byref(model), ctypes.byref(self.plan),
                                         ctypes.byref(ms)))
        self.setup_ms = ms.value

    def reset(self):
        _check(lib().rfba3b_gibbs_reset(self.plan, None))

    def run(self, samples=1, sample_iterations=1, variance=None) -> float:
        """`samples` next() steps of every chain; returns the kernel time (ms)."""
        for m in self.meta:
            if m["out"] is not None and samples > m["slots"]:
                raise B3Error("chain keeps %d sample slots, %d requested" % (m["slots"], samples))
        var = None
        if variance is not None:
            var = np.ascontiguousarray(np.broadcast_to(np.asarray(variance, np.float64), (samples,)))
        ms = c_float(0)
        _check(lib().rfba3b_gibbs_run(self.plan, samples, sample_iterations,
                                      _p(var, c_double) if var is not None else None, None, ctypes.byref(ms)))
        return ms.value

    def sample(self, i):
        m = self.meta[i]
        return m["sample"].to_array((m["K"], m["T"]), np.uint8)

    def samples(self, i, n=None):
        m = self.meta[i]
        n = m["slots"] if n is None else n
        return m["out"].to_array((n, m["K"], m["T"]), np.uint8)

    def sample_intensities(self, i):
        """Current sample intensities + base of chain i, [T, P]."""
        m = self.meta[i]
        out = DeviceBuffer(m["T"] * m["P"] * 8)
        try:
            _check(lib().rfba3b_gibbs_sample_intensities(self.plan, i, out.ptr, None))
            return out.to_array((m["T"], m["P"]), np.float64)
        finally:
            out.close()

    def spot_table(self, i):
        m = self.meta[i]
        return m["spot"].to_array((m["K"], m["P"]), np.float64)

    def draws(self, i):
        v = ctypes.c_ulonglong(0)
        _check(lib().rfba3b_gibbs_draws(self.plan, i, ctypes.byref(v)))
        return v.value

    def close(self):
        if self.plan:
            lib().rfba3b_gibbs_destroy(self.plan)
            self.plan = c_void_p()
        for b in self._bufs:
            b.close()
        self._bufs = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- workload
def synth_block(dims_xyz, T, K, seed, I=1304.0, sigma=1.3, zfactor=2.0, b0=100.0, variance=100.0, A=None,
                pi=(0.5, 0.5, 0.0)):
    """Seeded block: (data [T,Z,Y,X] float32, true params [K,5], states [K,T])."""
    from ctypes import c_uint16  # noqa: F401
    L = _lib.load("rfba_synth")
    L.rfba_synth_3b_block.argtypes = [c_int, c_int, c_int, c_int, c_int, c_uint64, c_double, c_double, c_double,
                                      c_double, c_double, POINTER(c_double), POINTER(c_double), POINTER(c_float),
                                      POINTER(c_double), POINTER(c_uint8)]
    L.rfba_synth_3b_block.restype = c_int
    X, Y, Z = dims_xyz
    A = blink_matrix() if A is None else np.asarray(A)
    A = np.ascontiguousarray(A, np.float64)
    pi = np.ascontiguousarray(pi, np.float64)
    data = np.empty((T, Z, Y, X), np.float32)
    params = np.empty((K, 5), np.float64)
    states = np.empty((K, T), np.uint8)
    rc = L.rfba_synth_3b_block(X, Y, Z, T, K, seed, I, sigma, zfactor, b0, variance, _p(A, c_double),
                               _p(pi, c_double), _p(data, c_float), _p(params, c_double), _p(states, c_uint8))
    if rc:
        raise B3Error("generator failed")
    return data, params, states


def synth_blocks(dims_list, T, K_list, seed0, I=1304.0, sigma=1.3, zfactor=2.0, b0=100.0, variance=100.0, A=None,
                 pi=(0.5, 0.5, 0.0), threads=None):
    """Many seeded blocks at once (host threads over blocks; block i uses
    seed seed0 + i and spots truncated at 5 sigma).  Returns a list of
    (data [T,Z,Y,X] float32, true params [K,5], states [K,T]) views."""
    from .synth import default_threads
    L = _lib.load("rfba_synth")
    lp = POINTER(ctypes.c_longlong)
    L.rfba_synth_3b_blocks.argtypes = [c_int, POINTER(c_int), c_int, POINTER(c_int), c_uint64, c_double,
                                       c_double, c_double, c_double, c_double, POINTER(c_double), POINTER(c_double),
                                       POINTER(c_float), lp, POINTER(c_double), lp, POINTER(c_uint8), lp, c_int]
    L.rfba_synth_3b_blocks.restype = c_int
    n = len(dims_list)
    dims = np.ascontiguousarray(np.asarray(dims_list, np.int32).reshape(n, 3))
    K = np.ascontiguousarray(np.asarray(K_list, np.int32))
    P = dims.prod(1).astype(np.int64)
    doff = np.concatenate([[0], np.cumsum(P * T)[:-1]]).astype(np.int64)
    soff = np.concatenate([[0], np.cumsum(K)[:-1]]).astype(np.int64)
    toff = soff * T
    data = np.empty(int((P * T).sum()), np.float32)
    params = np.empty((int(K.sum()), 5), np.float64)
    states = np.empty(int(K.sum()) * T, np.uint8)
    A = np.ascontiguousarray(blink_matrix() if A is None else A, np.float64)
    pi = np.ascontiguousarray(pi, np.float64)
    rc = L.rfba_synth_3b_blocks(n, _p(dims, c_int), T, _p(K, c_int), seed0, I, sigma, zfactor, b0, variance,
                                _p(A, c_double), _p(pi, c_double), _p(data, c_float), _p(doff, ctypes.c_longlong),
                                _p(params, c_double), _p(soff, ctypes.c_longlong), _p(states, c_uint8),
                                _p(toff, ctypes.c_longlong), threads or default_threads())
    if rc:
        raise B3Error("generator failed")
    out = []
    for i in range(n):
        X, Y, Z = (int(v) for v in dims[i])
        d = data[doff[i]:doff[i] + P[i] * T].reshape(T, Z, Y, X)
        out.append((d, params[soff[i]:soff[i] + K[i]], states[toff[i]:toff[i] + K[i] * T].reshape(int(K[i]), T)))
    return out


def perturb(params, Q, seed, pos=0.5, rel_int=0.2):
    """Q seeded perturbations of all spots: positions +-pos px, intensity +-rel_int."""
    rng = np.random.default_rng(seed)
    K = params.shape[0]
    out = np.repeat(params[None], Q, 0).copy()
    out[:, :, 2:5] += rng.uniform(-pos, pos, (Q, K, 3))
    out[:, :, 0] *= 1 + rng.uniform(-rel_int, rel_int, (Q, K))
    return out
