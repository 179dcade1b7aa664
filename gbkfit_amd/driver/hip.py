"""The ``hip`` driver: a drop-in sibling of gbkfit's ``host`` and ``cuda``
drivers (src/gbkfit/driver/drivers/{host,cuda}/driver.py) running the hot
path on an MI355X through libgbkhip.

Device memory is owned by torch-ROCm tensors (buffer ownership only; no torch
compute is on the evaluation path).  Host memory is pinned numpy arrays.
Backend methods keep the reference's Python signatures and marshal pointers
and reversed shapes exactly as src/gbkfit/driver/drivers/_detail/native.py
does, but call the C ABI of include/gbkhip.h.
"""
import ctypes

import numpy as np

from .. import _native as N
from .backend import (DriverBackendDModel, DriverBackendFFT,
                      DriverBackendGModel, DriverBackendObjective,
                      DriverBackends)
from .core import Driver, register_driver

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t
        _torch = t
    return _torch


def _tdtype(dtype):
    t = torch()
    dt = np.dtype(dtype)
    table = {np.dtype(np.float32): t.float32, np.dtype(np.float64): t.float64,
             np.dtype(np.int32): t.int32, np.dtype(np.int64): t.int64,
             np.dtype(np.uint8): t.uint8, np.dtype(np.bool_): t.bool,
             np.dtype(np.complex64): t.complex64,
             np.dtype(np.int16): t.int16}
    if dt not in table:
        raise RuntimeError(f"unsupported dtype {dt}")
    return table[dt]


def _shape(shape):
    if isinstance(shape, (int, np.integer)):
        return (int(shape),)
    return tuple(int(s) for s in shape)


class NativeMemoryHip:
    """Pointer / size / shape access to device tensors (the role of
    NativeMemoryCuda, drivers/cuda/_backend.py:14-30).  None maps to 0."""

    @staticmethod
    def ptr(x):
        if x is None:
            return None
        t = torch()
        if isinstance(x, t.Tensor):
            if not x.is_cuda:
                raise RuntimeError("expected a device tensor")
            if not x.is_contiguous():
                raise RuntimeError("expected a contiguous device tensor")
            return x.data_ptr()
        raise RuntimeError(f"expected a device tensor, got {type(x)}")

    @staticmethod
    def size(x):
        return 0 if x is None else int(x.numel())

    @staticmethod
    def shape(x):
        return None if x is None else tuple(int(s) for s in x.shape)


def _ptr(x):
    return NativeMemoryHip.ptr(x)


class DriverHip(Driver):
    """Driver of type 'hip' on one device (default cuda:0)."""

    @staticmethod
    def type():
        return 'hip'

    def __init__(self, device=0):
        t = torch()
        if not t.cuda.is_available():
            raise RuntimeError("the hip driver needs a GPU (torch.cuda)")
        self._device = t.device('cuda', device)
        self._ctx = None
        self._ctx_stream = None
        self._backends = None

    @property
    def device(self):
        return self._device

    def context(self):
        """The libgbkhip context of this driver, bound to torch's current
        stream so library kernels and tensor fills stay ordered."""
        t = torch()
        if self._ctx is None:
            with t.cuda.device(self._device):
                self._ctx = N.Context(self._device.index or 0)
                self._ctx_stream = None
        s = t.cuda.current_stream(self._device).cuda_stream
        if s != self._ctx_stream:          # rebind when the caller switched
            self._ctx.set_stream(s)
            self._ctx_stream = s
        return self._ctx

    # ---- memory ----------------------------------------------------------
    def mem_alloc_s(self, shape, dtype):
        return self.mem_alloc_h(shape, dtype), self.mem_alloc_d(shape, dtype)

    def mem_alloc_h(self, shape, dtype):
        t = torch()
        h = t.empty(_shape(shape), dtype=_tdtype(dtype), pin_memory=True)
        return h.numpy()

    def mem_alloc_d(self, shape, dtype):
        t = torch()
        return t.empty(_shape(shape), dtype=_tdtype(dtype),
                       device=self._device)

    def mem_copy_h2d(self, h_src, d_dst=None):
        t = torch()
        src = t.from_numpy(np.ascontiguousarray(h_src))
        if d_dst is None:
            d_dst = t.empty(src.shape, dtype=src.dtype, device=self._device)
        d_dst.copy_(src.reshape(d_dst.shape))
        return d_dst

    def mem_copy_d2h(self, d_src, h_dst=None):
        t = torch()
        if h_dst is None:
            h_dst = self.mem_alloc_h(tuple(d_src.shape),
                                     t.empty(0, dtype=d_src.dtype).numpy()
                                     .dtype)
        t.from_numpy(h_dst).copy_(d_src.reshape(h_dst.shape).cpu())
        return h_dst

    def mem_fill(self, x, value):
        x.fill_(value)

    # ---- math (torch elementwise; off the hot path) ------------------------
    def math_abs(self, x, out=None):
        return torch().abs(x, out=out)

    def math_sum(self, x, out=None):
        r = torch().nansum(x).reshape((1,) * x.dim())
        if out is not None:
            out.copy_(r)
            return out
        return r

    def math_add(self, x1, x2, out=None):
        return torch().add(x1, x2, out=out)

    def math_sub(self, x1, x2, out=None):
        return torch().sub(x1, x2, out=out)

    def math_mul(self, x1, x2, out=None):
        return torch().mul(x1, x2, out=out)

    def math_div(self, x1, x2, out=None):
        return torch().div(x1, x2, out=out)

    def math_pow(self, x1, x2, out=None):
        return torch().pow(x1, x2, out=out)

    def backends(self):
        if self._backends is None:
            self._backends = DriverBackendsHip(self)
        return self._backends


register_driver(DriverHip)


def _check_dtype(dtype):
    if np.dtype(dtype) != np.dtype(np.float32):
        raise RuntimeError(
            f"could not create hip backend; the requested dtype is not "
            f"supported (requested: {np.dtype(dtype).name}; supported: "
            f"float32)")


class DriverBackendsHip(DriverBackends):

    def __init__(self, driver):
        self._driver = driver

    def fft(self, dtype, padding='reference'):
        _check_dtype(dtype)
        return DriverBackendFFTHip(self._driver, padding)

    def dmodel(self, dtype):
        _check_dtype(dtype)
        return DriverBackendDModelHip(self._driver)

    def gmodel(self, dtype):
        _check_dtype(dtype)
        return DriverBackendGModelHip(self._driver)

    def objective(self, dtype):
        _check_dtype(dtype)
        return DriverBackendObjectiveHip(self._driver)


class DriverBackendFFTHip(DriverBackendFFT):
    """FFTf32 equivalent (host/src/bindings.cpp:35-40)."""

    def __init__(self, driver, padding='reference'):
        super().__init__(32, padding)   # threshold: _detail/native.py:59
        self._driver = driver

    def dtype(self):
        return np.float32

    def _ctx(self):
        return self._driver.context().handle

    def fft_r2c(self, data_r, data_c):
        N.call('gbkhip_fft_r2c', self._ctx(),
               N._i3(NativeMemoryHip.shape(data_r)[::-1]), _ptr(data_r),
               _ptr(data_c))

    def fft_c2r(self, data_c, data_r):
        N.call('gbkhip_fft_c2r', self._ctx(),
               N._i3(NativeMemoryHip.shape(data_r)[::-1]), _ptr(data_c),
               _ptr(data_r))

    def fft_convolve(self, data1_r, data1_c, data2_c):
        N.call('gbkhip_fft_convolve', self._ctx(),
               N._i3(NativeMemoryHip.shape(data1_r)[::-1]), _ptr(data1_r),
               _ptr(data1_c), _ptr(data2_c))

    def fft_convolve_cached(self, data1_r, data2_r):
        # the tensor version counter invalidates the cached spectrum when
        # data2_r is rewritten in place
        N.call('gbkhip_fft_convolve_cached', self._ctx(),
               N._i3(NativeMemoryHip.shape(data1_r)[::-1]), _ptr(data1_r),
               _ptr(data2_r), int(data2_r._version))


class DriverBackendDModelHip(DriverBackendDModel):
    """DModelf32 equivalent (host/src/bindings.cpp:16-20)."""

    def __init__(self, driver):
        self._driver = driver

    def dtype(self):
        return np.float32

    def _ctx(self):
        return self._driver.context().handle

    def dcube_downscale(self, scale, edge_hi, cube_hi, cube_lo):
        N.call('gbkhip_dcube_downscale', self._ctx(), N._i3(scale),
               N._i3(edge_hi), N._i3(NativeMemoryHip.shape(cube_hi)[::-1]),
               N._i3(NativeMemoryHip.shape(cube_lo)[::-1]), _ptr(cube_hi),
               _ptr(cube_lo))

    def dcube_mask(self, cutoff, apply, mcube, dcube, wcube):
        N.call('gbkhip_dcube_mask', self._ctx(), float(cutoff), int(apply),
               N._i3(NativeMemoryHip.shape(dcube)[::-1]), _ptr(dcube),
               _ptr(mcube), _ptr(wcube))

    def mmaps_moments(self, size, step, zero, dcube, wcube, cutoff, orders,
                      mmaps_d, mmaps_w, mmaps_m):
        N.call('gbkhip_mmaps_moments', self._ctx(), N._i3(size), N._f3(step),
               N._f3(zero), _ptr(dcube), _ptr(wcube), float(cutoff),
               NativeMemoryHip.size(orders), _ptr(orders), _ptr(mmaps_d),
               _ptr(mmaps_m), _ptr(mmaps_w if wcube is not None else None))


def _traits(uids, cvalues, ccounts, pvalues, pcounts):
    t = N.Traits()
    if uids is None:
        return t
    t.n = NativeMemoryHip.size(uids)
    t.uids = _ptr(uids)
    t.cvalues = _ptr(cvalues)
    t.ccounts = _ptr(ccounts)
    t.pvalues = _ptr(pvalues)
    t.pcounts = _ptr(pcounts)
    return t


def make_disk(loose, tilted, rnodes, vsys, xpos, ypos, posa, incl, trait_args,
              opacity):
    """Packs the reference's positional trait arguments (9 kinds x 5 arrays,
    _detail/native.py:245-317) into a gbkhip_disk."""
    d = N.Disk()
    d.loose, d.tilted = int(bool(loose)), int(bool(tilted))
    d.nrnodes = NativeMemoryHip.size(rnodes)
    d.rnodes = _ptr(rnodes)
    d.vsys, d.xpos, d.ypos = _ptr(vsys), _ptr(xpos), _ptr(ypos)
    d.posa, d.incl = _ptr(posa), _ptr(incl)
    for i, k in enumerate(N.TRAIT_KINDS):
        setattr(d, k, _traits(*trait_args[5 * i:5 * i + 5]))
    d.opacity = _ptr(opacity)
    return d


def make_grid(spat_size, spat_step, spat_zero, spec_size, spec_step,
              spec_zero):
    g = N.Grid()
    g.spat_size[:] = [int(v) for v in spat_size]
    g.spat_step[:] = [float(v) for v in spat_step]
    g.spat_zero[:] = [float(v) for v in spat_zero]
    g.spec_size = int(spec_size)
    g.spec_step = float(spec_step)
    g.spec_zero = float(spec_zero)
    return g


def make_outputs(**arrays):
    o = N.Outputs()
    for k in N.OUTPUT_KEYS:
        setattr(o, k, _ptr(arrays.get(k)))
    return o


class DriverBackendGModelHip(DriverBackendGModel):
    """GModelf32 equivalent (host/src/bindings.cpp:22-27)."""

    def __init__(self, driver):
        self._driver = driver
        self.seed = 42      # Philox key of mcdisk clouds

    def dtype(self):
        return np.float32

    def _ctx(self):
        return self._driver.context().handle

    def convolve(self, data1, data2, result):
        assert NativeMemoryHip.shape(data1) == NativeMemoryHip.shape(result)
        N.call('gbkhip_convolve', self._ctx(),
               N._i3(NativeMemoryHip.shape(data1)[::-1]),
               N._i3(NativeMemoryHip.shape(data2)[::-1]), _ptr(data1),
               _ptr(data2), _ptr(result))

    def wcube_evaluate(self, spat_size, spec_size, spat_data, spec_data):
        N.call('gbkhip_wcube_evaluate', self._ctx(), N._i3(spat_size),
               int(spec_size), _ptr(spat_data), _ptr(spec_data))

    def smdisk_evaluate(
            self,
            loose, tilted, rnodes,
            vsys, xpos, ypos, posa, incl,
            rpt_uids, rpt_cvalues, rpt_ccounts, rpt_pvalues, rpt_pcounts,
            rht_uids, rht_cvalues, rht_ccounts, rht_pvalues, rht_pcounts,
            vpt_uids, vpt_cvalues, vpt_ccounts, vpt_pvalues, vpt_pcounts,
            vht_uids, vht_cvalues, vht_ccounts, vht_pvalues, vht_pcounts,
            dpt_uids, dpt_cvalues, dpt_ccounts, dpt_pvalues, dpt_pcounts,
            dht_uids, dht_cvalues, dht_ccounts, dht_pvalues, dht_pcounts,
            zpt_uids, zpt_cvalues, zpt_ccounts, zpt_pvalues, zpt_pcounts,
            spt_uids, spt_cvalues, spt_ccounts, spt_pvalues, spt_pcounts,
            wpt_uids, wpt_cvalues, wpt_ccounts, wpt_pvalues, wpt_pcounts,
            opacity,
            spat_size, spat_step, spat_zero, spat_rota,
            spec_size, spec_step, spec_zero,
            image, scube,
            wdata, wdata_cmp,
            rdata, rdata_cmp,
            ordata, ordata_cmp,
            vdata_cmp, ddata_cmp):
        # spat_rota is accepted and not forwarded, as in the reference
        # (_detail/native.py:307-312)
        d = make_disk(loose, tilted, rnodes, vsys, xpos, ypos, posa, incl, (
            rpt_uids, rpt_cvalues, rpt_ccounts, rpt_pvalues, rpt_pcounts,
            rht_uids, rht_cvalues, rht_ccounts, rht_pvalues, rht_pcounts,
            vpt_uids, vpt_cvalues, vpt_ccounts, vpt_pvalues, vpt_pcounts,
            vht_uids, vht_cvalues, vht_ccounts, vht_pvalues, vht_pcounts,
            dpt_uids, dpt_cvalues, dpt_ccounts, dpt_pvalues, dpt_pcounts,
            dht_uids, dht_cvalues, dht_ccounts, dht_pvalues, dht_pcounts,
            zpt_uids, zpt_cvalues, zpt_ccounts, zpt_pvalues, zpt_pcounts,
            spt_uids, spt_cvalues, spt_ccounts, spt_pvalues, spt_pcounts,
            wpt_uids, wpt_cvalues, wpt_ccounts, wpt_pvalues, wpt_pcounts),
            opacity)
        g = make_grid(spat_size, spat_step, spat_zero, spec_size, spec_step,
                      spec_zero)
        o = make_outputs(image=image, scube=scube, wdata=wdata,
                         wdata_cmp=wdata_cmp, rdata=rdata,
                         rdata_cmp=rdata_cmp, ordata=ordata,
                         ordata_cmp=ordata_cmp, vdata_cmp=vdata_cmp,
                         ddata_cmp=ddata_cmp)
        N.call('gbkhip_smdisk_evaluate', self._ctx(), ctypes.byref(d),
               ctypes.byref(g), ctypes.byref(o))

    def mcdisk_evaluate(
            self,
            cflux, nclouds, ncloudspt, hasordint,
            loose, tilted, rnodes,
            vsys, xpos, ypos, posa, incl,
            rpt_uids, rpt_cvalues, rpt_ccounts, rpt_pvalues, rpt_pcounts,
            rht_uids, rht_cvalues, rht_ccounts, rht_pvalues, rht_pcounts,
            vpt_uids, vpt_cvalues, vpt_ccounts, vpt_pvalues, vpt_pcounts,
            vht_uids, vht_cvalues, vht_ccounts, vht_pvalues, vht_pcounts,
            dpt_uids, dpt_cvalues, dpt_ccounts, dpt_pvalues, dpt_pcounts,
            dht_uids, dht_cvalues, dht_ccounts, dht_pvalues, dht_pcounts,
            zpt_uids, zpt_cvalues, zpt_ccounts, zpt_pvalues, zpt_pcounts,
            spt_uids, spt_cvalues, spt_ccounts, spt_pvalues, spt_pcounts,
            wpt_uids, wpt_cvalues, wpt_ccounts, wpt_pvalues, wpt_pcounts,
            opacity,
            spat_size, spat_step, spat_zero, spat_rota,
            spec_size, spec_step, spec_zero,
            image, scube,
            wdata, wdata_cmp,
            rdata, rdata_cmp,
            ordata, ordata_cmp,
            vdata_cmp, ddata_cmp):
        d = make_disk(loose, tilted, rnodes, vsys, xpos, ypos, posa, incl, (
            rpt_uids, rpt_cvalues, rpt_ccounts, rpt_pvalues, rpt_pcounts,
            rht_uids, rht_cvalues, rht_ccounts, rht_pvalues, rht_pcounts,
            vpt_uids, vpt_cvalues, vpt_ccounts, vpt_pvalues, vpt_pcounts,
            vht_uids, vht_cvalues, vht_ccounts, vht_pvalues, vht_pcounts,
            dpt_uids, dpt_cvalues, dpt_ccounts, dpt_pvalues, dpt_pcounts,
            dht_uids, dht_cvalues, dht_ccounts, dht_pvalues, dht_pcounts,
            zpt_uids, zpt_cvalues, zpt_ccounts, zpt_pvalues, zpt_pcounts,
            spt_uids, spt_cvalues, spt_ccounts, spt_pvalues, spt_pcounts,
            wpt_uids, wpt_cvalues, wpt_ccounts, wpt_pvalues, wpt_pcounts),
            opacity)
        g = make_grid(spat_size, spat_step, spat_zero, spec_size, spec_step,
                      spec_zero)
        o = make_outputs(image=image, scube=scube, wdata=wdata,
                         wdata_cmp=wdata_cmp, rdata=rdata,
                         rdata_cmp=rdata_cmp, ordata=ordata,
                         ordata_cmp=ordata_cmp, vdata_cmp=vdata_cmp,
                         ddata_cmp=ddata_cmp)
        N.call('gbkhip_mcdisk_evaluate', self._ctx(), float(cflux),
               int(nclouds), _ptr(ncloudspt), NativeMemoryHip.size(ncloudspt),
               _ptr(hasordint), ctypes.byref(d), ctypes.byref(g),
               ctypes.byref(o), ctypes.c_uint64(self.seed))


class DriverBackendObjectiveHip(DriverBackendObjective):
    """Objectivef32 equivalent (host/src/bindings.cpp:29-33), with the
    residual_sum the CUDA driver lacks (cuda/src/bindings.cpp:28-31)."""

    def __init__(self, driver):
        self._driver = driver

    def dtype(self):
        return np.float32

    def _ctx(self):
        return self._driver.context().handle

    def residual(self, obs_d, obs_e, obs_m, mdl_d, mdl_w, mdl_m, weight, res):
        N.call('gbkhip_residual', self._ctx(), _ptr(obs_d), _ptr(obs_e),
               _ptr(obs_m), _ptr(mdl_d), _ptr(mdl_w), _ptr(mdl_m),
               NativeMemoryHip.size(obs_d), float(weight), _ptr(res))

    def residual_sum(self, residual, squared, sum_):
        N.call('gbkhip_residual_sum', self._ctx(), _ptr(residual),
               NativeMemoryHip.size(residual), int(bool(squared)), _ptr(sum_))

    def count_elements(self, array1, array2, threshold, counts):
        N.call('gbkhip_count_pixels', self._ctx(), _ptr(array1), _ptr(array2),
               NativeMemoryHip.size(array1), float(threshold), _ptr(counts))
