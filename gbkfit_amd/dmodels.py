"""Data models: the spectral-cube orchestration around the gmodel.

Mirror of src/gbkfit/model/dmodels/_dcube.py (DCube: hi-res size, FFT
padding, zero point, PSF x LSF cube, convolve -> downscale -> mask),
scube.py (DModelSCube) and mmaps.py (DModelMMaps), with the intent of the
constructors the reference's HEAD breaks (13 positional arguments into a
12-argument DCube, scube.py:61-63, mmaps.py:95-98) implemented.
"""
import logging

import numpy as np

from .psflsf import LSFPoint, PSFPoint

_log = logging.getLogger(__name__)


class DCube:

    def __init__(self, size, step, rpix, rval, rota, scale, psf, lsf,
                 smooth_weights, mask_cutoff, mask_apply, dtype,
                 padding='reference'):
        if mask_apply and mask_cutoff is None:
            mask_apply = False
        if smooth_weights and not (psf or lsf):
            smooth_weights = False
        self._size_lo = tuple(int(v) for v in size)
        self._step_lo = tuple(float(v) for v in step)
        self._zero_lo = tuple(rval[i] - rpix[i] * step[i] for i in range(3))
        self._rota = rota
        self._scale = tuple(int(v) for v in scale)
        self._psf, self._lsf = psf, lsf
        self._smooth_weights = smooth_weights
        self._mask_cutoff = mask_cutoff
        self._mask_apply = mask_apply
        self._dtype = np.dtype(dtype)
        self._padding = padding
        self._size_hi = self._step_hi = self._zero_hi = self._edge_hi = None
        self._dcube_lo = self._dcube_hi = None
        self._wcube_lo = self._wcube_hi = None
        self._mcube_lo = None
        self._pcube_hi = None
        self._driver = None

    def size(self):
        return self._size_lo

    def step(self):
        return self._step_lo

    def zero(self):
        return self._zero_lo

    def rota(self):
        return self._rota

    def scale(self):
        return self._scale

    def dcube(self):
        return self._dcube_lo

    def wcube(self):
        return self._wcube_lo

    def mcube(self):
        return self._mcube_lo

    def scratch_size(self):
        return self._size_hi

    def scratch_step(self):
        return self._step_hi

    def scratch_zero(self):
        return self._zero_hi

    def scratch_edge(self):
        return self._edge_hi

    def scratch_dcube(self):
        return self._dcube_hi

    def scratch_wcube(self):
        return self._wcube_hi

    def pcube(self):
        return self._pcube_hi

    def psf(self):
        return self._psf

    def lsf(self):
        return self._lsf

    def dtype(self):
        return self._dtype

    def geometry(self, backend_fft):
        """Hi-res size, step, edge and zero (_dcube.py:142-179)."""
        size_lo, step_lo, zero_lo = self._size_lo, self._step_lo, \
            self._zero_lo
        scale, psf, lsf = self._scale, self._psf, self._lsf
        size_hi = tuple(size_lo[i] * scale[i] for i in range(3))
        step_hi = tuple(step_lo[i] / scale[i] for i in range(3))
        edge_hi = (0, 0, 0)
        if psf or lsf:
            psz = psf.size(step_hi[:2]) if psf else (1, 1)
            lsz = lsf.size(step_hi[2]) if lsf else 1
            size_hi, edge_hi = backend_fft.fft_convolution_shape(
                size_hi, psz + (lsz,))
        zero_hi = tuple(zero_lo[i] - step_lo[i] / 2 - (edge_hi[i] - 0.5) *
                        step_hi[i] for i in range(3))
        return size_hi, step_hi, edge_hi, zero_hi

    def make_pcube(self, size_hi, step_hi, backend_fft):
        """PSF x LSF cube at the padded size, centre rolled to (0,0,0)
        (_dcube.py:191-207)."""
        off = [0 if s % 2 else -1 for s in size_hi]
        psf = self._psf or PSFPoint()
        lsf = self._lsf or LSFPoint()
        p = psf.asarray(step_hi[:2], size_hi[:2], off[:2])
        q = lsf.asarray(step_hi[2], size_hi[2], off[2])
        cube = (p * q[:, None, None]).astype(self._dtype)
        return backend_fft.fft_convolution_shift(cube)

    def prepare(self, driver, has_weights):
        backend_fft = driver.backends().fft(self._dtype, self._padding)
        size_hi, step_hi, edge_hi, zero_hi = self.geometry(backend_fft)
        size_lo = self._size_lo
        if self._psf or self._lsf:
            self._pcube_hi = driver.mem_copy_h2d(
                self.make_pcube(size_hi, step_hi, backend_fft))
        shape_lo, shape_hi = size_lo[::-1], tuple(size_hi)[::-1]
        self._dcube_lo = driver.mem_alloc_d(shape_lo, self._dtype)
        self._dcube_hi = self._dcube_lo
        driver.mem_fill(self._dcube_lo, 0)
        if tuple(size_lo) != tuple(size_hi):
            self._dcube_hi = driver.mem_alloc_d(shape_hi, self._dtype)
            driver.mem_fill(self._dcube_hi, 0)
        if has_weights:
            self._wcube_lo = driver.mem_alloc_d(shape_lo, self._dtype)
            self._wcube_hi = self._wcube_lo
            driver.mem_fill(self._wcube_lo, 1)
            if tuple(size_lo) != tuple(size_hi):
                self._wcube_hi = driver.mem_alloc_d(shape_hi, self._dtype)
                driver.mem_fill(self._wcube_hi, 1)
        if self._mask_cutoff is not None:
            self._mcube_lo = driver.mem_alloc_d(shape_lo, self._dtype)
            driver.mem_fill(self._mcube_lo, 1)
        self._size_hi, self._step_hi = tuple(size_hi), step_hi
        self._zero_hi, self._edge_hi = zero_hi, tuple(edge_hi)
        self._driver = driver
        self._backend_fft = backend_fft
        self._backend_dmodel = driver.backends().dmodel(self._dtype)

    def evaluate(self, out_extra=None):
        """Convolve, downscale, mask (_dcube.py:245-313)."""
        be_fft, be_dm = self._backend_fft, self._backend_dmodel
        has_w = self._wcube_lo is not None
        if self._psf or self._lsf:
            be_fft.fft_convolve_cached(self._dcube_hi, self._pcube_hi)
            if has_w:
                be_fft.fft_convolve_cached(self._wcube_hi, self._pcube_hi)
        if self._dcube_lo is not self._dcube_hi:
            be_dm.dcube_downscale(self._scale, self._edge_hi, self._dcube_hi,
                                  self._dcube_lo)
            if has_w and self._smooth_weights:
                be_dm.dcube_downscale(self._scale, self._edge_hi,
                                      self._wcube_hi, self._wcube_lo)
        if self._mask_cutoff is not None:
            be_dm.dcube_mask(self._mask_cutoff, self._mask_apply,
                             self._mcube_lo, self._dcube_lo, self._wcube_lo)
        if out_extra is not None:
            d = self._driver
            out_extra.update(dcube_lo=d.mem_copy_d2h(self._dcube_lo),
                             dcube_hi=d.mem_copy_d2h(self._dcube_hi))


def _full3(size, step, rpix, rval, scale):
    size = tuple(size)
    if rpix is None:
        rpix = tuple(s / 2 - 0.5 for s in size)
    return size, tuple(step), tuple(rpix), tuple(rval), tuple(scale)


class DModel:
    """DModel base (src/gbkfit/model/core.py:19-73): prepared lazily on the
    first evaluate for a given (driver, gmodel)."""

    def __init__(self):
        self._driver = None
        self._gmodel = None

    def prepare(self, driver, gmodel):
        self._driver, self._gmodel = driver, gmodel
        self._prepare_impl(gmodel)

    def evaluate(self, driver, gmodel, params, out_dmodel_extra=None,
                 out_gmodel_extra=None):
        if driver is not self._driver or gmodel is not self._gmodel:
            self.prepare(driver, gmodel)
        return self._evaluate_impl(params, out_dmodel_extra, out_gmodel_extra)

    def npix(self):
        return int(np.prod(self.size()))


class DModelSCube(DModel):
    """Spectral cube data model (scube.py:17-122)."""

    @staticmethod
    def type():
        return 'scube'

    def __init__(self, size, step=(1, 1, 1), rpix=None, rval=(0, 0, 0),
                 rota=0, scale=(1, 1, 1), psf=None, lsf=None, weight=1,
                 mask_cutoff=None, mask_create=False, mask_apply=False,
                 dtype='float32', padding='reference'):
        super().__init__()
        size, step, rpix, rval, scale = _full3(size, step, rpix, rval, scale)
        self._weight = weight
        self._dcube = DCube(size, step, rpix, rval, rota, scale, psf, lsf,
                            False, mask_cutoff if (mask_create or mask_apply)
                            else None, mask_apply, dtype, padding)

    def keys(self):
        return ['scube']

    def size(self):
        return self._dcube.size()

    def step(self):
        return self._dcube.step()

    def zero(self):
        return self._dcube.zero()

    def dtype(self):
        return self._dcube.dtype()

    def dcube(self):
        return self._dcube

    def _prepare_impl(self, gmodel):
        self._dcube.prepare(self._driver, gmodel.has_weights())

    def _evaluate_impl(self, params, out_dmodel_extra, out_gmodel_extra):
        d, g, c = self._driver, self._gmodel, self._dcube
        d.mem_fill(c.scratch_dcube(), 0)
        g.evaluate_scube(d, params, c.scratch_dcube(), c.scratch_wcube(),
                         c.scratch_size(), c.scratch_step(), c.scratch_zero(),
                         c.rota(), c.dtype(), out_gmodel_extra)
        c.evaluate(out_dmodel_extra)
        return dict(scube=dict(d=c.dcube(), m=c.mcube(), w=c.wcube()))


class DModelMMaps(DModel):
    """Moment maps of the model cube (mmaps.py:23-201)."""

    @staticmethod
    def type():
        return 'mmaps'

    def __init__(self, size, step=(1, 1), rpix=None, rval=(0, 0), rota=0,
                 scale=(1, 1), psf=None, lsf=None, mask_cutoff=1e-6,
                 orders=(0, 1, 2), dtype='float32', padding='reference'):
        super().__init__()
        orders = tuple(sorted(set(int(o) for o in orders)))
        if any(o < 0 or o > 7 for o in orders):
            raise RuntimeError("moment orders must be between 0 and 7")
        if mask_cutoff is None:
            raise RuntimeError("masking cannot be disabled for mmaps")
        step = tuple(step)
        size = tuple(size)
        if len(step) == 2:
            step = step + (1,)
        if len(size) == 2:
            from .psflsf import roundu_odd
            size = size + (int(roundu_odd(1000 / step[2])),)
        if rpix is not None and len(rpix) == 2:
            rpix = tuple(rpix) + (size[2] / 2 - 0.5,)
        rval = tuple(rval)
        if len(rval) == 2:
            rval = rval + (0,)
        scale = tuple(scale)
        if len(scale) == 2:
            scale = scale + (1,)
        size, step, rpix, rval, scale = _full3(size, step, rpix, rval, scale)
        self._orders = orders
        self._mask_cutoff = mask_cutoff
        self._dcube = DCube(size, step, rpix, rval, rota, scale, psf, lsf,
                            False, None, False, dtype, padding)

    def keys(self):
        return tuple(f'mmap{i}' for i in self._orders)

    def size(self):
        return self._dcube.size()[:2]

    def step(self):
        return self._dcube.step()[:2]

    def zero(self):
        return self._dcube.zero()[:2]

    def dtype(self):
        return self._dcube.dtype()

    def orders(self):
        return self._orders

    def dcube(self):
        return self._dcube

    def _prepare_impl(self, gmodel):
        d = self._driver
        n = len(self._orders)
        ny, nx = self.size()[::-1]
        self._mmaps_o = d.mem_copy_h2d(np.array(self._orders, np.int32))
        self._mmaps_d = d.mem_alloc_d((n, ny, nx), self.dtype())
        self._mmaps_m = d.mem_alloc_d((ny, nx), self.dtype())
        self._mmaps_w = d.mem_alloc_d((n, ny, nx), self.dtype())
        d.mem_fill(self._mmaps_d, np.nan)
        d.mem_fill(self._mmaps_m, 0)
        d.mem_fill(self._mmaps_w, 1)
        self._dcube.prepare(d, gmodel.has_weights())
        self._backend = d.backends().dmodel(self.dtype())

    def _evaluate_impl(self, params, out_dmodel_extra, out_gmodel_extra):
        d, g, c = self._driver, self._gmodel, self._dcube
        d.mem_fill(c.scratch_dcube(), 0)
        g.evaluate_scube(d, params, c.scratch_dcube(), c.scratch_wcube(),
                         c.scratch_size(), c.scratch_step(), c.scratch_zero(),
                         c.rota(), c.dtype(), out_gmodel_extra)
        c.evaluate(out_dmodel_extra)
        self._backend.mmaps_moments(
            c.size(), c.step(), c.zero(), c.dcube(), c.wcube(),
            self._mask_cutoff, self._mmaps_o, self._mmaps_d, self._mmaps_w,
            self._mmaps_m)
        return {key: dict(d=self._mmaps_d[i], m=self._mmaps_m,
                          w=self._mmaps_w[i])
                for i, key in enumerate(self.keys())}


class DModelImage(DModel):
    """Image data model (image.py:17-124): a one-channel cube (spectral
    step 0) convolved with the PSF only, for the GModelIntensity2D/3D
    gmodels (their evaluate_image)."""

    @staticmethod
    def type():
        return 'image'

    def __init__(self, size, step=(1, 1), rpix=None, rval=(0, 0), rota=0,
                 scale=(1, 1), psf=None, mask_cutoff=None, mask_apply=False,
                 dtype='float32', padding='reference'):
        super().__init__()
        size = tuple(int(v) for v in size)
        if rpix is None:
            rpix = tuple(s / 2 - 0.5 for s in size)
        self._dcube = DCube(size + (1,), tuple(step) + (0,),
                            tuple(rpix) + (0,), tuple(rval) + (0,), rota,
                            tuple(scale) + (1,), psf, None, False,
                            mask_cutoff, mask_apply, dtype, padding)

    def keys(self):
        return ['image']

    def size(self):
        return self._dcube.size()[:2]

    def step(self):
        return self._dcube.step()[:2]

    def zero(self):
        return self._dcube.zero()[:2]

    def rota(self):
        return self._dcube.rota()

    def scale(self):
        return self._dcube.scale()[:2]

    def psf(self):
        return self._dcube.psf()

    def dtype(self):
        return self._dcube.dtype()

    def dcube(self):
        return self._dcube

    def _prepare_impl(self, gmodel):
        self._dcube.prepare(self._driver, gmodel.has_weights())

    def _evaluate_impl(self, params, out_dmodel_extra, out_gmodel_extra):
        d, g, c = self._driver, self._gmodel, self._dcube
        d.mem_fill(c.scratch_dcube(), 0)
        g.evaluate_image(d, params, c.scratch_dcube(), c.scratch_wcube(),
                         c.scratch_size()[:2], c.scratch_step()[:2],
                         c.scratch_zero()[:2], c.rota(), c.dtype(),
                         out_gmodel_extra)
        c.evaluate(out_dmodel_extra)
        m, w = c.mcube(), c.wcube()
        return dict(image=dict(d=c.dcube()[0],
                               m=None if m is None else m[0],
                               w=None if w is None else w[0]))


class DModelLSlit(DModel):
    """Long-slit data model (lslit.py:17-123): a spectral cube one spatial
    pixel high (the slit; its width is the y step), returned as the
    (channel, position) plane.

    The reference's evaluate returns dcube()[0][:, 0, :], the first channel
    indexed as if it were a cube, and the mask as the weights; here the
    plane is cube[:, 0, :] of the data, mask and weight cubes.  The default
    reference pixel is the centre of the slit and of the spectrum, in the
    middle of the single row."""

    @staticmethod
    def type():
        return 'lslit'

    def __init__(self, size, step=(1, 1, 1), rpix=None, rval=(0, 0, 0),
                 rota=0, scale=(1, 1, 1), psf=None, lsf=None, weight=1,
                 mask_cutoff=None, mask_create=False, mask_apply=False,
                 dtype='float32', padding='reference'):
        super().__init__()
        if len(size) != 2:
            raise RuntimeError("lslit size is (positions, channels)")
        size = (int(size[0]), 1, int(size[1]))
        if rpix is None:
            rpix = (size[0] / 2 - 0.5, 0.0, size[2] / 2 - 0.5)
        self._weight = weight
        self._dcube = DCube(size, tuple(step), tuple(rpix), tuple(rval),
                            rota, tuple(scale), psf, lsf, False,
                            mask_cutoff if (mask_create or mask_apply)
                            else None, mask_apply, dtype, padding)

    def keys(self):
        return ['lslit']

    def size(self):
        return self._dcube.size()

    def step(self):
        return self._dcube.step()

    def zero(self):
        return self._dcube.zero()

    def rota(self):
        return self._dcube.rota()

    def scale(self):
        return self._dcube.scale()

    def psf(self):
        return self._dcube.psf()

    def lsf(self):
        return self._dcube.lsf()

    def dtype(self):
        return self._dcube.dtype()

    def dcube(self):
        return self._dcube

    def _prepare_impl(self, gmodel):
        self._dcube.prepare(self._driver, gmodel.has_weights())

    def _evaluate_impl(self, params, out_dmodel_extra, out_gmodel_extra):
        d, g, c = self._driver, self._gmodel, self._dcube
        d.mem_fill(c.scratch_dcube(), 0)
        g.evaluate_scube(d, params, c.scratch_dcube(), c.scratch_wcube(),
                         c.scratch_size(), c.scratch_step(), c.scratch_zero(),
                         c.rota(), c.dtype(), out_gmodel_extra)
        c.evaluate(out_dmodel_extra)
        m, w = c.mcube(), c.wcube()
        return dict(lslit=dict(d=c.dcube()[:, 0, :],
                               m=None if m is None else m[:, 0, :],
                               w=None if w is None else w[:, 0, :]))
