"""Point and line spread functions (mirror of src/gbkfit/psflsf/core.py,
psfs/psfs.py and lsfs/lsfs.py for the analytic shapes).

Sizes are odd: roundu_odd(2 * cutoff_factor * width / step) (core.py:20-25,
68-78); arrays are evaluated on a grid centred at size//2 + offset and
normalised to unit sum; Gaussian profiles are cut at 8 sigma (the LSF cut is
one-sided, lsfs.py:96, as in the reference).
"""
import math

import numpy as np


def roundu_odd(x):
    """src/gbkfit/math/math.py:32-34."""
    e = 2 * math.ceil(0.5 * x)
    return e + 1 if e - 1 < x else e - 1


def gauss_fwhm_to_sigma(fwhm):
    return fwhm / math.sqrt(8 * math.log(2))


def _grid_2d(size, step, offset, ratio, posa):
    """psfs.py:23-37."""
    cx = size[0] // 2 + offset[0]
    cy = size[1] // 2 + offset[1]
    x = ((np.arange(size[0]) - cx) * step[0])[None, :]
    y = ((np.arange(size[1]) - cy) * step[1])[:, None]
    t = np.radians(posa)
    xr = x * np.cos(t) + y * np.sin(t)
    yr = -x * np.sin(t) + y * np.cos(t)
    return np.sqrt(xr * xr + yr * yr / (ratio * ratio))


def _grid_1d(size, step, offset):
    c = size // 2 + offset
    return (np.arange(size) - c) * step


class PSF:
    def size(self, step, offset=(0, 0)):
        b = self._size_impl(step)
        return (int(roundu_odd(b[0] + offset[0])),
                int(roundu_odd(b[1] + offset[1])))

    def asarray(self, step, size=None, offset=(0, 0)):
        if size is None:
            size = self.size(step, offset)
        elif (size[0] + offset[0]) % 2 == 0 or (size[1] + offset[1]) % 2 == 0:
            raise RuntimeError("invalid PSF size: size + offset must be odd")
        return self._asarray_impl(step, size, offset)


class LSF:
    def size(self, step, offset=0):
        return int(roundu_odd(self._size_impl(step) + offset))

    def asarray(self, step, size=None, offset=0):
        if size is None:
            size = self.size(step, offset)
        elif (size + offset) % 2 == 0:
            raise RuntimeError("invalid LSF size: size + offset must be odd")
        return self._asarray_impl(step, size, offset)


class PSFPoint(PSF):
    def _size_impl(self, step):
        return 1, 1

    def _asarray_impl(self, step, size, offset):
        d = np.zeros((size[1], size[0]))
        d[size[1] // 2 + offset[1], size[0] // 2 + offset[0]] = 1
        return d


class PSFGauss(PSF):
    CUTOFF = 8

    def __init__(self, sigma, ratio=1.0, posa=0.0):
        self.sigma, self.ratio, self.posa = float(sigma), float(ratio), \
            float(posa)

    def _size_impl(self, step):
        return (2 * self.CUTOFF * self.sigma / step[0],
                2 * self.CUTOFF * self.sigma / step[1])

    def _asarray_impl(self, step, size, offset):
        r = _grid_2d(size, step, offset, self.ratio, self.posa)
        d = np.exp(-r * r / (2 * self.sigma * self.sigma))
        d[r > self.CUTOFF * self.sigma] = 0
        return d / d.sum()


class LSFPoint(LSF):
    def _size_impl(self, step):
        return 1

    def _asarray_impl(self, step, size, offset):
        d = np.zeros(size)
        d[size // 2 + offset] = 1
        return d


class LSFGauss(LSF):
    CUTOFF = 8

    def __init__(self, sigma):
        self.sigma = float(sigma)

    def _size_impl(self, step):
        return 2 * self.CUTOFF * self.sigma / step

    def _asarray_impl(self, step, size, offset):
        z = _grid_1d(size, step, offset)
        d = np.exp(-z * z / (2 * self.sigma * self.sigma))
        d[z > self.CUTOFF * self.sigma] = 0
        return d / d.sum()
