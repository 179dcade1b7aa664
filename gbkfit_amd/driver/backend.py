"""Backend interfaces (mirror of src/gbkfit/driver/backend.py:18-158),
including the FFT padding rules the data cubes are sized with."""
import abc

import numpy as np


def roundu_po2(x):
    """Smallest power of two >= x (src/gbkfit/math/math.py:53-57)."""
    p = 1
    while p < x:
        p *= 2
    return p


def roundu_multiple(x, multiple):
    """src/gbkfit/math/math.py:42-44."""
    r = x % multiple
    return x - r + multiple if r else x


def smooth235(n):
    """Smallest 2/3/5-smooth integer >= n."""
    m = max(1, int(n))
    while True:
        k = m
        for f in (2, 3, 5):
            while k % f == 0:
                k //= f
        if k == 1:
            return m
        m += 1


class DriverBackends(abc.ABC):

    @abc.abstractmethod
    def fft(self, dtype):
        pass

    @abc.abstractmethod
    def dmodel(self, dtype):
        pass

    @abc.abstractmethod
    def gmodel(self, dtype):
        pass

    @abc.abstractmethod
    def objective(self, dtype):
        pass


class DriverBackendFFT(abc.ABC):

    @staticmethod
    def fft_complex_shape(shape):
        shape = list(shape)
        return tuple(shape[:-1] + [shape[-1] // 2 + 1])

    @staticmethod
    def fft_convolution_shift(data):
        """Rolls the kernel centre to index 0 (backend.py:43-46)."""
        axis = tuple(range(data.ndim))
        return np.roll(data, np.array(data.shape) // 2 + 1, axis=axis)

    def __init__(self, fft_shape_threshold, padding='reference'):
        self._fft_shape_threshold = fft_shape_threshold
        if padding not in ('reference', 'smooth235'):
            raise RuntimeError("padding must be 'reference' or 'smooth235'")
        self._padding = padding

    def fft_optimal_shape(self, shape):
        """Reference rule (backend.py:51-59): powers of two up to the
        threshold, multiples of the threshold above it.  With
        padding='smooth235' the smallest 2/3/5-smooth size instead (a
        deliberate option: it changes the circular-wrap margin)."""
        new_shape = []
        for dim in shape:
            dim = int(dim)
            if self._padding == 'smooth235':
                new_shape.append(smooth235(dim))
                continue
            t = self._fft_shape_threshold
            p2 = roundu_po2(dim)
            new_shape.append(p2 if p2 <= t else roundu_multiple(dim, t))
        return np.array(new_shape)

    def fft_convolution_shape(self, data1_shape, data2_shape):
        shape = np.array(data1_shape) + np.array(data2_shape) - 1
        shape = self.fft_optimal_shape(shape)
        margin = np.array(data2_shape) // 2
        return (tuple(int(v) for v in shape), tuple(int(v) for v in margin))

    @abc.abstractmethod
    def fft_r2c(self, data_r, data_c):
        pass

    @abc.abstractmethod
    def fft_c2r(self, data_c, data_r):
        pass

    @abc.abstractmethod
    def fft_convolve(self, data1_r, data1_c, data2_c):
        pass

    @abc.abstractmethod
    def fft_convolve_cached(self, data1_r, data2_r):
        pass


class DriverBackendDModel(abc.ABC):

    @abc.abstractmethod
    def dcube_downscale(self, scale, edge_hi, cube_hi, cube_lo):
        pass

    @abc.abstractmethod
    def dcube_mask(self, cutoff, apply, mcube, dcube, wcube):
        pass

    @abc.abstractmethod
    def mmaps_moments(self, size, step, zero, dcube, wcube, cutoff, orders,
                      mmaps_d, mmaps_w, mmaps_m):
        pass


class DriverBackendGModel(abc.ABC):

    @abc.abstractmethod
    def mcdisk_evaluate(self, cflux, nclouds, ncloudspt, hasordint, loose,
                        tilted, rnodes, vsys, xpos, ypos, posa, incl,
                        *args):
        pass

    @abc.abstractmethod
    def smdisk_evaluate(self, loose, tilted, rnodes, vsys, xpos, ypos, posa,
                        incl, *args):
        pass


class DriverBackendObjective(abc.ABC):

    @abc.abstractmethod
    def residual(self, obs_d, obs_e, obs_m, mdl_d, mdl_w, mdl_m, weight,
                 res):
        pass

    @abc.abstractmethod
    def residual_sum(self, residual, squared, sum_):
        pass
