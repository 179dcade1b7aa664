"""FP64 CPU oracle of gbkfit's model-evaluation hot path.

TEST INFRASTRUCTURE ONLY: tests/, ``__graft_entry__.smoke()`` and bench.py's
cpu_baseline leg may import this package, and only as the checker.  The
product package ``gbkfit_amd`` never imports it.

Parts:
  * ``liboracle.so`` (oracle.c): smdisk, mcdisk (Philox cloud stream),
    downscale, mask, moments, residual, compensated residual sum, brute-force
    circular convolution -- restated from the reference's header-only core.
  * this module: the numpy FP64 pieces -- sub-ring nodes and trait packing
    (src/gbkfit/model/gmodels/_disk.py:42-156), FFT padding rules
    (src/gbkfit/driver/backend.py:37-81), PSF/LSF cubes
    (src/gbkfit/psflsf/psfs/psfs.py:23-37,107-120, lsfs/lsfs.py:23-26,88-97),
    the FFT convolution closed form (host/include/gbkfit/host/fft.hpp:199-207)
    and the DCube orchestration (src/gbkfit/model/dmodels/_dcube.py:128-313).
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

# --------------------------------------------------------------------------
# ctypes structures (oracle.h)
# --------------------------------------------------------------------------
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class ORTraits(ctypes.Structure):
    _fields_ = [('n', ctypes.c_int), ('uids', _ip), ('cvalues', _dp),
                ('ccounts', _ip), ('pvalues', _dp), ('pcounts', _ip)]


class ORDisk(ctypes.Structure):
    _fields_ = ([('loose', ctypes.c_int), ('tilted', ctypes.c_int),
                 ('nrnodes', ctypes.c_int), ('rnodes', _dp), ('vsys', _dp),
                 ('xpos', _dp), ('ypos', _dp), ('posa', _dp), ('incl', _dp)]
                + [(k, ORTraits) for k in
                   ('rp', 'rh', 'vp', 'vh', 'dp', 'dh', 'zp', 'sp', 'wp')]
                + [('opacity', _dp)])


class ORGrid(ctypes.Structure):
    _fields_ = [('spat_size', ctypes.c_int * 3),
                ('spat_step', ctypes.c_double * 3),
                ('spat_zero', ctypes.c_double * 3),
                ('spec_size', ctypes.c_int), ('spec_step', ctypes.c_double),
                ('spec_zero', ctypes.c_double)]


OUT_KEYS = ('image', 'scube', 'wdata', 'wdata_cmp', 'rdata', 'rdata_cmp',
            'ordata', 'ordata_cmp', 'vdata_cmp', 'ddata_cmp', 'amb')


class OROut(ctypes.Structure):
    _fields_ = [(k, _dp) for k in OUT_KEYS]


class ORStats(ctypes.Structure):
    _fields_ = [('voxels_on_disk', ctypes.c_longlong),
                ('channel_deposits', ctypes.c_longlong)]


def lib():
    """Load liboracle.so (built by ``make -C oracle`` / ``build()``)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, 'liboracle.so')
        if not os.path.exists(path):
            raise RuntimeError(
                f"oracle library missing at {path}; run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.or_smdisk.argtypes = [ctypes.POINTER(ORDisk), ctypes.POINTER(ORGrid),
                                ctypes.POINTER(OROut), ctypes.POINTER(ORStats)]
        L.or_mcdisk.argtypes = [ctypes.c_double, ctypes.c_int, _ip, ctypes.c_int,
                                ctypes.POINTER(ctypes.c_ubyte),
                                ctypes.POINTER(ORDisk), ctypes.POINTER(ORGrid),
                                ctypes.POINTER(OROut), ctypes.c_ulonglong,
                                ctypes.POINTER(ORStats)]
        L.or_cloud_uniform.restype = ctypes.c_double
        L.or_cloud_uniform.argtypes = [ctypes.c_ulonglong, ctypes.c_uint,
                                       ctypes.c_uint]
        L.or_philox4x32.argtypes = [ctypes.POINTER(ctypes.c_uint),
                                    ctypes.POINTER(ctypes.c_uint),
                                    ctypes.POINTER(ctypes.c_uint)]
        L.or_downscale.argtypes = [_ip, _ip, _ip, _ip, _dp, _dp]
        L.or_mask.argtypes = [ctypes.c_double, ctypes.c_int, _ip, _dp, _dp, _dp]
        L.or_moments.argtypes = [_ip, _dp, _dp, _dp, _dp, ctypes.c_double,
                                 ctypes.c_int, _ip, _dp, _dp, _dp]
        L.or_residual.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp,
                                  ctypes.c_longlong, ctypes.c_double, _dp]
        L.or_residual_sum.restype = ctypes.c_double
        L.or_residual_sum.argtypes = [_dp, ctypes.c_longlong, ctypes.c_int]
        L.or_circular_convolve.argtypes = [_ip, _dp, _dp, _dp]
        L.or_wcube.argtypes = [_ip, ctypes.c_int, _dp, _dp]
        _LIB = L
    return _LIB


def _p(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


# --------------------------------------------------------------------------
# Disk description and trait packing
# --------------------------------------------------------------------------
TRAIT_KINDS = ('rp', 'rh', 'vp', 'vh', 'dp', 'dh', 'zp', 'sp', 'wp')


def subrnodes(rnodes, rstep=None):
    """Sub-ring nodes (src/gbkfit/model/gmodels/_disk.py:151-156) with the
    default step min(1, min(diff)/2) (gmodels/_detail.py:51-52)."""
    rnodes = np.asarray(rnodes, dtype=np.float64)
    if rstep is None:
        rstep = min(1.0, float(np.min(np.diff(rnodes))) / 2)
    s = np.arange(rnodes[0] + rstep / 2, rnodes[-1], rstep)
    s = np.insert(s, 0, rnodes[0])
    s = np.append(s, s[-1] + rstep / 2)
    return s, rstep


class Trait:
    """One trait: uid, constants, smooth params, node-wise params (one array
    of len(rnodes) per node-wise parameter, interpolated onto sub-nodes)."""

    def __init__(self, uid, consts=(), params=(), nwparams=()):
        self.uid = int(uid)
        self.consts = tuple(float(c) for c in consts)
        self.params = tuple(float(p) for p in params)
        self.nwparams = tuple(np.asarray(p, dtype=np.float64)
                              for p in nwparams)


class DiskSpec:
    """Everything the native smdisk/mcdisk call receives, in FP64."""

    def __init__(self, rnodes, traits, xpos=0.0, ypos=0.0, posa=0.0, incl=0.0,
                 vsys=0.0, rstep=None, loose=False, tilted=False,
                 has_vsys=True, opacity=None):
        self.nodes = np.asarray(rnodes, dtype=np.float64)
        # spatial opacity cube (nz, ny, nx) of the grid, or None
        self.opacity = None if opacity is None else \
            np.asarray(opacity, dtype=np.float64)
        self.sub, self.rstep = subrnodes(self.nodes, rstep)
        self.loose = bool(loose)
        self.tilted = bool(tilted)

        def common(v, nw):
            v = np.atleast_1d(np.asarray(v, dtype=np.float64))
            if nw:
                if v.size == 1:
                    v = np.full(self.nodes.size, v[0])
                return np.interp(self.sub, self.nodes, v)
            return v[:1].copy()

        self.vsys = common(vsys, loose) if has_vsys else None
        self.xpos = common(xpos, loose)
        self.ypos = common(ypos, loose)
        self.posa = common(posa, tilted)
        self.incl = common(incl, tilted)
        self.traits = {}
        for k in TRAIT_KINDS:
            tl = traits.get(k)
            if not tl:
                self.traits[k] = None
                continue
            uids, cv, cc, pv, pc = [], [], [], [], []
            for t in tl:
                uids.append(t.uid)
                cv += list(t.consts)
                cc.append(len(t.consts))
                vals = list(t.params)
                for arr in t.nwparams:
                    vals += list(np.interp(self.sub, self.nodes, arr))
                pv += vals
                pc.append(len(vals))
            self.traits[k] = (np.array(uids, np.int32),
                              np.array(cv if cv else [0.0], np.float64),
                              np.array(cc, np.int32),
                              np.array(pv if pv else [0.0], np.float64),
                              np.array(pc, np.int32))

    def is_thin(self):
        return self.traits['rh'] is None

    def rounded_f32(self):
        """Copy with every real array rounded to FP32 -- the values an FP32
        native entry point receives."""
        import copy
        c = copy.copy(self)
        r = lambda a: None if a is None else \
            np.asarray(a, np.float32).astype(np.float64)
        c.sub = r(self.sub)
        c.vsys, c.xpos, c.ypos = r(self.vsys), r(self.xpos), r(self.ypos)
        c.posa, c.incl = r(self.posa), r(self.incl)
        c.opacity = r(self.opacity)
        c.traits = {}
        for k, tr in self.traits.items():
            c.traits[k] = None if tr is None else \
                (tr[0], r(tr[1]), tr[2], r(tr[3]), tr[4])
        return c

    def _struct(self):
        keep = []

        def arr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            keep.append(a)
            return _p(a)

        d = ORDisk()
        d.loose, d.tilted = int(self.loose), int(self.tilted)
        d.nrnodes = self.sub.size
        d.rnodes = arr(self.sub)
        d.vsys = arr(self.vsys)
        d.xpos, d.ypos = arr(self.xpos), arr(self.ypos)
        d.posa, d.incl = arr(self.posa), arr(self.incl)
        for k in TRAIT_KINDS:
            t = ORTraits()
            tr = self.traits[k]
            if tr is not None:
                uids, cv, cc, pv, pc = tr
                keep.extend([uids, cv, cc, pv, pc])
                t.n = uids.size
                t.uids, t.ccounts, t.pcounts = _p(uids, _ip), _p(cc, _ip), \
                    _p(pc, _ip)
                t.cvalues, t.pvalues = _p(cv), _p(pv)
            setattr(d, k, t)
        d.opacity = arr(self.opacity)
        return d, keep


def make_grid(spat_size, spat_step, spat_zero, spec_size, spec_step,
              spec_zero):
    g = ORGrid()
    g.spat_size[:] = [int(v) for v in spat_size]
    g.spat_step[:] = [float(v) for v in spat_step]
    g.spat_zero[:] = [float(v) for v in spat_zero]
    g.spec_size = int(spec_size)
    g.spec_step = float(spec_step)
    g.spec_zero = float(spec_zero)
    return g


def _outputs(grid, want):
    nx, ny, nz = grid.spat_size
    shapes = dict(image=(ny, nx), scube=(grid.spec_size, ny, nx),
                  amb=(grid.spec_size, ny, nx))
    res = {}
    o = OROut()
    for k in want:
        res[k] = np.zeros(shapes.get(k, (nz, ny, nx)), np.float64)
        setattr(o, k, _p(res[k]))
    return o, res


def smdisk(disk, grid, want=('scube', 'amb')):
    """Evaluate smdisk on `grid`; returns (outputs dict, stats dict)."""
    d, keep = disk._struct()
    o, res = _outputs(grid, want)
    st = ORStats()
    lib().or_smdisk(ctypes.byref(d), ctypes.byref(grid), ctypes.byref(o),
                    ctypes.byref(st))
    del keep
    return res, dict(voxels_on_disk=st.voxels_on_disk,
                     channel_deposits=st.channel_deposits)


def mcdisk(disk, grid, cflux, ncloudscsum, hasordint, seed,
           want=('scube', 'amb')):
    d, keep = disk._struct()
    o, res = _outputs(grid, want)
    st = ORStats()
    csum = np.ascontiguousarray(ncloudscsum, dtype=np.int32)
    hoi = np.ascontiguousarray(hasordint, dtype=np.uint8)
    nclouds = int(csum[-1]) if csum.size else 0
    lib().or_mcdisk(float(cflux), nclouds, _p(csum, _ip), csum.size,
                    hoi.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)),
                    ctypes.byref(d), ctypes.byref(grid), ctypes.byref(o),
                    ctypes.c_ulonglong(seed), ctypes.byref(st))
    del keep
    return res, dict(voxels_on_disk=st.voxels_on_disk,
                     channel_deposits=st.channel_deposits)


def cloud_uniform(seed, cid, i):
    return lib().or_cloud_uniform(ctypes.c_ulonglong(seed), cid, i)


def philox4x32(ctr, key):
    c = (ctypes.c_uint * 4)(*ctr)
    k = (ctypes.c_uint * 2)(*key)
    o = (ctypes.c_uint * 4)()
    lib().or_philox4x32(c, k, o)
    return list(o)


def _i3(v):
    return np.ascontiguousarray(v, dtype=np.int32)


def downscale(scale, offset, src, dst_size):
    src = np.ascontiguousarray(src, dtype=np.float64)
    dst = np.zeros(tuple(dst_size)[::-1], np.float64)
    lib().or_downscale(_p(_i3(scale), _ip), _p(_i3(offset), _ip),
                       _p(_i3(src.shape[::-1]), _ip),
                       _p(_i3(dst_size), _ip), _p(src), _p(dst))
    return dst


def mask(cutoff, apply, dcube, want_mcube=True):
    d = np.array(dcube, dtype=np.float64, copy=True)
    m = np.zeros_like(d) if want_mcube else None
    lib().or_mask(float(cutoff), int(apply), _p(_i3(d.shape[::-1]), _ip),
                  _p(d), _p(m), None)
    return d, m


def moments(dcube, step, zero, orders, cutoff, wcube=None):
    """Returns (mmaps_d (norders, ny, nx), mmaps_m (ny, nx), mmaps_w)."""
    d = np.ascontiguousarray(dcube, dtype=np.float64)
    nz, ny, nx = d.shape
    orders = np.ascontiguousarray(sorted(set(orders)), dtype=np.int32)
    md = np.full((orders.size, ny, nx), np.nan)
    mm = np.zeros((ny, nx))
    mw = np.ones((orders.size, ny, nx))
    w = None if wcube is None else np.ascontiguousarray(wcube, np.float64)
    lib().or_moments(_p(_i3((nx, ny, nz)), _ip),
                     _p(np.asarray(step, np.float64)),
                     _p(np.asarray(zero, np.float64)), _p(d), _p(w),
                     float(cutoff), orders.size, _p(orders, _ip),
                     _p(md), _p(mm), _p(mw))
    return md, mm, mw


def residual(obs_d, obs_e, obs_m, mdl_d, mdl_w, mdl_m, weight):
    arrs = [None if a is None else np.ascontiguousarray(a, np.float64).ravel()
            for a in (obs_d, obs_e, obs_m, mdl_d, mdl_w, mdl_m)]
    res = np.zeros_like(arrs[0])
    lib().or_residual(*[_p(a) for a in arrs], arrs[0].size, float(weight),
                      _p(res))
    return res


def residual_sum(r, squared=True):
    r = np.ascontiguousarray(r, np.float64).ravel()
    return lib().or_residual_sum(_p(r), r.size, int(squared))


def wcube(spat, spec_size):
    """Spectral weight cube from a spatial (nz, ny, nx) one
    (core/gmodel/gmodels.hpp:52-82)."""
    spat = np.ascontiguousarray(spat, np.float64)
    nz, ny, nx = spat.shape
    out = np.zeros((spec_size, ny, nx))
    lib().or_wcube(_p(_i3((nx, ny, nz)), _ip), int(spec_size), _p(spat),
                   _p(out))
    return out


def circular_convolve(a, k):
    a = np.ascontiguousarray(a, np.float64)
    k = np.ascontiguousarray(k, np.float64)
    out = np.zeros_like(a)
    lib().or_circular_convolve(_p(_i3(a.shape[::-1]), _ip), _p(a), _p(k),
                               _p(out))
    return out


# --------------------------------------------------------------------------
# numpy FP64 pieces: padding, PSF/LSF, FFT convolution, DCube orchestration
# --------------------------------------------------------------------------
def roundu_po2(x):
    p = 1
    while p < x:
        p *= 2
    return p


def roundu_multiple(x, m):
    return x - x % m + m if x % m else x


def roundu_odd(x):
    """gbkfit/math/math.py:32-34."""
    e = 2 * math.ceil(0.5 * x)
    return e + 1 if e - 1 < x else e - 1


def fft_optimal_shape(shape, threshold=32):
    """src/gbkfit/driver/backend.py:51-59."""
    out = []
    for d in shape:
        p2 = roundu_po2(d)
        out.append(p2 if p2 <= threshold else roundu_multiple(d, threshold))
    return tuple(out)


def smooth235(n):
    """Smallest 2/3/5-smooth integer >= n (the alternative padding rule)."""
    m = max(1, n)
    while True:
        k = m
        for f in (2, 3, 5):
            while k % f == 0:
                k //= f
        if k == 1:
            return m
        m += 1


def psf_gauss_size(sigma, step):
    return (int(roundu_odd(2 * 8 * sigma / step[0])),
            int(roundu_odd(2 * 8 * sigma / step[1])))


def lsf_gauss_size(sigma, step):
    return int(roundu_odd(2 * 8 * sigma / step))


def psf_gauss_array(sigma, step, size, offset=(0, 0)):
    """PSFGauss._asarray_impl (psfs.py:114-120) with ratio 1, posa 0."""
    cx = size[0] // 2 + offset[0]
    cy = size[1] // 2 + offset[1]
    x = (np.arange(size[0]) - cx) * step[0]
    y = (np.arange(size[1]) - cy) * step[1]
    x, y = x[None, :], y[:, None]
    # lh rotation by 0 (gbkfit/math/math.py:85-90): (x, y) -> (x, -y)... the
    # radius is invariant.
    r = np.sqrt(x * x + y * y)
    data = np.exp(-r * r / (2 * sigma * sigma))
    data[r > 8 * sigma] = 0
    return data / np.sum(data)


def lsf_gauss_array(sigma, step, size, offset=0):
    """LSFGauss._asarray_impl (lsfs.py:91-97); the cut is one-sided."""
    c = size // 2 + offset
    z = (np.arange(size) - c) * step
    data = np.exp(-z * z / (2 * sigma * sigma))
    data[z > 8 * sigma] = 0
    return data / np.sum(data)


class DCubeGeom:
    """DCube.prepare geometry (src/gbkfit/model/dmodels/_dcube.py:128-243)."""

    def __init__(self, size, step, rpix=None, rval=(0, 0, 0), scale=(1, 1, 1),
                 psf_sigma=None, lsf_sigma=None, pad='reference'):
        size = tuple(int(v) for v in size)
        step = tuple(float(v) for v in step)
        if rpix is None:
            rpix = tuple(s / 2 - 0.5 for s in size)
        self.size_lo, self.step_lo, self.scale = size, step, tuple(scale)
        self.zero_lo = tuple(rval[i] - rpix[i] * step[i] for i in range(3))
        size_hi = tuple(size[i] * scale[i] for i in range(3))
        self.step_hi = tuple(step[i] / scale[i] for i in range(3))
        self.psf_sigma, self.lsf_sigma = psf_sigma, lsf_sigma
        edge = (0, 0, 0)
        has_conv = psf_sigma is not None or lsf_sigma is not None
        if has_conv:
            psz = psf_gauss_size(psf_sigma, self.step_hi[:2]) \
                if psf_sigma is not None else (1, 1)
            lsz = lsf_gauss_size(lsf_sigma, self.step_hi[2]) \
                if lsf_sigma is not None else 1
            ksz = psz + (lsz,)
            conv = tuple(size_hi[i] + ksz[i] - 1 for i in range(3))
            if pad == 'reference':
                size_hi = fft_optimal_shape(conv)
            else:
                size_hi = tuple(smooth235(c) for c in conv)
            edge = tuple(k // 2 for k in ksz)
        self.size_hi = size_hi
        self.edge_hi = edge
        self.zero_hi = tuple(
            self.zero_lo[i] - step[i] / 2 - (edge[i] - 0.5) * self.step_hi[i]
            for i in range(3))
        self.has_conv = has_conv

    def pcube(self, dtype=np.float32):
        """PSF x LSF cube at the padded size, rolled so its centre is at
        (0,0,0) (_dcube.py:191-209, backend.py:43-46)."""
        nx, ny, nz = self.size_hi
        off = [0 if s % 2 else -1 for s in self.size_hi]
        if self.psf_sigma is not None:
            psf = psf_gauss_array(self.psf_sigma, self.step_hi[:2], (nx, ny),
                                  off[:2])
        else:
            psf = np.zeros((ny, nx))
            psf[ny // 2 + off[1], nx // 2 + off[0]] = 1
        if self.lsf_sigma is not None:
            lsf = lsf_gauss_array(self.lsf_sigma, self.step_hi[2], nz, off[2])
        else:
            lsf = np.zeros(nz)
            lsf[nz // 2 + off[2]] = 1
        cube = (psf * lsf[:, None, None]).astype(dtype)
        shift = np.array(cube.shape) // 2 + 1
        return np.roll(cube, shift, axis=(0, 1, 2))


def _fft_workers():
    return max(1, min(16, os.cpu_count() or 1))


def fft_convolve(cube, pcube):
    """irfftn(rfftn(cube) * rfftn(pcube)) in FP64 (fft.hpp:199-207);
    scipy's pocketfft over the host's threads (numpy's single-threaded
    pocketfft gives the same transform)."""
    import scipy.fft
    a = np.asarray(cube, np.float64)
    k = np.asarray(pcube, np.float64)
    w = _fft_workers()
    f = scipy.fft.rfftn(a, workers=w)
    f *= scipy.fft.rfftn(k, workers=w)
    return scipy.fft.irfftn(f, s=a.shape, axes=tuple(range(a.ndim)),
                            workers=w)


def fft_convolve_abs(amb, pcube):
    """Bound propagation of an ambiguity cube through the convolution."""
    return np.abs(fft_convolve(amb, np.abs(np.asarray(pcube, np.float64))))


def evaluate_scube(geom, scube_hi, amb_hi=None, mask_cutoff=None,
                   mask_apply=False):
    """DCube.evaluate (_dcube.py:245-313): convolve, downscale, mask.
    Returns (dcube_lo, mcube_lo, amb_lo)."""
    hi = scube_hi
    amb = amb_hi
    if geom.has_conv:
        pc = geom.pcube()
        hi = fft_convolve(hi, pc)
        if amb is not None:
            amb = fft_convolve_abs(amb, pc)
    lo = hi
    amb_lo = amb
    if tuple(geom.size_lo) != tuple(geom.size_hi):
        lo = downscale(geom.scale, geom.edge_hi, hi, geom.size_lo)
        if amb is not None:
            amb_lo = downscale(geom.scale, geom.edge_hi, amb, geom.size_lo)
    mcube = None
    if mask_cutoff is not None:
        lo, mcube = mask(mask_cutoff, mask_apply, lo)
    return lo, mcube, amb_lo


def spatial_z_grid(geom, size_z=None):
    """Kinematics3D spatial z default (kinematics_3d.py:104-109)."""
    sx, sy = geom.size_hi[0], geom.size_hi[1]
    wide = int(sx < sy)
    nz = size_z if size_z is not None else geom.size_hi[wide]
    step_z = geom.step_hi[wide]
    zero_z = -(nz / 2 - 0.5) * step_z
    return nz, step_z, zero_z


def grid_for(geom, thin, size_z=None):
    """The smdisk/mcdisk grid a DModelSCube hands its gmodel."""
    if thin:
        sz, stz, zz = 1, 0.0, 0.0
    else:
        sz, stz, zz = spatial_z_grid(geom, size_z)
    return make_grid(
        (geom.size_hi[0], geom.size_hi[1], sz),
        (geom.step_hi[0], geom.step_hi[1], stz),
        (geom.zero_hi[0], geom.zero_hi[1], zz),
        geom.size_hi[2], geom.step_hi[2], geom.zero_hi[2])


def f32(v):
    """Round a grid value to FP32, as the native FP32 entry points see it."""
    return float(np.float32(v))


def grid_f32(grid):
    """Grid with every real value rounded to FP32 (what the device sees)."""
    return make_grid(list(grid.spat_size), [f32(v) for v in grid.spat_step],
                     [f32(v) for v in grid.spat_zero], grid.spec_size,
                     f32(grid.spec_step), f32(grid.spec_zero))
