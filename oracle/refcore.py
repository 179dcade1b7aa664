"""ctypes access to the reference-core harness (oracle/_ref/libgbkref.so).

TEST INFRASTRUCTURE ONLY.  The harness is compiled from the reference's own
header-only core (src/gbkfit/driver/native/gbkfit/include/gbkfit/*) by
oracle/Makefile; it is used to write golden vectors (tests/golden/) and as the
"reference" CPU baseline.  It is absent wherever the reference tree is.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}

KINDS = ('rp', 'rh', 'vp', 'vh', 'dp', 'dh', 'zp', 'sp', 'wp')


def path():
    return os.path.join(_HERE, '_ref', 'libgbkref.so')


def available():
    return os.path.exists(path())


def _types(ct):
    fp = ctypes.POINTER(ct)
    ip = ctypes.POINTER(ctypes.c_int)

    class Traits(ctypes.Structure):
        _fields_ = [('n', ctypes.c_int), ('uids', ip), ('cvalues', fp),
                    ('ccounts', ip), ('pvalues', fp), ('pcounts', ip)]

    class Disk(ctypes.Structure):
        _fields_ = ([('loose', ctypes.c_int), ('tilted', ctypes.c_int),
                     ('nrnodes', ctypes.c_int), ('rnodes', fp), ('vsys', fp),
                     ('xpos', fp), ('ypos', fp), ('posa', fp), ('incl', fp)]
                    + [(k, Traits) for k in KINDS] + [('opacity', fp)])

    class Grid(ctypes.Structure):
        _fields_ = [('spat_size', ctypes.c_int * 3),
                    ('spat_step', ct * 3), ('spat_zero', ct * 3),
                    ('spec_size', ctypes.c_int), ('spec_step', ct),
                    ('spec_zero', ct)]

    class Out(ctypes.Structure):
        _fields_ = [(k, fp) for k in (
            'image', 'scube', 'wdata', 'wdata_cmp', 'rdata', 'rdata_cmp',
            'ordata', 'ordata_cmp', 'vdata_cmp', 'ddata_cmp')]

    return Traits, Disk, Grid, Out


class _Ref:
    def __init__(self, dtype):
        self.np = np.dtype(dtype)
        self.ct = ctypes.c_float if self.np == np.float32 else ctypes.c_double
        self.sfx = 'f32' if self.np == np.float32 else 'f64'
        self.Traits, self.Disk, self.Grid, self.Out = _types(self.ct)
        self.L = ctypes.CDLL(path())
        fp = ctypes.POINTER(self.ct)
        ip = ctypes.POINTER(ctypes.c_int)
        f = lambda n: getattr(self.L, f'ref_{n}_{self.sfx}')
        for n in ('smdisk', 'smdisk_omp'):
            f(n).argtypes = [ctypes.POINTER(self.Disk),
                             ctypes.POINTER(self.Grid),
                             ctypes.POINTER(self.Out)]
        f('mcdisk').argtypes = [self.ct, ctypes.c_int, ip, ctypes.c_int,
                                ctypes.POINTER(ctypes.c_ubyte),
                                ctypes.POINTER(self.Disk),
                                ctypes.POINTER(self.Grid),
                                ctypes.POINTER(self.Out), ctypes.c_int]
        f('downscale').argtypes = [ip, ip, ip, ip, fp, fp, ctypes.c_int]
        f('mask').argtypes = [self.ct, ctypes.c_int, ip, fp, fp, fp,
                              ctypes.c_int]
        f('moments').argtypes = [ip, fp, fp, fp, fp, self.ct, ctypes.c_int,
                                 ip, fp, fp, fp, ctypes.c_int]
        f('residual').argtypes = [fp, fp, fp, fp, fp, fp, ctypes.c_int,
                                  self.ct, fp, ctypes.c_int]
        f('wcube').argtypes = [ip, ctypes.c_int, fp, fp]
        f('residual_sum').restype = self.ct
        f('residual_sum').argtypes = [fp, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int]
        self.f = f

    def p(self, a):
        if a is None:
            return None
        return a.ctypes.data_as(ctypes.POINTER(self.ct))

    def arr(self, a, keep):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=self.np)
        keep.append(a)
        return self.p(a)

    def disk(self, spec, keep):
        """Build the harness disk struct from an oracle.DiskSpec."""
        ip = ctypes.POINTER(ctypes.c_int)
        d = self.Disk()
        d.loose, d.tilted = int(spec.loose), int(spec.tilted)
        d.nrnodes = spec.sub.size
        d.rnodes = self.arr(spec.sub, keep)
        d.vsys = self.arr(spec.vsys, keep)
        d.xpos, d.ypos = self.arr(spec.xpos, keep), self.arr(spec.ypos, keep)
        d.posa, d.incl = self.arr(spec.posa, keep), self.arr(spec.incl, keep)
        for k in KINDS:
            t = self.Traits()
            tr = spec.traits[k]
            if tr is not None:
                uids, cv, cc, pv, pc = tr
                keep.extend([uids, cc, pc])
                t.n = uids.size
                t.uids = uids.ctypes.data_as(ip)
                t.ccounts = cc.ctypes.data_as(ip)
                t.pcounts = pc.ctypes.data_as(ip)
                t.cvalues = self.arr(cv, keep)
                t.pvalues = self.arr(pv, keep)
            setattr(d, k, t)
        d.opacity = self.arr(spec.opacity, keep)
        return d

    def grid(self, g):
        G = self.Grid()
        G.spat_size[:] = list(g.spat_size)
        G.spat_step[:] = list(g.spat_step)
        G.spat_zero[:] = list(g.spat_zero)
        G.spec_size = g.spec_size
        G.spec_step = g.spec_step
        G.spec_zero = g.spec_zero
        return G

    def _out(self, g, want):
        nx, ny, nz = g.spat_size
        shapes = dict(image=(ny, nx), scube=(g.spec_size, ny, nx))
        res, o = {}, self.Out()
        for k in want:
            res[k] = np.zeros(shapes.get(k, (nz, ny, nx)), self.np)
            if k in ('vdata_cmp', 'ddata_cmp'):
                res[k][:] = np.nan
            setattr(o, k, self.p(res[k]))
        return o, res

    def smdisk(self, spec, g, want=('scube',), parallel=False):
        keep = []
        d = self.disk(spec, keep)
        G = self.grid(g)
        o, res = self._out(g, want)
        self.f('smdisk_omp' if parallel else 'smdisk')(
            ctypes.byref(d), ctypes.byref(G), ctypes.byref(o))
        return res

    def mcdisk(self, spec, g, cflux, csum, hasordint, seed,
               want=('scube',)):
        keep = []
        d = self.disk(spec, keep)
        G = self.grid(g)
        o, res = self._out(g, want)
        csum = np.ascontiguousarray(csum, np.int32)
        hoi = np.ascontiguousarray(hasordint, np.uint8)
        self.f('mcdisk')(self.ct(cflux), int(csum[-1]),
                         csum.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                         csum.size,
                         hoi.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)),
                         ctypes.byref(d), ctypes.byref(G), ctypes.byref(o),
                         int(seed))
        return res

    def _i3(self, v):
        a = np.ascontiguousarray(v, np.int32)
        return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))

    def downscale(self, scale, offset, src, dst_size, parallel=False):
        src = np.ascontiguousarray(src, self.np)
        dst = np.zeros(tuple(dst_size)[::-1], self.np)
        a, b, c, d = (self._i3(scale), self._i3(offset),
                      self._i3(src.shape[::-1]), self._i3(dst_size))
        self.f('downscale')(a[1], b[1], c[1], d[1], self.p(src), self.p(dst),
                            int(parallel))
        return dst

    def mask(self, cutoff, apply, dcube, parallel=False):
        d = np.array(dcube, dtype=self.np, copy=True)
        m = np.zeros_like(d)
        s = self._i3(d.shape[::-1])
        self.f('mask')(self.ct(cutoff), int(apply), s[1], self.p(d),
                       self.p(m), None, int(parallel))
        return d, m

    def moments(self, dcube, step, zero, orders, cutoff, wcube=None,
                parallel=False):
        d = np.ascontiguousarray(dcube, self.np)
        nz, ny, nx = d.shape
        o = np.ascontiguousarray(sorted(set(orders)), np.int32)
        md = np.full((o.size, ny, nx), np.nan, self.np)
        mm = np.zeros((ny, nx), self.np)
        mw = np.ones((o.size, ny, nx), self.np)
        s = self._i3((nx, ny, nz))
        st = np.ascontiguousarray(step, self.np)
        ze = np.ascontiguousarray(zero, self.np)
        w = None if wcube is None else np.ascontiguousarray(wcube, self.np)
        self.f('moments')(s[1], self.p(st), self.p(ze), self.p(d), self.p(w),
                          self.ct(cutoff), o.size,
                          o.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                          self.p(md), self.p(mm), self.p(mw), int(parallel))
        return md, mm, mw

    def residual(self, obs_d, obs_e, obs_m, mdl_d, mdl_w, mdl_m, weight,
                 parallel=False):
        arrs = [None if a is None else
                np.ascontiguousarray(a, self.np).ravel()
                for a in (obs_d, obs_e, obs_m, mdl_d, mdl_w, mdl_m)]
        res = np.zeros_like(arrs[0])
        self.f('residual')(*[self.p(a) for a in arrs], arrs[0].size,
                           self.ct(weight), self.p(res), int(parallel))
        return res

    def residual_sum(self, r, squared=True, parallel=False):
        r = np.ascontiguousarray(r, self.np).ravel()
        return self.f('residual_sum')(self.p(r), r.size, int(squared),
                                      int(parallel))

    def wcube(self, spat, spec_size):
        spat = np.ascontiguousarray(spat, self.np)
        nz, ny, nx = spat.shape
        out = np.zeros((spec_size, ny, nx), self.np)
        size = (ctypes.c_int * 3)(nx, ny, nz)
        self.f('wcube')(size, int(spec_size), self.p(spat), self.p(out))
        return out


def ref(dtype=np.float32):
    key = np.dtype(dtype)
    if key not in _LIBS:
        if not available():
            raise RuntimeError("reference-core harness not built")
        _LIBS[key] = _Ref(key)
    return _LIBS[key]


def num_threads():
    L = ctypes.CDLL(path())
    return L.ref_num_threads()
