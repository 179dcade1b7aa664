"""Galaxy models: disks evaluated into spectral cubes.

Mirror of the reference callers of the GModel backend:
  * Disk parameter handling and packing (src/gbkfit/model/gmodels/_disk.py:
    16-631): sub-ring nodes, linear interpolation of node-wise parameters
    onto them, flat trait arrays (uids / consts / params and their counts);
  * SMDisk / MCDisk (_smdisk.py:38-96, _mcdisk.py:60-189);
  * spectral components SpectralSMDisk2D/3D and SpectralMCDisk3D
    (spectral_smdisk_2d.py, spectral_smdisk_3d.py, spectral_mcdisk_3d.py);
  * brightness components BrightnessSMDisk2D/3D and BrightnessMCDisk3D
    (brightness_smdisk_2d.py, brightness_smdisk_3d.py,
    brightness_mcdisk_3d.py) and opacity components OpacitySMDisk3D /
    OpacityMCDisk3D (opacity_smdisk_3d.py, opacity_mcdisk_3d.py);
  * GModelKinematics2D/3D (kinematics_2d.py, kinematics_3d.py) and
    GModelIntensity2D/3D (intensity_2d.py, intensity_3d.py).
Parameter names follow the reference: 'xpos', 'ypos', 'posa', 'incl',
'vsys' and '<kind>[i]_<name>' per trait (miscutils.py:70-95), with 'cmp<i>_'
prefixed for components after the first and 'ocmp_', 'ocmp<i>_' for opacity
components (_detail.py:339-342).

The 3-D spatial weight and opacity cubes are zeroed before every evaluation:
the reference allocates them uninitialised once and adds into them on every
call (kinematics_3d.py:112-118, intensity_3d.py:111-117; the smdisk kernels
add, gmodels.hpp:933-948), so its weights and opacities depend on what the
memory held.
"""
import itertools

import numpy as np

from . import traits as T

_TRAIT_SLOTS = ('rp', 'rh', 'vp', 'vh', 'dp', 'dh', 'zp', 'sp', 'wp')


def subrnodes(rnodes, rstep):
    """_disk.py:151-156."""
    s = np.arange(rnodes[0] + rstep / 2, rnodes[-1], rstep)
    s = np.insert(s, 0, rnodes[0])
    s = np.append(s, s[-1] + rstep / 2)
    return s


def parse_rnodes(rnmin=None, rnmax=None, rnsep=None, rnlen=None, rnodes=None,
                 rstep=None):
    """gmodels/_detail.py:13-66 (node list, arange or linspace form)."""
    listed = rnodes is not None
    arange = None not in (rnmin, rnmax, rnsep)
    linsp = None not in (rnmin, rnmax, rnlen)
    if [listed, arange, linsp].count(True) != 1:
        raise RuntimeError("exactly one of rnodes, (rnmin, rnmax, rnsep), "
                           "(rnmin, rnmax, rnlen) must be given")
    if arange:
        rnodes = np.arange(rnmin, rnmax + rnsep, rnsep)
    elif linsp:
        rnodes = np.linspace(rnmin, rnmax, rnlen)
    rnodes = np.asarray(rnodes, np.float64)
    if rnodes.size < 2:
        raise RuntimeError("at least two rnodes must be provided")
    d = np.diff(rnodes)
    if np.any(d <= 0) or np.any(rnodes < 0):
        raise RuntimeError("rnodes must be positive, unique and ascending")
    if rstep is None:
        rstep = min(1.0, float(d.min()) / 2)
    if rstep <= 0 or rstep > d.min() / 2:
        raise RuntimeError("rstep must be in (0, min(diff(rnodes)) / 2]")
    return rnodes, float(rstep)


def _names(traits_, prefix):
    """Per trait: dict original name -> full parameter name."""
    out = []
    for i, tr in enumerate(traits_):
        full = prefix + (str(i) if i else '') + '_'
        names = [n for n, _ in tr.params_sm] + list(tr.params_nw)
        out.append({n: full + n for n in names})
    return out


class Disk:
    """Shared machinery of SMDisk and MCDisk (_disk.py:128-631)."""

    def __init__(self, loose, tilted, rnodes, rstep, traits_):
        self.loose = bool(loose)
        self.tilted = bool(tilted)
        self.rnodes = np.asarray(rnodes, np.float64)
        self.rstep = float(rstep)
        self.subrnodes = subrnodes(self.rnodes, self.rstep)
        self.traits = {k: tuple(traits_.get(k, ())) for k in _TRAIT_SLOTS}
        self.is_opacity = bool(self.traits['rp']) and \
            self.traits['rp'][0].kind == 'opt'
        rp, rh = ('opt', 'oht') if self.is_opacity else ('bpt', 'bht')
        self.prefixes = dict(rp=rp, rh=rh, vp='vpt', vh='vht', dp='dpt',
                             dh='dht', zp='zpt', sp='spt', wp='wpt')
        self.names = {k: _names(self.traits[k], self.prefixes[k])
                      for k in _TRAIT_SLOTS}
        nn = self.rnodes.size
        pd = {}
        if self.traits['vp']:
            pd['vsys'] = nn if loose else 1
        pd['xpos'] = nn if loose else 1
        pd['ypos'] = nn if loose else 1
        pd['posa'] = nn if tilted else 1
        pd['incl'] = nn if tilted else 1
        for k in _TRAIT_SLOTS:
            for tr, names in zip(self.traits[k], self.names[k]):
                for n, size in tr.params_sm:
                    pd[names[n]] = size
                for n in tr.params_nw:
                    pd[names[n]] = nn
        self._pdescs = pd
        self._dev = None
        self._driver = None

    def pdescs(self):
        return dict(self._pdescs)

    def _interp(self, values):
        v = np.asarray(values, np.float64)
        if v.size == 1:
            v = np.full(self.rnodes.size, float(v))
        if v.size != self.rnodes.size:
            raise RuntimeError("node-wise parameter has the wrong length")
        lo, hi = self.rnodes[0], self.rnodes[-1]
        if self.subrnodes[0] < lo - 1e-9 or self.subrnodes[-1] > hi + 1e-9:
            raise RuntimeError("sub-ring nodes outside the node range")
        return np.interp(self.subrnodes, self.rnodes, v)

    def pack(self, params):
        """Host arrays for one parameter set: common parameters and, per
        trait kind, (uids, cvalues, ccounts, pvalues, pcounts)
        (_disk.py:42-125,515-553)."""
        def common(name, nw):
            if name not in self._pdescs:
                return None
            v = params[name]
            if nw:
                return self._interp(v).astype(np.float32)
            return np.atleast_1d(np.asarray(v, np.float32))[:1].copy()
        out = dict(vsys=common('vsys', self.loose),
                   xpos=common('xpos', self.loose),
                   ypos=common('ypos', self.loose),
                   posa=common('posa', self.tilted),
                   incl=common('incl', self.tilted))
        for k in _TRAIT_SLOTS:
            tl = self.traits[k]
            if not tl:
                out[k] = None
                continue
            uids, cv, cc, pv, pc = [], [], [], [], []
            for tr, names in zip(tl, self.names[k]):
                uids.append(tr.uid())
                cv += list(tr.consts())
                cc.append(len(tr.consts()))
                vals = []
                for n, size in tr.params_sm:
                    vals += list(np.atleast_1d(np.asarray(params[names[n]],
                                                          np.float64)))[:size]
                for n in tr.params_nw:
                    vals += list(self._interp(params[names[n]]))
                pv += vals
                pc.append(len(vals))
            out[k] = (np.array(uids, np.int32),
                      np.array(cv or [0.0], np.float32),
                      np.array(cc, np.int32),
                      np.array(pv or [0.0], np.float32),
                      np.array(pc, np.int32))
        return out

    def _upload(self, driver, packed):
        """Copies the packed arrays of one parameter set to the device in one
        transfer: they lie back to back (16-byte aligned) in a pinned host
        staging buffer and in one device buffer whose views are the native
        arguments.  The copy runs on the current stream, before the kernels
        that read it; the staging buffer is rewritten only once the previous
        copy has read it."""
        from .driver.hip import torch as _torch, _tdtype
        t = _torch()
        if self._driver is not driver or self._dev is None:
            self._driver = driver
            self._dev = {}
            self._dev['rnodes'] = driver.mem_copy_h2d(
                self.subrnodes.astype(np.float32))
            self._layout = None
        dev = self._dev
        items = [(k, packed[k])
                 for k in ('vsys', 'xpos', 'ypos', 'posa', 'incl')]
        for k in _TRAIT_SLOTS:
            tr = packed[k]
            for j, name in enumerate(('uids', 'cvalues', 'ccounts', 'pvalues',
                                      'pcounts')):
                items.append((f'{k}_{name}', None if tr is None else tr[j]))
        layout = tuple((k, None if a is None else (a.shape, a.dtype.str))
                       for k, a in items)
        if layout != self._layout:
            offs, at = {}, 0
            for k, a in items:
                if a is not None:
                    offs[k] = at
                    at += (a.nbytes + 15) // 16 * 16
            total = max(at, 16)
            self._hbuf = t.empty(total, dtype=t.uint8, pin_memory=True)
            self._dbuf = t.empty(total, dtype=t.uint8, device=driver.device)
            self._hnp = self._hbuf.numpy()
            self._offs = offs
            self._copied = None
            for k, a in items:
                if a is None:
                    dev[k] = None
                    continue
                o = offs[k]
                dev[k] = self._dbuf[o:o + a.nbytes].view(
                    _tdtype(a.dtype)).view(a.shape)
            self._layout = layout
        if self._copied is not None:
            self._copied.synchronize()
        for k, a in items:
            if a is not None:
                o = self._offs[k]
                self._hnp[o:o + a.nbytes] = \
                    np.ascontiguousarray(a).view(np.uint8).ravel()
        self._dbuf.copy_(self._hbuf, non_blocking=True)
        self._copied = t.cuda.Event()
        self._copied.record()
        return dev

    def native_args(self, driver, params):
        """The positional disk arguments of smdisk/mcdisk_evaluate
        (loose .. wpt_pcounts)."""
        dev = self._upload(driver, self.pack(params))
        args = [self.loose, self.tilted, dev['rnodes'], dev['vsys'],
                dev['xpos'], dev['ypos'], dev['posa'], dev['incl']]
        for k in _TRAIT_SLOTS:
            for name in ('uids', 'cvalues', 'ccounts', 'pvalues', 'pcounts'):
                args.append(dev[f'{k}_{name}'])
        return args


class SMDisk(Disk):
    """_smdisk.py:8-96."""

    def evaluate(self, driver, params, odata, image, scube, wdata, rdata,
                 ordata, spat_size, spat_step, spat_zero, spat_rota,
                 spec_size, spec_step, spec_zero, dtype, out_extra=None):
        backend = driver.backends().gmodel(dtype)
        backend.smdisk_evaluate(
            *self.native_args(driver, params), odata,
            spat_size, spat_step, spat_zero, spat_rota,
            spec_size, spec_step, spec_zero,
            image, scube, wdata, None, rdata, None, ordata, None, None, None)


class MCDisk(Disk):
    """_mcdisk.py:16-189: the cloud budget is the trait integral per ring
    (or whole disk) divided by the flux of one cloud."""

    def __init__(self, cflux, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.cflux = float(cflux)
        self._budget_dev = None

    def cloud_budget(self, params):
        counts = []
        rings = self.subrnodes[1:-1]
        for tr, names in zip(self.traits['rp'], self.names['rp']):
            tp = {}
            for n, _ in tr.params_sm:
                tp[n] = params[names[n]]
            for n in tr.params_nw:
                tp[n] = self._interp(params[names[n]])[1:-1]
            integral = tr.integrate(tp, rings)
            counts.extend(np.atleast_1d(integral / self.cflux)
                          .astype(np.int32))
        csum = np.array(list(itertools.accumulate(counts)), np.int32)
        hoi = np.array([tr.has_analytical_integral()
                        for tr in self.traits['rp']], np.uint8)
        return csum, hoi

    def evaluate(self, driver, params, odata, image, scube, wdata, rdata,
                 ordata, spat_size, spat_step, spat_zero, spat_rota,
                 spec_size, spec_step, spec_zero, dtype, out_extra=None):
        backend = driver.backends().gmodel(dtype)
        csum, hoi = self.cloud_budget(params)
        d_csum = driver.mem_copy_h2d(csum)
        d_hoi = driver.mem_copy_h2d(hoi)
        backend.mcdisk_evaluate(
            self.cflux, int(csum[-1]) if csum.size else 0, d_csum, d_hoi,
            *self.native_args(driver, params), odata,
            spat_size, spat_step, spat_zero, spat_rota,
            spec_size, spec_step, spec_zero,
            image, scube, wdata, None, rdata, None, ordata, None, None, None)


def _traits_of(kind, value):
    if value is None:
        return ()
    if isinstance(value, (list, tuple)):
        return tuple(T.load(kind, v) if isinstance(v, dict) else v
                     for v in value)
    return (T.load(kind, value) if isinstance(value, dict) else value,)


class _Component:
    def pdescs(self):
        return self.disk.pdescs()

    def has_weights(self):
        return bool(self.disk.traits['wp'])


class SpectralSMDisk2D(_Component):
    """Thin smooth disk (spectral_smdisk_2d.py:60-148): evaluated on a single
    spatial plane (size_z 1, step_z 0, zero_z 0)."""

    def __init__(self, loose, tilted, bptraits, vptraits, dptraits,
                 sptraits=None, wptraits=None, rnmin=None, rnmax=None,
                 rnsep=None, rnlen=None, rnodes=None, rstep=None,
                 interp='linear'):
        if interp != 'linear':
            raise RuntimeError("only linear interpolation is supported")
        rn, rs = parse_rnodes(rnmin, rnmax, rnsep, rnlen, rnodes, rstep)
        tr = dict(rp=_traits_of('bpt', bptraits), vp=_traits_of('vpt', vptraits),
                  dp=_traits_of('dpt', dptraits), sp=_traits_of('spt', sptraits),
                  wp=_traits_of('wpt', wptraits))
        if not (tr['rp'] and tr['vp'] and tr['dp']):
            raise RuntimeError("bptraits, vptraits and dptraits are required")
        self.disk = SMDisk(loose, tilted, rn, rs, tr)
        self.thin = True

    def evaluate(self, driver, params, scube, wdata, bdata, spat_size,
                 spat_step, spat_zero, spat_rota, spec_size, spec_step,
                 spec_zero, dtype, out_extra=None):
        self.disk.evaluate(driver, params, None, None, scube, wdata, bdata,
                           None, tuple(spat_size) + (1,),
                           tuple(spat_step) + (0,), tuple(spat_zero) + (0,),
                           spat_rota, spec_size, spec_step, spec_zero, dtype,
                           out_extra)


def _thick_traits(bptraits, bhtraits, vptraits, vhtraits, dptraits, dhtraits,
                  zptraits, sptraits, wptraits):
    """parse_component_s3d_traits (gmodels/_detail.py:183-266): missing
    velocity / dispersion height traits default to 'one'."""
    tr = dict(rp=_traits_of('bpt', bptraits), rh=_traits_of('bht', bhtraits),
              vp=_traits_of('vpt', vptraits), dp=_traits_of('dpt', dptraits),
              zp=_traits_of('zpt', zptraits), sp=_traits_of('spt', sptraits),
              wp=_traits_of('wpt', wptraits))
    if not (tr['rp'] and tr['rh'] and tr['vp'] and tr['dp']):
        raise RuntimeError("bptraits, bhtraits, vptraits and dptraits are "
                           "required")
    tr['vh'] = _traits_of('vht', vhtraits) or tuple(
        T.make('vht', 'one') for _ in tr['vp'])
    tr['dh'] = _traits_of('dht', dhtraits) or tuple(
        T.make('dht', 'one') for _ in tr['dp'])
    for a, b in (('rp', 'rh'), ('vp', 'vh'), ('dp', 'dh')):
        if len(tr[a]) != len(tr[b]):
            raise RuntimeError(f"{a} and {b} trait counts differ")
    return tr


class SpectralSMDisk3D(_Component):
    """Thick smooth disk (spectral_smdisk_3d.py:62-129)."""

    def __init__(self, loose, tilted, bptraits, vptraits, dptraits, bhtraits,
                 vhtraits=None, dhtraits=None, zptraits=None, sptraits=None,
                 wptraits=None, rnmin=None, rnmax=None, rnsep=None,
                 rnlen=None, rnodes=None, rstep=None, interp='linear'):
        if interp != 'linear':
            raise RuntimeError("only linear interpolation is supported")
        rn, rs = parse_rnodes(rnmin, rnmax, rnsep, rnlen, rnodes, rstep)
        self.disk = SMDisk(loose, tilted, rn, rs, _thick_traits(
            bptraits, bhtraits, vptraits, vhtraits, dptraits, dhtraits,
            zptraits, sptraits, wptraits))
        self.thin = False

    def evaluate(self, driver, params, odata, scube, wdata, bdata, obdata,
                 spat_size, spat_step, spat_zero, spat_rota, spec_size,
                 spec_step, spec_zero, dtype, out_extra=None):
        self.disk.evaluate(driver, params, odata, None, scube, wdata, bdata,
                           obdata, spat_size, spat_step, spat_zero, spat_rota,
                           spec_size, spec_step, spec_zero, dtype, out_extra)


class SpectralMCDisk3D(_Component):
    """Monte Carlo cloud disk (spectral_mcdisk_3d.py:14-133)."""

    def __init__(self, cflux, loose, tilted, bptraits, vptraits, dptraits,
                 bhtraits, vhtraits=None, dhtraits=None, zptraits=None,
                 sptraits=None, wptraits=None, rnmin=None, rnmax=None,
                 rnsep=None, rnlen=None, rnodes=None, rstep=None,
                 interp='linear'):
        if interp != 'linear':
            raise RuntimeError("only linear interpolation is supported")
        rn, rs = parse_rnodes(rnmin, rnmax, rnsep, rnlen, rnodes, rstep)
        self.disk = MCDisk(cflux, loose, tilted, rn, rs, _thick_traits(
            bptraits, bhtraits, vptraits, vhtraits, dptraits, dhtraits,
            zptraits, sptraits, wptraits))
        self.thin = False

    def evaluate(self, driver, params, odata, scube, wdata, bdata, obdata,
                 spat_size, spat_step, spat_zero, spat_rota, spec_size,
                 spec_step, spec_zero, dtype, out_extra=None):
        self.disk.evaluate(driver, params, odata, None, scube, wdata, bdata,
                           obdata, spat_size, spat_step, spat_zero, spat_rota,
                           spec_size, spec_step, spec_zero, dtype, out_extra)


def _b3d_traits(bptraits, bhtraits, zptraits, sptraits, wptraits,
                kinds=('bpt', 'bht')):
    """parse_component_b3d_traits / parse_component_o3d_traits
    (gmodels/_detail.py): density polar and height traits, one height trait
    per polar one."""
    rp, rh = kinds
    tr = dict(rp=_traits_of(rp, bptraits), rh=_traits_of(rh, bhtraits),
              zp=_traits_of('zpt', zptraits), sp=_traits_of('spt', sptraits),
              wp=_traits_of('wpt', wptraits))
    if not (tr['rp'] and tr['rh']):
        raise RuntimeError(f"{rp}traits and {rh}traits are required")
    if len(tr['rp']) != len(tr['rh']):
        raise RuntimeError(f"{rp} and {rh} trait counts differ")
    return tr


class BrightnessSMDisk2D(_Component):
    """Thin smooth disk of surface brightness (brightness_smdisk_2d.py:
    50-106): density polar, selection and weight traits, no kinematics,
    evaluated into an image (spectral size 1)."""

    def __init__(self, loose, tilted, bptraits, sptraits=None, wptraits=None,
                 rnmin=None, rnmax=None, rnsep=None, rnlen=None, rnodes=None,
                 rstep=None, interp='linear'):
        if interp != 'linear':
            raise RuntimeError("only linear interpolation is supported")
        rn, rs = parse_rnodes(rnmin, rnmax, rnsep, rnlen, rnodes, rstep)
        tr = dict(rp=_traits_of('bpt', bptraits),
                  sp=_traits_of('spt', sptraits),
                  wp=_traits_of('wpt', wptraits))
        if not tr['rp']:
            raise RuntimeError("bptraits are required")
        self.disk = SMDisk(loose, tilted, rn, rs, tr)
        self.thin = True

    def evaluate(self, driver, params, image, wdata, bdata, spat_size,
                 spat_step, spat_zero, spat_rota, dtype, out_extra=None):
        self.disk.evaluate(driver, params, None, image, None, wdata, bdata,
                           None, tuple(spat_size) + (1,),
                           tuple(spat_step) + (0,), tuple(spat_zero) + (0,),
                           spat_rota, 1, 0, 0, dtype, out_extra)


class BrightnessSMDisk3D(_Component):
    """Thick smooth disk of brightness (brightness_smdisk_3d.py:62-123):
    the image is the line-of-sight sum through the opacity cube."""

    def __init__(self, loose, tilted, bptraits, bhtraits, zptraits=None,
                 sptraits=None, wptraits=None, rnmin=None, rnmax=None,
                 rnsep=None, rnlen=None, rnodes=None, rstep=None,
                 interp='linear'):
        if interp != 'linear':
            raise RuntimeError("only linear interpolation is supported")
        rn, rs = parse_rnodes(rnmin, rnmax, rnsep, rnlen, rnodes, rstep)
        self.disk = SMDisk(loose, tilted, rn, rs, _b3d_traits(
            bptraits, bhtraits, zptraits, sptraits, wptraits))
        self.thin = False

    def evaluate(self, driver, params, odata, image, wdata, bdata, obdata,
                 spat_size, spat_step, spat_zero, spat_rota, dtype,
                 out_extra=None):
        self.disk.evaluate(driver, params, odata, image, None, wdata, bdata,
                           obdata, spat_size, spat_step, spat_zero,
                           spat_rota, 1, 0, 0, dtype, out_extra)


class BrightnessMCDisk3D(_Component):
    """Monte Carlo cloud disk of brightness (brightness_mcdisk_3d.py:
    55-105)."""

    def __init__(self, cflux, loose, tilted, bptraits, bhtraits,
                 zptraits=None, sptraits=None, wptraits=None, rnmin=None,
                 rnmax=None, rnsep=None, rnlen=None, rnodes=None, rstep=None,
                 interp='linear'):
        if interp != 'linear':
            raise RuntimeError("only linear interpolation is supported")
        rn, rs = parse_rnodes(rnmin, rnmax, rnsep, rnlen, rnodes, rstep)
        self.disk = MCDisk(cflux, loose, tilted, rn, rs, _b3d_traits(
            bptraits, bhtraits, zptraits, sptraits, wptraits))
        self.thin = False

    evaluate = BrightnessSMDisk3D.evaluate


class OpacitySMDisk3D(_Component):
    """Smooth disk of opacity (opacity_smdisk_3d.py:52-127): its density
    (opacity polar and height traits) is added into the 3-D opacity cube,
    which the brightness or spectral components then read."""

    def __init__(self, loose, tilted, optraits, ohtraits, zptraits=None,
                 sptraits=None, wptraits=None, rnmin=None, rnmax=None,
                 rnsep=None, rnlen=None, rnodes=None, rstep=None,
                 interp='linear'):
        if interp != 'linear':
            raise RuntimeError("only linear interpolation is supported")
        rn, rs = parse_rnodes(rnmin, rnmax, rnsep, rnlen, rnodes, rstep)
        self.disk = SMDisk(loose, tilted, rn, rs, _b3d_traits(
            optraits, ohtraits, zptraits, sptraits, wptraits,
            kinds=('opt', 'oht')))
        self.thin = False

    def evaluate(self, driver, params, odata, spat_size, spat_step,
                 spat_zero, spat_rota, dtype, out_extra=None):
        self.disk.evaluate(driver, params, None, None, None, None, odata,
                           None, spat_size, spat_step, spat_zero, spat_rota,
                           1, 0, 0, dtype, out_extra)


class OpacityMCDisk3D(_Component):
    """Monte Carlo cloud disk of opacity (opacity_mcdisk_3d.py:54-120)."""

    def __init__(self, cflux, loose, tilted, optraits, ohtraits,
                 zptraits=None, sptraits=None, wptraits=None, rnmin=None,
                 rnmax=None, rnsep=None, rnlen=None, rnodes=None, rstep=None,
                 interp='linear'):
        if interp != 'linear':
            raise RuntimeError("only linear interpolation is supported")
        rn, rs = parse_rnodes(rnmin, rnmax, rnsep, rnlen, rnodes, rstep)
        self.disk = MCDisk(cflux, loose, tilted, rn, rs, _b3d_traits(
            optraits, ohtraits, zptraits, sptraits, wptraits,
            kinds=('opt', 'oht')))
        self.thin = False

    evaluate = OpacitySMDisk3D.evaluate


def _merge_params(components, prefix='cmp', zero_prefix=False):
    """Component parameters: '<prefix><i>_'-prefixed after the first, and
    '<prefix>_' on the first with zero_prefix (miscutils.py:70-95,
    gmodels/_detail.py:320-342)."""
    merged, maps = {}, []
    for i, c in enumerate(components):
        full_prefix = (prefix if i or zero_prefix else '') + \
            (str(i) if i else '')
        full_prefix += '_' if full_prefix else ''
        m = {}
        for n, size in c.pdescs().items():
            m[n] = full_prefix + n
            merged[full_prefix + n] = size
        maps.append(m)
    return merged, maps


def _tuplify(components):
    if components is None:
        return ()
    return tuple(components) if isinstance(components, (list, tuple)) \
        else (components,)


def _sub_params(params, c, m):
    return {n: params[m[n]] for n in c.pdescs()}


class _GModel:
    """Components and opacity components, their merged parameters, and the
    3-D spatial grid of the thick gmodels: z extent of the larger spatial
    axis with zero in the middle unless given (kinematics_3d.py:97-109,
    intensity_3d.py:95-107)."""

    def __init__(self, components, opacity_components=None, size_z=None,
                 step_z=None, zero_z=None):
        self.components = _tuplify(components)
        if not self.components:
            raise RuntimeError("at least one component must be configured")
        self.ocomponents = _tuplify(opacity_components)
        self._zspec = (size_z, step_z, zero_z)
        p, self._maps = _merge_params(self.components)
        op, self._omaps = _merge_params(self.ocomponents, 'ocmp', True)
        self._pdescs = {**p, **op}
        self._wdata = self._odata = None
        self._shape = None

    def pdescs(self):
        return dict(self._pdescs)

    def has_weights(self):
        return any(c.has_weights() for c in self.components)

    def spatial_grid(self, size, step, zero):
        sz, stz, zz = self._zspec
        wide = int(size[0] < size[1])
        if sz is None:
            sz = size[wide]
        if stz is None:
            stz = step[wide]
        if zz is None:
            zz = -(sz / 2 - 0.5) * stz
        return ((int(size[0]), int(size[1]), int(sz)),
                (step[0], step[1], stz), (zero[0], zero[1], zz))

    def _spatial(self, driver, spat_size, dtype, want_w):
        """Zeroed 3-D weight (when weighted) and opacity (when there are
        opacity components) cubes of the spatial grid."""
        shape = (tuple(spat_size)[::-1], dtype)
        if shape != self._shape:
            self._shape = shape
            self._wdata = self._odata = None
        if want_w:
            if self._wdata is None:
                self._wdata = driver.mem_alloc_d(*shape)
            driver.mem_fill(self._wdata, 0)
        if self.ocomponents:
            if self._odata is None:
                self._odata = driver.mem_alloc_d(*shape)
            driver.mem_fill(self._odata, 0)
        return (self._wdata if want_w else None), \
            (self._odata if self.ocomponents else None)

    def _opacity(self, driver, params, odata, spat_size, spat_step,
                 spat_zero, rota, dtype, out_extra):
        for c, m in zip(self.ocomponents, self._omaps):
            c.evaluate(driver, _sub_params(params, c, m), odata, spat_size,
                       spat_step, spat_zero, rota, dtype, out_extra)


class GModelKinematics2D(_GModel):
    """kinematics_2d.py: thin components on the data model's spatial grid."""

    def __init__(self, components):
        super().__init__(components)

    def evaluate_scube(self, driver, params, scube_d, scube_w, size, step,
                       zero, rota, dtype, out_extra=None):
        spat_size, spat_step, spat_zero = size[:2], step[:2], zero[:2]
        wdata, _ = self._spatial(driver, tuple(spat_size) + (1,), dtype,
                                 scube_w is not None)
        for c, m in zip(self.components, self._maps):
            c.evaluate(driver, _sub_params(params, c, m), scube_d, wdata,
                       None, spat_size, spat_step, spat_zero, rota, size[2],
                       step[2], zero[2], dtype, out_extra)
        if scube_w is not None:
            driver.backends().gmodel(dtype).wcube_evaluate(
                tuple(spat_size) + (1,), size[2], wdata, scube_w)


class GModelKinematics3D(_GModel):
    """kinematics_3d.py:28-187: thick spectral components, and opacity
    components evaluated first into the opacity cube they read, on a 3-D
    spatial grid."""

    def __init__(self, components, opacity_components=None, size_z=None,
                 step_z=None, zero_z=None):
        super().__init__(components, opacity_components, size_z, step_z,
                         zero_z)

    def evaluate_scube(self, driver, params, scube_d, scube_w, size, step,
                       zero, rota, dtype, out_extra=None):
        spat_size, spat_step, spat_zero = self.spatial_grid(size, step, zero)
        wdata, odata = self._spatial(driver, spat_size, dtype,
                                     scube_w is not None)
        self._opacity(driver, params, odata, spat_size, spat_step, spat_zero,
                      rota, dtype, out_extra)
        for c, m in zip(self.components, self._maps):
            c.evaluate(driver, _sub_params(params, c, m), odata, scube_d,
                       wdata, None, None, spat_size, spat_step, spat_zero,
                       rota, size[2], step[2], zero[2], dtype, out_extra)
        if scube_w is not None:
            driver.backends().gmodel(dtype).wcube_evaluate(
                spat_size, size[2], wdata, scube_w)


class GModelIntensity2D(_GModel):
    """intensity_2d.py: thin brightness components summed into an image on
    the data model's spatial grid; the weight image is the spaxel's weight
    (wcube with one channel)."""

    def __init__(self, components):
        super().__init__(components)

    def evaluate_image(self, driver, params, image_d, image_w, size, step,
                       zero, rota, dtype, out_extra=None):
        spat_size, spat_step, spat_zero = size[:2], step[:2], zero[:2]
        wdata, _ = self._spatial(driver, tuple(spat_size) + (1,), dtype,
                                 image_w is not None)
        for c, m in zip(self.components, self._maps):
            c.evaluate(driver, _sub_params(params, c, m), image_d, wdata,
                       None, spat_size, spat_step, spat_zero, rota, dtype,
                       out_extra)
        if image_w is not None:
            driver.backends().gmodel(dtype).wcube_evaluate(
                tuple(spat_size) + (1,), 1, wdata, image_w)


class GModelIntensity3D(_GModel):
    """intensity_3d.py:31-199: thick brightness components (and opacity
    components evaluated first) on a 3-D spatial grid, summed along the line
    of sight into an image."""

    def __init__(self, components, opacity_components=None, size_z=None,
                 step_z=None, zero_z=None):
        super().__init__(components, opacity_components, size_z, step_z,
                         zero_z)

    def evaluate_image(self, driver, params, image_d, image_w, size, step,
                       zero, rota, dtype, out_extra=None):
        spat_size, spat_step, spat_zero = self.spatial_grid(size[:2],
                                                            step[:2],
                                                            zero[:2])
        wdata, odata = self._spatial(driver, spat_size, dtype,
                                     image_w is not None)
        self._opacity(driver, params, odata, spat_size, spat_step, spat_zero,
                      rota, dtype, out_extra)
        for c, m in zip(self.components, self._maps):
            c.evaluate(driver, _sub_params(params, c, m), odata, image_d,
                       wdata, None, None, spat_size, spat_step, spat_zero,
                       rota, dtype, out_extra)
        if image_w is not None:
            driver.backends().gmodel(dtype).wcube_evaluate(
                spat_size, 1, wdata, image_w)
