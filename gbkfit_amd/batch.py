"""Batched and served evaluation of many parameter sets of one galaxy model.

The reference evaluates one parameter set per likelihood call (its fitters
call Objective.log_likelihood per walker, src/gbkfit/fitting/fitters/emcee/
sampler.py:34-57).  ``BatchEvaluator`` evaluates N sets of one disk model
(DModelSCube + Kinematics2D/3D with one SpectralSMDisk or SpectralMCDisk
component) against one observed cube in a single native call
(gbkhip_batch_evaluate): the batched model, the convolution with the shared
PSF x LSF kernel (direct separable filters, a spectral-first fused FFT, or an
FFT of the padded cube) and a fused downscale + residual + FP64 chi^2 pass,
captured in a HIP graph.
"""
import ctypes

import numpy as np

from . import _native as N
from .driver.hip import NativeMemoryHip, make_grid

_c = ctypes


class BTraits(_c.Structure):
    _fields_ = [('n', _c.c_int), ('uids', _c.POINTER(_c.c_int)),
                ('ccounts', _c.POINTER(_c.c_int)),
                ('pcounts', _c.POINTER(_c.c_int)),
                ('off_cvalues', _c.c_int), ('off_pvalues', _c.c_int)]


class BDesc(_c.Structure):
    _fields_ = [('loose', _c.c_int), ('tilted', _c.c_int),
                ('nrnodes', _c.c_int), ('rnodes', _c.POINTER(_c.c_float)),
                ('vsys_count', _c.c_int), ('off_vsys', _c.c_int),
                ('off_xpos', _c.c_int), ('off_ypos', _c.c_int),
                ('off_posa', _c.c_int), ('off_incl', _c.c_int),
                ('traits', BTraits * 9), ('pstride', _c.c_int),
                ('grid', N.Grid), ('size_lo', _c.c_int * 3),
                ('size_hi', _c.c_int * 3), ('scale', _c.c_int * 3),
                ('edge_hi', _c.c_int * 3), ('has_conv', _c.c_int),
                ('pcube', _c.c_void_p), ('obs', _c.c_void_p),
                ('err', _c.c_void_p), ('mask', _c.c_void_p),
                ('err_scalar', _c.c_float), ('weight', _c.c_float),
                ('max_batch', _c.c_int), ('path', _c.c_int),
                ('mcdisk', _c.c_int), ('cflux', _c.c_float),
                ('csum_len', _c.c_int), ('hoi', _c.c_void_p),
                ('seed', _c.c_ulonglong), ('nextra', _c.c_int),
                ('extra', _c.c_void_p)]


N.register_signatures({
    'gbkhip_batch_create': ([_c.c_void_p, _c.POINTER(BDesc),
                             _c.POINTER(_c.c_void_p)], _c.c_int),
    'gbkhip_batch_destroy': ([_c.c_void_p], _c.c_int),
    'gbkhip_batch_evaluate': ([_c.c_void_p, _c.c_void_p, _c.c_int,
                               _c.c_void_p, _c.c_void_p, _c.c_int], _c.c_int),
    'gbkhip_batch_evaluate_clouds': ([_c.c_void_p, _c.c_void_p, _c.c_void_p,
                                      _c.c_void_p, _c.c_int, _c.c_void_p,
                                      _c.c_void_p, _c.c_int], _c.c_int),
    'gbkhip_mcdisk_fixed_scale': ([_c.c_float, _c.c_int, _c.c_void_p,
                                   _c.POINTER(_c.c_double)], _c.c_int),
    'gbkhip_batch_info': ([_c.c_void_p, _c.POINTER(_c.c_int)], _c.c_int),
    'gbkhip_batch_taps': ([_c.c_void_p, _c.POINTER(_c.c_int)], _c.c_int),
    'gbkhip_batch_set_tiles': ([_c.c_void_p, _c.c_int, _c.c_int], _c.c_int),
    'gbkhip_batch_reserve_clouds': ([_c.c_void_p, _c.c_longlong], _c.c_int),
    'gbkhip_batch_check': ([_c.c_void_p, _c.POINTER(_c.c_int)], _c.c_int),
    'gbkhip_batch_tile_stats': ([_c.c_void_p, _c.POINTER(_c.c_longlong)],
                                _c.c_int),
    'gbkhip_batch_line_stats': ([_c.c_void_p, _c.POINTER(_c.c_longlong)],
                                _c.c_int),
    'gbkhip_group_create': ([_c.c_void_p, _c.POINTER(_c.c_void_p), _c.c_int,
                             _c.POINTER(_c.c_void_p)], _c.c_int),
    'gbkhip_group_destroy': ([_c.c_void_p], _c.c_int),
    'gbkhip_group_evaluate': ([_c.c_void_p, _c.POINTER(_c.c_void_p),
                               _c.POINTER(_c.c_void_p),
                               _c.POINTER(_c.c_void_p),
                               _c.POINTER(_c.c_int),
                               _c.POINTER(_c.c_void_p),
                               _c.POINTER(_c.c_void_p), _c.c_int], _c.c_int),
    'gbkhip_batch_kernel_times': ([_c.c_void_p, _c.POINTER(_c.c_double),
                                   _c.c_int, _c.POINTER(_c.c_int)], _c.c_int),
    'gbkhip_batch_set_timing': ([_c.c_void_p, _c.c_int], _c.c_int),
    'gbkhip_batch_set_stream': ([_c.c_void_p, _c.c_void_p], _c.c_int),
    'gbkhip_batch_arena_bytes': ([_c.c_void_p,
                                  _c.POINTER(_c.c_ulonglong)], _c.c_int),
    'gbkhip_batch_bind_work': ([_c.c_void_p, _c.c_void_p, _c.c_ulonglong],
                               _c.c_int),
    'gbkhip_batch_set_mmaps': ([_c.c_void_p, _c.c_int, _c.c_void_p,
                                _c.c_float, _c.c_float, _c.c_float,
                                _c.c_void_p, _c.c_void_p, _c.c_void_p,
                                _c.c_void_p, _c.c_void_p], _c.c_int),
    'gbkhip_batch_release': ([_c.c_void_p], _c.c_int),
    'gbkhip_batch_memory': ([_c.c_void_p, _c.POINTER(_c.c_ulonglong),
                             _c.POINTER(_c.c_ulonglong),
                             _c.POINTER(_c.c_int)], _c.c_int),
    'gbkhip_batch_stage_times': ([_c.c_void_p, _c.POINTER(_c.c_double),
                                  _c.POINTER(_c.c_int)], _c.c_int),
})

_SLOTS = ('rp', 'rh', 'vp', 'vh', 'dp', 'dh', 'zp', 'sp', 'wp')


def _interp_matrix(nodes, sub):
    """W with W @ values == np.interp(sub, nodes, values)."""
    W = np.zeros((sub.size, nodes.size))
    j = np.clip(np.searchsorted(nodes, sub, side='right') - 1, 0,
                nodes.size - 2)
    t = (sub - nodes[j]) / (nodes[j + 1] - nodes[j])
    rows = np.arange(sub.size)
    W[rows, j] = 1 - t
    W[rows, j + 1] = t
    return W


class SetPacker:
    """Flattens parameter dicts of one Disk into fixed-layout float vectors:
    [vsys][xpos][ypos][posa][incl] then per trait slot its constants and its
    parameters (the arrays Disk.pack produces, concatenated)."""

    def __init__(self, disk, template_params):
        packed = disk.pack(template_params)
        self.disk = disk
        self._W = None
        off = 0
        self.offsets = {}
        for k in ('vsys', 'xpos', 'ypos', 'posa', 'incl'):
            a = packed[k]
            self.offsets[k] = (off, 0 if a is None else a.size)
            off += 0 if a is None else a.size
        self.trait_offsets = {}
        self.trait_ints = {}
        for k in _SLOTS:
            tr = packed[k]
            if tr is None:
                continue
            uids, cv, cc, pv, pc = tr
            self.trait_offsets[k] = (off, off + cv.size)
            self.trait_ints[k] = (uids.copy(), cc.copy(), pc.copy())
            off += cv.size + pv.size
        self.pstride = off

    def pack_arrays(self, arrays, nsets):
        """Vectorised packing: `arrays` maps parameter name -> array of
        shape (nsets,) (scalar parameters) or (nsets, n) (vector / node-wise
        ones).  Node-wise values are interpolated with one matrix product."""
        d = self.disk
        out = np.zeros((nsets, self.pstride), np.float32)
        if self._W is None:                   # the nodes are fixed per disk
            self._W = _interp_matrix(d.rnodes, d.subrnodes)
        W = self._W

        def col(name, nw):
            v = np.asarray(arrays[name], np.float64).reshape(nsets, -1)
            if nw:
                if v.shape[1] == 1:
                    v = np.repeat(v, d.rnodes.size, axis=1)
                return v @ W.T
            return v
        for k, nw in (('vsys', d.loose), ('xpos', d.loose),
                      ('ypos', d.loose), ('posa', d.tilted),
                      ('incl', d.tilted)):
            o, n = self.offsets[k]
            if n:
                out[:, o:o + n] = col(k, nw)[:, :n]
        for k, (oc, op) in self.trait_offsets.items():
            pos = op
            cpos = oc
            for tr, names in zip(d.traits[k], d.names[k]):
                nc = len(tr.consts())
                out[:, cpos:cpos + nc] = np.asarray(tr.consts(), np.float32)
                cpos += nc
                for n, size in tr.params_sm:
                    out[:, pos:pos + size] = col(names[n], False)[:, :size]
                    pos += size
                for n in tr.params_nw:
                    ns = d.subrnodes.size
                    out[:, pos:pos + ns] = col(names[n], True)
                    pos += ns
        return out

    def pack(self, params_list):
        out = np.zeros((len(params_list), self.pstride), np.float32)
        for i, params in enumerate(params_list):
            packed = self.disk.pack(params)
            for k, (o, n) in self.offsets.items():
                if n:
                    out[i, o:o + n] = packed[k]
            for k, (oc, op) in self.trait_offsets.items():
                _, cv, _, pv, _ = packed[k]
                out[i, oc:oc + cv.size] = cv
                out[i, op:op + pv.size] = pv
        return out


class ModelPacker:
    """Parameter vectors of a gmodel with several disk components: each
    component's SetPacker vector side by side, component i's part starting
    at bases[i].  Parameter dicts carry the gmodel's merged names ('xpos',
    'cmp1_xpos', ...; gmodels/_detail.py:320-342); the attributes of the
    first component's packer (disk, offsets, trait_offsets, trait_ints)
    are passed through."""

    def __init__(self, gmodel, template_params):
        self.components = gmodel.components
        self.maps = gmodel._maps
        self.packers, self.bases = [], []
        base = 0
        for c, m in zip(self.components, self.maps):
            pk = SetPacker(c.disk, self._sub(template_params, c, m))
            self.packers.append(pk)
            self.bases.append(base)
            base += pk.pstride
        self.pstride = base
        first = self.packers[0]
        self.disk = first.disk
        self.offsets = first.offsets
        self.trait_offsets = first.trait_offsets
        self.trait_ints = first.trait_ints

    @staticmethod
    def _sub(params, c, m):
        return {n: params[m[n]] for n in c.pdescs()}

    def pack(self, params_list):
        return np.concatenate(
            [pk.pack([self._sub(p, c, m) for p in params_list])
             for pk, c, m in zip(self.packers, self.components, self.maps)],
            axis=1)

    def pack_arrays(self, arrays, nsets):
        return np.concatenate(
            [pk.pack_arrays(self._sub(arrays, c, m), nsets)
             for pk, c, m in zip(self.packers, self.components, self.maps)],
            axis=1)


class BatchEvaluator:
    """chi^2 of many parameter sets of one disk model against one cube.

    dmodel: DModelSCube (its DCube geometry, PSF and LSF are used);
    gmodel: GModelKinematics2D/3D whose first component is a SpectralSMDisk
            or SpectralMCDisk and whose further components (if any) are
            SpectralSMDisks, summed into one cube as the reference sums
            components (gmodels/_detail.py:405-422; thick smdisk, mcdisk and
            several components: direct or padded-cube pipeline from a
            deposited hi-res cube; mcdisk: each set's cloud budget computed
            on the host as the reference does,
            src/gbkfit/model/gmodels/_mcdisk.py:16-189);
    obs:    observed low-res cube (numpy, z-y-x); error: scalar or cube;
            mask: cube or None.  With a DModelMMaps data model the objective
            is its moment maps (mmaps.py:136-201, objective.py:279-308):
            obs maps each key 'mmap<order>' to its (y, x) data map (a dict,
            or a gbkfit_amd.dataset.DatasetMMaps whose items also give the
            errors and masks), error is a scalar or a dict of scalars or
            maps, mask a dict of maps or None, weight a scalar or a dict of
            per-key weights.
    """

    def __init__(self, driver, dmodel, gmodel, obs, error=None, mask=None,
                 weight=1.0, max_batch=64, template_params=None,
                 path='auto'):
        self.driver = driver
        self.mmaps = type(dmodel).__name__ == 'DModelMMaps'
        comps = gmodel.components
        kinds = [type(getattr(c, 'disk', None)).__name__ for c in comps]
        if kinds[0] not in ('SMDisk', 'MCDisk') or \
                any(k != 'SMDisk' for k in kinds[1:]) or \
                not all(hasattr(c.disk, 'traits') for c in comps):
            raise RuntimeError(
                "BatchEvaluator needs a first smdisk or mcdisk component "
                "and smdisk components after it")
        if getattr(gmodel, 'ocomponents', ()) or \
                not hasattr(gmodel, 'evaluate_scube'):
            raise RuntimeError("BatchEvaluator takes spectral gmodels "
                               "without opacity components")
        disk = comps[0].disk
        self.mc = type(disk).__name__ == 'MCDisk'
        cube = dmodel.dcube()
        be_fft = driver.backends().fft(cube.dtype(), cube._padding)
        size_hi, step_hi, edge_hi, zero_hi = cube.geometry(be_fft)
        thin = getattr(comps[0], 'thin', True)
        # thick disks and mcdisk deposit a hi-res cube that the direct
        # pipeline's x pass reads (thin disks deposit into its LDS tiles)
        self.cube_source = self.mc or not thin
        if thin:
            spat = ((size_hi[0], size_hi[1], 1), (step_hi[0], step_hi[1], 0),
                    (zero_hi[0], zero_hi[1], 0))
        else:
            spat = gmodel.spatial_grid(size_hi, step_hi, zero_hi)
        self.grid = make_grid(spat[0], spat[1], spat[2], size_hi[2],
                              step_hi[2], zero_hi[2])
        if template_params is None:
            template_params = {n: np.zeros(s) if s > 1 else 0.0
                               for n, s in gmodel.pdescs().items()}
        self.packer = SetPacker(disk, template_params) if len(comps) == 1 \
            else ModelPacker(gmodel, template_params)
        t = driver_torch()
        dev = driver.device
        self._keep = []
        d = BDesc()
        d.loose, d.tilted = int(disk.loose), int(disk.tilted)
        rn = np.ascontiguousarray(disk.subrnodes, np.float32)
        self._keep.append(rn)
        d.nrnodes = rn.size
        d.rnodes = rn.ctypes.data_as(_c.POINTER(_c.c_float))
        o = self.packer.offsets
        d.vsys_count = o['vsys'][1]
        d.off_vsys, d.off_xpos, d.off_ypos = o['vsys'][0], o['xpos'][0], \
            o['ypos'][0]
        d.off_posa, d.off_incl = o['posa'][0], o['incl'][0]
        for i, k in enumerate(_SLOTS):
            bt = BTraits()
            if k in self.packer.trait_ints:
                uids, cc, pc = self.packer.trait_ints[k]
                self._keep.extend([uids, cc, pc])
                bt.n = uids.size
                bt.uids = uids.ctypes.data_as(_c.POINTER(_c.c_int))
                bt.ccounts = cc.ctypes.data_as(_c.POINTER(_c.c_int))
                bt.pcounts = pc.ctypes.data_as(_c.POINTER(_c.c_int))
                bt.off_cvalues, bt.off_pvalues = self.packer.trait_offsets[k]
            d.traits[i] = bt
        d.pstride = self.packer.pstride
        if len(comps) > 1:
            self._extra = (BDesc * (len(comps) - 1))()
            for j in range(1, len(comps)):
                self._fill_disk(self._extra[j - 1], comps[j].disk,
                                self.packer.packers[j],
                                self.packer.bases[j])
            d.nextra = len(comps) - 1
            d.extra = _c.cast(self._extra, _c.c_void_p)
        d.grid = self.grid
        d.size_lo[:] = list(cube.size())
        d.size_hi[:] = list(size_hi)
        d.scale[:] = list(cube.scale())
        d.edge_hi[:] = list(edge_hi)
        d.has_conv = int(bool(cube.psf() or cube.lsf()))
        if d.has_conv:
            self._pcube = driver.mem_copy_h2d(
                cube.make_pcube(size_hi, step_hi, be_fft))
            d.pcube = self._pcube.data_ptr()
        if self.mmaps:
            mm_args = self._mmap_data(dmodel, obs, error, mask, weight)
            obs, error, mask, weight = None, None, None, 1.0
        else:
            obs = np.array(obs, np.float32)
        err = None if error is None or np.isscalar(error) else \
            np.array(error, np.float32)
        if mask is not None:
            # masked pixels contribute a zero residual even where their data
            # or error hold the NaN that marks them (the objective's
            # residual kernel does the same): zero data, unit error there
            mask = np.asarray(mask, np.float32)
            obs[mask == 0] = 0.0
            if err is not None:
                err[mask == 0] = 1.0
        if obs is not None:
            self._obs = driver.mem_copy_h2d(obs)
            d.obs = self._obs.data_ptr()
        if err is None:
            d.err = None
            d.err_scalar = 1.0 if error is None else float(error)
        else:
            self._err = driver.mem_copy_h2d(err)
            d.err = self._err.data_ptr()
            d.err_scalar = 1.0
        if mask is not None:
            self._mask = driver.mem_copy_h2d(mask)
            d.mask = self._mask.data_ptr()
        d.weight = float(weight)
        d.max_batch = int(max_batch)
        paths = dict(auto=0, padded=1, fused=2, direct=3)
        if path not in paths:
            raise ValueError(f"path must be one of {sorted(paths)}")
        d.path = paths[path]
        self._csum_dev = self._qscale_dev = None
        if self.mc:
            csum0, hoi = disk.cloud_budget(template_params)
            self._hoi = driver.mem_copy_h2d(np.ascontiguousarray(hoi,
                                                                 np.uint8))
            d.mcdisk = 1
            d.cflux = disk.cflux
            d.csum_len = int(csum0.size)
            d.hoi = self._hoi.data_ptr()
            d.seed = int(driver.backends().gmodel(np.float32).seed)
            self.disk = disk
        self.desc = d
        self.size_lo = tuple(cube.size())
        self.size_hi = tuple(size_hi)
        h = _c.c_void_p()
        # the batch refers to the context: keep it alive as long as we are
        self._ctx = driver.context()
        N.call('gbkhip_batch_create', self._ctx.handle,
               _c.byref(d), _c.byref(h))
        self.handle = h
        if self.mmaps:
            self._set_mmaps(*mm_args)
        self._mc_cap = 0
        if self.mc:   # capacity for the template's budget, with headroom
            self.reserve_clouds(int(1.5 * max(int(csum0[-1]), 1)))
        self._params_dev = None
        self._chi2_dev = None
        self._cubes_dev = None
        del t, dev

    def _fill_disk(self, e, disk, pk, base):
        """The disk fields of a further component's desc: its sub-ring
        nodes, traits and offsets (shifted by `base`) in the set vector."""
        e.loose, e.tilted = int(disk.loose), int(disk.tilted)
        rn = np.ascontiguousarray(disk.subrnodes, np.float32)
        self._keep.append(rn)
        e.nrnodes = rn.size
        e.rnodes = rn.ctypes.data_as(_c.POINTER(_c.c_float))
        o = pk.offsets
        e.vsys_count = o['vsys'][1]
        e.off_vsys = base + o['vsys'][0]
        e.off_xpos = base + o['xpos'][0]
        e.off_ypos = base + o['ypos'][0]
        e.off_posa = base + o['posa'][0]
        e.off_incl = base + o['incl'][0]
        for i, k in enumerate(_SLOTS):
            bt = BTraits()
            if k in pk.trait_ints:
                uids, cc, pc = pk.trait_ints[k]
                self._keep.extend([uids, cc, pc])
                bt.n = uids.size
                bt.uids = uids.ctypes.data_as(_c.POINTER(_c.c_int))
                bt.ccounts = cc.ctypes.data_as(_c.POINTER(_c.c_int))
                bt.pcounts = pc.ctypes.data_as(_c.POINTER(_c.c_int))
                oc, op = pk.trait_offsets[k]
                bt.off_cvalues, bt.off_pvalues = base + oc, base + op
            e.traits[i] = bt

    def _mmap_data(self, dmodel, obs, error, mask, weight):
        """Stacked (norders, y, x) data, error and mask maps of the
        moment-map objective, in the data model's key order."""
        keys = list(dmodel.keys())
        ny, nx = dmodel.size()[::-1]

        def item(src, key):
            if src is None:
                return None
            if isinstance(src, dict):
                return src.get(key)
            return src[key] if key in src else None

        d_maps, e_maps, m_maps, e_sc, w = [], [], [], [], []
        for key in keys:
            it = item(obs, key)
            if it is None:
                raise RuntimeError(f"no data map for '{key}'")
            if hasattr(it, 'error') and hasattr(it, 'mask'):   # Data item
                dd, ee, mm = it.data(), it.error(), it.mask()
            else:
                dd, ee, mm = it, None, None
            e = error if np.isscalar(error) or error is None \
                else item(error, key)
            if e is None:
                e = ee
            m = mm if mask is None else item(mask, key)
            dd = np.array(dd, np.float32)
            if dd.shape != (ny, nx):
                raise RuntimeError(f"data map '{key}' has shape {dd.shape}, "
                                   f"not {(ny, nx)}")
            if e is None or np.isscalar(e):
                e_sc.append(1.0 if e is None else float(e))
                e_maps.append(None)
            else:
                e_sc.append(1.0)
                e_maps.append(np.array(e, np.float32))
            m_maps.append(None if m is None else np.asarray(m, np.float32))
            d_maps.append(dd)
            w.append(float(weight if np.isscalar(weight)
                           else weight.get(key, 1.0)))
        any_e = any(x is not None for x in e_maps)
        any_m = any(x is not None for x in m_maps)
        D = np.stack(d_maps)
        E = np.stack([x if x is not None else np.full((ny, nx), e_sc[i],
                                                      np.float32)
                      for i, x in enumerate(e_maps)]) if any_e else None
        M = np.stack([x if x is not None else np.ones((ny, nx), np.float32)
                      for x in m_maps]) if any_m else None
        if M is not None:
            # masked pixels contribute a zero residual even where their data
            # or error hold the NaN that marks them
            D[M == 0] = 0.0
            if E is not None:
                E[M == 0] = 1.0
        cube = dmodel.dcube()
        return (np.array(dmodel.orders(), np.int32), D, E,
                np.array(e_sc, np.float32), M, np.array(w, np.float32),
                float(dmodel._mask_cutoff), float(cube.zero()[2]),
                float(cube.step()[2]))

    def _set_mmaps(self, orders, D, E, esc, M, W, cutoff, zero_z, step_z):
        dv = self.driver
        self._mm_obs = dv.mem_copy_h2d(D)
        self._mm_err = None if E is None else dv.mem_copy_h2d(E)
        self._mm_mask = None if M is None else dv.mem_copy_h2d(M)
        self._mm_host = (orders, esc, W)
        N.call('gbkhip_batch_set_mmaps', self.handle, len(orders),
               orders.ctypes.data, _c.c_float(cutoff), _c.c_float(zero_z),
               _c.c_float(step_z), self._mm_obs.data_ptr(),
               None if E is None else self._mm_err.data_ptr(),
               esc.ctypes.data,
               None if M is None else self._mm_mask.data_ptr(),
               W.ctypes.data)

    def upload(self, params_list):
        """Packs parameter dicts into the persistent device buffer (mcdisk:
        with each set's cloud budget and fixed-point scale)."""
        host = self.packer.pack(params_list)
        if self.mc:
            self._upload_clouds(params_list)
        return self.upload_packed(host)

    def reserve_clouds(self, clouds):
        """mcdisk: per-set cloud capacity of the binned deposits (grown,
        never shrunk; growing frees the work buffers and cached graphs)."""
        if clouds > self._mc_cap:
            N.call('gbkhip_batch_reserve_clouds', self.handle,
                   _c.c_longlong(int(clouds)))
            self._mc_cap = int(clouds)

    def cloud_budgets(self, params_list):
        """mcdisk: each set's cumulative cloud counts (int32, nsets x
        csum_len) and fixed-point scale (FP64), as evaluate_clouds takes
        them."""
        n = len(params_list)
        csum = np.zeros((n, self.desc.csum_len), np.int32)
        qs = np.zeros(n, np.float64)
        out = _c.c_double()
        for i, p in enumerate(params_list):
            c, _ = self.disk.cloud_budget(p)
            csum[i] = c
            N.call('gbkhip_mcdisk_fixed_scale', _c.c_float(self.desc.cflux),
                   int(c[-1]), _c.byref(self.grid), _c.byref(out))
            qs[i] = out.value
        if n:
            top = int(csum[:, -1].max())
            if top > self._mc_cap:
                self.reserve_clouds(int(1.25 * top))
        return csum, qs

    def _upload_clouds(self, params_list):
        t = driver_torch()
        n = len(params_list)
        csum, qs = self.cloud_budgets(params_list)
        dev = self.driver.device
        if self._csum_dev is None or self._csum_dev.shape != csum.shape:
            self._csum_dev = t.empty(csum.shape, dtype=t.int32, device=dev)
            self._qscale_dev = t.empty(n, dtype=t.float64, device=dev)
        self._csum_dev.copy_(t.from_numpy(csum))
        self._qscale_dev.copy_(t.from_numpy(qs))

    def upload_packed(self, host):
        t = driver_torch()
        n = host.shape[0]
        if self._params_dev is None or self._params_dev.shape[0] != n:
            self._params_dev = t.empty(host.shape, dtype=t.float32,
                                       device=self.driver.device)
            self._chi2_dev = t.empty(n, dtype=t.float64,
                                     device=self.driver.device)
            self._cubes_dev = None
        self._params_dev.copy_(t.from_numpy(host))
        return n

    def run(self, nsets, want_cubes=False, use_graph=True):
        """Evaluates the uploaded sets (asynchronous on the driver stream)."""
        t = driver_torch()
        if want_cubes and (self._cubes_dev is None or
                           self._cubes_dev.shape[0] != nsets):
            self._cubes_dev = t.empty((nsets,) + self.size_lo[::-1],
                                      dtype=t.float32,
                                      device=self.driver.device)
        # evaluate on torch's current stream, where the uploads and the
        # downloads of evaluate() are ordered: rebind the context to it (its
        # handle is 0 for the default stream, which set_stream reads as "the
        # context's stream")
        self.driver.context()
        self.set_stream(t.cuda.current_stream())
        cubes = self._cubes_dev.data_ptr() if want_cubes else None
        if self.mc:
            N.call('gbkhip_batch_evaluate_clouds', self.handle,
                   self._params_dev.data_ptr(), self._csum_dev.data_ptr(),
                   self._qscale_dev.data_ptr(), int(nsets),
                   self._chi2_dev.data_ptr(), cubes, int(bool(use_graph)))
            return
        N.call('gbkhip_batch_evaluate', self.handle,
               self._params_dev.data_ptr(), int(nsets),
               self._chi2_dev.data_ptr(), cubes, int(bool(use_graph)))

    def check(self):
        """Raises when an evaluation since the last check had an mcdisk set
        with more clouds than the reserved capacity (waits for the batch's
        stream)."""
        v = _c.c_int()
        N.call('gbkhip_batch_check', self.handle, _c.byref(v))
        if v.value:
            raise N.GbkHipError(
                1, 'gbkhip_batch_check', "an mcdisk cloud budget exceeded "
                "the reserved capacity (reserve_clouds); the results of the "
                "evaluations since the last check are not valid")

    def evaluate(self, params_list, want_cubes=False, use_graph=True):
        """chi^2 per set (numpy FP64), and the model cubes if asked."""
        n = self.upload(params_list)
        self.run(n, want_cubes, use_graph)
        chi2 = self._chi2_dev.cpu().numpy().copy()
        if self.mc:
            self.check()
        if want_cubes:
            return chi2, self._cubes_dev.cpu().numpy().copy()
        return chi2

    def info(self):
        """dict(pipeline='padded'|'fused'|'direct', fft_size=(x, y, z),
        nparts=int): the batched pipeline in use and its transform sizes."""
        v = (_c.c_int * 5)()
        N.call('gbkhip_batch_info', self.handle, v)
        pipe = ('padded', 'fused', 'direct')[v[0]]
        return dict(pipeline=pipe, fused=pipe == 'fused',
                    fft_size=(v[1], v[2], v[3]), nparts=v[4])

    def taps(self):
        """Direct pipeline: (taps per output along x, y, z; inputs skipped
        before the first tap along x, y, z)."""
        v = (_c.c_int * 6)()
        N.call('gbkhip_batch_taps', self.handle, v)
        return tuple(v[:3]), tuple(v[3:])

    def set_tiles(self, lanes=2, ysplit=1):
        """Direct pipeline tiles: chunks in flight (1 or 2) and the split of
        a y-pass column strip's outputs over blocks (1, 2 or 4)."""
        N.call('gbkhip_batch_set_tiles', self.handle, int(lanes), int(ysplit))
        self.tiles = (int(lanes), int(ysplit))

    def tile_stats(self):
        """Direct pipeline, last chunk on the first lane: (live, total)
        x-pass tiles, y-pass output groups and z-pass blocks; the others
        are all-zero and skipped."""
        v = (_c.c_longlong * 6)()
        N.call('gbkhip_batch_tile_stats', self.handle, v)
        return (v[0], v[1]), (v[2], v[3]), (v[4], v[5])

    def line_stats(self):
        """Direct pipeline of a thin disk, last chunk on the first lane:
        (live, total) hi-res spaxels in the 64-spaxel row segments some line
        reaches (the only line parameters written and read)."""
        v = (_c.c_longlong * 2)()
        N.call('gbkhip_batch_line_stats', self.handle, v)
        return v[0], v[1]

    def set_timing(self, enable):
        N.call('gbkhip_batch_set_timing', self.handle, int(bool(enable)))

    def stage_times(self):
        """(smdisk, convolution, downscale+chi2) milliseconds summed over
        the chunks of the last evaluation, and the chunk count."""
        ms = (_c.c_double * 3)()
        n = _c.c_int()
        N.call('gbkhip_batch_stage_times', self.handle, ms, _c.byref(n))
        return tuple(ms), n.value

    def set_stream(self, stream):
        """Stream of later evaluations (a torch.cuda.Stream, a raw handle or
        None for the context's stream)."""
        h = getattr(stream, 'cuda_stream', stream)
        N.call('gbkhip_batch_set_stream', self.handle,
               _c.c_void_p(int(h) if h else 0))

    def release(self):
        """Frees the work buffers; the next evaluation allocates them."""
        N.call('gbkhip_batch_release', self.handle)

    def arena_bytes(self):
        """Bytes of a work arena this batch can run with (bind_work)."""
        v = _c.c_ulonglong()
        N.call('gbkhip_batch_arena_bytes', self.handle, _c.byref(v))
        return v.value

    def bind_work(self, arena):
        """Runs the next evaluations with their work buffers in `arena` (a
        device tensor of at least arena_bytes() bytes that no other batch
        uses at the same time), or in buffers of their own (None).  Graphs
        captured with an arena stay cached for it."""
        if arena is None:
            N.call('gbkhip_batch_bind_work', self.handle, None, 0)
            return
        nbytes = arena.numel() * arena.element_size()
        N.call('gbkhip_batch_bind_work', self.handle, arena.data_ptr(),
               _c.c_ulonglong(nbytes))

    def memory(self):
        """dict(resident=bytes, work=bytes, allocated=bool)."""
        r, w, a = _c.c_ulonglong(), _c.c_ulonglong(), _c.c_int()
        N.call('gbkhip_batch_memory', self.handle, _c.byref(r), _c.byref(w),
               _c.byref(a))
        return dict(resident=r.value, work=w.value, allocated=bool(a.value))

    FUSED_KERNELS = ('batch_lines_thin', 'fz_deposit_r2c', 'fx_lines (y)',
                     'fx_rows_conv', 'fx_lines (inverse y)',
                     'fz_c2r_down_chi2', 'batch_chi2_final')
    PADDED_STAGES = ('smdisk', 'fft_convolution', 'downscale_chi2')
    DIRECT_KERNELS = ('batch_lines_thin', 'sep_xpass', 'sep_ypass',
                      'sep_zpass', 'batch_chi2_final')
    DIRECT_CUBE_KERNELS = ('model_deposit', 'sep_xpass_cube', 'sep_ypass',
                           'sep_zpass', 'batch_chi2_final')

    def kernel_times(self):
        """{kernel: milliseconds} of the last timed evaluation, summed over
        its chunks (stages on the padded-cube path)."""
        ms = (_c.c_double * 8)()
        n = _c.c_int()
        N.call('gbkhip_batch_kernel_times', self.handle, ms, 8, _c.byref(n))
        direct = (self.DIRECT_CUBE_KERNELS if self.cube_source
                  else self.DIRECT_KERNELS)
        names = {7: self.FUSED_KERNELS, 5: direct}.get(n.value,
                                                       self.PADDED_STAGES)
        return {k: ms[i] for i, k in enumerate(names[:n.value])}

    def close(self):
        if getattr(self, 'handle', None) and not N.shutting_down():
            N.lib().gbkhip_batch_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def driver_torch():
    import torch
    return torch


__all__ = ['BatchEvaluator', 'SetPacker', 'NativeMemoryHip']


# ---------------------------------------------------------------------------
# Autotuner and ragged groups
# ---------------------------------------------------------------------------
_TUNE_CACHE = {}


def _tune_key(dmodel, gmodel, mc):
    cube = dmodel.dcube()
    comp = gmodel.components[0]
    return (tuple(cube.size()), tuple(cube.scale()), tuple(cube.step()),
            repr(cube.psf()), repr(cube.lsf()), cube._padding, bool(mc),
            bool(getattr(comp, 'thin', True)))


def autotune(driver, dmodel, gmodel, obs, sample_params, error=None,
             mask=None, weight=1.0, max_batches=(16, 32, 64, 128, 256),
             reps=3,
             budget_bytes=None, template_params=None, cache=True):
    """Chooses the batched pipeline (direct / fused / padded), the chunk size
    and the direct pipeline's tiles (lanes, ysplit) for one galaxy model by
    timing each candidate on `sample_params` (graph replays after a warm-up),
    and returns (BatchEvaluator of the fastest, report).  Candidates whose
    work buffers exceed budget_bytes are skipped.  Choices are cached per
    geometry / kernel / model kind, so galaxies of one shape tune once."""
    import time
    t = driver_torch()
    mc = type(gmodel.components[0].disk).__name__ == 'MCDisk'
    key = _tune_key(dmodel, gmodel, mc)
    n = len(sample_params)
    tp = template_params if template_params is not None else sample_params[0]

    def make(path, mb):
        return BatchEvaluator(driver, dmodel, gmodel, obs, error=error,
                              mask=mask, weight=weight, max_batch=mb,
                              template_params=tp, path=path)

    if cache and key in _TUNE_CACHE:
        path, mb, tiles = _TUNE_CACHE[key]['choice']
        ev = make(path, min(mb, max(1, n)))
        if ev.info()['pipeline'] == 'direct':
            ev.set_tiles(*tiles)
        return ev, dict(_TUNE_CACHE[key], cached=True)

    def timed(ev):
        ev.evaluate(sample_params)               # warm-up + graph capture
        t.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            ev.evaluate(sample_params)
        t.cuda.synchronize()
        return (time.perf_counter() - t0) / (reps * n)

    mbs = sorted({min(max(1, n), m) for m in max_batches})
    trials, best = [], None            # best: (s per set, path, mb, tiles)
    for path in (('direct', 'padded') if mc
                 else ('direct', 'fused', 'padded')):
        try:
            probe = make(path, 1)
        except N.GbkHipError:
            continue                           # pipeline does not apply
        per_set = probe.memory()['work']
        probe.close()
        for mb in mbs:
            if budget_bytes is not None and per_set * mb > budget_bytes:
                continue
            ev = make(path, mb)
            try:
                direct = ev.info()['pipeline'] == 'direct'
                for tiles in ([(2, 1), (1, 1), (2, 2), (2, 4)] if direct
                              else [(1, 1)]):
                    if direct:
                        ev.set_tiles(*tiles)
                    sec = timed(ev)
                    trials.append(dict(path=path, max_batch=mb, tiles=tiles,
                                       s_per_set=sec))
                    if best is None or sec < best[0]:
                        best = (sec, path, mb, tiles)
            finally:
                ev.close()
    if best is None:
        raise RuntimeError("no batched pipeline fits the memory budget")
    ev = make(best[1], best[2])
    if ev.info()['pipeline'] == 'direct':
        ev.set_tiles(*best[3])
    report = dict(choice=(best[1], best[2], best[3]), s_per_set=best[0],
                  trials=trials)
    if cache:
        _TUNE_CACHE[key] = report
    return ev, dict(report, cached=False)


class RaggedGroup:
    """Galaxies of different cube sizes, models (smdisk / mcdisk) and data,
    each a BatchEvaluator (typically from autotune), evaluated together in
    one native call (gbkhip_group_evaluate): one HIP graph launch per set of
    buffers, members spread over parallel streams.  Each member's chi^2
    equals its own evaluate()."""

    def __init__(self, evaluators):
        self.members = list(evaluators)
        if not self.members:
            raise ValueError("empty group")
        drv = self.members[0].driver
        self.driver = drv
        self._ctx = drv.context()
        arr = (_c.c_void_p * len(self.members))(
            *[m.handle.value for m in self.members])
        h = _c.c_void_p()
        N.call('gbkhip_group_create', self._ctx.handle, arr,
               len(self.members), _c.byref(h))
        self.handle = h

    def evaluate(self, params_lists, want_cubes=False, use_graph=True,
                 download_cubes=True):
        """params_lists[i]: the parameter dicts of member i.  Returns the
        list of chi^2 arrays (and of model cube arrays if asked; with
        download_cubes False the cubes stay on the device, in each member's
        cube buffer, and None is returned for them)."""
        t = driver_torch()
        n = len(self.members)
        if len(params_lists) != n:
            raise ValueError("one parameter list per member")
        P = (_c.c_void_p * n)()
        C = (_c.c_void_p * n)()
        Q = (_c.c_void_p * n)()
        S = (_c.c_int * n)()
        X = (_c.c_void_p * n)()
        K = (_c.c_void_p * n)()
        any_mc = False
        for i, (ev, pl) in enumerate(zip(self.members, params_lists)):
            k = ev.upload(pl) if len(pl) else 0
            S[i] = k
            if not k:
                continue
            P[i] = ev._params_dev.data_ptr()
            X[i] = ev._chi2_dev.data_ptr()
            if ev.mc:
                any_mc = True
                C[i] = ev._csum_dev.data_ptr()
                Q[i] = ev._qscale_dev.data_ptr()
            if want_cubes:
                if ev._cubes_dev is None or ev._cubes_dev.shape[0] != k:
                    ev._cubes_dev = t.empty((k,) + ev.size_lo[::-1],
                                            dtype=t.float32,
                                            device=ev.driver.device)
                K[i] = ev._cubes_dev.data_ptr()
        # the group launches on the context stream: bind it to torch's
        # current stream, where the uploads above are ordered
        self.driver.context()
        N.call('gbkhip_group_evaluate', self.handle, P,
               C if any_mc else None, Q if any_mc else None, S, X,
               K if want_cubes else None, int(bool(use_graph)))
        chi2 = [ev._chi2_dev.cpu().numpy().copy() if S[i] else
                np.zeros(0) for i, ev in enumerate(self.members)]
        for ev in self.members:
            if ev.mc:
                ev.check()
        if want_cubes:
            cubes = [ev._cubes_dev.cpu().numpy().copy()
                     if S[i] and download_cubes else None
                     for i, ev in enumerate(self.members)]
            return chi2, cubes
        return chi2

    def close(self):
        if getattr(self, 'handle', None):
            N.call('gbkhip_group_destroy', self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
