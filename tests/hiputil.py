"""Test helpers: move an oracle.DiskSpec to device tensors and call the hip
backends with the reference's positional signatures."""
import numpy as np

import oracle

KINDS = oracle.TRAIT_KINDS


def driver():
    from gbkfit_amd.driver.hip import DriverHip
    return DriverHip()


def dev(drv, a, dtype=np.float32):
    if a is None:
        return None
    return drv.mem_copy_h2d(np.ascontiguousarray(a, dtype=dtype))


def disk_args(drv, spec):
    """(loose, tilted, rnodes, vsys, xpos, ypos, posa, incl, 45 trait arrays,
    opacity) as device tensors."""
    args = [spec.loose, spec.tilted, dev(drv, spec.sub), dev(drv, spec.vsys),
            dev(drv, spec.xpos), dev(drv, spec.ypos), dev(drv, spec.posa),
            dev(drv, spec.incl)]
    for k in KINDS:
        tr = spec.traits[k]
        if tr is None:
            args += [None] * 5
        else:
            uids, cv, cc, pv, pc = tr
            args += [dev(drv, uids, np.int32), dev(drv, cv), dev(drv, cc,
                     np.int32), dev(drv, pv), dev(drv, pc, np.int32)]
    args.append(dev(drv, getattr(spec, 'opacity', None)))   # opacity
    return args


def grid_args(grid):
    return ((grid.spat_size[0], grid.spat_size[1], grid.spat_size[2]),
            (grid.spat_step[0], grid.spat_step[1], grid.spat_step[2]),
            (grid.spat_zero[0], grid.spat_zero[1], grid.spat_zero[2]),
            0.0, grid.spec_size, grid.spec_step, grid.spec_zero)


OUT_ORDER = ('image', 'scube', 'wdata', 'wdata_cmp', 'rdata', 'rdata_cmp',
             'ordata', 'ordata_cmp', 'vdata_cmp', 'ddata_cmp')


def alloc_outputs(drv, grid, want, fill=None):
    nx, ny, nz = grid.spat_size
    shapes = dict(image=(ny, nx), scube=(grid.spec_size, ny, nx))
    out = {}
    for k in want:
        t = drv.mem_alloc_d(shapes.get(k, (nz, ny, nx)), np.float32)
        f = (fill or {}).get(k, 0.0)
        drv.mem_fill(t, f)
        out[k] = t
    return out


def smdisk(drv, spec, grid, want=('scube',), fill=None):
    backend = drv.backends().gmodel(np.float32)
    outs = alloc_outputs(drv, grid, want, fill)
    args = disk_args(drv, spec) + list(grid_args(grid)) + \
        [outs.get(k) for k in OUT_ORDER]
    backend.smdisk_evaluate(*args)
    return {k: v.cpu().numpy().astype(np.float64) for k, v in outs.items()}


def mcdisk(drv, spec, grid, cflux, csum, hoi, seed, want=('scube',)):
    backend = drv.backends().gmodel(np.float32)
    backend.seed = seed
    outs = alloc_outputs(drv, grid, want)
    csum_d = dev(drv, csum, np.int32)
    hoi_d = dev(drv, np.asarray(hoi, np.uint8), np.uint8)
    args = [cflux, int(csum[-1]), csum_d, hoi_d] + disk_args(drv, spec) + \
        list(grid_args(grid)) + [outs.get(k) for k in OUT_ORDER]
    backend.mcdisk_evaluate(*args)
    return {k: v.cpu().numpy().astype(np.float64) for k, v in outs.items()}


def normwise(a, ref):
    """max|a - ref| / max|ref| over finite entries of ref."""
    cplx = np.iscomplexobj(ref) or np.iscomplexobj(a)
    dt = np.complex128 if cplx else np.float64
    ref = np.asarray(ref, dt)
    a = np.asarray(a, dt)
    m = np.isfinite(ref)
    scale = np.max(np.abs(ref[m])) if m.any() else 0.0
    if scale == 0:
        return float(np.max(np.abs(a[m]))) if m.any() else 0.0
    return float(np.max(np.abs(a[m] - ref[m])) / scale)


def within(a, ref, tol, amb=None):
    """|a - ref| <= tol * max|ref| + amb, element-wise; returns the number of
    violations and the normwise error outside ambiguous voxels."""
    ref = np.asarray(ref, np.float64)
    a = np.asarray(a, np.float64)
    scale = np.max(np.abs(ref)) if ref.size else 0.0
    bound = tol * scale + (0 if amb is None else amb)
    bad = np.abs(a - ref) > bound
    return int(bad.sum()), normwise(a, ref)
