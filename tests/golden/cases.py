"""Case definitions shared by make_golden.py (which writes the fixtures from
the reference's own core, compiled by oracle/Makefile) and the tests that
read them.  Every trait uid the native smdisk/mcdisk signatures accept
appears in at least one case; at most four traits per kind (the reference's
TRAIT_NUM_MAX, core/constants.hpp:8)."""
import numpy as np

from oracle import DCubeGeom, DiskSpec, Trait as T, grid_for

RN = np.arange(0.0, 6.5, 1.0)           # 7 ring nodes, 0 .. 6
NN = RN.size


def nw(a, b):
    """A node-wise parameter varying linearly from a to b."""
    return np.linspace(a, b, NN)


ONE = T(1, consts=(0,))                 # VHTraitOne / DHTraitOne
GEOM = DCubeGeom((8, 8, 16), (1, 1, 10), scale=(2, 2, 1), psf_sigma=0.5,
                 lsf_sigma=5.0)             # padded hi-res grid 32 x 32 x 32


def _thin(name, traits, **kw):
    return dict(name=name, thin=True, traits=traits, kw=kw)


def _thick(name, traits, size_z=16, **kw):
    t = dict(traits)
    for p, h in (('vp', 'vh'), ('dp', 'dh')):
        if t.get(p):
            t.setdefault(h, [ONE] * len(t[p]))
    return dict(name=name, thin=False, traits=t, kw=kw, size_z=size_z)


SMDISK_CASES = [
    _thin('thin_basic', dict(
        rp=[T(2, params=(1.0, 4.0))],
        vp=[T(2, params=(3.0, 200.0))],
        dp=[T(1, params=(20.0,))]), posa=30, incl=45, vsys=5.0),
    _thick('thick_sech2', dict(
        rp=[T(2, params=(1.0, 4.0))],
        rh=[T(7, consts=(0, 0), params=(1.0,))],
        vp=[T(5, params=(3.0, 220.0))],
        dp=[T(1, params=(15.0,))]), posa=120, incl=60),
    _thin('thin_profiles_a', dict(
        rp=[T(1, params=(0.3,)), T(3, params=(1.0, 3.0)),
            T(4, params=(0.8, 3.0, 1.5)), T(5, params=(0.5, 2.0))],
        vp=[T(3, params=(2.0, 150.0)), T(4, params=(2.0, 180.0, 2.0, 0.5)),
            T(6, params=(2.5, 120.0)), T(1, params=(30.0,))],
        dp=[T(2, params=(10.0, 5.0)), T(3, params=(8.0, 6.0))]),
        posa=200, incl=35),
    _thin('thin_profiles_b', dict(
        rp=[T(6, params=(1.0, 3.0, 2.0)), T(7, params=(0.7, 4.0))],
        vp=[T(7, params=(3.0, 100.0, 0.1)), T(8, params=(3.0, 150.0, 0.5, 2.0))],
        dp=[T(4, params=(10.0, 4.0, 1.5)), T(5, params=(9.0, 3.0)),
            T(6, params=(8.0, 3.0, 1.5)), T(7, params=(7.0, 5.0))]),
        posa=75, incl=70, vsys=-10.0),
    _thin('thin_nodewise', dict(
        rp=[T(101, nwparams=(nw(1.0, 0.2),)),
            T(102, consts=(2,), nwparams=(nw(0.3, 0.1), nw(10.0, 40.0))),
            T(103, nwparams=(nw(0.5, 0.2), nw(20.0, 60.0), nw(1.0, 3.0)))],
        vp=[T(101, nwparams=(nw(50.0, 200.0),)),
            T(102, consts=(1,), nwparams=(nw(10.0, 20.0), nw(0.0, 30.0))),
            T(103, nwparams=(nw(5.0, 15.0),)),
            T(104, consts=(2,), nwparams=(nw(3.0, 6.0), nw(10.0, 20.0)))],
        dp=[T(101, nwparams=(nw(30.0, 10.0),)),
            T(102, consts=(0,), nwparams=(nw(5.0, 2.0),)),
            T(103, nwparams=(nw(3.0, 1.0), nw(0.0, 90.0), nw(2.0, 4.0)))],
        sp=[T(1, params=(10.0, 300.0))]),
        posa=310, incl=50, vsys=3.0),
    _thick('thick_heights', dict(
        rp=[T(2, params=(1.0, 4.0)), T(3, params=(0.5, 3.0)),
            T(1, params=(0.05,)), T(5, params=(0.3, 2.0))],
        rh=[T(1, consts=(0, 0), params=(1.5,)),
            T(2, consts=(3, 1), nwparams=(nw(0.5, 1.5),)),
            T(3, consts=(2, 0), params=(1.0,)),
            T(5, consts=(0, 1), nwparams=(nw(0.8, 1.2),))],
        vp=[T(105, nwparams=(nw(5.0, 10.0),)),
            T(106, consts=(1,), nwparams=(nw(3.0, 6.0), nw(0.0, 40.0))),
            T(107, nwparams=(nw(2.0, 4.0),)),
            T(108, consts=(2,), nwparams=(nw(1.0, 2.0), nw(5.0, 10.0)))],
        dp=[T(1, params=(12.0,))],
        zp=[T(101, nwparams=(nw(0.0, 0.5),)),
            T(102, consts=(1,), nwparams=(nw(0.1, 0.3), nw(0.0, 90.0)))],
        sp=[T(101, nwparams=(nw(0.0, 40.0), nw(300.0, 200.0)))]),
        posa=20, incl=65),
    _thick('thick_heights_b', dict(
        rp=[T(2, params=(1.0, 4.0)), T(7, params=(0.6, 3.0)),
            T(6, params=(0.4, 3.0, 2.5))],
        rh=[T(4, consts=(0, 0), params=(1.0, 1.5)),
            T(7, consts=(2, 0), params=(1.2,)),
            T(6, consts=(0, 0), params=(1.0, 2.0))],
        vp=[T(2, params=(2.0, 210.0))],
        dp=[T(1, params=(25.0,))]),
        posa=160, incl=40),
    _thin('thin_mixtures', dict(
        rp=[T(51, consts=(2,), params=(2.0, 5.0, 30.0, 90.0, 1.0, 0.8,
                                       2.0, 1.5, 0.7, 0.5, 10.0, 20.0)),
            T(52, consts=(1,), params=(4.0, 180.0, 0.6, 1.0, 0.6, 45.0)),
            T(53, consts=(1,), params=(3.0, 270.0, 0.5, 1.2, 1.8, 0.9, 0.0)),
            T(54, consts=(1,), params=(5.0, 0.0, 0.4, 1.5, 2.2, 0.7, 30.0))],
        vp=[T(2, params=(3.0, 180.0))],
        dp=[T(52, consts=(1,), params=(1.0, 0.0, 10.0, 3.0, 0.8, 0.0)),
            T(1, params=(5.0,))]),
        posa=0, incl=30),
    _thin('thin_loose_tilted', dict(
        rp=[T(2, params=(1.0, 4.0))],
        vp=[T(2, params=(3.0, 200.0))],
        dp=[T(1, params=(20.0,))]),
        loose=True, tilted=True, xpos=nw(0.0, 0.8), ypos=nw(0.2, -0.3),
        posa=nw(30.0, 50.0), incl=nw(45.0, 55.0), vsys=nw(0.0, 0.0)),
    _thick('thick_loose_tilted', dict(
        rp=[T(2, params=(1.0, 4.0))],
        rh=[T(7, consts=(0, 0), params=(1.0,))],
        vp=[T(2, params=(3.0, 200.0))],
        dp=[T(1, params=(20.0,))]),
        loose=True, tilted=True, xpos=nw(-0.3, 0.3), ypos=nw(0.0, 0.5),
        posa=nw(100.0, 90.0), incl=nw(60.0, 50.0), vsys=nw(10.0, 10.0)),
]

def opacity_cube(case):
    """A seeded spatial opacity cube (nz, ny, nx) for a case's grid: values
    in [0, 0.25) over the disk's extent, zero elsewhere."""
    g = build_grid(case)
    nx, ny, nz = g.spat_size
    rng = np.random.default_rng(2024)
    op = rng.uniform(0.0, 0.25, (nz, ny, nx))
    op[:, :, :nx // 8] = 0.0
    return op


# Opacity (a spatial cube that dims every voxel by the voxels in front of
# it, gmodels.hpp:903-914; thick disks only, thin ones have one voxel per
# column) and weights (wp traits, wdata outputs, gmodels.hpp:52-82,
# traits.hpp:1859-1885; the only weight trait, axis_range, is 1).
EXTRA_CASES = [
    _thick('thick_opacity', dict(
        rp=[T(2, params=(1.0, 4.0))],
        rh=[T(7, consts=(0, 0), params=(1.0,))],
        vp=[T(2, params=(3.0, 200.0))],
        dp=[T(1, params=(20.0,))]), posa=60, incl=55),
    _thin('thin_weights', dict(
        rp=[T(2, params=(1.0, 4.0))],
        vp=[T(2, params=(3.0, 200.0))],
        dp=[T(1, params=(20.0,))],
        wp=[T(1, consts=(1, 30.0, 0.5))]), posa=30, incl=45),
    _thick('thick_weights', dict(
        rp=[T(2, params=(1.0, 4.0))],
        rh=[T(7, consts=(0, 0), params=(1.0,))],
        vp=[T(2, params=(3.0, 200.0))],
        dp=[T(1, params=(20.0,))],
        wp=[T(1, consts=(0, 10.0, 2.0))]), posa=120, incl=60),
    # an opacity disk (opacity polar traits uid 1000 + n, opacity height
    # traits amplitude x shape, traits.hpp:1437-1514), its density into
    # rdata (opacity_smdisk_3d.py:116-127); the second height trait is
    # truncated at 2 scales
    _thick('opacity_disk', dict(
        rp=[T(1002, params=(0.1, 4.0)), T(1003, params=(0.05, 3.0))],
        rh=[T(1007, consts=(0, 0), params=(1.0, 1.0)),
            T(1003, consts=(2.0, 0), params=(0.5, 0.8))]),
        posa=75, incl=65),
    # brightness-only disks (no kinematic traits) into an image
    # (brightness_smdisk_2d.py / _3d.py: spectral size 1)
    _thin('bright_thin', dict(
        rp=[T(2, params=(1.0, 4.0)), T(3, params=(0.5, 2.0))]),
        posa=210, incl=40),
    _thick('bright_thick', dict(
        rp=[T(3, params=(1.0, 3.0))],
        rh=[T(2, consts=(0, 0), params=(0.7,))]), posa=300, incl=70),
]
EXTRA_WANT = {
    'thick_opacity': ('scube', 'image', 'rdata', 'rdata_cmp', 'ordata',
                      'ordata_cmp'),
    'thin_weights': ('scube', 'wdata', 'wdata_cmp'),
    'thick_weights': ('scube', 'image', 'wdata', 'wdata_cmp'),
    'opacity_disk': ('rdata', 'rdata_cmp'),
    'bright_thin': ('image',),
    'bright_thick': ('image', 'rdata', 'ordata'),
}


# mcdisk cases: thick disks whose density traits are sampled per ring.
MCDISK_CASES = [
    dict(name='mc_sech2', traits=dict(
        rp=[T(2, params=(1.0, 4.0))],
        rh=[T(7, consts=(0, 0), params=(1.0,))],
        vp=[T(5, params=(3.0, 220.0))], vh=[ONE],
        dp=[T(1, params=(15.0,))], dh=[ONE]), posa=30, incl=50,
        nclouds=60000, size_z=16),
    dict(name='mc_uniform_gauss', traits=dict(
        rp=[T(1, params=(0.2,)), T(3, params=(1.0, 3.0))],
        rh=[T(3, consts=(0, 0), params=(1.0,)),
            T(2, consts=(0, 0), params=(0.8,))],
        vp=[T(2, params=(2.0, 180.0))], vh=[ONE],
        dp=[T(1, params=(20.0,))], dh=[ONE]), posa=140, incl=35,
        nclouds=60000, size_z=16),
]


def build_spec(case):
    kw = dict(case['kw']) if 'kw' in case else {
        k: case[k] for k in ('posa', 'incl') if k in case}
    if case.get('name') == 'thick_opacity':
        kw['opacity'] = opacity_cube(case)
    return DiskSpec(RN, case['traits'], **kw)


def build_grid(case):
    return grid_for(GEOM, thin=case.get('thin', False),
                    size_z=case.get('size_z'))


def ring_integrals(spec, tr):
    """Cloud budget per ring for one density trait (gmodels/traits.py:188-275,
    _mcdisk.py:180-200): analytical traits give one value, others one value
    per ring centred on the inner sub-nodes."""
    rings = spec.sub[1:-1]
    uid = tr.uid
    q = tr.params
    rsep = rings[1] - rings[0]
    if uid == 1:   # uniform: analytical integral over the whole disk
        rmin = rings[0] - 0.5 * rsep
        rmax = rings[-1] + 0.5 * rsep
        return np.atleast_1d(np.pi * q[0] * (rmax * rmax - rmin * rmin)), True
    amp = {2: lambda r: q[0] * np.exp(-np.abs(r) / q[1]) * (r >= 0),
           3: lambda r: q[0] * np.exp(-r * r / (2 * q[1] * q[1]))}[uid](rings)
    area = np.pi * ((rings + rsep / 2) ** 2 - (rings - rsep / 2) ** 2)
    return amp * area, False


def cloud_budget(spec, traits, nclouds_target):
    """cflux, per-ring counts cumsum and has-analytical-integral flags."""
    ints, flags = [], []
    for tr in traits['rp']:
        v, h = ring_integrals(spec, tr)
        ints.append(v)
        flags.append(h)
    total = sum(float(np.sum(v)) for v in ints)
    cflux = total / nclouds_target
    counts = np.concatenate([(v / cflux).astype(np.int32) for v in ints])
    return cflux, np.cumsum(counts).astype(np.int32), np.array(flags, np.uint8)
