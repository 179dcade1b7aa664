"""Disk traits: the radial / vertical profiles a disk component is built from.

Mirror of src/gbkfit/model/gmodels/traits.py: the same trait kinds, type
names, UIDs (the wire format the native kernels dispatch on, traits.py:16-148
= core/gmodel/traits.hpp:7-135), parameter names and order (smooth parameters
first, then node-wise ones), and constants (``consts()`` is the tuple of the
constructor's stored options, traits.py:299-326).
"""
import numpy as np

# kind -> type -> (uid, smooth parameter names, node-wise parameter names)
_P = {
    'uniform': (1, ('a',), ()),
    'exponential': (2, ('a', 's'), ()),
    'gauss': (3, ('a', 's'), ()),
    'ggauss': (4, ('a', 's', 'b'), ()),
    'lorentz': (5, ('a', 's'), ()),
    'moffat': (6, ('a', 's', 'b'), ()),
    'sech2': (7, ('a', 's'), ()),
}
_MIX = {
    'mixture_exponential': (51, ('r', 't', 'a', 's', 'q', 'p')),
    'mixture_gauss': (52, ('r', 't', 'a', 's', 'q', 'p')),
    'mixture_ggauss': (53, ('r', 't', 'a', 's', 'b', 'q', 'p')),
    'mixture_moffat': (54, ('r', 't', 'a', 's', 'b', 'q', 'p')),
}
_H1 = {'uniform': 1, 'exponential': 2, 'gauss': 3, 'lorentz': 5, 'sech2': 7}
_H2 = {'ggauss': 4, 'moffat': 6}
_VP = {
    'tan_uniform': (1, ('vt',)), 'tan_arctan': (2, ('rt', 'vt')),
    'tan_boissier': (3, ('rt', 'vt')), 'tan_epinat': (4, ('rt', 'vt', 'a', 'g')),
    'tan_lramp': (5, ('rt', 'vt')), 'tan_tanh': (6, ('rt', 'vt')),
    'tan_polyex': (7, ('rt', 'vt', 'a')), 'tan_rix': (8, ('rt', 'vt', 'b', 'g')),
}
_VP_NW = {'nw_tan_uniform': (101, 'vt'), 'nw_rad_uniform': (103, 'vr'),
          'nw_ver_uniform': (105, 'vv'), 'nw_los_uniform': (107, 'vl')}
_VP_NWH = {'nw_tan_harmonic': 102, 'nw_rad_harmonic': 104,
           'nw_ver_harmonic': 106, 'nw_los_harmonic': 108}

KINDS = ('bpt', 'bht', 'vpt', 'vht', 'dpt', 'dht', 'zpt', 'spt', 'wpt',
         'opt', 'oht')


class Trait:
    """A configured trait.  ``params_sm`` are (name, size) pairs of smooth
    parameters; ``params_nw`` names node-wise parameters (one value per
    ring node, interpolated onto sub-ring nodes when packed)."""

    def __init__(self, kind, type_, uid, consts=(), params_sm=(),
                 params_nw=(), options=None):
        self.kind = kind
        self._type = type_
        self._uid = uid
        self._consts = tuple(float(c) for c in consts)
        self.params_sm = tuple(params_sm)
        self.params_nw = tuple(params_nw)
        self.options = dict(options or {})

    def type(self):
        return self._type

    def uid(self):
        return self._uid

    def consts(self):
        return self._consts

    def has_analytical_integral(self):
        """Density traits integrated as one whole-disk value
        (traits.py:500-501,1559-1560): only the uniform profiles."""
        return self._type == 'uniform' and self.kind in ('bpt', 'opt')

    def integrate(self, params, rings):
        """Flux per ring (or total) for the mcdisk cloud budget
        (traits.py:188-275)."""
        t = self._type
        rings = np.asarray(rings, np.float64)
        rsep = rings[1] - rings[0]
        if t == 'uniform':
            rmin = rings[0] - 0.5 * rsep
            rmax = rings[-1] + 0.5 * rsep
            return np.pi * params['a'] * (rmax * rmax - rmin * rmin)
        a = params.get('a')
        s = params.get('s')
        b = params.get('b')
        if t == 'exponential':
            amp = (rings >= 0) * (a * np.exp(-np.abs(rings) / s))
        elif t == 'gauss':
            amp = a * np.exp(-rings * rings / (2 * s * s))
        elif t == 'ggauss':
            amp = a * np.exp(-np.power(np.abs(rings) / s, b))
        elif t == 'lorentz':
            amp = a * s * s / (rings * rings + s * s)
        elif t == 'moffat':
            amp = a / np.power(1 + (rings / s) ** 2, b)
        elif t == 'sech2':
            amp = a / np.cosh(rings / s) ** 2
        elif t in ('nw_uniform', 'nw_harmonic'):
            amp = np.asarray(params['a'], np.float64)
        else:
            raise NotImplementedError(
                f"trait '{t}' has no integral (not supported by mcdisk)")
        return amp * np.pi * ((rings + rsep / 2) ** 2 - (rings - rsep / 2) ** 2)

    def __repr__(self):
        return f"Trait({self.kind}:{self._type}, uid={self._uid})"


def _polar(kind, type_, offset=0, **opts):
    if type_ in _P:
        uid, sm, _ = _P[type_]
        return Trait(kind, type_, uid + offset, (), [(n, 1) for n in sm])
    if type_ in _MIX:
        uid, names = _MIX[type_]
        nb = int(opts['nblobs'])
        return Trait(kind, type_, uid + offset, (nb,), [(n, nb) for n in names],
                     options=dict(nblobs=nb))
    if type_ == 'nw_uniform':
        return Trait(kind, type_, 101 + offset, (), (), ('a',))
    if type_ == 'nw_harmonic':
        k = int(opts['order'])
        return Trait(kind, type_, 102 + offset, (k,), (),
                     ('a', 'p') if k > 0 else ('a',), dict(order=k))
    if type_ == 'nw_distortion':
        return Trait(kind, type_, 103 + offset, (), (), ('a', 'p', 's'))
    raise RuntimeError(f"unknown {kind} trait type '{type_}'")


def make(kind, type_, **opts):
    """Builds a trait from its kind and type name, e.g.
    ``make('bpt', 'exponential')``, ``make('bht', 'sech2', rnodes=False)``."""
    if kind in ('bpt', 'dpt'):
        return _polar(kind, type_, 0, **opts)
    if kind == 'opt':
        return _polar(kind, type_, 1000, **opts)
    if kind in ('bht', 'oht'):
        rnodes = bool(opts.get('rnodes', False))
        trunc = float(opts.get('trunc', 0))
        if kind == 'bht':
            if type_ in _H1:
                uid, names = _H1[type_], ('s',)
            elif type_ in _H2:
                uid, names = _H2[type_], ('s', 'b')
            else:
                raise RuntimeError(f"unknown bht trait type '{type_}'")
        else:
            if type_ in _H1:
                uid, names = 1000 + _H1[type_], ('a', 's')
            elif type_ in _H2:
                uid, names = 1000 + _H2[type_], ('a', 's', 'b')
            else:
                raise RuntimeError(f"unknown oht trait type '{type_}'")
        # consts = (trunc, rnodes): traits.py:328-420,803-838
        sm = () if rnodes else [(n, 1) for n in names]
        nwp = names if rnodes else ()
        return Trait(kind, type_, uid, (trunc, float(rnodes)), sm, nwp,
                     dict(rnodes=rnodes, trunc=trunc))
    if kind == 'vpt':
        if type_ in _VP:
            uid, names = _VP[type_]
            return Trait(kind, type_, uid, (), [(n, 1) for n in names])
        if type_ in _VP_NW:
            uid, name = _VP_NW[type_]
            return Trait(kind, type_, uid, (), (), (name,))
        if type_ in _VP_NWH:
            k = int(opts['order'])
            return Trait(kind, type_, _VP_NWH[type_], (k,), (),
                         ('a', 'p') if k > 0 else ('a',), dict(order=k))
        raise RuntimeError(f"unknown vpt trait type '{type_}'")
    if kind in ('vht', 'dht'):
        if type_ != 'one':
            raise RuntimeError(f"unknown {kind} trait type '{type_}'")
        return Trait(kind, 'one', 1, (0.0,))
    if kind == 'zpt':
        if type_ == 'nw_uniform':
            return Trait(kind, type_, 101, (), (), ('a',))
        if type_ == 'nw_harmonic':
            k = int(opts['order'])
            return Trait(kind, type_, 102, (k,), (),
                         ('a', 'p') if k > 0 else ('a',), dict(order=k))
        raise RuntimeError(f"unknown zpt trait type '{type_}'")
    if kind == 'spt':
        if type_ == 'azrange':
            return Trait(kind, type_, 1, (), [('p', 1), ('s', 1)])
        if type_ == 'nw_azrange':
            return Trait(kind, type_, 101, (), (), ('p', 's'))
        raise RuntimeError(f"unknown spt trait type '{type_}'")
    if kind == 'wpt':
        if type_ == 'axis_range':
            return Trait(kind, type_, 1, (float(opts['axis']),
                                          float(opts['angle']),
                                          float(opts['weight'])))
        raise RuntimeError(f"unknown wpt trait type '{type_}'")
    raise RuntimeError(f"unknown trait kind '{kind}'")


def load(kind, info):
    """Builds a trait from a ``{'type': ..., **options}`` dict."""
    info = dict(info)
    return make(kind, info.pop('type'), **info)
