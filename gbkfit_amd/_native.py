"""ctypes binding of libgbkhip.so (C ABI declared in include/gbkhip.h).

The library is built in-tree by ``gbkfit_amd/csrc/build.py`` (or
``__graft_entry__.build()``).  Loading fails loudly when it is missing: there
is no CPU fallback behind this package.
"""
import atexit
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libgbkhip.so')

_c = ctypes
_fp = _c.POINTER(_c.c_float)
_ip = _c.POINTER(_c.c_int)
_vp = _c.c_void_p


class GbkHipError(RuntimeError):
    """A libgbkhip call returned an error status."""

    def __init__(self, code, func, msg):
        super().__init__(f"{func} failed (code {code}): {msg}")
        self.code = code


class Traits(_c.Structure):
    _fields_ = [('n', _c.c_int), ('uids', _vp), ('cvalues', _vp),
                ('ccounts', _vp), ('pvalues', _vp), ('pcounts', _vp)]


TRAIT_KINDS = ('rp', 'rh', 'vp', 'vh', 'dp', 'dh', 'zp', 'sp', 'wp')


class Disk(_c.Structure):
    _fields_ = ([('loose', _c.c_int), ('tilted', _c.c_int),
                 ('nrnodes', _c.c_int), ('rnodes', _vp), ('vsys', _vp),
                 ('xpos', _vp), ('ypos', _vp), ('posa', _vp), ('incl', _vp)]
                + [(k, Traits) for k in TRAIT_KINDS] + [('opacity', _vp)])


class Grid(_c.Structure):
    _fields_ = [('spat_size', _c.c_int * 3), ('spat_step', _c.c_float * 3),
                ('spat_zero', _c.c_float * 3), ('spec_size', _c.c_int),
                ('spec_step', _c.c_float), ('spec_zero', _c.c_float)]


OUTPUT_KEYS = ('image', 'scube', 'wdata', 'wdata_cmp', 'rdata', 'rdata_cmp',
               'ordata', 'ordata_cmp', 'vdata_cmp', 'ddata_cmp')


class Outputs(_c.Structure):
    _fields_ = [(k, _vp) for k in OUTPUT_KEYS]


def _i3(v):
    return (_c.c_int * 3)(*[int(x) for x in v])


def _f3(v):
    return (_c.c_float * 3)(*[float(x) for x in v])


_SIGS = {
    'gbkhip_version': ([], _c.c_int),
    'gbkhip_last_error': ([], _c.c_char_p),
    'gbkhip_ctx_create': ([_c.c_int, _c.POINTER(_vp)], _c.c_int),
    'gbkhip_ctx_destroy': ([_vp], _c.c_int),
    'gbkhip_ctx_set_stream': ([_vp, _vp], _c.c_int),
    'gbkhip_ctx_synchronize': ([_vp], _c.c_int),
    'gbkhip_fft_r2c': ([_vp, _ip, _vp, _vp], _c.c_int),
    'gbkhip_fft_c2r': ([_vp, _ip, _vp, _vp], _c.c_int),
    'gbkhip_fft_convolve': ([_vp, _ip, _vp, _vp, _vp], _c.c_int),
    'gbkhip_fft_convolve_cached': ([_vp, _ip, _vp, _vp, _c.c_uint64],
                                   _c.c_int),
    'gbkhip_fft_clear_cache': ([_vp], _c.c_int),
    'gbkhip_dcube_downscale': ([_vp, _ip, _ip, _ip, _ip, _vp, _vp], _c.c_int),
    'gbkhip_dcube_mask': ([_vp, _c.c_float, _c.c_int, _ip, _vp, _vp, _vp],
                          _c.c_int),
    'gbkhip_mmaps_moments': ([_vp, _ip, _fp, _fp, _vp, _vp, _c.c_float,
                              _c.c_int, _vp, _vp, _vp, _vp], _c.c_int),
    'gbkhip_mmaps_moments_batched': ([_vp, _ip, _fp, _fp, _vp, _c.c_float,
                                      _c.c_int, _vp, _vp, _vp, _c.c_int],
                                     _c.c_int),
    'gbkhip_smdisk_evaluate': ([_vp, _c.POINTER(Disk), _c.POINTER(Grid),
                                _c.POINTER(Outputs)], _c.c_int),
    'gbkhip_mcdisk_evaluate': ([_vp, _c.c_float, _c.c_int, _vp, _c.c_int, _vp,
                                _c.POINTER(Disk), _c.POINTER(Grid),
                                _c.POINTER(Outputs), _c.c_uint64], _c.c_int),
    'gbkhip_mcdisk_evaluate_noalloc': ([_vp, _c.c_float, _c.c_int, _vp,
                                        _c.c_int, _vp, _c.POINTER(Disk),
                                        _c.POINTER(Grid), _c.POINTER(Outputs),
                                        _c.c_uint64], _c.c_int),
    'gbkhip_wcube_evaluate': ([_vp, _ip, _c.c_int, _vp, _vp], _c.c_int),
    'gbkhip_convolve': ([_vp, _ip, _ip, _vp, _vp, _vp], _c.c_int),
    'gbkhip_residual': ([_vp, _vp, _vp, _vp, _vp, _vp, _vp, _c.c_longlong,
                         _c.c_float, _vp], _c.c_int),
    'gbkhip_residual_sum': ([_vp, _vp, _c.c_longlong, _c.c_int, _vp],
                            _c.c_int),
    'gbkhip_count_pixels': ([_vp, _vp, _vp, _c.c_longlong, _c.c_float, _vp],
                            _c.c_int),
}

_LIB = None
_SHUTTING_DOWN = [False]


@atexit.register
def _mark_shutdown():
    # Native handles are not released during interpreter shutdown: the
    # order in which Python finalises objects then is arbitrary, and the
    # process is about to release every device resource anyway.
    _SHUTTING_DOWN[0] = True


def shutting_down():
    return _SHUTTING_DOWN[0]


def lib():
    """The loaded library; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libgbkhip.so not found at {LIB_PATH}; build it with "
                f"`python gbkfit_amd/csrc/build.py`")
        L = _c.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        for name, (args, res) in _EXTRA_SIGS.items():
            if hasattr(L, name):
                f = getattr(L, name)
                f.argtypes = args
                f.restype = res
        _LIB = L
    return _LIB


_EXTRA_SIGS = {}


def register_signatures(sigs):
    """Adds signatures of further entry points (batched API)."""
    _EXTRA_SIGS.update(sigs)
    if _LIB is not None:
        for name, (args, res) in sigs.items():
            if hasattr(_LIB, name):
                f = getattr(_LIB, name)
                f.argtypes = args
                f.restype = res


def call(name, *args):
    """Calls an entry point and raises GbkHipError on a non-zero status."""
    f = getattr(lib(), name)
    rc = f(*args)
    if rc != 0:
        msg = lib().gbkhip_last_error()
        raise GbkHipError(rc, name, msg.decode() if msg else '')
    return rc


def exported_symbols():
    """Names declared in include/gbkhip.h that the library must export."""
    return sorted(_SIGS) + sorted(_EXTRA_SIGS)


class Context:
    """Owns a gbkhip_ctx on one device; work goes to the given stream."""

    def __init__(self, device=0, stream=None):
        h = _vp()
        call('gbkhip_ctx_create', int(device), _c.byref(h))
        self.handle = h
        self.device = device
        if stream is not None:
            self.set_stream(stream)

    def set_stream(self, stream):
        call('gbkhip_ctx_set_stream', self.handle, _vp(int(stream or 0)))

    def synchronize(self):
        call('gbkhip_ctx_synchronize', self.handle)

    def close(self):
        if self.handle and not shutting_down():
            lib().gbkhip_ctx_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
