"""The C ABI library loads and exports every entry point include/gbkhip.h
declares (no device calls: this runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, 'include')):
        if fn.endswith('.h'):
            txt = open(os.path.join(ROOT, 'include', fn)).read()
            names |= set(re.findall(r'\b(gbkhip_[a-z0-9_]+)\s*\(', txt))
    return names


def test_library_exports_every_declared_symbol():
    from gbkfit_amd import _native
    lib = _native.lib()
    names = declared()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_declarations():
    from gbkfit_amd import _native
    import gbkfit_amd.batch  # noqa: F401  (registers the batch entry points)
    _native.lib()
    bound = set(_native.exported_symbols())
    assert declared() <= bound, sorted(declared() - bound)


def test_version_and_error_string():
    from gbkfit_amd import _native
    lib = _native.lib()
    assert lib.gbkhip_version() >= 10000
    # a call with a null context fails cleanly with a message
    rc = lib.gbkhip_ctx_synchronize(None)
    assert rc != 0
    assert b'context' in lib.gbkhip_last_error()


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    from gbkfit_amd import _native
    monkeypatch.setattr(_native, 'LIB_PATH', str(tmp_path / 'nope.so'))
    monkeypatch.setattr(_native, '_LIB', None)
    try:
        _native.lib()
    except RuntimeError as e:
        assert 'not found' in str(e)
    else:
        raise AssertionError('expected a RuntimeError')
