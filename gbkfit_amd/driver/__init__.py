"""Driver layer: gbkfit's Driver / DriverBackends interface and the 'hip'
driver (src/gbkfit/driver/core.py, backend.py, drivers/*)."""
from .backend import (DriverBackendDModel, DriverBackendFFT,  # noqa: F401
                      DriverBackendGModel, DriverBackendObjective,
                      DriverBackends)
from .core import Driver, load_driver, register_driver  # noqa: F401


def DriverHip(*args, **kwargs):
    """Lazily imports the hip driver (it needs torch)."""
    from .hip import DriverHip as _D
    return _D(*args, **kwargs)


def _register():
    try:
        from . import hip  # noqa: F401
    except Exception as e:  # pragma: no cover - torch missing
        import logging
        logging.getLogger(__name__).warning(
            "hip driver not registered: %s", e)


_register()
