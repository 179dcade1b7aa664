"""Driver interface (mirror of src/gbkfit/driver/core.py:26-89).

A driver owns memory allocation, host<->device copies, a few elementwise math
helpers and the factory of the four native backends (FFT, DModel, GModel,
Objective).  Drivers register under their ``type()`` string.
"""
import abc


class Driver(abc.ABC):

    @staticmethod
    @abc.abstractmethod
    def type():
        pass

    @classmethod
    def load(cls, info, *args, **kwargs):
        return cls()

    def dump(self):
        return dict(type=self.type())

    @abc.abstractmethod
    def mem_alloc_s(self, shape, dtype):
        """Returns a (host, device) pair of arrays of the same shape."""

    @abc.abstractmethod
    def mem_alloc_h(self, shape, dtype):
        pass

    @abc.abstractmethod
    def mem_alloc_d(self, shape, dtype):
        pass

    @abc.abstractmethod
    def mem_copy_h2d(self, h_src, d_dst=None):
        pass

    @abc.abstractmethod
    def mem_copy_d2h(self, d_src, h_dst=None):
        pass

    @abc.abstractmethod
    def mem_fill(self, x, value):
        pass

    @abc.abstractmethod
    def math_abs(self, x, out=None):
        pass

    @abc.abstractmethod
    def math_sum(self, x, out=None):
        pass

    @abc.abstractmethod
    def math_add(self, x1, x2, out=None):
        pass

    @abc.abstractmethod
    def math_sub(self, x1, x2, out=None):
        pass

    @abc.abstractmethod
    def math_mul(self, x1, x2, out=None):
        pass

    @abc.abstractmethod
    def math_div(self, x1, x2, out=None):
        pass

    @abc.abstractmethod
    def math_pow(self, x1, x2, out=None):
        pass

    @abc.abstractmethod
    def backends(self):
        pass


_REGISTRY = {}


def register_driver(cls):
    """Registers a Driver subclass under its type() (unique, like
    src/gbkfit/utils/parseutils.py:375-379)."""
    t = cls.type()
    if t in _REGISTRY and _REGISTRY[t] is not cls:
        raise RuntimeError(f"driver type '{t}' already registered")
    _REGISTRY[t] = cls
    return cls


def load_driver(info):
    """Builds a driver from a ``{'type': ...}`` dict (the config form
    ``driver: {type: hip}``)."""
    t = info.get('type') if isinstance(info, dict) else info
    if t not in _REGISTRY:
        raise RuntimeError(
            f"unknown driver type '{t}'; registered: {sorted(_REGISTRY)}")
    return _REGISTRY[t].load(info)
