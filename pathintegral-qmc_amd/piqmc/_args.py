"""Argument conversion mirroring the Cython typed-memoryview unpacking of the
reference (piqmc/qmc.c:2649-2655): float64 buffers of a fixed rank, writable
where the function mutates them, C ints and a C float for the scalars."""
import numpy as np


def f64_view(obj, ndim, name, writable=False):
    if not isinstance(obj, np.ndarray):
        try:
            memoryview(obj)
        except TypeError:
            raise TypeError("a bytes-like object is required, not '%s'" % type(obj).__name__)
        obj = np.asarray(obj)
    if obj.dtype != np.float64:
        raise ValueError("Buffer dtype mismatch, expected 'float_t' but got '%s' (argument %s)"
                         % (obj.dtype, name))
    if obj.ndim != ndim:
        raise ValueError("Buffer has wrong number of dimensions (expected %d, got %d)"
                         % (ndim, obj.ndim))
    if writable and not obj.flags.writeable:
        raise ValueError("buffer source array is read-only")
    return obj


def c_int(v):
    return int(v)


def c_float(v):
    return float(np.float32(float(v)))


def elem_strides(a):
    return tuple(s // a.itemsize for s in a.strides)
