"""Per-pixel cubature rules (src/quadrature.c:21-43) from the C library."""
from ctypes import POINTER, c_double, c_float

import numpy as np

from . import _lib as L


def names():
    out, i = [], 0
    while True:
        n = L.load().lensed_quad_name(i)
        if not n:
            return out
        out.append(n.decode())
        i += 1


def table(rule="g3k7"):
    """Unscaled rule on the unit pixel, double: (pts [n, 2], wht [n], err [n])."""
    lib = L.load()
    n = lib.lensed_quad_size(rule.encode())
    if n < 0:
        raise KeyError(rule)
    pts, wht, err = np.zeros((n, 2)), np.zeros(n), np.zeros(n)
    lib.lensed_quad_table(rule.encode(), pts.ctypes.data_as(POINTER(c_double)),
                          wht.ctypes.data_as(POINTER(c_double)), err.ctypes.data_as(POINTER(c_double)))
    return pts, wht, err


def rule(name="g3k7", sx=1.0, sy=1.0):
    """Nodes scaled by the pixel scale and weights, float32 as the device uses them."""
    lib = L.load()
    n = lib.lensed_quad_size(name.encode())
    if n < 0:
        raise KeyError(name)
    qq, ww = np.zeros((n, 2), np.float32), np.zeros((n, 2), np.float32)
    lib.lensed_quad_rule(name.encode(), qq.ctypes.data_as(POINTER(c_float)), ww.ctypes.data_as(POINTER(c_float)),
                         sx, sy)
    return qq, ww
