"""Observation data and output layers over the C ABI (liblensed_amd.so).

Lensed reads its image, weights, mask and PSF from FITS files
(src/lensed.c:406-549, src/data.c) and writes the dumper's layers IMG, RES,
RAW, ERR, WHT and PVL to `<root>.fits` (src/nested.c:133-306).  The library
reads and writes the FITS subset those files use; this module wraps it.
"""
from ctypes import POINTER, byref, c_char_p, c_double, c_float, c_size_t

import numpy as np

from . import _lib as L

LAYERS = ("IMG", "RES", "RAW", "ERR", "WHT", "PVL")


class DataError(RuntimeError):
    pass


def read_fits(spec):
    """Image of a FITS specification (`file.fits`, `file.fits[IMG]`,
    `file.fits[x1:x2,y1:y2]`, ...) -> (float32 array [height, width],
    pixel coordinate system (rx, ry, sx, sy))."""
    lib = L.load()
    w, h = c_size_t(), c_size_t()
    ptr = POINTER(c_float)()
    pcs = (c_double * 4)()
    if lib.lensed_fits_read(str(spec).encode(), byref(w), byref(h), byref(ptr), pcs) != 0:
        raise DataError(L.last_error())
    try:
        arr = np.ctypeslib.as_array(ptr, shape=(h.value, w.value)).copy() if w.value * h.value else \
            np.zeros((h.value, w.value), np.float32)
    finally:
        lib.lensed_fits_free(ptr)
    return arr, tuple(pcs)


def write_fits(path, layers):
    """layers: dict name -> 2-D array (all the same shape), or a list of
    (name, array); written as float IMAGE extensions in that order."""
    items = list(layers.items()) if isinstance(layers, dict) else list(layers)
    arrs = [np.ascontiguousarray(a, dtype=np.float32) for _, a in items]
    h, w = arrs[0].shape
    if any(a.shape != (h, w) for a in arrs):
        raise DataError("all layers must have the same shape")
    ptrs = (POINTER(c_float) * len(arrs))(*[a.ctypes.data_as(POINTER(c_float)) for a in arrs])
    names = (c_char_p * len(arrs))(*[str(n).encode() for n, _ in items])
    if L.load().lensed_fits_write(str(path).encode(), w, h, len(arrs), ptrs, names) != 0:
        raise DataError(L.last_error())


def find_mode(values, mask=None):
    """(mode, fwhm) of the pixel values, skipping pixels where mask == 0."""
    v = np.ascontiguousarray(values, dtype=np.float32).ravel()
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.float32).ravel()
    mode, fwhm = c_double(), c_double()
    L.load().lensed_find_mode(v.size, v.ctypes.data_as(POINTER(c_float)),
                              None if m is None else m.ctypes.data_as(POINTER(c_float)), byref(mode), byref(fwhm))
    return mode.value, fwhm.value


class Observation:
    """The prepared inputs of one system: image, weight, PSF, pixel
    coordinate system, and the checks Lensed reports."""

    def __init__(self, image, weight=None, gain=None, offset=0.0, bscale=0.0, xweight=None, mask=None, psf=None):
        def opt(v):
            return None if v is None else str(v).encode()
        o = L.DataOptions(image=opt(image), weight=opt(weight), gain=opt(gain), offset=float(offset),
                          bscale=float(bscale), xweight=opt(xweight), mask=opt(mask), psf=opt(psf))
        d = L.Data()
        if L.load().lensed_data_read(byref(o), byref(d)) != 0:
            raise DataError(L.last_error())
        self._take(d)

    @classmethod
    def _from_struct(cls, d):
        obs = cls.__new__(cls)
        obs._take(d)
        return obs

    def _take(self, d):
        """Copy a lensed_data struct and free it."""
        lib = L.load()
        try:
            shape = (d.height, d.width)
            self.image = np.ctypeslib.as_array(d.image, shape=shape).copy()
            self.weight = np.ctypeslib.as_array(d.weight, shape=shape).copy()
            self.psf = np.ctypeslib.as_array(d.psf, shape=(d.psfh, d.psfw)).copy() if d.psf else None
            self.pcs = tuple(d.pcs)
            self.masked = d.masked
            self.mode, self.fwhm = d.mode, d.fwhm
            self.warnings = d.warnings
        finally:
            lib.lensed_data_free(byref(d))

    @property
    def warning_names(self):
        out = []
        if self.warnings & L.LENSED_WARN_SKY_NOT_ZERO:
            out.append("background sky is not zero")
        if self.warnings & L.LENSED_WARN_NO_OFFSET:
            out.append("background sky is zero but no offset was given")
        if self.warnings & L.LENSED_WARN_NEGATIVE_WEIGHT:
            out.append("negative values in weights")
        return out
