"""Python access to the CPU oracle (oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the CPU baseline,
never as the thing measured or shipped.

Parameter order per object is the order of each reference object's params
table (objects/<name>.cl); defaults are the sky gradient's (sky.cl:6-7).
"""
import ctypes
import os
import subprocess
from ctypes import POINTER, c_char_p, c_double, c_float, c_int

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")

PARAMS = {
    "sis": ["x", "y", "r"],
    "sis_plus_shear": ["x", "y", "r", "g1", "g2"],
    "nsis": ["x", "y", "r", "rc"],
    "point_mass": ["x", "y", "r"],
    "sie": ["x", "y", "r", "q", "pa"],
    "sie_plus_shear": ["x", "y", "r", "q", "pa", "g1", "g2"],
    "nsie": ["x", "y", "r", "rc", "q", "pa"],
    "epl": ["x", "y", "r", "t", "q", "pa"],
    "epl_plus_shear": ["x", "y", "r", "t", "q", "pa", "g1", "g2"],
    "sersic": ["x", "y", "r", "mag", "n", "q", "pa"],
    "sersic-old": ["x", "y", "r", "mag", "n", "q", "pa"],
    "gauss": ["x", "y", "sigma", "mag", "q", "pa"],
    "exponential": ["x", "y", "rs", "mag", "q", "pa"],
    "devauc": ["x", "y", "r", "mag", "q", "pa"],
    "sky": ["bg", "dx", "dy"],
}
DEFAULTS = {("sky", "dx"): 0.0, ("sky", "dy"): 0.0}

_libs = {}


def build():
    """Compile the oracle libraries (make in oracle/)."""
    subprocess.check_call(["make", "-s", "-C", HERE], stdout=subprocess.DEVNULL)


def _load(name):
    if name not in _libs:
        path = os.path.join(BUILD, name)
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        dp, fp, ip = POINTER(c_double), POINTER(c_float), POINTER(c_int)
        lib.orc_kind.argtypes = [c_char_p]
        lib.orc_kind.restype = c_int
        for sfx, rp in (("_d", dp), ("_f", fp)):
            getattr(lib, "orc_render" + sfx).argtypes = [c_int, ip, ip, rp, c_int, c_int, rp, c_int, rp, rp, rp, rp]
            getattr(lib, "orc_convolve" + sfx).argtypes = [c_int, c_int, rp, c_int, c_int, rp, rp]
            getattr(lib, "orc_deflect" + sfx).argtypes = [c_int, ip, rp, c_int, c_int, rp, rp]
            getattr(lib, "orc_epl_omega" + sfx).argtypes = [c_int, rp, rp, rp, c_int, rp]
        lib.orc_chi2_f32terms.argtypes = [c_int, fp, fp, fp]
        lib.orc_chi2_f32terms.restype = c_double
        lib.orc_chi2_d.argtypes = [c_int, dp, dp, dp]
        lib.orc_chi2_d.restype = c_double
        lib.orc_set_threads.argtypes = [c_int]
        lib.orc_set_threads.restype = c_int
        lib.orc_sersic_b.argtypes = [c_double]
        lib.orc_sersic_b.restype = c_double
        lib.orc_gauss_quantile.argtypes = [c_double]
        lib.orc_gauss_quantile.restype = c_double
        _libs[name] = lib
    return _libs[name]


def lib(fast=False):
    return _load("liboracle_fast.so" if fast else "liboracle.so")


def _ptr(a, t):
    return a.ctypes.data_as(POINTER(t))


def param_vector(objects, values):
    """Object-order parameter vector from {"id.param": value} (defaults filled in)."""
    out = []
    for oid, name in objects:
        for p in PARAMS[name]:
            key = f"{oid}.{p}"
            if key in values:
                out.append(float(values[key]))
            elif (name, p) in DEFAULTS:
                out.append(DEFAULTS[(name, p)])
            else:
                raise KeyError(key)
    return np.array(out)


def kinds(objects):
    k = [lib().orc_kind(name.encode()) for _, name in objects]
    if min(k) < 0:
        raise KeyError(objects)
    return np.array(k, dtype=np.int32)


def render(objects, params, shape, qq, ww, pcs=(1.0, 1.0, 1.0, 1.0), dtype=np.float64, ipp=None, fast=False):
    """Quadrature render (value, error) of kernel/lensed.cl:9-38 on the CPU."""
    h, w = shape
    rt = c_double if dtype == np.float64 else c_float
    sfx = "_d" if dtype == np.float64 else "_f"
    k = kinds(objects)
    ip = np.zeros(len(objects), np.int32) if ipp is None else np.asarray(ipp, np.int32)
    p = np.ascontiguousarray(params, dtype=dtype)
    pc = np.ascontiguousarray(pcs, dtype=dtype)
    q = np.ascontiguousarray(qq, dtype=dtype).ravel()
    wt = np.ascontiguousarray(ww, dtype=dtype).ravel()
    value = np.zeros(h * w, dtype=dtype)
    error = np.zeros(h * w, dtype=dtype)
    getattr(lib(fast), "orc_render" + sfx)(len(objects), _ptr(k, c_int), _ptr(ip, c_int), _ptr(p, rt), w, h,
                                           _ptr(pc, rt), len(q) // 2, _ptr(q, rt), _ptr(wt, rt),
                                           _ptr(value, rt), _ptr(error, rt))
    return value.reshape(h, w), error.reshape(h, w)


def convolve(image, psf, dtype=np.float64, fast=False):
    """Direct clamp-to-edge convolution of kernel/lensed.cl:56-103."""
    rt = c_double if dtype == np.float64 else c_float
    sfx = "_d" if dtype == np.float64 else "_f"
    a = np.ascontiguousarray(image, dtype=dtype)
    p = np.ascontiguousarray(psf, dtype=dtype)
    out = np.zeros_like(a)
    getattr(lib(fast), "orc_convolve" + sfx)(a.shape[1], a.shape[0], _ptr(a, rt), p.shape[1], p.shape[0],
                                             _ptr(p, rt), _ptr(out, rt))
    return out


def deflect(objects, params, which, xy, dtype=np.float64):
    rt = c_double if dtype == np.float64 else c_float
    sfx = "_d" if dtype == np.float64 else "_f"
    k = kinds(objects)
    p = np.ascontiguousarray(params, dtype=dtype)
    pts = np.ascontiguousarray(xy, dtype=dtype).reshape(-1, 2)
    out = np.zeros_like(pts)
    getattr(lib(), "orc_deflect" + sfx)(len(objects), _ptr(k, c_int), _ptr(p, rt), which, len(pts),
                                         _ptr(pts, rt), _ptr(out, rt))
    return out


def epl_omega(q, t, phi, nterms=10, dtype=np.float64):
    """Angular factor of the EPL series (epl.cl:47-61) at (q, t, phi)."""
    rt = c_double if dtype == np.float64 else c_float
    sfx = "_d" if dtype == np.float64 else "_f"
    q, t, phi = [np.ascontiguousarray(np.broadcast_to(v, np.broadcast(q, t, phi).shape), dtype=dtype).ravel()
                 for v in (q, t, phi)]
    out = np.zeros((len(q), 2), dtype=dtype)
    getattr(lib(), "orc_epl_omega" + sfx)(len(q), _ptr(q, rt), _ptr(t, rt), _ptr(phi, rt), nterms, _ptr(out, rt))
    return out


def chi2_f32terms(image, weight, model):
    a = [np.ascontiguousarray(v, dtype=np.float32).ravel() for v in (image, weight, model)]
    return lib().orc_chi2_f32terms(len(a[0]), *[_ptr(v, c_float) for v in a])


def chi2_d(image, weight, model):
    a = [np.ascontiguousarray(v, dtype=np.float64).ravel() for v in (image, weight, model)]
    return lib().orc_chi2_d(len(a[0]), *[_ptr(v, c_double) for v in a])


def set_threads(n, fast=True):
    """OpenMP threads of the oracle's pixel loops; returns the count in use."""
    return lib(fast).orc_set_threads(int(n))


def sersic_b(n):
    """b(n) of the Sersic profile (objects/sersic.cl:32-33), FP64."""
    return lib().orc_sersic_b(float(n))


def gauss_quantile(u):
    return lib().orc_gauss_quantile(float(u))


def loglike(objects, params, image, weight, qq, ww, psf=None, pcs=(1.0, 1.0, 1.0, 1.0), dtype=np.float64):
    """Full forward model -> logL = -chi^2/2, in `dtype` throughout."""
    pcs = list(pcs)
    if psf is not None:
        if psf.shape[1] % 2 == 0:
            pcs[0] += 0.5
        if psf.shape[0] % 2 == 0:
            pcs[1] += 0.5
    value, _ = render(objects, params, image.shape, qq, ww, pcs, dtype)
    model = convolve(value, psf, dtype) if psf is not None else value
    if dtype == np.float32:
        return -0.5 * chi2_f32terms(image, weight, model), model
    return -0.5 * chi2_d(image, weight, model), model
