"""Compile the generated model code (lensed_model_source) with the host C++
compiler and render on the CPU, to check plugin translation and the
generated compute()/set_params() against the oracle without a GPU."""
import ctypes
import hashlib
import os
import subprocess
import tempfile
from ctypes import POINTER, byref, c_float, c_int

import numpy as np

from lensed_amd import _lib as L

HARNESS = r'''
#include <cstring>
extern "C" int host_render(const float* params, int W, int H, const float* pcs, int nq,
                           const float* qq, const float* ww, float* value, float* error) {
    static lcl::uint data[4096];
    memset(data, 0, sizeof data);
    lcl::set_params_0(data, params);
    for (int k = 0; k < W * H; ++k) {
        float cx = pcs[0] + pcs[2] * (float)(k % W), cy = pcs[1] + pcs[3] * (float)(k / W);
        float f = 0, e = 0;
        for (int n = 0; n < nq; ++n) {
            float v = lcl::compute_0(data, lcl::float2(cx + qq[2 * n], cy + qq[2 * n + 1]), nullptr);
            f += ww[2 * n] * v;
            e += ww[2 * n + 1] * v;
        }
        value[k] = f;
        error[k] = e;
    }
    return 0;
}
'''

_cache = {}


def load_objects(objects, ipp=()):
    lib = L.load()
    arr = (L.Object * len(objects))()
    for i, (oid, name) in enumerate(objects):
        if lib.lensed_object_load(byref(arr[i]), oid.encode(), name.encode()) != 0:
            raise RuntimeError(L.last_error())
        if oid in ipp:
            arr[i].pars[0].ipp = 1
            arr[i].pars[1].ipp = 1
    return arr


def compile_model(objects, ipp=()):
    arr = load_objects(objects, ipp)
    ptr = L.load().lensed_model_source(len(objects), arr)
    if not ptr:
        raise RuntimeError(L.last_error())
    src = L.take_string(ptr) + HARNESS
    key = hashlib.sha1(src.encode()).hexdigest()[:16]
    if key not in _cache:
        d = tempfile.mkdtemp(prefix="lensed_host_")
        cpp, so = os.path.join(d, "m.cpp"), os.path.join(d, "m.so")
        open(cpp, "w").write(src)
        subprocess.check_call(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-w", "-o", so, cpp])
        lib = ctypes.CDLL(so)
        fp = POINTER(c_float)
        lib.host_render.argtypes = [fp, c_int, c_int, fp, c_int, fp, fp, fp, fp]
        _cache[key] = lib
    return _cache[key]


def render(objects, params, shape, qq, ww, pcs=(1.0, 1.0, 1.0, 1.0), ipp=()):
    lib = compile_model(objects, ipp)
    h, w = shape
    fp = POINTER(c_float)
    p = np.ascontiguousarray(params, np.float32)
    pc = np.ascontiguousarray(pcs, np.float32)
    q = np.ascontiguousarray(qq, np.float32).ravel()
    wt = np.ascontiguousarray(ww, np.float32).ravel()
    value = np.zeros(h * w, np.float32)
    error = np.zeros(h * w, np.float32)
    lib.host_render(p.ctypes.data_as(fp), w, h, pc.ctypes.data_as(fp), len(q) // 2, q.ctypes.data_as(fp),
                    wt.ctypes.data_as(fp), value.ctypes.data_as(fp), error.ctypes.data_as(fp))
    return value.reshape(h, w), error.reshape(h, w)
