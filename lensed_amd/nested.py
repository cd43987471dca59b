"""Nested sampling over the C ABI: the caller of the likelihood path.

Lensed runs MultiNest on its `loglike` callback (src/lensed.c:1236-1288).
liblensed_amd.so carries a nested sampler behind the same `run()` entry
point, plus a batched form that proposes many candidates per likelihood call
so one launch evaluates them all (`run_context`).  This module exposes
both to Python; the sampling itself is native code.
"""
import ctypes
from ctypes import POINTER, byref, c_double, c_int

import numpy as np

from . import _lib as L


class NestedResult:
    """Evidence and posterior summary of one run (parameters in sampler order)."""

    def __init__(self, res, mean, sigma, ml):
        self.logz = res.logz
        self.logz_err = res.logz_err
        self.info = res.info
        self.max_loglike = res.max_loglike
        self.nevals = res.nevals
        self.niter = res.niter
        self.mean = np.array(mean)
        self.sigma = np.array(sigma)
        self.ml = np.array(ml)

    def __repr__(self):
        return (f"NestedResult(logz={self.logz:.4f}+-{self.logz_err:.4f}, niter={self.niter}, "
                f"nevals={self.nevals})")


def options(nlive=300, tol=0.1, efr=0.8, ceff=False, batch=1, seed=0, updint=0, maxiter=0, resume=False,
            outfile=False, wrap=None, root=None, feedback=False):
    o = L.NsOptions(nlive=nlive, tol=tol, efr=efr, ceff=int(ceff), batch=batch, seed=seed, updint=updint,
                    maxiter=maxiter, resume=int(resume), outfile=int(outfile), feedback=int(feedback))
    keep = []
    if wrap is not None:
        w = (c_int * len(wrap))(*[int(bool(v)) for v in wrap])
        o.wrap = ctypes.cast(w, POINTER(c_int))
        keep.append(w)
    if root is not None:
        r = str(root).encode()
        o.root = r
        keep.append(r)
    return o, keep


def sample(loglike, ndim, npar=None, **kw):
    """Nested sampling of a Python batch likelihood.

    loglike(cubes[n, ndim]) -> (logl[n], phys[n, npar]) or logl[n]; the
    sampler calls it with `batch` candidates at a time.
    """
    npar = ndim if npar is None else npar
    lib = L.load()

    def cb(cubes, n, phys, lnew, _ctx):
        u = np.ctypeslib.as_array(cubes, shape=(n, ndim)).copy()
        out = loglike(u)
        if isinstance(out, tuple):
            ll, ph = out
        else:
            ll, ph = out, u
        np.ctypeslib.as_array(lnew, shape=(n,))[:] = ll
        np.ctypeslib.as_array(phys, shape=(n, npar))[:] = np.asarray(ph).reshape(n, npar)

    fn = L.BATCH_FN(cb)
    o, keep = options(**kw)
    mean, sigma, ml = (c_double * npar)(), (c_double * npar)(), (c_double * npar)()
    res = L.NsResult()
    rc = lib.lensed_nested_sample(ndim, npar, fn, None, byref(o), mean, sigma, ml, byref(res))
    del keep
    if rc != 0:
        raise RuntimeError(L.last_error())
    return NestedResult(res, mean, sigma, ml)


def run_context(ctx, s=0, **kw):
    """Batched nested sampling of system s of a lensed_amd.lensed.Context."""
    lib = L.load()
    npar = ctx.npars(s)
    o, keep = options(**kw)
    mean, sigma, ml = (c_double * npar)(), (c_double * npar)(), (c_double * npar)()
    res = L.NsResult()
    rc = lib.lensed_nested_run(ctx.ptr, s, byref(o), mean, sigma, ml, byref(res))
    del keep
    if rc != 0:
        raise RuntimeError(L.last_error())
    return NestedResult(res, mean, sigma, ml)


def run_multinest_abi(loglike, ndim, npar, dumper=None, nlive=300, tol=0.1, efr=0.8, ceff=False, updint=100,
                      seed=0, wrap=None, root=None, resume=False, outfile=False, maxiter=0, context=None,
                      feedback=False):
    """Call the library's MultiNest-compatible run() with ctypes callbacks.

    loglike(cube, ndim, npar, lnew, ctx) and dumper(...) have MultiNest's C
    signatures (pointer arguments), as src/nested.h:3-6 declares them.
    """
    lib = L.load()
    lfn = loglike if isinstance(loglike, L.LOGLIKE_FN) else L.LOGLIKE_FN(loglike)
    if dumper is None:
        dfn = ctypes.cast(None, L.DUMPER_FN)
    else:
        dfn = dumper if isinstance(dumper, L.DUMPER_FN) else L.DUMPER_FN(dumper)
    w = (c_int * npar)(*([0] * npar if wrap is None else [int(bool(v)) for v in wrap]))
    rootbuf = ctypes.create_string_buffer((root or "").encode(), 100)
    lib.run(1, 1, int(ceff), nlive, tol, efr, ndim, npar, ndim, 100, updint, -1e90, rootbuf, seed, w, int(feedback),
            int(resume), int(outfile), 1, -np.finfo(np.float64).max, maxiter, lfn, dfn, context)
