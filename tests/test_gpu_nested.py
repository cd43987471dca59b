"""Nested sampling end to end on the GPU likelihood.

The plain run is Lensed's own call (src/lensed.c:1270-1275): run() with the
library's loglike and dumper, one candidate per likelihood call.  The
batched run proposes many candidates per launch (lensed_nested_run).  The
two are independent samplings of the same posterior, so they must agree
within their sampling noise.

C1 (SIE + Sersic, 100 x 100) with the configuration's 50 live points and
bounds that enclose the live points: the log-evidence within four combined
standard errors, posterior means within 0.6 sigma (the 0.3 sigma that two
runs of 200 live points meet, widened by sqrt(200 / 50) for the sampling
noise of 50), and the mock's truth inside the posterior.

C3 (EPL + shear, Gaussian source, 256 x 256, 21 x 21 PSF) with Lensed's
default sampler settings (constant efficiency mode, target acceptance 0.05,
src/input/options.c:168-179) and 400 live points: the batched run with the
EPL emulator and 4096 candidates per launch against the plain run, posterior
means within 0.1 sigma.  Constant efficiency mode shrinks the bound below
the live points' hull, so its evidence is not compared.

C5 (EPL + shear, Sersic bulge and disk, sky, 512 x 512, 31 x 31 Moffat
PSF, 23 free parameters) with Lensed's default sampler settings and 400 live
points: the batched run with the EPL emulator, the recheck of accepted
points and 2048 candidates per launch against the plain run.  Both runs use
the same sampler, so their evidences carry the same constant-efficiency bias
and must agree within their noise: three combined standard errors.  Three
batched runs of different seeds differ among themselves by 0.03-0.05 sigma
in the median parameter and 0.09-0.16 sigma in the worst of the 23: the
bars are 0.1 sigma for the median and 0.25 sigma for the worst.
"""
import ctypes
import math
import os
import time
from ctypes import byref, c_double

import numpy as np
import pytest

from lensed_amd import _lib as L
from lensed_amd import nested
from lensed_amd import workloads as W
from lensed_amd._lib import LENSED_EMULATE, LENSED_GRAPH
from lensed_amd.lensed import Context, LensedObject, SystemSpec

pytestmark = pytest.mark.gpu


def gpu_render_cb(pids):
    def render(objects, params, shape, psf):
        pri = {k: repr(float(v)) for k, v in zip(pids, params)}
        spec = SystemSpec(objects, pri, np.zeros(shape, np.float32), np.ones(shape, np.float32), psf=psf)
        with Context(spec, max_batch=1) as c:
            return c.render(params)["image"].astype(np.float64)
    return render


@pytest.fixture(scope="module")
def c1_ctx():
    wl = W.c1()
    objs = [LensedObject(oid, name) for oid, name in wl.objects]
    pids = [f"{o.id}.{p}" for o in objs for p in o.params]
    wl.make_data(gpu_render_cb(pids), pids, mag_keys=("source.mag",))
    ctx = Context(wl.spec(), max_batch=1024)
    yield wl, pids, ctx
    ctx.close()


def results(ctx):
    n = ctx.npars()
    mean, sigma, ml, mp = (c_double * n)(), (c_double * n)(), (c_double * n)(), (c_double * n)()
    lz, lze, mx = c_double(), c_double(), c_double()
    assert L.load().lensed_results(ctx.ptr, mean, sigma, ml, mp, byref(lz), byref(lze), byref(mx)) == 0
    return np.array(mean), np.array(sigma), lz.value, lze.value, mx.value


def test_plain_and_batched_nested_runs_agree(c1_ctx):
    wl, pids, ctx = c1_ctx
    lib = L.load()
    npar, ndim = ctx.npars(), ctx.ndims()
    pmap = ctx.pmap()
    print(f"C1 nested: ndim {ndim} npar {npar}", flush=True)

    t0 = time.perf_counter()
    nested.run_multinest_abi(ctypes.cast(lib.loglike, L.LOGLIKE_FN), ndim, npar,
                             ctypes.cast(lib.dumper, L.DUMPER_FN), nlive=50, tol=0.5, efr=0.8, ceff=False,
                             updint=1000, seed=21, context=ctx.ptr)
    t_plain = time.perf_counter() - t0
    mean_p, sig_p, logz_p, logz_err_p, maxl_p = results(ctx)
    print(f"plain run(): {t_plain:.1f} s, logZ {logz_p:.3f} +- {logz_err_p:.3f}, max logL {maxl_p:.2f}", flush=True)

    t0 = time.perf_counter()
    rb = nested.run_context(ctx, 0, nlive=50, tol=0.5, efr=0.8, batch=1024, seed=22)
    t_batch = time.perf_counter() - t0
    print(f"batched: {t_batch:.1f} s, {rb}", flush=True)
    mean_b = np.zeros(npar)
    sig_b = np.zeros(npar)
    mean_b[pmap] = rb.mean
    sig_b[pmap] = rb.sigma

    err = math.hypot(logz_err_p, rb.logz_err)
    assert abs(logz_p - rb.logz) < 4 * err + 0.1, (logz_p, rb.logz, err)
    dev = np.abs(mean_p - mean_b) / sig_p
    print("mean difference / sigma:", dict(zip(pids, np.round(dev, 3))), flush=True)
    assert np.all(dev < 0.6), dev
    assert np.all(np.abs(sig_b / sig_p - 1) < 0.6)
    truth = wl.truth_vector(pids)
    assert np.all(np.abs(mean_b - truth) < 5 * sig_b + 1e-6), (mean_b - truth) / sig_b
    assert abs(rb.max_loglike - maxl_p) < 5


EMULATOR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lensed_amd", "data",
                        "epl_mlp.bin")


def test_c3_emulated_batched_posterior_matches_plain_run():
    wl = W.c3()
    objs = [LensedObject(oid, name) for oid, name in wl.objects]
    pids = [f"{o.id}.{p}" for o in objs for p in o.params]
    wl.make_data(gpu_render_cb(pids), pids, mag_keys=("source.mag",))
    kw = dict(nlive=400, tol=0.1, efr=0.05, ceff=True)
    lib = L.load()
    with Context(wl.spec(), max_batch=1) as ctx:
        npar, ndim = ctx.npars(), ctx.ndims()
        # position angles are periodic, as Lensed's examples declare them
        assert [pids[k] for k, w in zip(ctx.pmap(), ctx.wrap()) if w] == ["lens.pa", "source.pa"]
        t0 = time.perf_counter()
        nested.run_multinest_abi(ctypes.cast(lib.loglike, L.LOGLIKE_FN), ndim, npar,
                                 ctypes.cast(lib.dumper, L.DUMPER_FN), updint=100000, seed=31, wrap=ctx.wrap(),
                                 context=ctx.ptr, **kw)
        t_plain = time.perf_counter() - t0
        mean_p, sig_p, logz_p, logz_err_p, maxl_p = results(ctx)
    print(f"C3 plain run(): {t_plain:.1f} s, logZ {logz_p:.2f}, max logL {maxl_p:.2f}", flush=True)
    with Context(wl.spec(), max_batch=4096, flags=LENSED_GRAPH | LENSED_EMULATE, emulator=EMULATOR) as ctx:
        pmap = ctx.pmap()
        t0 = time.perf_counter()
        rb = nested.run_context(ctx, 0, batch=4096, seed=33, **kw)
        t_batch = time.perf_counter() - t0
    print(f"C3 batched + emulated: {t_batch:.1f} s, {rb}", flush=True)
    mean_b = np.zeros(npar)
    mean_b[pmap] = rb.mean
    dev = np.abs(mean_p - mean_b) / sig_p
    print("mean difference / sigma:", dict(zip(pids, np.round(dev, 3))), flush=True)
    assert np.all(dev < 0.1), dev
    assert abs(rb.max_loglike - maxl_p) < 5



def test_c5_emulated_batched_run_matches_plain_run():
    wl = W.c5()
    objs = [LensedObject(oid, name) for oid, name in wl.objects]
    pids = [f"{o.id}.{p}" for o in objs for p in o.params]
    wl.make_data(gpu_render_cb(pids), pids, mag_keys=("bulge.mag", "disk.mag"))
    kw = dict(nlive=400, tol=0.1, efr=0.05, ceff=True)
    lib = L.load()
    with Context(wl.spec(), max_batch=1) as ctx:
        npar, ndim = ctx.npars(), ctx.ndims()
        assert ndim == 23
        t0 = time.perf_counter()
        nested.run_multinest_abi(ctypes.cast(lib.loglike, L.LOGLIKE_FN), ndim, npar,
                                 ctypes.cast(lib.dumper, L.DUMPER_FN), updint=100000, seed=51, wrap=ctx.wrap(),
                                 context=ctx.ptr, **kw)
        t_plain = time.perf_counter() - t0
        mean_p, sig_p, logz_p, logz_err_p, maxl_p = results(ctx)
    print(f"C5 plain run(): {t_plain:.1f} s, logZ {logz_p:.2f} +- {logz_err_p:.2f}, max logL {maxl_p:.2f}",
          flush=True)
    with Context(wl.spec(), max_batch=2048, flags=LENSED_GRAPH | LENSED_EMULATE, emulator=EMULATOR) as ctx:
        pmap = ctx.pmap()
        t0 = time.perf_counter()
        rb = nested.run_context(ctx, 0, batch=2048, seed=53, **kw)
        t_batch = time.perf_counter() - t0
        rechecked, changed, max_d = ctx.emulator_rechecks()
    print(f"C5 batched + emulated: {t_batch:.1f} s, {rb}, max logL {rb.max_loglike:.2f}; rechecked {rechecked}, "
          f"changed {changed}", flush=True)
    mean_b = np.zeros(npar)
    mean_b[pmap] = rb.mean
    free = sig_p > 0  # (the sky gradients are fixed)
    dev = np.abs(mean_p - mean_b)[free] / sig_p[free]
    print("mean difference / sigma:", dict(zip(np.array(pids)[free], np.round(dev, 3))), flush=True)
    err = math.hypot(logz_err_p, rb.logz_err)
    assert abs(logz_p - rb.logz) < 3 * err, (logz_p, rb.logz, err)
    assert np.median(dev) < 0.1 and np.max(dev) < 0.25, dev
    assert abs(rb.max_loglike - maxl_p) < 5
    # the mock's truth inside the posterior (position angles aside: a
    # periodic posterior across 0/180 has no meaningful linear mean)
    truth = wl.truth_vector(pids)
    lin = free & np.array([not k.endswith(".pa") for k in pids])
    assert np.all(np.abs(truth - mean_p)[lin] < 5 * sig_p[lin]), dict(
        zip(np.array(pids)[lin], np.round((np.abs(truth - mean_p) / np.where(sig_p > 0, sig_p, 1))[lin], 2)))
