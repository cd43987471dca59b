"""The entry point that allocates nothing (lensed_loglike_batch_noalloc),
through the C ABI, against lensed_loglike_batch on the same candidates: bit
for bit, on two calls in a row (the second with fewer candidates, so it
runs on what the first left in the static scratch), for a culled EPL-table
system and an uncullled SIE one, and a clear error beyond the scratch."""
import ctypes
from ctypes import POINTER, c_double

import numpy as np
import pytest

from lensed_amd import _lib as L
from lensed_amd import workloads as W
from lensed_amd._lib import LENSED_GRAPH, LENSED_TABLE
from lensed_amd.lensed import Context
from oracle import oracle as O
from test_gpu_parity import oracle_render_cb

pytestmark = pytest.mark.gpu

DP = POINTER(c_double)


def noalloc(ctx, cubes, s=0):
    cubes = np.ascontiguousarray(cubes, dtype=np.float64)
    n = cubes.shape[0]
    phys = np.zeros((n, ctx.npars(s)))
    lnew = np.zeros(n)
    rc = L.load().lensed_loglike_batch_noalloc(ctypes.c_void_p(ctx.ptr), s, cubes.ctypes.data_as(DP), n,
                                              phys.ctypes.data_as(DP), lnew.ctypes.data_as(DP))
    return rc, lnew, phys


def workload(make):
    wl = make()
    pids = [f"{o}.{p}" for o, n in wl.objects for p in O.PARAMS[n]]
    wl.make_data(oracle_render_cb, pids, mag_keys=("source.mag",))
    return wl


@pytest.mark.parametrize("make,flags,counts", [(W.c3, LENSED_TABLE | LENSED_GRAPH, (400, 123)),
                                               (W.c1, 0, (256, 17))])
def test_noalloc_equals_batch(make, flags, counts):
    wl = workload(make)
    with Context(wl.spec(), max_batch=max(counts), flags=flags) as ctx:
        nd = ctx.ndims()
        for k, n in enumerate(counts):
            cubes = np.random.default_rng(100 + k).random((n, nd))
            ref, pref = ctx.loglike_batch(cubes)
            rc, got, pgot = noalloc(ctx, cubes)
            assert rc == 0, L.last_error()
            assert np.all(np.isfinite(got)), k
            assert np.array_equal(got, ref), (k, np.max(np.abs(got - ref)))
            assert np.array_equal(pgot, pref), k
        # the batch path again after the static one: unchanged
        again, _ = ctx.loglike_batch(cubes)
        assert np.array_equal(again, ref)


def test_noalloc_rejects_what_does_not_fit():
    wl = workload(W.c3)
    with Context(wl.spec(), max_batch=600, flags=LENSED_TABLE) as ctx:
        cubes = np.random.default_rng(7).random((600, ctx.ndims()))
        rc, _, _ = noalloc(ctx, cubes)
        assert rc == -1 and "static scratch" in L.last_error()
        rc, _, _ = noalloc(ctx, cubes[:10])
        assert rc == 0
    with Context(wl.spec(), max_batch=8, flags=LENSED_TABLE) as ctx:
        rc, _, _ = noalloc(ctx, np.random.default_rng(8).random((9, ctx.ndims())))
        assert rc == -1 and "max_batch" in L.last_error()
