"""The dumper's output layers on the GPU (src/nested.c:169-226).

With an output file set, every dumper() call renders the maximum-likelihood
parameters and writes IMG (PSF-convolved model), RES (data - model), RAW
(model before the PSF), ERR (quadrature error / model), WHT (weights) and
PVL (erfc(sqrt(chi^2 / 2))) as named extensions.  IMG and RAW are checked
against the FP64 oracle to the model-image bar (1e-4 relative where the
model exceeds 1e-3 of its peak), ERR against the oracle's error plane to
the bar the reference's own ERR layers meet (2e-5); the derived layers are
checked against their definitions applied to those planes.
"""
import ctypes
from ctypes import POINTER, c_double, c_int

import numpy as np
import pytest
from scipy.special import erfc

from lensed_amd import _lib as L
from lensed_amd import data as D
from lensed_amd import workloads as W
from lensed_amd.lensed import Context
from oracle import oracle as O
from test_gpu_parity import PTS, WW64, oracle_model, oracle_render_cb, rel_err

pytestmark = pytest.mark.gpu


def test_dumper_writes_the_six_layers(tmp_path):
    wl = W.c2()
    wl.shape = (64, 72)
    for k, v in (("lens.x", 32.5), ("lens.y", 32.5), ("bulge.x", 33.5), ("bulge.y", 31.5), ("disk.x", 33.5),
                 ("disk.y", 31.5), ("lens.r", 10.0)):
        wl.truth[k] = v
    pids = [f"{o}.{p}" for o, n in wl.objects for p in O.PARAMS[n]]
    wl.priors.update({"lens.x": "unif 30 35", "lens.y": "unif 30 35", "bulge.x": "unif 30 37",
                      "bulge.y": "unif 28 35", "disk.x": "unif 30 37", "disk.y": "unif 28 35",
                      "lens.r": "unif 8 12"})
    wl.make_data(oracle_render_cb, pids, mag_keys=("bulge.mag", "disk.mag"))
    path = str(tmp_path / "model.fits")
    with Context(wl.spec(), max_batch=4) as ctx:
        n, pmap = ctx.npars(), ctx.pmap()
        ctx.set_output(path)
        truth = wl.truth_vector(pids)
        # constraints [MEAN | SIGMA | ML | MAP] in sampler order, as MultiNest passes them
        ml = np.array([truth[pmap[i]] for i in range(n)])
        cons = np.concatenate([ml, np.ones(n), ml, ml])
        cp = cons.ctypes.data_as(POINTER(c_double))
        dp = ctypes.pointer(cp)
        ns, nl, npar = c_int(1), c_int(1), c_int(n)
        z = c_double(0.0)
        L.load().dumper(ctypes.byref(ns), ctypes.byref(nl), ctypes.byref(npar), dp, dp, dp, ctypes.byref(z),
                        ctypes.byref(z), ctypes.byref(z), ctypes.byref(z), ctx.ptr)
    layers = {name: D.read_fits(f"{path}[{name}]")[0] for name in D.LAYERS}
    raw_exact = oracle_model(wl.objects, truth, wl.shape)
    img_exact = oracle_model(wl.objects, truth, wl.shape, wl.psf)
    assert rel_err(layers["IMG"], img_exact)[0] <= 1e-4
    assert rel_err(layers["RAW"], raw_exact)[0] <= 1e-4
    assert np.array_equal(layers["RES"], wl.image - layers["IMG"])
    assert np.array_equal(layers["WHT"], wl.weight)
    chi2 = wl.weight * (layers["IMG"].astype(np.float64) - wl.image) ** 2
    assert np.allclose(layers["PVL"], erfc(np.sqrt(0.5 * chi2)), atol=1e-5)
    val, err = O.render(wl.objects, truth, wl.shape, PTS, WW64, [1.0, 1.0, 1.0, 1.0])
    mask = np.abs(val) > 1e-3 * np.abs(val).max()
    assert np.max(np.abs(layers["ERR"][mask] - (err / val)[mask])) <= 2e-5
