"""The column-sweep convolution (conv_rows = 1: one wave per 64-column strip
runs down the image) against the tile kernel, on the GPU.

The sweep forms every output with the tile kernel's FMA chain and sums chi^2
over the same 8 x 64 blocks in the same order, so every logL and every
dumper plane must be the same bits: culled (EPL table, Gaussian source) and
unculled (SIE) systems, images whose sides are not multiples of 64 or 8,
odd and even PSF sides, several systems of different sizes and PSFs in one
launch, and the fallback to the tile where the sweep does not apply (a PSF
of higher rank, or wider than 21).  The FP64 oracle checks the sweep's planes directly on a
small case.
"""
import numpy as np
import pytest

from lensed_amd import workloads as W
from lensed_amd._lib import LENSED_GRAPH, LENSED_TABLE
from lensed_amd.lensed import Context, SystemSpec
from oracle import oracle as O
from test_gpu_parity import oracle_model, oracle_render_cb, rel_err

pytestmark = pytest.mark.gpu

SWEEP = 1


def pids_of(objects):
    return [f"{oid}.{p}" for oid, n in objects for p in O.PARAMS[n]]


def both_ways(specs, flags, run, max_batch):
    """run(ctx) under the 32-row tile and under the sweep; the sweep must
    have been taken (the resolved tuning says so)."""
    with Context(specs, max_batch=max_batch, flags=flags) as ctx:
        ctx.set_tuning(conv_rows=8)
        tile = run(ctx)
        assert ctx.tuning()["conv_rows"] == 8
        ctx.set_tuning(conv_rows=SWEEP)
        sweep = run(ctx)
        assert ctx.tuning()["conv_rows"] == SWEEP
    return tile, sweep


def c3_like(shape, psf, seed):
    wl = W.c3()
    wl.shape = shape
    wl.psf = psf
    wl.seed = seed
    cx, cy = shape[1] / 2 + 0.5, shape[0] / 2 + 0.5
    for k, v in (("lens.x", cx), ("lens.y", cy), ("source.x", cx + 1.5), ("source.y", cy - 1.0)):
        wl.truth[k] = v
    for k, (lo, hi) in (("lens.x", (cx - 2, cx + 2)), ("lens.y", (cy - 2, cy + 2)),
                        ("source.x", (cx - 3.5, cx + 6.5)), ("source.y", (cy - 6.0, cy + 4.0))):
        wl.priors[k] = f"unif {lo} {hi}"
    wl.make_data(oracle_render_cb, pids_of(wl.objects), mag_keys=("source.mag",))
    return wl


@pytest.mark.parametrize("shape,psf", [((256, 256), W.gaussian_psf(21, 3.0)),
                                       ((200, 136), W.gaussian_psf(21, 3.0)),
                                       ((100, 150), W.gaussian_psf(20, 3.0)),
                                       ((72, 40), W.gaussian_psf(15, 2.5))])
def test_sweep_equals_tile_culled(shape, psf):
    """EPL + shear (table, pixel cull) and a Gaussian source: prior draws and
    the truth, logL bit for bit; the dumper planes at the truth too."""
    wl = c3_like(shape, psf, seed=3)
    truth = wl.truth_vector(pids_of(wl.objects))
    cubes = np.random.default_rng(17).random((512, 14))

    def run(ctx):
        lb, _ = ctx.loglike_batch(cubes)
        lt = ctx.loglike_phys(truth[None])[0]
        return lb, lt, ctx.render(truth)

    (lb0, lt0, p0), (lb1, lt1, p1) = both_ways(wl.spec(), LENSED_TABLE | LENSED_GRAPH, run, 512)
    assert np.array_equal(lb0, lb1), np.max(np.abs(lb0 - lb1))
    assert lt0 == lt1
    for k in ("image", "chi2"):
        assert np.array_equal(p0[k], p1[k]), k


@pytest.mark.parametrize("shape", [(256, 256), (100, 150)])
def test_sweep_dark_blocks_take_the_data_partials(shape):
    """Sources from far past the image's edges to its centre: candidates whose
    lit rows sit at the top, at the bottom, between, or nowhere.  The sweep
    takes a dark end block's chi^2 partial from the data's own (computed once
    per context, for a model of 0) and sweeps only the rows between; logL bit
    for bit with the tile kernel, which computes every block."""
    wl = c3_like(shape, W.gaussian_psf(21, 3.0), seed=5)
    h, w = shape
    wl.priors["source.x"] = f"unif {-0.6 * w} {1.6 * w}"
    wl.priors["source.y"] = f"unif {-0.6 * h} {1.6 * h}"
    cubes = np.random.default_rng(29).random((512, 14))

    def run(ctx):
        return ctx.loglike_batch(cubes)[0]

    l0, l1 = both_ways(wl.spec(), LENSED_TABLE | LENSED_GRAPH, run, 512)
    assert np.all(np.isfinite(l0))
    assert np.array_equal(l0, l1), np.max(np.abs(l0 - l1))


def test_sweep_equals_tile_unculled_c2():
    """C2 (SIE + shear, bulge + disk + sky, 200 x 200, 21 x 21 PSF): no
    pixel cull, ragged strips and blocks; 1024 prior draws."""
    wl = W.c2()
    pids = pids_of(wl.objects)
    wl.make_data(oracle_render_cb, pids, mag_keys=("bulge.mag", "disk.mag"))
    with Context(wl.spec(), max_batch=1, flags=0) as c:
        ndim = c.ndims()
    cubes = np.random.default_rng(23).random((1024, ndim))

    def run(ctx):
        return ctx.loglike_batch(cubes)[0]

    l0, l1 = both_ways(wl.spec(), LENSED_GRAPH, run, 1024)
    assert np.array_equal(l0, l1), np.max(np.abs(l0 - l1))


def test_sweep_equals_tile_mixed_systems():
    """Four systems of different sizes and PSF sides in one launch, culled
    and unculled: each system's logL bit for bit."""
    specs, phys = [], []
    rng = np.random.default_rng(29)
    for (h, w), psf, lens in (((64, 64), W.gaussian_psf(15, 2.5), "epl_plus_shear"),
                              ((128, 200), W.gaussian_psf(20, 3.0), "sie"),
                              ((256, 256), W.gaussian_psf(21, 3.0), "epl_plus_shear"),
                              ((512, 96), W.gaussian_psf(15, 2.5), "epl_plus_shear")):
        wl = c3_like((h, w), psf, seed=int(h + w))
        if lens == "sie":
            wl.objects = [("lens", "sie"), ("source", "gauss")]
            for k in ("lens.t", "lens.g1", "lens.g2"):
                wl.truth.pop(k)
                wl.priors.pop(k)
            wl.make_data(oracle_render_cb, pids_of(wl.objects), mag_keys=("source.mag",))
        specs.append(wl.spec())
        ids = pids_of(wl.objects)
        lo = np.array([wl.prior_range(k)[0] for k in ids])
        hi = np.array([wl.prior_range(k)[1] for k in ids])
        phys.append(lo + rng.random((128, len(ids))) * (hi - lo))

    def run(ctx):
        return ctx.loglike_phys(phys)

    l0, l1 = both_ways(specs, LENSED_TABLE | LENSED_GRAPH, run, 128)
    for s in range(len(specs)):
        assert np.array_equal(l0[s], l1[s]), (s, np.max(np.abs(l0[s] - l1[s])))


def test_sweep_against_the_fp64_oracle():
    """The sweep's convolved model and logL at the truth of a ragged culled
    case against the FP64 oracle (model within 1e-4 relative, logL 1e-2)."""
    wl = c3_like((120, 90), W.gaussian_psf(21, 3.0), seed=5)
    truth = wl.truth_vector(pids_of(wl.objects))
    with Context(wl.spec(), max_batch=1, flags=LENSED_TABLE) as ctx:
        ctx.set_tuning(conv_rows=SWEEP)
        planes = ctx.render(truth)
        lt = ctx.loglike_phys(truth[None])[0]
        assert ctx.tuning()["conv_rows"] == SWEEP
    model = oracle_model(wl.objects, truth, wl.shape, wl.psf)
    rel, _ = rel_err(planes["image"], model)
    assert rel <= 1e-4, rel
    assert abs(lt - (-0.5 * O.chi2_d(wl.image, wl.weight, model))) <= 1e-2


@pytest.mark.parametrize("psf", [W.moffat_psf(15, 3.0, 3.0), W.gaussian_psf(31, 3.0)])
def test_sweep_falls_back_where_it_does_not_apply(psf):
    """A Moffat PSF (rank above 1) and a Gaussian wider than 21 keep the tile
    kernel under conv_rows = 1, with unchanged results."""
    wl = c3_like((96, 96), psf, seed=7)
    cubes = np.random.default_rng(31).random((64, 14))
    with Context(wl.spec(), max_batch=64, flags=LENSED_TABLE) as ctx:
        ctx.set_tuning(conv_rows=8)
        ref, _ = ctx.loglike_batch(cubes)
        ctx.set_tuning(conv_rows=SWEEP)
        got, _ = ctx.loglike_batch(cubes)
        assert ctx.tuning()["conv_rows"] != SWEEP
    assert np.array_equal(ref, got)
