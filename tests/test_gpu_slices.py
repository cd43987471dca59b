"""Slices of the reference path run on the GPU against the FP64 oracle:
image-plane priors, every quadrature rule, the 200 x 200 batched launch of
1024 prior draws, a lens plugin whose data layout the tabulated form cannot
read, the render's walk over culled items, and the launch-shape knobs.

Bars as in test_gpu_parity: model images within 1e-4 relative of the FP64
oracle where the model exceeds 1e-3 of its peak (for prior draws, or twice
the FP32 restatement's own error where that is larger: a compact Sersic
cusp at pixel scale); |delta logL| <= 1e-2 near the truth and 1e-2 + 1e-5
|logL| for arbitrary prior draws (FP32 model rounding at |logL| up to 1e7);
results under different launch shapes and culling choices equal bit for
bit.
"""
import os

import numpy as np
import pytest

from lensed_amd import quadrature as Q
from lensed_amd import workloads as W
from lensed_amd._lib import LENSED_GRAPH, LENSED_NO_CULL, LENSED_TABLE
from lensed_amd.lensed import Context, SystemSpec
from oracle import oracle as O
from test_gpu_parity import oracle_model, oracle_render_cb, rel_err

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PTS, WHT, ERR = Q.table("g3k7")
WW64 = np.stack([WHT, ERR], 1)


def pids_of(objects):
    return [f"{oid}.{p}" for oid, n in objects for p in O.PARAMS[n]]



@pytest.mark.parametrize("lens,flags", [("sie", 0), ("epl_plus_shear", 0), ("epl_plus_shear", LENSED_TABLE)])
def test_image_plane_prior_on_gpu(lens, flags):
    """`image` priors on the source position (src/kernel.c:141-150,490-575):
    the sampled position is where the source appears in the image, traced
    back through the lens by the generated set_params before the source's
    own set().  GPU planes and logL against the FP64 oracle's traced
    render, and against the same source placed at the traced position."""
    objs = [("lens", lens), ("src", "gauss")]
    lensp = {"sie": [40.5, 39.5, 12.0, 0.7, 25.0],
             "epl_plus_shear": [40.5, 39.5, 12.0, 1.15, 0.7, 25.0, 0.03, -0.02]}[lens]
    ximg = [49.0, 44.0]  # a point on the image plane near the Einstein ring
    p = np.array(lensp + ximg + [1.5, 0.0, 0.8, 30.0])
    pids = pids_of(objs)
    shape = (80, 80)
    traced, _ = O.render(objs, p, shape, PTS, WW64, ipp=[0, 1])
    rng = np.random.default_rng(11)
    image = (traced + rng.normal(0, 0.05, shape)).astype(np.float32)
    weight = np.full(shape, 400.0, np.float32)
    pri = {k: f"unif {v - 2} {v + 2}" for k, v in zip(pids, p)}
    pri["lens.q"] = "unif 0.5 0.9"
    pri["src.x"] = f"image unif {ximg[0] - 3} {ximg[0] + 3}"
    pri["src.y"] = f"image unif {ximg[1] - 3} {ximg[1] + 3}"
    with Context(SystemSpec(objs, pri, image, weight), max_batch=64, flags=flags) as ctx:
        planes = ctx.render(p)
        rel, peak_rel = rel_err(planes["value"], traced)
        assert rel <= 1e-4 and peak_rel <= 3e-5, (rel, peak_rel)
        # the same image with the source set at the traced source-plane position
        a = O.deflect(objs, np.concatenate([p[:len(lensp)], [0, 0, 1.5, 0.0, 0.8, 30.0]]), 0, np.array(ximg))[0]
        direct, _ = O.render(objs, np.concatenate([p[:len(lensp)], np.array(ximg) - a, p[len(lensp) + 2:]]), shape,
                             PTS, WW64)
        assert np.max(np.abs(direct - traced)) <= 1e-12 * np.abs(traced).max()
        # logL of the truth and of near-truth draws through the batch path
        near = p[None] + 0.02 * rng.normal(size=(16, len(p)))
        near[0] = p
        lnew = ctx.loglike_phys(near)
        for i in (0, 1, 2, 3):
            model, _ = O.render(objs, near[i], shape, PTS, WW64,
                                ipp=[0, 1])
            lref = -0.5 * O.chi2_d(image, weight, model)
            assert abs(lnew[i] - lref) <= 1e-2, (i, lnew[i], lref)
        assert abs(-2 * lnew[0] / image.size - 1) < 0.05
        # the batch path transforms unit cubes into image positions, in range
        cubes = np.random.default_rng(3).random((64, ctx.ndims()))
        lb, phys = ctx.loglike_batch(cubes)
        ix = ctx.pmap().index(pids.index("src.x"))
        assert np.all(np.abs(phys[:, ix] - ximg[0]) <= 3) and np.all(np.isfinite(lb))


@pytest.mark.parametrize("rule", ["point", "sub2", "sub4", "gm75", "g5k11", "g7k15"])
def test_quadrature_rule_on_gpu(rule):
    """The non-default rules of src/quadrature.c:21-30 on the GPU: value and
    error planes against the FP64 oracle with the same rule's FP64 nodes,
    and logL at the truth and near it."""
    objs = [("lens", "sie"), ("src", "sersic")]
    p = np.array([50.5, 50.5, 20.0, 0.8, 30.0, 51.3, 49.6, 4.0, 0.0, 2.5, 0.7, 60.0])
    shape = (100, 100)
    pts, wht, err = Q.table(rule)
    exact, exact_err = O.render(objs, p, shape, pts, np.stack([wht, err], 1))
    rng = np.random.default_rng(21)
    image = (exact + rng.normal(0, 0.05, shape)).astype(np.float32)
    weight = np.full(shape, 400.0, np.float32)
    pri = {k: f"unif {v * 0.98 - 0.1} {v * 1.02 + 0.1}" for k, v in zip(pids_of(objs), p)}
    with Context(SystemSpec(objs, pri, image, weight, rule=rule), max_batch=8) as ctx:
        planes = ctx.render(p)
        rel, peak_rel = rel_err(planes["value"], exact)
        assert rel <= 1e-4 and peak_rel <= 3e-5, (rule, rel, peak_rel)
        if np.abs(exact_err).max() > 0:  # the one-node rule has no error estimate
            assert np.max(np.abs(planes["error"] - exact_err)) <= 3e-5 * np.abs(exact).max(), rule
        else:
            assert not np.any(planes["error"])
        near = p[None] * (1 + 1e-4 * rng.normal(size=(8, len(p))))
        near[0] = p
        lnew = ctx.loglike_phys(near)
        for i in (0, 1, 2):
            model, _ = O.render(objs, near[i], shape, pts, np.stack([wht, err], 1))
            lref = -0.5 * O.chi2_d(image, weight, model)
            assert abs(lnew[i] - lref) <= 1e-2, (rule, i, lnew[i], lref)


@pytest.fixture(scope="module")
def c2():
    wl = W.c2()
    pids = pids_of(wl.objects)
    wl.make_data(oracle_render_cb, pids, mag_keys=("bulge.mag", "disk.mag"))
    return wl, pids


def test_c2_batched_launch_against_fp64_oracle(c2):
    """200 x 200, SIE + shear, Sersic bulge + exponential disk + sky, 21 x 21
    PSF: one launch of 1024 candidates drawn uniformly from the priors
    (seeded), a seeded sample of 16 of them plus the truth checked against
    the FP64 oracle (logL, and model images for four of them)."""
    wl, pids = c2
    K = 1024
    with Context(wl.spec(), max_batch=K, flags=LENSED_GRAPH) as ctx:
        cubes = wl.candidates(ctx.ndims(), K)
        lnew, phys_s = ctx.loglike_batch(cubes)
        assert np.all(np.isfinite(lnew)) and np.all(lnew <= 0)
        pmap = ctx.pmap()
        phys = np.zeros((K, len(pids)))
        phys[:, pmap] = phys_s
        sample = np.sort(np.random.default_rng(102).choice(K, 16, replace=False))
        truth = wl.truth_vector(pids)
        lt = ctx.loglike_phys(truth[None])[0]
        worst = 0.0
        for j, i in enumerate(sample):
            model = oracle_model(wl.objects, phys[i], wl.shape, wl.psf)
            lref = -0.5 * O.chi2_d(wl.image, wl.weight, model)
            worst = max(worst, abs(lnew[i] - lref) / (1 + abs(lref)))
            assert abs(lnew[i] - lref) <= 1e-2 + 1e-5 * abs(lref), (i, lnew[i], lref)
            if j < 4:
                # a prior draw can put a compact Sersic cusp at pixel scale,
                # where rounding the ray positions to FP32 alone moves the
                # model by more than 1e-4: the bar is then the FP32
                # restatement's own error (the reference's precision), twice
                planes = ctx.render(phys[i])
                rel, _ = rel_err(planes["image"], model)
                rel32, _ = rel_err(oracle_model(wl.objects, phys[i], wl.shape, wl.psf, np.float32), model)
                assert rel <= max(1e-4, 2 * rel32), (i, rel, rel32)
        model = oracle_model(wl.objects, truth, wl.shape, wl.psf)
        lref = -0.5 * O.chi2_d(wl.image, wl.weight, model)
        assert abs(lt - lref) <= 1e-2, (lt, lref)
        assert abs(-2 * lt / wl.image.size - 1) < 0.03
        print(f"C2: 1024 prior draws, 16 + truth vs FP64: worst |dlogL|/(1+|logL|) {worst:.2e}, "
              f"truth dlogL {lt - lref:+.2e}", flush=True)


def test_untabulatable_epl_layout_keeps_the_series(tmp_path, monkeypatch):
    """An EPL plugin whose data block lacks a field the tabulated form reads
    (here the centre renamed) asks for LENSED_TABLE: it builds and runs the
    per-ray series, bit for bit what the context without the flag gives."""
    own = open(os.path.join(ROOT, "lensed_amd", "objects", "epl_plus_shear.cl")).read()
    (tmp_path / "epl_own_layout.cl").write_text(
        own.replace("float2 c;", "float2 centre;").replace("this->c ", "this->centre ")
        .replace("this->c;", "this->centre;").replace("this->c =", "this->centre ="))
    monkeypatch.setenv("LENSED_OBJECTS_PATH", str(tmp_path))
    objs = [("lens", "epl_own_layout"), ("src", "gauss")]
    p = [32.5, 31.5, 9.0, 1.2, 0.7, 35.0, 0.04, -0.03, 33.0, 31.0, 2.0, 0.0, 0.8, 10.0]
    pri = {k: f"unif {v - 0.5} {v + 0.5}" for k, v in zip(pids_of([("lens", "epl_plus_shear"), ("src", "gauss")]), p)}
    pri["lens.q"] = "unif 0.5 0.9"
    image = np.random.default_rng(4).normal(0, 0.05, (64, 64)).astype(np.float32)
    spec = SystemSpec(objs, pri, image, np.full_like(image, 400.0))
    cubes = np.random.default_rng(5).random((32, 14))
    with Context(spec, max_batch=32) as plain:
        lp, _ = plain.loglike_batch(cubes)
    with Context(spec, max_batch=32, flags=LENSED_TABLE) as tab:
        lt, _ = tab.loglike_batch(cubes)
    assert np.all(np.isfinite(lp)) and np.array_equal(lp, lt)
    # the package's own layout under the same name takes the table (close, not equal)
    monkeypatch.delenv("LENSED_OBJECTS_PATH")
    spec2 = SystemSpec([("lens", "epl_plus_shear"), ("src", "gauss")], pri, image, np.full_like(image, 400.0))
    with Context(spec2, max_batch=32, flags=LENSED_TABLE) as tab2:
        lt2, _ = tab2.loglike_batch(cubes)
    assert not np.array_equal(lt2, lp) and np.all(np.abs(lt2 - lp) <= 1e-2 + 1e-5 * np.abs(lp))


@pytest.fixture(scope="module")
def c3():
    wl = W.c3()
    pids = pids_of(wl.objects)
    wl.make_data(oracle_render_cb, pids, mag_keys=("source.mag",))
    return wl, pids


def test_walk_over_culled_items_is_exact(c3):
    """The render's other branch after a pixel cull: every item walked, a
    culled system's dark squares skipped by their flag (list_fraction < 0
    forces it where the list would run).  Two batches in a row in one
    layout, so the second cull trusts its record of written squares: logL
    and planes bit for bit those of tracing every ray."""
    wl, pids = c3
    cubes = [np.random.default_rng(s).random((512, 14)) for s in (31, 32)]
    truth = wl.truth_vector(pids)
    out = []
    for extra, walk in ((0, True), (0, False), (LENSED_NO_CULL, False)):
        with Context(wl.spec(), max_batch=512, flags=LENSED_TABLE | LENSED_GRAPH | extra) as ctx:
            if walk:
                ctx.set_tuning(list_fraction=-1.0)
            ls = [ctx.loglike_batch(c)[0] for c in cubes]
            stats = ctx.render_stats()[:3]
            if walk:
                assert ctx.tuning()["list_fraction"] == -1.0
            planes = ctx.render(truth)
            out.append((ls, stats, planes))
    (lw, sw, pw), (ll, sl, pl), (ln, sn, pn) = out
    for a, b, c in zip(lw, ll, ln):
        assert np.all(np.isfinite(a)) and np.array_equal(a, c) and np.array_equal(b, c)
    for k in pn:
        assert np.array_equal(pw[k], pn[k]) and np.array_equal(pl[k], pn[k]), k
    # the walk traces every quad of the squares holding a listed quad, the
    # list only those quads; neither traces them all
    assert sl[2] <= sw[2] < sn[2] and sw[2] % 4 == 0 and sw[1] == sl[1], (sw, sl, sn)


def test_launch_shapes_keep_results_bit_identical(c3):
    """Every launch-shape knob (the autotuner's search space) leaves the logL
    of every candidate unchanged, bit for bit: the column sweep (the default
    of this culled layout), tile heights 2, 8 and 16 rows per thread, both
    convolution item orders, grids from half to twice the resident count,
    list and walk."""
    wl, pids = c3
    cubes = np.random.default_rng(41).random((1024, 14))
    with Context(wl.spec(), max_batch=1024, flags=LENSED_TABLE | LENSED_GRAPH) as ctx:
        ref, _ = ctx.loglike_batch(cubes)
        base = ctx.tuning()
        assert base["conv_rows"] == 1 and base["conv_order"] == 1 and base["list_fraction"] == 0.75
        for knobs in (dict(conv_rows=8), dict(conv_rows=2), dict(conv_rows=16), dict(conv_rows=8, conv_order=2),
                      dict(conv_rows=16, conv_order=2),
                      dict(render_grid=0.5, cull_grid=2.0, conv_grid=2.0), dict(render_grid=2.0, conv_grid=0.5),
                      dict(list_fraction=2.0), dict(list_fraction=-1.0, conv_order=2)):
            ctx.set_tuning(**knobs)
            got, _ = ctx.loglike_batch(cubes)
            assert np.array_equal(got, ref), knobs
            t = ctx.tuning()
            for k, v in knobs.items():
                assert t[k] == v, (k, t)
        ctx.set_tuning()
        assert ctx.tuning()["conv_rows"] == 0
