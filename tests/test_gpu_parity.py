"""HIP path vs the CPU oracle, through the C ABI, on an MI355X.

Bars (written per test): model images within 1e-4 relative of the FP64
oracle wherever the model exceeds 1e-3 of its peak (and within 3e-5 of the
peak everywhere for the reference renders); |delta logL| <= 1e-2 against
the FP64 oracle near the truth, where logL differences decide acceptance;
for arbitrary prior draws, whose |logL| reaches 1e5-1e7, the FP32 model
rounding bounds the agreement at 1e-5 relative.  Plain and batched paths
are compared bit for bit (same kernels, fixed reduction order).
"""
import numpy as np
import pytest

from conftest import fixture_objects
from lensed_amd import quadrature as Q
from lensed_amd import workloads as W
from lensed_amd._lib import LENSED_GRAPH, LENSED_PROFILE
from lensed_amd.lensed import LOG_ZERO, Context, SystemSpec
from oracle import oracle as O

pytestmark = pytest.mark.gpu

PTS, WHT, ERR = Q.table("g3k7")
WW64 = np.stack([WHT, ERR], 1)


def oracle_model(objects, params, shape, psf=None, dtype=np.float64):
    pcs = [1.0, 1.0, 1.0, 1.0]
    if psf is not None:
        pcs[0] += 0.5 * (psf.shape[1] % 2 == 0)
        pcs[1] += 0.5 * (psf.shape[0] % 2 == 0)
    if dtype == np.float64:
        v, _ = O.render(objects, params, shape, PTS, WW64, pcs)
    else:
        qq, ww = Q.rule("g3k7")
        v, _ = O.render(objects, params, shape, qq, ww, pcs, dtype=np.float32)
    return O.convolve(v, psf, dtype) if psf is not None else v


def oracle_render_cb(objects, params, shape, psf):
    return oracle_model(objects, params, shape, psf)


def rel_err(got, exact):
    peak = np.abs(exact).max()
    mask = np.abs(exact) > 1e-3 * peak
    return np.max(np.abs(got - exact)[mask] / np.abs(exact[mask])), np.max(np.abs(got - exact)) / peak


@pytest.fixture(scope="module")
def fixture_ctx(golden):
    meta, arrays = golden
    names = sorted(meta)
    specs = []
    for t in names:
        m = meta[t]
        objs = fixture_objects(m)
        pri = {k: repr(v) for k, v in m["params"].items()}
        img = arrays[t + "/img"]
        specs.append(SystemSpec(objs, pri, img, np.full_like(img, m["weight"])))
    ctx = Context(specs, max_batch=2)
    yield names, specs, ctx
    ctx.close()


def test_gpu_reproduces_reference_renders(golden, fixture_ctx):
    meta, arrays = golden
    names, specs, ctx = fixture_ctx
    for s, t in enumerate(names):
        objs = fixture_objects(meta[t])
        p = O.param_vector(objs, meta[t]["params"])
        planes = ctx.render(p, s)
        img, err = arrays[t + "/img"], arrays[t + "/err"]
        peak = np.abs(img).max()
        mask = np.abs(img) > 1e-3 * peak
        assert np.max(np.abs(planes["value"] - img)) <= 3e-5 * peak, t
        exact = oracle_model(objs, p, img.shape)
        rel, _ = rel_err(planes["value"], exact)
        assert rel <= 1e-4, (t, rel)
        assert np.max(np.abs(planes["error"][mask] / planes["value"][mask] - err[mask])) <= 2e-5, t
        # no PSF: the image plane is the raw model, chi^2 per pixel w (m - d)^2
        assert np.array_equal(planes["image"], planes["value"])
        d = planes["value"] - img
        assert np.allclose(planes["chi2"], (img * 0 + meta[t]["weight"]) * d * d, rtol=1e-6, atol=1e-12)


def test_fixture_loglike_all_systems_one_launch(golden, fixture_ctx):
    meta, arrays = golden
    names, specs, ctx = fixture_ctx
    phys = [O.param_vector(fixture_objects(meta[t]), meta[t]["params"])[None, :] for t in names]
    out = ctx.loglike_phys(phys)
    for s, t in enumerate(names):
        objs = fixture_objects(meta[t])
        img = arrays[t + "/img"]
        model = oracle_model(objs, phys[s][0], img.shape)
        lref = -0.5 * O.chi2_d(img, np.full_like(img, meta[t]["weight"]), model)
        assert abs(out[s][0] - lref) <= 1e-2, (t, out[s][0], lref)


@pytest.fixture(scope="module")
def c1():
    wl = W.c1()
    pids = [f"{oid}.{p}" for oid, n in wl.objects for p in O.PARAMS[n]]
    wl.make_data(oracle_render_cb, pids, mag_keys=("source.mag",))
    return wl, pids


def test_c1_loglike_plain_batch_oracle(c1):
    wl, pids = c1
    with Context(wl.spec(), max_batch=64) as ctx:
        assert ctx.npars() == 12 and ctx.ndims() == 12
        cubes = wl.candidates(ctx.ndims(), 64)
        cubes[0] = [(wl.truth[k] - lo) / (hi - lo) for k, (lo, hi) in ((k, wl.prior_range(k)) for k in pids)]
        lb, pb = ctx.loglike_batch(cubes)
        for i in range(8):
            lp, pp = ctx.loglike(cubes[i])
            assert lp == lb[i]
            assert np.array_equal(pp, pb[i])
        pmap = ctx.pmap()
        for i in range(8):
            phys = np.zeros(12)
            phys[pmap] = pb[i]
            model = oracle_model(wl.objects, phys, wl.shape)
            lref = -0.5 * O.chi2_d(wl.image, wl.weight, model)
            if i == 0:  # the truth: chi^2 ~ N
                assert abs(lb[0] - lref) <= 1e-2, (lb[0], lref)
                assert abs(-2 * lb[0] / wl.image.size - 1) < 0.05
            assert abs(lb[i] - lref) <= 1e-2 + 1e-5 * abs(lref), (i, lb[i], lref)


def test_batched_priors_match_the_plain_transform(c1):
    """The batched path's prior transform gives the plain path's physical
    parameters and logL bit for bit whatever the priors: all uniform with a
    delta (one affine loop per candidate, staged straight as floats) and with
    normal priors (the per-parameter transform); a draw whose normal prior
    lands outside its parameter's bounds gets logZero on both paths."""
    wl, pids = c1
    variants = {
        "unif+delta": dict(wl.priors, **{"source.n": "2.5"}),
        "unif+norm": dict(wl.priors, **{"lens.q": "norm 0.8 0.1", "source.n": "norm 2.5 0.3"}),
    }
    for name, pri in variants.items():
        with Context(wl.spec(pri), max_batch=64) as ctx:
            nd = ctx.ndims()
            assert nd == (11 if name == "unif+delta" else 12)
            cubes = np.random.default_rng(7).random((40, nd))
            cubes[5, :] = 0.999999  # lens.q = 0.8 + 4.75 sigma > 1 under the normal prior
            lb, pb = ctx.loglike_batch(cubes)
            for i in range(len(cubes)):
                lp, pp = ctx.loglike(cubes[i])
                assert lp == lb[i], (name, i, lp, lb[i])
                assert np.array_equal(pp, pb[i]), (name, i)
            assert np.all(np.isfinite(np.delete(lb, 5))), name
            assert (lb[5] == LOG_ZERO) == (name == "unif+norm"), (name, lb[5])


@pytest.fixture(scope="module")
def c2():
    wl = W.c2()
    pids = [f"{oid}.{p}" for oid, n in wl.objects for p in O.PARAMS[n]]
    wl.make_data(oracle_render_cb, pids, mag_keys=("bulge.mag", "disk.mag"))
    return wl, pids


def test_c2_psf_convolution_and_loglike(c2):
    wl, pids = c2
    with Context(wl.spec(), max_batch=16) as ctx:
        phys = np.array([wl.truth.get(k, 0.0) for k in pids])
        planes = ctx.render(phys)
        conv = O.convolve(planes["value"].astype(np.float64), wl.psf.astype(np.float64))
        assert np.max(np.abs(planes["image"] - conv)) <= 2e-6 * np.abs(conv).max()
        exact = oracle_model(wl.objects, phys, wl.shape, wl.psf)
        rel, _ = rel_err(planes["image"], exact)
        assert rel <= 1e-4, rel
        lnew = ctx.loglike_phys(phys[None, :])[0]
        lref = -0.5 * O.chi2_d(wl.image, wl.weight, exact)
        assert abs(lnew - lref) <= 1e-2, (lnew, lref)


@pytest.mark.parametrize("shape,psf_shape", [((37, 53), (5, 3)), ((64, 64), (4, 4)), ((33, 70), (6, 5)),
                                             ((100, 40), (31, 31)), ((20, 20), (1, 1))])
def test_ragged_images_and_even_psfs(shape, psf_shape):
    rng = np.random.default_rng(sum(shape) + sum(psf_shape))
    objs = [("lens", "sie"), ("src", "gauss")]
    h, w = shape
    p = np.array([w / 2 + 0.3, h / 2 - 0.2, min(h, w) / 5, 0.7, 33.0, w / 2 + 1.0, h / 2, 2.0, 0.0, 0.8, 10.0])
    psf = rng.random(psf_shape).astype(np.float32)
    psf /= psf.sum()
    image = rng.normal(0, 0.1, shape).astype(np.float32)
    weight = rng.uniform(0, 100, shape).astype(np.float32)
    weight[0, :3] = 0  # masked pixels
    pri = {f"{oid}.{n}": repr(float(v)) for (oid, n), v in
           zip([(o, n) for o, name in objs for n in O.PARAMS[name]], p)}
    with Context(SystemSpec(objs, pri, image, weight, psf=psf), max_batch=4) as ctx:
        planes = ctx.render(p)
        pcs = [1.0 + 0.5 * (psf_shape[1] % 2 == 0), 1.0 + 0.5 * (psf_shape[0] % 2 == 0), 1.0, 1.0]
        raw, _ = O.render(objs, p, shape, PTS, WW64, pcs)
        rel, _ = rel_err(planes["value"], raw)
        assert rel <= 1e-4
        conv = O.convolve(planes["value"].astype(np.float64), psf.astype(np.float64))
        # FP32 sum of PW*PH taps in another order than the FP64 reference:
        # rounding grows like sqrt(taps) * eps; a wrong tap or index is O(1e-2)
        tol = 8 * np.sqrt(psf.size) * np.finfo(np.float32).eps
        assert np.max(np.abs(planes["image"] - conv)) <= tol * np.abs(conv).max()
        lnew, _ = ctx.loglike([])
        lref = -0.5 * O.chi2_f32terms(image, weight, planes["image"])
        assert abs(lnew - lref) <= 1e-6 * abs(lref) + 1e-6


def test_multi_system_launch_equals_single():
    rng = np.random.default_rng(5)
    specs, phys = [], []
    for n, lens, psf in [(64, "sie", None), (96, "epl", W.gaussian_psf(15, 3.0)), (128, "sie", W.gaussian_psf(31, 3.0))]:
        objs = [("lens", lens), ("src", "sersic")]
        names = [f"{o}.{p}" for o, nm in objs for p in O.PARAMS[nm]]
        base = {"lens.x": n / 2, "lens.y": n / 2, "lens.r": n / 6, "lens.q": 0.7, "lens.pa": 20.0, "lens.t": 1.1,
                "src.x": n / 2 + 1, "src.y": n / 2, "src.r": 3.0, "src.mag": 0.0, "src.n": 2.0, "src.q": 0.8,
                "src.pa": 5.0}
        pri = {k: f"unif {base[k] * 0.95 - 0.5} {base[k] * 1.05 + 0.5}" for k in names}
        pri["lens.q"] = "unif 0.5 0.9"
        img = rng.normal(0, 0.05, (n, n)).astype(np.float32)
        specs.append(SystemSpec(objs, pri, img, np.full_like(img, 400.0), psf=psf))
        phys.append(np.array([[base[k] * (1 + 0.01 * rng.normal()) for k in names] for _ in range(5)]))
    with Context(specs, max_batch=8) as ctx:
        together = ctx.loglike_phys(phys)
    for s, spec in enumerate(specs):
        with Context(spec, max_batch=8) as one:
            alone = one.loglike_phys(phys[s])
        assert np.array_equal(together[s], alone)


def test_graph_replay_matches_eager(c1):
    wl, pids = c1
    cubes = wl.candidates(12, 200)
    with Context(wl.spec(), max_batch=256) as eager:
        le, _ = eager.loglike_batch(cubes)
    with Context(wl.spec(), max_batch=256, flags=LENSED_GRAPH) as g:
        for rep in range(3):
            lg, _ = g.loglike_batch(cubes)
            assert np.array_equal(lg, le), rep
            lp, _ = g.loglike(cubes[rep])  # eager work between replays
            assert lp == le[rep]
        lg2, _ = g.loglike_batch(cubes[::-1].copy())
        assert np.array_equal(lg2, le[::-1])


def test_chunking_beyond_max_batch(c1):
    wl, _ = c1
    cubes = wl.candidates(12, 37)
    with Context(wl.spec(), max_batch=8) as ctx:
        l1, p1 = ctx.loglike_batch(cubes)
        l2 = np.array([ctx.loglike(c)[0] for c in cubes])
    assert np.array_equal(l1, l2)


def test_out_of_bounds_prior_gives_logzero(c1):
    wl, _ = c1
    pri = dict(wl.priors)
    pri["lens.q"] = "unif 0.5 1.5"  # partly outside the axis-ratio bounds (0, 1)
    with Context(wl.spec(pri), max_batch=64) as ctx:
        cubes = wl.candidates(12, 64)
        lnew, phys = ctx.loglike_batch(cubes)
        q = phys[:, ctx.pmap().index(3)]
        assert np.all(lnew[q > 1] == LOG_ZERO)
        assert np.all(np.isfinite(lnew[q <= 1]))
        assert (q > 1).any() and (q <= 1).any()


def test_nonfinite_parameters_do_not_poison_later_calls(c1):
    wl, pids = c1
    with Context(wl.spec(), max_batch=4) as ctx:
        good = np.array([wl.truth[k] for k in pids])
        l0 = ctx.loglike_phys(good[None])[0]
        bad = good.copy()
        bad[2] = np.nan
        ctx.loglike_phys(np.stack([bad, good]))
        assert ctx.loglike_phys(good[None])[0] == l0


def test_profile_stages(c1):
    wl, _ = c1
    with Context(wl.spec(), max_batch=16, flags=LENSED_PROFILE) as ctx:
        ctx.loglike_batch(wl.candidates(12, 16))
        ctx.loglike_batch(wl.candidates(12, 16))
        calls, ms = ctx.profile()
    assert calls == 2
    assert list(ms) == ["+params", "-params", "set_params", "render", "convolve", "loglike", "+loglike_mem",
                        "-loglike_mem"]
    assert ms["render"] > 0


@pytest.fixture(scope="module")
def c3():
    wl = W.c3()
    pids = [f"{oid}.{p}" for oid, n in wl.objects for p in O.PARAMS[n]]
    wl.make_data(oracle_render_cb, pids, mag_keys=("source.mag",))
    return wl, pids


def test_c3_full_size_batch_properties(c3):
    """256 x 256, 21x21 PSF, 4096 candidates per launch: size-independent
    properties of the whole batch plus oracle checks of a sample."""
    wl, pids = c3
    with Context(wl.spec(), max_batch=4096, flags=LENSED_GRAPH) as ctx:
        cubes = wl.candidates(ctx.ndims(), 4096)
        lnew, phys = ctx.loglike_batch(cubes)
        assert np.all(np.isfinite(lnew)) and np.all(lnew <= 0)
        perm = np.random.default_rng(0).permutation(4096)
        lperm, _ = ctx.loglike_batch(cubes[perm])
        assert np.array_equal(lperm, lnew[perm])  # no cross-candidate interference
        truth = np.array([wl.truth[k] for k in pids])
        lt = ctx.loglike_phys(truth[None])[0]
        n = wl.image.size
        assert abs(-2 * lt / n - 1) < 0.03  # chi^2/dof ~ 1 at the truth
        assert lt > lnew.max()
        pmap = ctx.pmap()
        for i in [0, 1, 2, 4095]:
            p = np.zeros(len(pids))
            p[pmap] = phys[i]
            model = oracle_model(wl.objects, p, wl.shape, wl.psf)
            lref = -0.5 * O.chi2_d(wl.image, wl.weight, model)
            assert abs(lnew[i] - lref) <= 1e-2 + 1e-5 * abs(lref), (i, lnew[i], lref)
        model = oracle_model(wl.objects, truth, wl.shape, wl.psf)
        assert abs(lt - (-0.5 * O.chi2_d(wl.image, wl.weight, model))) <= 1e-2
        planes = ctx.render(truth)
        rel, _ = rel_err(planes["image"], model)
        assert rel <= 1e-4, rel


def near_truth_phys(wl, pids, n, scale, seed=0):
    """Parameter vectors scattered around the truth (object order)."""
    rng = np.random.default_rng(seed)
    truth = np.array([wl.truth[k] for k in pids])
    out = truth + scale * rng.normal(size=(n, len(pids))) * np.maximum(np.abs(truth), 1.0)
    qi = pids.index("lens.q")
    out[:, qi] = np.clip(out[:, qi], 0.3, 1.0)
    return out


def test_c3_epl_table_matches_exact(c3):
    """The tabulated EPL angular factor against the per-ray series and the
    FP64 oracle: model images and logL near the truth."""
    from lensed_amd._lib import LENSED_TABLE
    wl, pids = c3
    phys = near_truth_phys(wl, pids, 64, 2e-4)
    cubes = wl.candidates(14, 256)
    with Context(wl.spec(), max_batch=256) as ex:
        le = ex.loglike_phys(phys)
        lc, _ = ex.loglike_batch(cubes)
    with Context(wl.spec(), max_batch=256, flags=LENSED_TABLE | LENSED_GRAPH) as tb:
        lt = tb.loglike_phys(phys)
        ltc, _ = tb.loglike_batch(cubes)
        single = tb.loglike(cubes[3])[0]
        truth = np.array([wl.truth[k] for k in pids])
        planes = tb.render(truth)
    assert np.max(np.abs(lt - le)) <= 1e-2, np.max(np.abs(lt - le))
    assert np.all(np.abs(ltc - lc) <= 1e-2 + 2e-5 * np.abs(lc))
    assert single == ltc[3]
    model = oracle_model(wl.objects, truth, wl.shape, wl.psf)
    rel, _ = rel_err(planes["image"], model)
    assert rel <= 1e-4, rel


def test_c3_culling_is_exact(c3):
    """Pixels whose rays provably all land beyond the source's cutoff are not
    traced (lensed_frame_cull): logL and model planes are bit for bit those
    of tracing every ray -- over prior draws, near the truth, and with the
    lens centre on a pixel centre (a ray through u = 0, t on both sides of
    1 and equal to it); and the culled launch is the faster one."""
    from lensed_amd._lib import LENSED_NO_CULL, LENSED_PROFILE, LENSED_TABLE
    wl, pids = c3
    phys = near_truth_phys(wl, pids, 64, 2e-4)
    cubes = wl.candidates(14, 1024)
    edge = np.array([wl.truth[k] for k in pids])[None].repeat(4, 0)
    edge[:, pids.index("lens.x")] = 128.0
    edge[:, pids.index("lens.y")] = 128.0
    edge[:, pids.index("lens.t")] = [0.6, 0.999, 1.0, 1.4]
    # a second batch of other candidates in the same launch layout: the cull
    # then writes zeros only where the first batch left nonzero values
    cubes2 = np.random.default_rng(5).random((1024, 14))
    res, render_ms, stats = [], [], []
    for extra in (0, LENSED_NO_CULL):
        with Context(wl.spec(), max_batch=1024, flags=LENSED_TABLE | LENSED_GRAPH | extra) as ctx:
            lc, _ = ctx.loglike_batch(cubes)
            stats.append(ctx.render_stats()[:3])
            lc2, _ = ctx.loglike_batch(cubes2)
            ln = ctx.loglike_phys(np.concatenate([phys, edge]))
            planes = [ctx.render(p) for p in (edge[0], edge[2], edge[3], phys[0])]
            lc3, _ = ctx.loglike_batch(cubes)
            res.append((np.concatenate([lc, lc2, lc3]), ln, planes))
        with Context(wl.spec(), max_batch=1024, flags=LENSED_TABLE | LENSED_PROFILE | extra) as ctx:
            ctx.loglike_batch(cubes)
            calls, ms = ctx.profile()
            render_ms.append(ms["render"] / calls)
    (lc, ln, pl), (lc0, ln0, pl0) = res
    assert np.all(np.isfinite(lc)) and np.all(np.isfinite(ln))
    assert np.array_equal(lc, lc0) and np.array_equal(ln, ln0)
    for a, b in zip(pl, pl0):
        for k in a:
            assert np.array_equal(a[k], b[k]), k
    print(f"render ms per launch: culled {render_ms[0]:.3f}, every ray {render_ms[1]:.3f}; "
          f"(items, kept, waves traced) {stats}", flush=True)
    assert render_ms[0] < render_ms[1], render_ms
    # 16 x 16 pixel items of sixteen 4 x 4 quads; without culling every quad
    # of every item is traced, with it a strict subset over prior draws
    items = 1024 * 256
    assert stats[1] == (items, items, 16 * items)
    assert stats[0][0] == items and 0 < stats[0][1] < items and stats[0][1] <= stats[0][2] <= 16 * stats[0][1]


@pytest.mark.parametrize("shape,psf_size", [((37, 53), 5), ((70, 33), 5), ((48, 40), 0)])
def test_culling_ragged_image_is_exact(shape, psf_size):
    """Images whose sides are not multiples of the 16-pixel render blocks:
    the edge blocks' pixels outside the image are neither traced nor
    written, and culled launches give bit for bit the logL and planes of
    tracing every ray, with and without a PSF."""
    from lensed_amd._lib import LENSED_NO_CULL, LENSED_TABLE
    h, w = shape
    cx, cy = (w + 1) / 2, (h + 1) / 2
    objs = [("lens", "epl_plus_shear"), ("source", "gauss")]
    pri = {"lens.x": f"unif {cx - 1} {cx + 1}", "lens.y": f"unif {cy - 1} {cy + 1}", "lens.r": "unif 4 8",
           "lens.t": "unif 0.5 1.5", "lens.q": "unif 0.3 1", "lens.pa": "unif 0 180", "lens.g1": "unif -0.1 0.1",
           "lens.g2": "unif -0.1 0.1", "source.x": f"unif {cx - 3} {cx + 3}", "source.y": f"unif {cy - 3} {cy + 3}",
           "source.sigma": "unif 0.3 0.8", "source.mag": "unif -5 5", "source.q": "unif 0.4 1",
           "source.pa": "unif 0 180"}
    rng = np.random.default_rng(h * w)
    image = rng.normal(0, 0.05, shape).astype(np.float32)
    weight = np.full(shape, 400.0, np.float32)
    psf = W.gaussian_psf(psf_size, 2.0) if psf_size else None
    cubes = rng.random((256, 14))
    res, stats = [], []
    for extra in (0, LENSED_NO_CULL):
        with Context(SystemSpec(objs, pri, image, weight, psf=psf), max_batch=256,
                     flags=LENSED_TABLE | LENSED_GRAPH | extra) as ctx:
            lc, phys = ctx.loglike_batch(cubes)
            stats.append(ctx.render_stats()[:3])
            p = np.zeros(14)
            p[ctx.pmap()] = phys[0]
            res.append((lc, ctx.render(p)))
    (lc, pl), (lc0, pl0) = res
    assert np.all(np.isfinite(lc)) and np.array_equal(lc, lc0)
    for k in pl:
        assert np.array_equal(pl[k], pl0[k]), k
    items = 256 * ((h + 15) // 16) * ((w + 15) // 16)
    assert stats[1] == (items, items, 16 * items)
    assert stats[0][0] == items and 0 < stats[0][1] < items, stats


def test_culling_two_systems_in_one_launch():
    """Two culled systems of different sizes (256 x 256 and 70 x 33) in one
    launch, two batches of other candidates in a row (the second trusts the
    cull's record of which squares the first wrote): logL bit for bit that
    of tracing every ray, for both systems and both batches."""
    from lensed_amd._lib import LENSED_NO_CULL, LENSED_TABLE
    big = W.c3()
    big.make_data(oracle_render_cb, [f"{oid}.{p}" for oid, n in big.objects for p in O.PARAMS[n]],
                  mag_keys=("source.mag",))
    h, w = 70, 33
    cx, cy = (w + 1) / 2, (h + 1) / 2
    objs = [("lens", "epl_plus_shear"), ("source", "gauss")]
    pri = {"lens.x": f"unif {cx - 1} {cx + 1}", "lens.y": f"unif {cy - 1} {cy + 1}", "lens.r": "unif 4 8",
           "lens.t": "unif 0.5 1.5", "lens.q": "unif 0.3 1", "lens.pa": "unif 0 180", "lens.g1": "unif -0.1 0.1",
           "lens.g2": "unif -0.1 0.1", "source.x": f"unif {cx - 3} {cx + 3}", "source.y": f"unif {cy - 3} {cy + 3}",
           "source.sigma": "unif 0.3 0.8", "source.mag": "unif -5 5", "source.q": "unif 0.4 1",
           "source.pa": "unif 0 180"}
    rng = np.random.default_rng(17)
    small = SystemSpec(objs, pri, rng.normal(0, 0.05, (h, w)).astype(np.float32),
                       np.full((h, w), 400.0, np.float32), psf=W.gaussian_psf(5, 2.0))
    specs = [big.spec(), small]
    out = []
    for extra in (0, LENSED_NO_CULL):
        with Context(specs, max_batch=256, flags=LENSED_TABLE | LENSED_GRAPH | extra) as ctx:
            batches = []
            for seed in (1, 2):
                r = np.random.default_rng(seed)
                phys = []
                for sidx in range(2):
                    cubes = r.random((256 if sidx == 0 else 192, ctx.ndims(sidx)))
                    _, p = ctx.loglike_batch(cubes, sidx)
                    full = np.zeros((len(p), ctx.npars(sidx)))
                    full[:, ctx.pmap(sidx)] = p
                    phys.append(full)
                batches.append(phys)
            # the two-system launches back to back, in the same layout
            out.append([ctx.loglike_phys(phys) for phys in batches])
    for (a0, a1), (b0, b1) in zip(zip(*out[0]), zip(*out[1])):
        assert np.all(np.isfinite(a0)) and np.array_equal(a0, b0)
        assert np.array_equal(a1, b1)
