"""The uneven multi-system batch (C4) and the 512 x 512 Moffat-PSF case (C5)
at full size on the GPU.

C4: 64 systems of 64..512 px, SIE or EPL + Sersic, 15 x 15 or 31 x 31 PSFs,
256 candidates each, all in one launch.  Each system's logL must equal, bit
for bit, what a context holding only that system computes (the launch mixes
item sizes but never work across systems), the FP64 oracle within 1e-2 at
the truth and near it on the small systems, and chi^2/dof ~ 1 at the truth.

C5: 512 x 512, EPL + shear, two Sersic components and sky, 31 x 31 Moffat
PSF (rank 9: the direct-convolution path), 2048 candidates per launch: the
C3 properties at this size (finite, no cross-candidate interference,
chi^2/dof ~ 1 at the truth, the FP64 oracle within 1e-2 for a sample of
candidates and the truth, model images within 1e-4 relative), and the EPL
table path, like the per-ray series, within 1e-2 of the FP64 oracle near
the truth.
"""
import time

import numpy as np
import pytest

from lensed_amd import workloads as W
from lensed_amd._lib import LENSED_GRAPH, LENSED_TABLE
from lensed_amd.lensed import Context, SystemSpec
from oracle import oracle as O
from test_gpu_parity import near_truth_phys, oracle_model, oracle_render_cb, rel_err

pytestmark = pytest.mark.gpu


def pids_of(wl):
    return [f"{oid}.{p}" for oid, n in wl.objects for p in O.PARAMS[n]]


def gpu_render_cb(objects, params, shape, psf):
    names = [f"{oid}.{p}" for oid, n in objects for p in O.PARAMS[n]]
    pri = {k: repr(float(v)) for k, v in zip(names, params)}
    spec = SystemSpec(objects, pri, np.zeros(shape, np.float32), np.ones(shape, np.float32), psf=psf)
    with Context(spec, max_batch=1) as c:
        return c.render(params)["image"].astype(np.float64)


@pytest.fixture(scope="module")
def c4():
    wls = W.c4()
    for wl in wls:
        wl.make_data(gpu_render_cb, pids_of(wl), mag_keys=("source.mag",))
    return wls


def test_c4_uneven_batch_in_one_launch(c4):
    wls = c4
    phys = [near_truth_phys(wl, pids_of(wl), 256, 2e-4, seed=s) for s, wl in enumerate(wls)]
    for s, wl in enumerate(wls):
        phys[s][0] = wl.truth_vector(pids_of(wl))
    with Context([wl.spec() for wl in wls], max_batch=256, flags=LENSED_GRAPH) as ctx:
        together = ctx.loglike_phys(phys)
        t0 = time.perf_counter()
        reps = 3
        for _ in range(reps):
            ctx.loglike_phys(phys)
        dt = (time.perf_counter() - t0) / reps
    npix = sum(wl.image.size for wl in wls)
    print(f"C4: 64 systems, {npix / 1e6:.2f} Mpx, 256 candidates each: {dt * 1e3:.1f} ms per launch, "
          f"{64 * 256 / dt:.0f} logL evals/s", flush=True)
    for s, wl in enumerate(wls):
        lt = together[s][0]
        n = wl.image.size
        assert np.all(np.isfinite(together[s])), s
        assert abs(-2 * lt / n - 1) < 6 * np.sqrt(2 / n), (s, -2 * lt / n)
    sizes = [wl.shape[0] for wl in wls]
    # one of each size and lens kind, alone: bit-identical logL
    picks = []
    for size in (64, 128, 256, 512):
        for lens in ("sie", "epl"):
            idx = [s for s, wl in enumerate(wls) if sizes[s] == size and wl.objects[0][1] == lens]
            picks += idx[:1]
    for s in picks:
        with Context(wls[s].spec(), max_batch=256) as one:
            alone = one.loglike_phys(phys[s])
        assert np.array_equal(together[s], alone), s
    # the FP64 oracle on the small systems: truth and two near-truth candidates
    for s in [s for s in range(64) if sizes[s] <= 128][:6]:
        wl = wls[s]
        for i in (0, 1, 2):
            model = oracle_model(wl.objects, phys[s][i], wl.shape, wl.psf)
            lref = -0.5 * O.chi2_d(wl.image, wl.weight, model)
            assert abs(together[s][i] - lref) <= 1e-2, (s, i, together[s][i], lref)


def test_c4_table_mode_matches_exact(c4):
    wls = c4
    phys = [near_truth_phys(wl, pids_of(wl), 32, 2e-4, seed=100 + s) for s, wl in enumerate(wls)]
    with Context([wl.spec() for wl in wls], max_batch=32) as ex:
        le = ex.loglike_phys(phys)
    with Context([wl.spec() for wl in wls], max_batch=32, flags=LENSED_TABLE | LENSED_GRAPH) as tb:
        lt = tb.loglike_phys(phys)
    for s, wl in enumerate(wls):
        if wl.objects[0][1] == "sie":
            assert np.array_equal(lt[s], le[s]), s  # no EPL: the same kernels
        else:
            assert np.max(np.abs(lt[s] - le[s])) <= 1e-2, (s, np.max(np.abs(lt[s] - le[s])))


@pytest.fixture(scope="module")
def c5():
    wl = W.c5()
    pids = pids_of(wl)
    wl.make_data(oracle_render_cb, pids, mag_keys=("bulge.mag", "disk.mag"))
    return wl, pids


def test_c5_full_size_batch(c5):
    wl, pids = c5
    K = 2048
    with Context(wl.spec(), max_batch=K, flags=LENSED_GRAPH) as ctx:
        cubes = wl.candidates(ctx.ndims(), K)
        lnew, phys = ctx.loglike_batch(cubes)
        t0 = time.perf_counter()
        ctx.loglike_batch(cubes)
        dt = time.perf_counter() - t0
        print(f"C5: 512x512, 31x31 Moffat, {K} candidates: {dt * 1e3:.1f} ms per launch, {K / dt:.0f} logL evals/s",
              flush=True)
        assert np.all(np.isfinite(lnew)) and np.all(lnew <= 0)
        perm = np.random.default_rng(1).permutation(K)
        lperm, _ = ctx.loglike_batch(cubes[perm])
        assert np.array_equal(lperm, lnew[perm])
        truth = wl.truth_vector(pids)
        lt = ctx.loglike_phys(truth[None])[0]
        n = wl.image.size
        assert abs(-2 * lt / n - 1) < 0.02
        assert lt > lnew.max()
        pmap = ctx.pmap()
        for i in (0, K - 1):
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


def test_c5_table_mode_near_truth(c5):
    """Near the truth both FP32 paths (per-ray series and table) are held to
    the FP64 oracle at 1e-2, and to each other at 1e-2 (the table's own
    interpolation error, 1.5e-8 on the angular factor, is below FP32
    rounding)."""
    wl, pids = c5
    phys = near_truth_phys(wl, pids, 64, 2e-4, seed=9)
    # the sky gradients (truth 0, in flux per pixel) stay at the truth: a
    # 2e-4 gradient over 512 px is 5 noise sigmas, not "near the truth"
    for k in ("sky.dx", "sky.dy"):
        phys[:, pids.index(k)] = wl.truth[k]
    with Context(wl.spec(), max_batch=64) as ex:
        le = ex.loglike_phys(phys)
    with Context(wl.spec(), max_batch=64, flags=LENSED_TABLE | LENSED_GRAPH) as tb:
        lt = tb.loglike_phys(phys)
    for i in (0, 1, 2, 3):
        model = oracle_model(wl.objects, phys[i], wl.shape, wl.psf)
        lref = -0.5 * O.chi2_d(wl.image, wl.weight, model)
        print(f"C5 candidate {i}: series {le[i] - lref:+.2e}, table {lt[i] - lref:+.2e} vs FP64", flush=True)
        assert abs(le[i] - lref) <= 1e-2 and abs(lt[i] - lref) <= 1e-2, (i, le[i] - lref, lt[i] - lref)
    assert np.max(np.abs(lt - le)) <= 1e-2, np.max(np.abs(lt - le))


def prior_draws(wl, n, seed):
    rng = np.random.default_rng(seed)
    return np.array([[rng.uniform(*wl.prior_range(k)) for k in pids_of(wl)] for _ in range(n)])


def test_c4_culling_is_exact(c4):
    """The pixel cull in a launch that mixes tested systems (tabulated EPL +
    Sersic) with systems it never tests (SIE, listed whole): every system's
    logL bit for bit that of tracing every ray, over prior draws.  A Sersic
    profile of these indices reaches exact zero only tens of effective
    radii out, beyond these images, so here every item stays listed."""
    from lensed_amd._lib import LENSED_NO_CULL
    wls = c4
    phys = [prior_draws(wl, 32, 200 + s) for s, wl in enumerate(wls)]
    out, stats = [], []
    for extra in (0, LENSED_NO_CULL):
        with Context([wl.spec() for wl in wls], max_batch=32, flags=LENSED_TABLE | LENSED_GRAPH | extra) as ctx:
            out.append(ctx.loglike_phys(phys))
            stats.append(ctx.render_stats()[:3])
    for s in range(len(wls)):
        assert np.all(np.isfinite(out[0][s])), s
        assert np.array_equal(out[0][s], out[1][s]), s
    items = sum(32 * (wl.shape[0] // 16) * (wl.shape[1] // 16) for wl in wls)
    assert stats[1] == (items, items, 16 * items)
    sie = sum(32 * (wl.shape[0] // 16) * (wl.shape[1] // 16) for wl in wls if wl.objects[0][1] == "sie")
    assert stats[0][0] == items and sie <= stats[0][1] <= items, (stats, sie)
