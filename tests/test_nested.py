"""The nested sampler behind run() and the batched driver (CPU likelihoods).

MultiNest itself is not part of the build and no reference test pins a
posterior or an evidence, so the sampler is pinned by what the mathematics
fixes: the evidence and moments of Gaussian likelihoods in closed form
(Z = (2 pi sigma^2)^(d/2) for a normalised-to-one-peak Gaussian well inside
the unit cube), the same for a mode across the seam of a periodic
parameter, and the MultiNest calling convention of src/lensed.c:1270-1275
(cube in/out, dumper constraints [MEAN | SIGMA | ML | MAP] x npar as
src/nested.c:137-166 reads them).
"""
import ctypes
import math
import os

import numpy as np
import pytest

from lensed_amd import _lib as L
from lensed_amd import nested

CENTRE = np.array([0.4, 0.5, 0.62])
SIGMA = np.array([0.05, 0.03, 0.04])


def gauss_logl(u):
    return -0.5 * np.sum(((u - CENTRE) / SIGMA) ** 2, axis=1)


def gauss_logz():
    return float(np.sum(0.5 * np.log(2 * np.pi * SIGMA**2)))


@pytest.mark.parametrize("batch", [1, 64])
def test_gaussian_evidence_and_moments(batch):
    r = nested.sample(gauss_logl, 3, nlive=400, tol=0.05, batch=batch, seed=3)
    assert abs(r.logz - gauss_logz()) < 4 * r.logz_err + 0.05, (r.logz, gauss_logz(), r.logz_err)
    assert np.all(np.abs(r.mean - CENTRE) < 0.25 * SIGMA), (r.mean, CENTRE)
    assert np.all(np.abs(r.sigma / SIGMA - 1) < 0.15), r.sigma / SIGMA
    assert np.all(np.abs(r.ml - CENTRE) < SIGMA)
    # the information H = -logZ - d/2 for this likelihood
    assert abs(r.info - (-gauss_logz() - 1.5)) < 0.3
    assert r.nevals >= r.niter


def test_batched_and_serial_agree_within_errors():
    a = nested.sample(gauss_logl, 3, nlive=400, tol=0.05, batch=1, seed=11)
    b = nested.sample(gauss_logl, 3, nlive=400, tol=0.05, batch=256, seed=12)
    err = math.hypot(a.logz_err, b.logz_err)
    assert abs(a.logz - b.logz) < 4 * err
    assert np.all(np.abs(a.mean - b.mean) < 0.3 * SIGMA)


def test_planted_mistake_in_likelihood_is_seen():
    # a likelihood wider by 10 % in one direction changes logZ by ln 1.1
    wide = SIGMA * np.array([1.1, 1, 1])
    r = nested.sample(lambda u: -0.5 * np.sum(((u - CENTRE) / wide) ** 2, axis=1), 3, nlive=400, tol=0.05,
                      batch=32, seed=5)
    assert abs(r.logz - gauss_logz()) > 0.05
    assert abs(r.logz - (gauss_logz() + math.log(1.1))) < 4 * r.logz_err + 0.05


def test_periodic_parameter_across_the_seam():
    c, s = 0.01, 0.04

    def ll(u):
        d = np.abs(u[:, 0] - c)
        d = np.minimum(d, 1 - d)
        return -0.5 * (d / s) ** 2

    r = nested.sample(ll, 1, nlive=300, tol=0.05, batch=16, seed=7, wrap=[1])
    assert abs(r.logz - 0.5 * math.log(2 * math.pi * s * s)) < 4 * r.logz_err + 0.05


def grid_logz(logl, n=2000):
    """ln of the mean of L over the unit square, by the midpoint rule."""
    g = (np.arange(n) + 0.5) / n
    x, y = np.meshgrid(g, g)
    ll = logl(np.stack([x.ravel(), y.ravel()], 1))
    return ll.max() + math.log(np.mean(np.exp(ll - ll.max())))


def eggbox(u):
    x = u * 10 * np.pi
    return (2 + np.cos(x[:, 0] / 2) * np.cos(x[:, 1] / 2)) ** 5


def two_shells(u):
    x = u * 12 - 6
    out = [-0.5 * ((np.hypot(x[:, 0] - c, x[:, 1]) - 2) / 0.1) ** 2 - 0.5 * math.log(2 * math.pi * 0.01)
           for c in (-3.5, 3.5)]
    return np.logaddexp(out[0], out[1])


@pytest.mark.parametrize("logl", [eggbox, two_shells], ids=["eggbox", "two_shells"])
def test_multimodal_and_curved_evidence(logl):
    """Feroz, Hobson & Bridges (2009) test problems: 18 modes, and two thin
    rings; the evidence against brute-force integration, and the bound made of
    several ellipsoids keeps the cost per iteration low (the median of three
    seeded runs: a single run of the eggbox can spend a phase with modes
    merged under one ellipsoid, at 25-30 evaluations per iteration, where
    the typical run needs 4)."""
    ref = grid_logz(logl)
    cost = []
    for seed in (4, 5, 6):
        r = nested.sample(logl, 2, nlive=500, tol=0.1, batch=256, seed=seed)
        assert abs(r.logz - ref) < 4 * r.logz_err + 0.05, (seed, r.logz, ref, r.logz_err)
        cost.append(r.nevals / r.niter)
    assert np.median(cost) < 20, cost


def correlated_gaussian(d=8):
    rng = np.random.default_rng(7)
    c = rng.uniform(0.35, 0.65, d)
    a = rng.normal(size=(d, d))
    cov = a @ a.T + d * np.eye(d)
    sd = rng.uniform(0.004, 0.02, d)
    cov = cov / np.sqrt(np.outer(np.diag(cov), np.diag(cov))) * np.outer(sd, sd)
    logz = 0.5 * np.linalg.slogdet(2 * np.pi * cov)[1]
    return c, sd, np.linalg.inv(cov), logz


@pytest.mark.parametrize("batch", [1, 1024])
def test_constant_efficiency_mode(batch):
    """Lensed's default sampler settings (ceff on, target acceptance 0.05,
    src/input/options.c:150-227): the acceptance settles at the target while
    the moments and, for this compact region, the evidence stay right."""
    c, sd, prec, logz = correlated_gaussian()
    r = nested.sample(lambda u: -0.5 * np.einsum("ij,jk,ik->i", u - c, prec, u - c), len(c), nlive=300, tol=0.1,
                      efr=0.05, ceff=True, batch=batch, seed=2)
    assert 0.04 < r.niter / r.nevals < 0.07, r.niter / r.nevals
    assert np.all(np.abs(r.mean - c) < 0.15 * sd), (r.mean - c) / sd
    assert np.all(np.abs(r.sigma / sd - 1) < 0.1), r.sigma / sd
    assert abs(r.logz - logz) < 4 * r.logz_err, (r.logz, logz, r.logz_err)


def test_constant_efficiency_thin_shell():
    """A thin spherical shell in 5-D: shrinking the bound towards its centre
    leaves the region, so steering the enlargement only ever downwards for
    more acceptance ends at almost none (such a run stalled at millions of
    evaluations per iteration).  Evidence: the shell's area x sqrt(2 pi) w."""
    d, radius, w = 5, 0.3, 0.002
    logz = math.log(2 * math.pi ** (d / 2) / math.gamma(d / 2) * radius ** (d - 1) * math.sqrt(2 * math.pi) * w)

    def shell(u):
        return -0.5 * ((np.sqrt(((u - 0.5) ** 2).sum(1)) - radius) / w) ** 2
    r = nested.sample(shell, d, nlive=400, tol=0.1, efr=0.05, ceff=True, batch=1024, seed=1)
    assert r.nevals / r.niter < 200, r.nevals / r.niter
    assert np.all(np.abs(r.mean - 0.5) < 0.02)
    assert abs(r.logz - logz) < 4 * r.logz_err, (r.logz, logz, r.logz_err)


def test_multinest_abi_run_and_dumper():
    """run() with MultiNest's callbacks: cube is unit in, physical out."""
    ndim, npar = 2, 3
    calls = {"loglike": 0, "dumper": 0, "last": None}

    def loglike(cube, pndim, pnpar, lnew, ctx):
        assert pndim[0] == ndim and pnpar[0] == npar
        u = np.array([cube[0], cube[1]])
        calls["loglike"] += 1
        lnew[0] = -0.5 * (((u[0] - 0.3) / 0.05) ** 2 + ((u[1] - 0.7) / 0.05) ** 2)
        cube[0] = 10 * u[0] - 5  # physical values written back, derived one last
        cube[1] = 4 * u[1]
        cube[2] = cube[0] + cube[1]

    def dumper(ns, nl, np_, physlive, post, cons, maxl, logz, inslogz, logzerr, ctx):
        n = np_[0]
        c = np.ctypeslib.as_array(cons[0], shape=(4 * n,)).copy()
        live = np.ctypeslib.as_array(physlive[0], shape=((n + 1) * nl[0],)).reshape(n + 1, nl[0])
        calls["dumper"] += 1
        calls["last"] = (c, maxl[0], logz[0], logzerr[0], live.copy(), ns[0])

    nested.run_multinest_abi(loglike, ndim, npar, dumper, nlive=300, tol=0.05, updint=200, seed=1)
    assert calls["dumper"] >= 2
    c, maxl, logz, logzerr, live, ns = calls["last"]
    mean, sig, ml = c[0:3], c[3:6], c[6:9]
    assert np.allclose(mean, [-2.0, 2.8, 0.8], atol=0.1)
    assert np.allclose(sig, [0.5, 0.2, math.hypot(0.5, 0.2)], rtol=0.15)
    assert ml[2] == pytest.approx(ml[0] + ml[1])
    assert maxl <= 0 and maxl > -1
    assert abs(logz - math.log(2 * math.pi * 0.05**2)) < 4 * logzerr + 0.05
    assert live.shape[1] == 300 and np.all(live[-1] <= 0)
    assert ns > 0 and calls["loglike"] > ns


def test_checkpoint_resume_and_posterior_file(tmp_path):
    root = str(tmp_path / "gauss-")
    part = nested.sample(gauss_logl, 3, nlive=200, tol=0.05, batch=8, seed=2, maxiter=400, updint=100,
                         root=root, outfile=True)
    assert part.niter == 400
    assert os.path.exists(root + "resume.dat")
    full = nested.sample(gauss_logl, 3, nlive=200, tol=0.05, batch=8, seed=2, resume=True, root=root,
                         outfile=True)
    assert full.niter > 400
    assert abs(full.logz - gauss_logz()) < 4 * full.logz_err + 0.05
    post = np.loadtxt(root + ".txt")
    assert post.shape[1] == 2 + 3
    assert post.shape[0] == full.niter
    # the dead points carry all but the final live points' share of the weight
    assert 0.5 < post[:, 0].sum() <= 1.0 + 1e-9
    assert np.allclose(post[:, 1], -2 * gauss_logl(post[:, 2:]))


def test_bad_arguments_raise():
    with pytest.raises(RuntimeError):
        nested.sample(gauss_logl, 3, nlive=2)


def test_abi_symbols_present():
    lib = L.load()
    for name in ("run", "lensed_nested_sample", "lensed_nested_run"):
        assert hasattr(lib, name)
    assert ctypes.sizeof(L.NsOptions) >= 48
