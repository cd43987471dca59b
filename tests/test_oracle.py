"""Pin the CPU oracle before trusting it.

Which check pins which oracle function:
  render + objects + compute   the reference's own golden images (16 fixtures,
                               IMG layer) and their quadrature-error layers (ERR)
  EPL series (epl, epl_plus_shear)
                               t = 1 equals the SIE closed form; q = 1 equals the
                               circular power law; the deflection field is
                               curl-free with divergence 2*kappa of the EPL
  SIE / NSIE                   q -> 1 limit is the SIS; curl-free field with the
                               SIE convergence; rc = 0 NSIE equals SIE
  SIS / NSIS / point mass      closed forms
  source normalisations        total flux = 10^(-0.4 mag) for circular profiles
  Sersic b(n)                  the root of its defining equation
                               Gamma(2n, b)/Gamma(2n) = 1/2 (scipy), within the
                               series' own error; n = 1/2 Sersic is a Gaussian
  set_params image-plane prior traced position equals x - alpha(x)
  convolve                     scipy convolution of an edge-padded image,
                               delta PSF = identity, odd and even PSF sizes
  chi^2                        numpy sums
  norm prior quantile          scipy's exact quantile within the A&S bound 4.5e-4
"""
import numpy as np
import pytest
import scipy.signal
import scipy.special
import scipy.stats

from conftest import fixture_objects
from lensed_amd import quadrature as Q
from oracle import oracle as O

G3K7 = Q.table("g3k7")
QQ = G3K7[0]
WW = np.stack([G3K7[1], G3K7[2]], 1)


def golden_cases(golden):
    meta, _ = golden
    return sorted(meta)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_oracle_reproduces_reference_renders(golden, dtype):
    meta, arrays = golden
    worst = {}
    for t, m in sorted(meta.items()):
        objs = fixture_objects(m)
        p = O.param_vector(objs, m["params"])
        value, error = O.render(objs, p, m["shape"], QQ, WW, dtype=dtype)
        img, err = arrays[t + "/img"], arrays[t + "/err"]
        peak = np.abs(img).max()
        mask = np.abs(img) > 1e-3 * peak
        rel = np.max(np.abs(value - img)[mask] / np.abs(img[mask]))
        worst[t] = rel
        # the reference renders are FP32 with fast-relaxed OpenCL maths;
        # atan/atanh of the SIE family differ most (up to 3.0e-5 of the peak)
        assert np.max(np.abs(value - img)) <= 3e-5 * peak, t
        assert rel <= 8e-5, (t, rel)
        # ERR layer = quadrature error / value
        assert np.max(np.abs(error[mask] / value[mask] - err[mask])) <= 2e-5, t


def test_current_sersic_differs_from_stale_fixture(golden):
    # the current MacArthur b(n) does not reproduce the stale source/sersic
    # fixture: a check that the old/new switch is real
    meta, arrays = golden
    m = meta["source/sersic"]
    objs = [("source", "sersic")]
    value, _ = O.render(objs, O.param_vector(objs, m["params"]), m["shape"], QQ, WW)
    img = arrays["source/sersic/img"]
    assert np.max(np.abs(value - img)) / np.abs(img).max() > 1e-4


def rng_points(n=200, seed=7, lo=20.0, hi=80.0):
    return np.random.default_rng(seed).uniform(lo, hi, size=(n, 2))


def test_epl_isothermal_equals_sie():
    pts = rng_points()
    for q, pa in [(0.8, 30.0), (0.5, 120.0), (0.3, 75.0)]:
        epl = O.deflect([("l", "epl")], [50.5, 49.0, 20.0, 1.0, q, pa], 0, pts)
        sie = O.deflect([("l", "sie")], [50.5, 49.0, 20.0, q, pa], 0, pts)
        f = (1 - q) / (1 + q)
        # the series stops at tenth order: truncation ~ f^11
        tol = 50 * f ** 11 + 1e-12
        assert np.max(np.abs(epl - sie)) / 20.0 <= tol, (q, np.max(np.abs(epl - sie)))


def test_epl_round_is_power_law():
    pts = rng_points()
    for t in (0.5, 0.8, 1.3, 1.5):
        a = O.deflect([("l", "epl")], [50.0, 50.0, 15.0, t, 1.0, 10.0], 0, pts)
        d = pts - 50.0
        r = np.hypot(d[:, 0], d[:, 1])
        expect = (15.0 * (r / 15.0) ** (1 - t))[:, None] * d / r[:, None]
        assert np.allclose(a, expect, rtol=1e-12, atol=1e-12)


def divergence_and_curl(objs, params, pts, h=1e-5):
    ax = lambda p: O.deflect(objs, params, 0, p)
    ex, ey = np.array([h, 0.0]), np.array([0.0, h])
    dxa = (ax(pts + ex) - ax(pts - ex)) / (2 * h)
    dya = (ax(pts + ey) - ax(pts - ey)) / (2 * h)
    return dxa[:, 0] + dya[:, 1], dya[:, 0] - dxa[:, 1]


def test_epl_field_is_gradient_with_epl_convergence():
    pts = rng_points(100, seed=3)
    cx, cy, b, q, pa = 50.5, 49.5, 18.0, 0.75, 35.0
    for t in (0.6, 1.0, 1.4):
        div, curl = divergence_and_curl([("l", "epl")], [cx, cy, b, t, q, pa], pts)
        c, s = np.cos(np.radians(pa)), np.sin(np.radians(pa))
        d = pts - [cx, cy]
        X, Y = c * d[:, 0] + s * d[:, 1], -s * d[:, 0] + c * d[:, 1]
        R = np.sqrt(q * X ** 2 + Y ** 2 / q)
        kappa = (2 - t) / 2 * (b / R) ** t
        assert np.max(np.abs(curl)) < 1e-6
        assert np.allclose(div, 2 * kappa, rtol=1e-5, atol=1e-7), t


def test_sie_limits_and_convergence():
    pts = rng_points(100, seed=5)
    sis = O.deflect([("l", "sis")], [50.0, 50.0, 12.0], 0, pts)
    sie = O.deflect([("l", "sie")], [50.0, 50.0, 12.0, 1 - 1e-9, 40.0], 0, pts)
    assert np.allclose(sie, sis, rtol=1e-6, atol=1e-7)
    nsie = O.deflect([("l", "nsie")], [50.0, 50.0, 12.0, 0.0, 0.6, 40.0], 0, pts)
    sie6 = O.deflect([("l", "sie")], [50.0, 50.0, 12.0, 0.6, 40.0], 0, pts)
    assert np.allclose(nsie, sie6, rtol=1e-12, atol=1e-12)
    div, curl = divergence_and_curl([("l", "sie")], [50.0, 50.0, 12.0, 0.6, 40.0], pts)
    assert np.max(np.abs(curl)) < 1e-6
    # isothermal: kappa = b sqrt(q) / (2 sqrt(q^2 X^2 + Y^2))
    c, s = np.cos(np.radians(40.0)), np.sin(np.radians(40.0))
    d = pts - 50.0
    X, Y = c * d[:, 0] + s * d[:, 1], -s * d[:, 0] + c * d[:, 1]
    kappa = 12.0 * np.sqrt(0.6) / (2 * np.sqrt(0.36 * X ** 2 + Y ** 2))
    assert np.allclose(div, 2 * kappa, rtol=1e-5)


def test_spherical_lenses_closed_form():
    pts = rng_points(50, seed=11)
    d = pts - [40.0, 60.0]
    r = np.hypot(d[:, 0], d[:, 1])[:, None]
    sis = O.deflect([("l", "sis")], [40.0, 60.0, 9.0], 0, pts)
    assert np.allclose(sis, 9.0 * d / r)
    pm = O.deflect([("l", "point_mass")], [40.0, 60.0, 9.0], 0, pts)
    assert np.allclose(pm, 81.0 * d / r ** 2)
    nsis = O.deflect([("l", "nsis")], [40.0, 60.0, 9.0, 2.0], 0, pts)
    assert np.allclose(nsis, 9.0 * d / (2.0 + r))
    sh = O.deflect([("l", "sis_plus_shear")], [40.0, 60.0, 9.0, 0.05, -0.03], 0, pts)
    shear = np.stack([0.05 * d[:, 0] - 0.03 * d[:, 1], -0.03 * d[:, 0] - 0.05 * d[:, 1]], 1)
    assert np.allclose(sh, 9.0 * d / r + shear)


@pytest.mark.parametrize("name,size", [("gauss", 3.0), ("exponential", 2.0), ("devauc", 0.5),
                                       ("sersic", 3.0), ("sersic-old", 3.0)])
def test_source_total_flux(name, size):
    # circular profiles well inside a 160 x 160 image: sum of pixel values
    # equals the luminosity 10^(-0.4 mag)
    mag = 1.5
    if name.startswith("sersic"):
        p = [80.5, 80.5, size, mag, 1.0, 1.0, 0.0]
    else:
        p = [80.5, 80.5, size, mag, 1.0, 0.0]
    value, _ = O.render([("s", name)], p, (160, 160), QQ, WW)
    tol = {"devauc": 3e-3}.get(name, 1e-4)
    assert value.sum() == pytest.approx(10 ** (-0.4 * mag), rel=tol)


def sersic_b_exact(n):
    """The root b of Gamma(2n, b) / Gamma(2n) = 1/2, the definition of b(n)
    (half the light within the effective radius, docs/sources.md:109-111)."""
    return scipy.special.gammainccinv(2 * n, 0.5)


def test_sersic_b_solves_its_defining_equation():
    # objects/sersic.cl:32-33 uses the asymptotic series of Ciotti & Bertin
    # (1999) for n > 0.36.  Its error against the exact root falls like n^-3
    # (its n^-3 coefficient 131/1148175 differs from the exact expansion's,
    # Choi's 281/3444525, by 3.3e-5), so the bound is 4.5e-5 / n^3 from n = 1
    # on and 3e-4 below; at n = 8..100 that is tight enough that changing any
    # coefficient's sign, or the leading ones by 1e-3 of their value, fails.
    for n in np.concatenate([np.linspace(0.37, 1.0, 30), np.linspace(1.0, 8.0, 50), [12.0, 20.0, 50.0, 100.0]]):
        err = O.sersic_b(n) - sersic_b_exact(n)
        bound = 4.5e-5 / n ** 3 if n >= 1 else 3e-4
        assert abs(err) <= bound, (n, err, bound)
    # n <= 0.36: the polynomial fit of MacArthur, Courteau & Holtzman (2003),
    # good to 1.2 % of b down to n = 0.2
    for n in np.linspace(0.2, 0.36, 17):
        err = O.sersic_b(n) - sersic_b_exact(n)
        assert abs(err) <= 0.012 * sersic_b_exact(n) and abs(err) <= 2e-3, (n, err)
    # the two branches meet within their errors
    assert abs(O.sersic_b(0.36) - O.sersic_b(0.3600001)) < 5e-4


def test_sersic_b_bound_catches_a_planted_error():
    # the check above, run on b(n) with one coefficient of the series
    # altered, fails: it pins the coefficients, not only the leading 2n - 1/3
    def series(n, c):
        return 2 * n - 1 / 3 + c[0] / n + c[1] / n ** 2 + c[2] / n ** 3 + c[3] / n ** 4

    good = [4 / 405, 46 / 25515, 131 / 1148175, -2194697 / 30690717750]
    ns = np.concatenate([np.linspace(1.0, 8.0, 50), [12.0, 20.0, 50.0, 100.0]])

    def passes(c):
        return all(abs(series(n, c) - sersic_b_exact(n)) <= 4.5e-5 / n ** 3 for n in ns)

    assert passes(good)
    assert abs(series(2.5, good) - O.sersic_b(2.5)) < 1e-13
    for k in range(4):
        bad = list(good)
        bad[k] = -bad[k]
        assert not passes(bad), k
    for k in range(2):
        bad = list(good)
        bad[k] *= 1 + 1e-3
        assert not passes(bad), k


def test_sersic_half_is_gaussian():
    # Sersic n = 1/2 is a Gaussian with sigma^2 = r^2/(2 b), b = b(1/2)
    objs = [("s", "sersic")]
    n, r = 0.5, 4.0
    v, _ = O.render(objs, [40.5, 40.5, r, 0.0, n, 1.0, 0.0], (80, 80), QQ, WW)
    sig = r / np.sqrt(2 * O.sersic_b(n))
    g, _ = O.render([("g", "gauss")], [40.5, 40.5, sig, 0.0, 1.0, 0.0], (80, 80), QQ, WW)
    assert np.allclose(v, g, rtol=1e-9, atol=1e-14)


def test_image_plane_prior_traces_back():
    objs = [("lens", "sie"), ("src", "gauss")]
    lens = [50.0, 50.0, 15.0, 0.7, 20.0]
    ximg = np.array([62.0, 55.0])
    alpha = O.deflect(objs, lens + [0, 0, 2.0, 0.0, 1.0, 0.0], 0, ximg)[0]
    beta = ximg - alpha
    a, _ = O.render(objs, lens + [ximg[0], ximg[1], 2.0, 0.0, 1.0, 0.0], (100, 100), QQ, WW, ipp=[0, 1])
    b, _ = O.render(objs, lens + [beta[0], beta[1], 2.0, 0.0, 1.0, 0.0], (100, 100), QQ, WW)
    assert np.allclose(a, b, rtol=1e-12, atol=1e-15)


def test_sky_gradient_and_unlensed_foreground():
    v, _ = O.render([("f", "sky")], [3.0, 0.1, -0.2], (7, 9), QQ, WW)
    y, x = np.mgrid[1:8, 1:10]
    assert np.allclose(v, 3.0 + 0.1 * (x - 1) - 0.2 * (y - 1))
    # a foreground listed after a lens is not deflected
    w, _ = O.render([("l", "sis"), ("f", "sky")], [4.0, 4.0, 2.0, 3.0, 0.1, -0.2], (7, 9), QQ, WW)
    assert np.allclose(w, v)


@pytest.mark.parametrize("pw,ph", [(1, 1), (3, 3), (5, 3), (4, 4), (21, 21), (6, 5)])
def test_convolve_matches_padded_scipy(pw, ph):
    rng = np.random.default_rng(pw * 100 + ph)
    img = rng.normal(size=(23, 31))
    psf = rng.random((ph, pw))
    psf /= psf.sum()
    out = O.convolve(img, psf)
    # output(y, x) = sum psf[j, i] in(clamp(y + cy - j), clamp(x + cx - i))
    cy, cx = ph - 1 - ph // 2, pw - 1 - pw // 2
    pad = np.pad(img, ((ph, ph), (pw, pw)), mode="edge")
    full = scipy.signal.convolve2d(pad, psf, mode="full")
    ref = full[ph + cy: ph + cy + img.shape[0], pw + cx: pw + cx + img.shape[1]]
    assert np.allclose(out, ref, rtol=1e-12, atol=1e-12)


def test_convolve_delta_is_identity():
    img = np.random.default_rng(1).normal(size=(16, 20))
    psf = np.zeros((5, 5))
    psf[2, 2] = 1.0
    assert np.array_equal(O.convolve(img, psf), img)


def test_chi2_sums():
    rng = np.random.default_rng(2)
    img, w, m = rng.normal(size=(3, 40, 50))
    w = np.abs(w)
    assert O.chi2_d(img, w, m) == pytest.approx(np.sum(w * (m - img) ** 2), rel=1e-13)
    f32 = [a.astype(np.float32) for a in (img, w, m)]
    terms = (f32[1] * (f32[2] - f32[0]) * (f32[2] - f32[0])).astype(np.float64)
    assert O.chi2_f32terms(*f32) == pytest.approx(terms.sum(), rel=1e-13)


def test_gauss_quantile_within_abramowitz_stegun_bound():
    u = np.linspace(1e-6, 1 - 1e-6, 2001)
    got = np.array([O.gauss_quantile(x) for x in u])
    assert np.max(np.abs(got - scipy.stats.norm.ppf(u))) < 4.5e-4
    assert O.gauss_quantile(0.5) == pytest.approx(0.0, abs=1e-5)
    assert O.gauss_quantile(0.2) == pytest.approx(-O.gauss_quantile(0.8), abs=1e-15)
