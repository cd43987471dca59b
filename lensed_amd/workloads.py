"""Synthetic workloads of the benchmark configurations.

Lensed units are pixels (docs/units.md:12-22); arcsec values of the
configurations are converted at 0.05"/px.  Mock data are the forward model
at the true parameters plus seeded Gaussian noise; the weight is 1/sigma^2
(the stated noise model, rather than Lensed's Poisson weights
gain/(image + offset) of src/data.c:314-330).  Seeds: mock noise 1..5 for
C1..C5, candidates 101..105.

`render(objects, params, shape, psf)` is supplied by the caller and returns
the noiseless model image (tests pass the CPU oracle, the benchmark the
device path), so this module computes nothing itself.
"""
import math

import numpy as np

from .lensed import SystemSpec

ARCSEC = 1.0 / 0.05  # pixels per arcsec


def gaussian_psf(size, fwhm):
    sig = fwhm / (2 * math.sqrt(2 * math.log(2)))
    c = (size - 1) / 2
    y, x = np.mgrid[0:size, 0:size]
    p = np.exp(-((x - c) ** 2 + (y - c) ** 2) / (2 * sig ** 2))
    return (p / p.sum()).astype(np.float32)  # unit sum, as src/data.c:363-369


def moffat_psf(size, fwhm, beta):
    alpha = fwhm / (2 * math.sqrt(2 ** (1 / beta) - 1))
    c = (size - 1) / 2
    y, x = np.mgrid[0:size, 0:size]
    p = (1 + ((x - c) ** 2 + (y - c) ** 2) / alpha ** 2) ** (-beta)
    return (p / p.sum()).astype(np.float32)


def param_order(objects, names_of):
    return [f"{oid}.{p}" for oid, name in objects for p in names_of(name)]


class Workload:
    """Objects, truth, priors and data of one system, plus candidate draws."""

    def __init__(self, name, objects, truth, priors, shape, sigma, psf=None, seed=1, cand_seed=101,
                 ncand=1, fixed=(), mag_window=None):
        self.name, self.objects, self.truth, self.priors = name, objects, dict(truth), dict(priors)
        self.shape, self.sigma, self.psf = shape, sigma, psf
        self.seed, self.cand_seed, self.ncand = seed, cand_seed, ncand
        self.fixed = set(fixed)
        self.mag_window = mag_window
        self.image = self.weight = None

    def truth_vector(self, param_ids):
        return np.array([self.truth[k] for k in param_ids])

    def make_data(self, render, param_ids, peak_sigmas=10.0, mag_keys=()):
        """Scale the listed magnitudes so the brightest pixel is peak_sigmas
        noise sigmas, then add seeded noise.  With a magnitude window w, the
        listed magnitudes' priors become uniform over the truth +- w."""
        p = self.truth_vector(param_ids)
        model = render(self.objects, p, self.shape, self.psf)
        if mag_keys:
            peak = float(np.max(model))
            dm = -2.5 * math.log10(peak_sigmas * self.sigma / peak)
            for k in mag_keys:
                self.truth[k] += dm
                if self.mag_window:
                    t, w = self.truth[k], self.mag_window
                    self.priors[k] = f"unif {t - w:.6f} {t + w:.6f}"
            p = self.truth_vector(param_ids)
            model = render(self.objects, p, self.shape, self.psf)
        rng = np.random.default_rng(self.seed)
        self.image = (model + rng.normal(0.0, self.sigma, self.shape)).astype(np.float32)
        self.weight = np.full(self.shape, 1.0 / self.sigma ** 2, np.float32)
        return self

    def spec(self, priors=None):
        pri = dict(self.priors if priors is None else priors)
        for k in self.fixed:
            pri[k] = repr(float(self.truth[k]))
        return SystemSpec(self.objects, pri, self.image, self.weight, psf=self.psf)

    def prior_range(self, key):
        """(lower, upper) of a `[wrap] [image] unif lower upper` prior."""
        words = [w for w in self.priors[key].split() if w not in ("wrap", "image")]
        assert words[0] == "unif", self.priors[key]
        return float(words[1]), float(words[2])

    def candidates(self, ndim, n=None):
        """Seeded unit-cube draws (MultiNest's input)."""
        rng = np.random.default_rng(self.cand_seed)
        return rng.random((n or self.ncand, ndim))


def c1():
    """SIE + Sersic, 100 x 100, g3k7, no PSF, sigma 0.05."""
    objects = [("lens", "sie"), ("source", "sersic")]
    truth = {"lens.x": 50.5, "lens.y": 50.5, "lens.r": 1.0 * ARCSEC, "lens.q": 0.8, "lens.pa": 30.0,
             "source.x": 51.3, "source.y": 49.6, "source.r": 0.2 * ARCSEC, "source.mag": 0.0,
             "source.n": 2.5, "source.q": 0.7, "source.pa": 60.0}
    priors = {"lens.x": "unif 49.5 51.5", "lens.y": "unif 49.5 51.5", "lens.r": "unif 18 22",
              "lens.q": "unif 0.6 0.99", "lens.pa": "unif 0 90", "source.x": "unif 49 53",
              "source.y": "unif 48 52", "source.r": "unif 2 6", "source.mag": "unif -8 8",
              "source.n": "unif 1 4", "source.q": "unif 0.4 1", "source.pa": "unif 0 180"}
    return Workload("C1", objects, truth, priors, (100, 100), 0.05, seed=1, cand_seed=101, ncand=1)


def c2():
    """SIE + shear, Sersic bulge + exponential disk, sky; 200 x 200; 21x21 PSF."""
    objects = [("lens", "sie_plus_shear"), ("bulge", "sersic"), ("disk", "exponential"), ("sky", "sky")]
    truth = {"lens.x": 100.5, "lens.y": 100.5, "lens.r": 20.0, "lens.q": 0.75, "lens.pa": 50.0,
             "lens.g1": 0.04, "lens.g2": -0.03,
             "bulge.x": 101.5, "bulge.y": 99.5, "bulge.r": 3.0, "bulge.mag": 0.0, "bulge.n": 3.0,
             "bulge.q": 0.8, "bulge.pa": 30.0,
             "disk.x": 101.5, "disk.y": 99.5, "disk.rs": 5.0, "disk.mag": 0.5, "disk.q": 0.5, "disk.pa": 100.0,
             "sky.bg": 0.01, "sky.dx": 0.0, "sky.dy": 0.0}
    priors = {"lens.x": "unif 99 102", "lens.y": "unif 99 102", "lens.r": "unif 15 25", "lens.q": "unif 0.4 0.99",
              "lens.pa": "unif 0 180", "lens.g1": "unif -0.1 0.1", "lens.g2": "unif -0.1 0.1",
              "bulge.x": "unif 98 105", "bulge.y": "unif 96 103", "bulge.r": "unif 1 6", "bulge.mag": "unif -5 5",
              "bulge.n": "unif 1 5", "bulge.q": "unif 0.4 1", "bulge.pa": "unif 0 180",
              "disk.x": "unif 98 105", "disk.y": "unif 96 103", "disk.rs": "unif 2 10", "disk.mag": "unif -5 5",
              "disk.q": "unif 0.2 1", "disk.pa": "unif 0 180", "sky.bg": "unif -0.1 0.1"}
    return Workload("C2", objects, truth, priors, (200, 200), 0.05, psf=gaussian_psf(21, 3.0), seed=2,
                    cand_seed=102, ncand=1024)


def c3():
    """EPL + shear, Gaussian source (sigma 0.1"), 256 x 256, 21x21 PSF, 4096 candidates."""
    objects = [("lens", "epl_plus_shear"), ("source", "gauss")]
    truth = {"lens.x": 128.5, "lens.y": 128.5, "lens.r": 20.0, "lens.t": 1.1, "lens.q": 0.7, "lens.pa": 40.0,
             "lens.g1": 0.03, "lens.g2": -0.02,
             "source.x": 130.0, "source.y": 127.5, "source.sigma": 0.1 * ARCSEC, "source.mag": 0.0,
             "source.q": 0.8, "source.pa": 20.0}
    priors = {"lens.x": "unif 126 131", "lens.y": "unif 126 131", "lens.r": "unif 15 25",
              "lens.t": "unif 0.5 1.5", "lens.q": "unif 0.3 1", "lens.pa": "wrap unif 0 180",
              "lens.g1": "unif -0.1 0.1", "lens.g2": "unif -0.1 0.1",
              "source.x": "unif 125 135", "source.y": "unif 122 132", "source.sigma": "unif 1.5 2.5",
              "source.mag": "unif -5 5", "source.q": "unif 0.4 1", "source.pa": "wrap unif 0 180"}
    return Workload("C3", objects, truth, priors, (256, 256), 0.05, psf=gaussian_psf(21, 3.0), seed=3,
                    cand_seed=103, ncand=4096)


def c5():
    """EPL + shear, two Sersic components and sky, 512 x 512, 31x31 Moffat PSF.

    The priors are those of a follow-up fit: a few pixels around the lens and
    source centres, sizes, indices and axis ratios over a factor of about
    two, magnitudes within 1.5 of the truth (set by make_data), position
    angles periodic over the whole half turn, as Lensed's examples declare
    them (examples/full_mock_psf.ini:42,49).  With magnitudes free over ten
    and positions over eight pixels, the bulk of the prior is a plateau of
    faint or misplaced sources (logL within ~1700 of the best fit, the sky
    alone fitted): after 6000 iterations of 400 live points the live set
    still spanned nearly the whole cube in 21 of 23 dimensions and the
    bounding ellipsoid accepted 3e-4 of its draws."""
    objects = [("lens", "epl_plus_shear"), ("bulge", "sersic"), ("disk", "sersic"), ("sky", "sky")]
    truth = {"lens.x": 256.5, "lens.y": 256.5, "lens.r": 40.0, "lens.t": 0.9, "lens.q": 0.75, "lens.pa": 70.0,
             "lens.g1": -0.02, "lens.g2": 0.04,
             "bulge.x": 258.0, "bulge.y": 255.0, "bulge.r": 4.0, "bulge.mag": 0.0, "bulge.n": 3.5,
             "bulge.q": 0.8, "bulge.pa": 10.0,
             "disk.x": 258.0, "disk.y": 255.0, "disk.r": 9.0, "disk.mag": 0.5, "disk.n": 1.0,
             "disk.q": 0.5, "disk.pa": 120.0, "sky.bg": 0.005, "sky.dx": 0.0, "sky.dy": 0.0}
    priors = {"lens.x": "unif 255.5 257.5", "lens.y": "unif 255.5 257.5", "lens.r": "unif 37 43",
              "lens.t": "unif 0.7 1.1", "lens.q": "unif 0.6 0.9", "lens.pa": "wrap unif 0 180",
              "lens.g1": "unif -0.1 0.1", "lens.g2": "unif -0.1 0.1",
              "bulge.x": "unif 257 259", "bulge.y": "unif 254 256", "bulge.r": "unif 2 6", "bulge.mag": "unif -5 5",
              "bulge.n": "unif 2.5 4.5", "bulge.q": "unif 0.6 1", "bulge.pa": "wrap unif 0 180",
              "disk.x": "unif 257 259", "disk.y": "unif 254 256", "disk.r": "unif 6 12", "disk.mag": "unif -5 5",
              "disk.n": "unif 0.6 1.6", "disk.q": "unif 0.3 0.7", "disk.pa": "wrap unif 0 180",
              "sky.bg": "unif -0.01 0.02"}
    return Workload("C5", objects, truth, priors, (512, 512), 0.02, psf=moffat_psf(31, 3.0, 3.0), seed=5,
                    cand_seed=105, ncand=2048, mag_window=1.5)


def c4(nsys=64, ncand=256):
    """64 independent systems of uneven size, SIE or EPL + Sersic, 15x15 or 31x31 PSF."""
    rng = np.random.default_rng(4)
    out = []
    for s in range(nsys):
        n = int(rng.choice([64, 128, 256, 512]))
        lens = "sie" if rng.random() < 0.5 else "epl"
        c = (n + 1) / 2
        objects = [("lens", lens), ("source", "sersic")]
        r = n / 8
        truth = {"lens.x": c + rng.uniform(-1, 1), "lens.y": c + rng.uniform(-1, 1), "lens.r": r,
                 "lens.q": rng.uniform(0.5, 0.9), "lens.pa": rng.uniform(0, 180),
                 "source.x": c + rng.uniform(-2, 2), "source.y": c + rng.uniform(-2, 2),
                 "source.r": r / 5, "source.mag": 0.0, "source.n": rng.uniform(1, 4),
                 "source.q": rng.uniform(0.5, 1), "source.pa": rng.uniform(0, 180)}
        priors = {"lens.x": f"unif {c - 2} {c + 2}", "lens.y": f"unif {c - 2} {c + 2}",
                  "lens.r": f"unif {0.8 * r} {1.2 * r}", "lens.q": "unif 0.3 0.99", "lens.pa": "unif 0 180",
                  "source.x": f"unif {c - 4} {c + 4}", "source.y": f"unif {c - 4} {c + 4}",
                  "source.r": f"unif {r / 10} {r / 2.5}", "source.mag": "unif -5 5", "source.n": "unif 1 4",
                  "source.q": "unif 0.4 1", "source.pa": "unif 0 180"}
        if lens == "epl":
            truth["lens.t"] = rng.uniform(0.7, 1.3)
            priors["lens.t"] = "unif 0.5 1.5"
        psf = gaussian_psf(15 if rng.random() < 0.5 else 31, 3.0)
        out.append(Workload(f"C4.{s}", objects, truth, priors, (n, n), 0.05, psf=psf, seed=400 + s,
                            cand_seed=1400 + s, ncand=ncand))
    return out


CONFIGS = {"C1": c1, "C2": c2, "C3": c3, "C5": c5}
