"""The tabulated EPL angular factor: the cubic-Hermite scheme of
lensed_amd/kernels/epl_table.h (two octants in tau = min/max of |u|, 128
intervals each, FP64 nodes) restated in numpy and checked against the
oracle's tenth-order series, including the reflection symmetries that map
every quadrant to the first."""
import numpy as np

from oracle import oracle as O

NT = 128


def omega_and_derivative(q, t, psi):
    f, T = (1 - q) / (1 + q), 2 - t
    c, ox, oy, dx, dy = 1.0, 0.0, 0.0, 0.0, 0.0
    for k in range(11):
        if k:
            c *= -f * (2 * k - T) / (2 * k + T)
        m = 2 * k + 1
        ox += c * np.cos(m * psi)
        oy += c * np.sin(m * psi)
        dx -= c * m * np.sin(m * psi)
        dy += c * m * np.cos(m * psi)
    return ox, oy, dx, dy


def table(q, t):
    tab = np.zeros((2, NT, 2, 4))
    for half in range(2):
        tau = np.arange(NT + 1) / NT
        a = np.arctan(tau)
        psi = np.pi / 2 - a if half else a
        ox, oy, gx, gy = omega_and_derivative(q, t, psi)
        dpsi = (-1.0 if half else 1.0) / (1 + tau ** 2) / NT
        for comp, (v, g) in enumerate(((ox, gx * dpsi), (oy, gy * dpsi))):
            v0, v1, d0, d1 = v[:-1], v[1:], g[:-1], g[1:]
            tab[half, :, comp] = np.stack([v0, d0, 3 * (v1 - v0) - 2 * d0 - d1, 2 * (v0 - v1) + d0 + d1], 1)
    return tab.astype(np.float32)


def lookup(tab, u):
    au, av = np.abs(u[:, 0]), np.abs(u[:, 1])
    tau = np.minimum(au, av) / np.maximum(au, av)
    s = tau * NT
    i = np.minimum(s.astype(int), NT - 1)
    f = s - i
    half = (av > au).astype(int)
    c = tab[half, i]
    ox = c[:, 0, 0] + f * (c[:, 0, 1] + f * (c[:, 0, 2] + f * c[:, 0, 3]))
    oy = c[:, 1, 0] + f * (c[:, 1, 1] + f * (c[:, 1, 2] + f * c[:, 1, 3]))
    return np.stack([np.copysign(ox, u[:, 0]), np.copysign(oy, u[:, 1])], 1)


def test_table_interpolation_matches_series():
    rng = np.random.default_rng(0)
    worst = 0.0
    for q, t in [(0.3, 0.5), (0.3, 1.5), (0.6, 1.0), (0.95, 1.2), (1.0, 0.8), (0.31, 1.49)]:
        tab = table(q, t)
        phi = rng.uniform(-np.pi, np.pi, 4000)
        u = np.stack([np.cos(phi), np.sin(phi)], 1) * rng.uniform(0.1, 3.0, (4000, 1))
        got = lookup(tab, u)
        exact = O.epl_omega(q, t, phi)
        worst = max(worst, np.max(np.abs(got - exact)))
    # float32 coefficients: a few ulp of |Omega| <= ~3
    assert worst < 2e-6, worst


def test_series_derivative_is_consistent():
    # the derivative used for the Hermite nodes is the derivative of the series
    q, t, h = 0.4, 1.2, 1e-6
    psi = np.linspace(0.01, 1.5, 50)
    _, _, dx, dy = omega_and_derivative(q, t, psi)
    a = O.epl_omega(q, t, psi + h)
    b = O.epl_omega(q, t, psi - h)
    assert np.allclose((a - b)[:, 0] / (2 * h), dx, atol=1e-7)
    assert np.allclose((a - b)[:, 1] / (2 * h), dy, atol=1e-7)
