"""The tabulated EPL angular factor: the cubic-Hermite scheme of
lensed_amd/kernels/epl_table.h (the diamond angle p = |u_y| / (|u_x| +
|u_y|), 256 intervals, FP64 nodes of W = g^h Omega with g = p^2 + (1-p)^2,
h = (1-t)/2, so that |u|^(1-t) Omega = (|u_x| + |u_y|)^(1-t) W) restated in
numpy and checked against the oracle's tenth-order series, including the
reflection symmetries that map every quadrant to the first."""
import numpy as np
import pytest

from oracle import oracle as O

NT = 256


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
    h = (1 - t) / 2
    p = np.arange(NT + 1) / NT
    g = p ** 2 + (1 - p) ** 2
    ox, oy, gx, gy = omega_and_derivative(q, t, np.arctan2(p, 1 - p))
    gh, lg = g ** h, h * (4 * p - 2) / g
    tab = np.zeros((2, NT + 1, 4))
    for comp, (o, go) in enumerate(((ox, gx), (oy, gy))):
        v, d = gh * o, gh * (lg * o + go / g) / NT  # W and dW/df
        v0, v1, d0, d1 = v[:-1], v[1:], d[:-1], d[1:]
        tab[comp, :NT] = np.stack([v0, d0, 3 * (v1 - v0) - 2 * d0 - d1, 2 * (v0 - v1) + d0 + d1], 1)
        tab[comp, NT, 0] = v[-1]
    return tab.astype(np.float32)


def deflection_vector(tab, u, t):
    """|u|^(1-t) Omega from the table: s^(1-t) W(p), s = |u_x| + |u_y|."""
    au, av = np.abs(u[:, 0]), np.abs(u[:, 1])
    s = au + av
    x = av / s * NT
    i = x.astype(int)
    f = x - i
    c = tab[:, i]
    w = c[:, :, 0] + f * (c[:, :, 1] + f * (c[:, :, 2] + f * c[:, :, 3]))
    return s[:, None] ** (1 - t) * np.stack([np.copysign(w[0], u[:, 0]), np.copysign(w[1], u[:, 1])], 1)


def test_table_interpolation_matches_series():
    rng = np.random.default_rng(0)
    worst = 0.0
    for q, t in [(0.3, 0.5), (0.3, 1.5), (0.6, 1.0), (0.95, 1.2), (1.0, 0.8), (0.31, 1.49), (0.25, 1.9)]:
        tab = table(q, t)
        phi = rng.uniform(-np.pi, np.pi, 4000)
        r = rng.uniform(0.1, 3.0, 4000)
        u = np.stack([np.cos(phi), np.sin(phi)], 1) * r[:, None]
        got = deflection_vector(tab, u, t)
        exact = O.epl_omega(q, t, phi) * (r ** (1 - t))[:, None]
        worst = max(worst, np.max(np.abs(got - exact) / (r ** (1 - t))[:, None]))
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
    # and the tabulated W(p) = g^h Omega(psi(p)): its f-derivative a1 against
    # a central difference of the table's own values a0 one interval apart
    tab = table(q, t).astype(np.float64)
    fd = (tab[:, 2:, 0] - tab[:, :-2, 0]) / 2
    assert np.allclose(tab[:, 1:-1, 1], fd, atol=3e-5)


# ------------------------------------------------------------------------
# the product's table code (kernels/epl_table.h inside the generated model
# source) compiled for the host: lookups, the tau = 1 end row, u = 0, and the
# deflection from the pair (F (p - c) + F q, S (p - c) + S q) as the render
# kernel forms it, against the FP64 oracle's deflection

import ctypes  # noqa: E402
import hashlib  # noqa: E402
import os  # noqa: E402
import subprocess  # noqa: E402
import tempfile  # noqa: E402
from ctypes import POINTER, byref, c_double, c_float, c_int  # noqa: E402

from lensed_amd import _lib as L  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HARNESS = r'''
#define __device__
namespace hostseries {
using float4 = lcl::float4;
using ::sincos; using ::atan;
inline float4 make_float4(float a, float b, float c, float d) { return float4(a, b, c, d); }
@SERIES@
}
static void build_table(double q, double t, lcl::float4* tab) {
    for (int i = 0; i <= LENSED_EPL_NT; ++i) {
        double x0, y0, dx0, dy0, x1, y1, dx1, dy1;
        lcl::float4& ox = tab[LENSED_EPL_AT(0, i)];
        lcl::float4& oy = tab[LENSED_EPL_AT(1, i)];
        if (i == LENSED_EPL_NT) {
            hostseries::lensed_epl_node(q, t, 1.0, x1, y1, dx1, dy1);
            ox = lcl::float4((float)x1, 0.f, 0.f, 0.f);
            oy = lcl::float4((float)y1, 0.f, 0.f, 0.f);
            continue;
        }
        hostseries::lensed_epl_node(q, t, (double)i / LENSED_EPL_NT, x0, y0, dx0, dy0);
        hostseries::lensed_epl_node(q, t, (double)(i + 1) / LENSED_EPL_NT, x1, y1, dx1, dy1);
        ox = hostseries::lensed_hermite(x0, dx0, x1, dx1);
        oy = hostseries::lensed_hermite(y0, dy0, y1, dy1);
    }
}
// the lens-frame deflection vector |u|^(1-t) Omega of lensed_epl_table_vector_pre
struct HostEpl { float hexp; };
extern "C" void host_vector(double q, double t, const float* u, int n, float* out) {
    static lcl::float4 tab[LENSED_EPL_TABLE_F4];
    build_table(q, t, tab);
    const HostEpl e{(float)(0.5 * (1 - t))};
    for (int k = 0; k < n; ++k) {
        lcl::float2 v = lcl::lensed_epl_table_vector_pre(&e, lcl::float4(u[2 * k], u[2 * k + 1], 0.f, 0.f), tab);
        out[2 * k] = v.x; out[2 * k + 1] = v.y;
    }
}
// source-plane position of the ray at pixel p plus node q behind object 0
// (epl_plus_shear), from the split pair as the generated ray code forms it
extern "C" void host_trace(const float* params, const float* p, const float* qn, int n, float* out) {
    static lcl::uint data[4096];
    lcl::set_params_0(data, params);
    const auto* s = (const lcl::data_epl_plus_shear*)data;
    static lcl::float4 tab[LENSED_EPL_TABLE_F4];
    build_table(params[4], params[3], tab);
    for (int k = 0; k < n; ++k) {
        lcl::float2 px(p[2 * k], p[2 * k + 1]), qq(qn[2 * k], qn[2 * k + 1]);
        lcl::float4 b = lcl::lensed_epl_pre_shear(s, px - s->c, px), d = lcl::lensed_epl_pre_shear(s, qq, qq);
        lcl::float4 pre(b.x + d.x, b.y + d.y, b.z + d.z, b.w + d.w);
        lcl::float2 y = lcl::float2(pre.z, pre.w) - lcl::lensed_epl_table_deflection_pre(s, pre, tab);
        out[2 * k] = y.x; out[2 * k + 1] = y.y;
    }
}
'''


def _host_lib():
    lib = L.load()
    arr = (L.Object * 2)()
    for i, (oid, name) in enumerate([("lens", "epl_plus_shear"), ("src", "gauss")]):
        assert lib.lensed_object_load(byref(arr[i]), oid.encode(), name.encode()) == 0
    src = L.take_string(lib.lensed_model_source(2, arr))
    series = open(os.path.join(ROOT, "lensed_amd", "kernels", "epl_series.h")).read().replace("#pragma once", "")
    src += HARNESS.replace("@SERIES@", series)
    d = tempfile.mkdtemp(prefix="lensed_tab_")
    cpp, so = os.path.join(d, "t.cpp"), os.path.join(d, "t.so")
    open(cpp, "w").write(src)
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-w", "-o", so, cpp])
    h = ctypes.CDLL(so)
    fp = POINTER(c_float)
    h.host_vector.argtypes = [c_double, c_double, fp, c_int, fp]
    h.host_trace.argtypes = [fp, fp, fp, c_int, fp]
    return h


def test_product_table_lookup_on_the_host():
    h = _host_lib()
    rng = np.random.default_rng(4)
    fp = POINTER(c_float)
    for q, t in [(0.3, 0.5), (0.7, 1.1), (0.95, 1.45), (1.0, 1.0)]:
        phi = rng.uniform(-np.pi, np.pi, 3000)
        r = rng.uniform(0.05, 5.0, 3000)
        u = np.stack([r * np.cos(phi), r * np.sin(phi)], 1)
        # diagonals (p = 1/2), axes (p = 0 and the end row p = 1)
        u[:8] = [[1, 1], [-2, 2], [3, -3], [-0.5, -0.5], [1, 0], [0, -2], [-3, 0], [0, 4]]
        phi[:8] = np.arctan2(u[:8, 1], u[:8, 0])
        r[:8] = np.hypot(u[:8, 0], u[:8, 1])
        u = u.astype(np.float32)
        out = np.zeros_like(u)
        h.host_vector(q, t, u.ctypes.data_as(fp), len(u), out.ctypes.data_as(fp))
        exact = O.epl_omega(q, t, phi) * (r ** (1 - t))[:, None]
        err = np.max(np.abs(out - exact) / (r ** (1 - t))[:, None])
        assert err < 3e-6, (q, t, err)
    # u = 0: the deflection of the series, 0 for t < 1 and not finite beyond
    z = np.zeros((1, 2), np.float32)
    h.host_vector(0.6, 0.8, z.ctypes.data_as(fp), 1, z.ctypes.data_as(fp))
    assert np.all(z == 0)
    z = np.zeros((1, 2), np.float32)
    h.host_vector(0.6, 1.2, z.ctypes.data_as(fp), 1, z.ctypes.data_as(fp))
    assert not np.all(np.isfinite(z))


def test_split_pair_ray_tracing_matches_oracle():
    """The source-plane position x - S (x - c) - a_EPL(x) of a ray formed from
    pixel and node parts, against x - alpha(x) of the FP64 oracle."""
    h = _host_lib()
    rng = np.random.default_rng(5)
    fp = POINTER(c_float)
    objs = [("lens", "epl_plus_shear"), ("src", "gauss")]
    params = np.array([50.3, 49.7, 12.0, 1.2, 0.65, 35.0, 0.03, -0.02, 50.0, 50.0, 2.0, 0.0, 0.8, 10.0])
    p = rng.uniform(30, 70, (2000, 2)).astype(np.float32)
    qn = rng.uniform(-0.5, 0.5, (2000, 2)).astype(np.float32)
    out = np.zeros_like(p)
    h.host_trace(params.astype(np.float32).ctypes.data_as(fp), p.ctypes.data_as(fp), qn.ctypes.data_as(fp),
                 len(p), out.ctypes.data_as(fp))
    x = p.astype(np.float64) + qn.astype(np.float64)
    alpha = O.deflect(objs, params, 0, x)
    scale = np.max(np.abs(alpha))
    assert np.max(np.abs(out - (x - alpha))) < 2e-5 * scale, np.max(np.abs(out - (x - alpha))) / scale


# ------------------------------------------------------------------------
# the table form of the generated ray code (lensed_model_source_tabulated):
# with one framed source behind the lens, the source sees
# u = T (x - S (x - c) - c_s) - (T bwd) v; whole rays against the oracle

TABLE_HARNESS = r'''
// pixels the cull skips, and how many of their rays (nodes q[n]) give a
// brightness the device's FP32 arithmetic counts (must be none): with
// subnormals flushed, as the kernels compile, only a normal value adds to a
// pixel, so a brightness below FLT_MIN is 0 there
extern "C" void host_cull(const float* params, int w, int h, const float* q, int nq, float rho, int* out) {
    static lcl::uint data[4096];
    lcl::set_params_0(data, params);
    static lcl::float4 tab[LENSED_EPL_TABLE_F4];
    build_table(params[4], params[3], tab);
    lcl::float4 cc[LENSED_NPRE_CULL];
    lcl::cull_consts_0(data, cc);
    int skipped = 0, lit = 0;
    for (int k = 0; k < w * h; ++k) {
        const float cx = 1.0f + (float)(k % w), cy = 1.0f + (float)(k / w);
        const lcl::float4 b = lcl::epl_pre_0(data, lcl::float2(cx, cy), true);
        if (!lcl::cull_0(data, b, tab, cc, rho)) continue;
        ++skipped;
        for (int n = 0; n < nq; ++n) {
            const lcl::float4 d = lcl::epl_pre_0(data, lcl::float2(q[2 * n], q[2 * n + 1]), false);
            const lcl::float4 pre(b.x + d.x, b.y + d.y, b.z + d.z, b.w + d.w);
            if (std::fabs(lcl::compute_0(data, lcl::float2(cx + q[2 * n], cy + q[2 * n + 1]), tab, pre)) >= FLT_MIN)
                ++lit;
        }
    }
    out[0] = skipped;
    out[1] = lit;
}
extern "C" void host_rays(const float* params, int w, int h, float dx, float dy, float* out) {
    static lcl::uint data[4096];
    lcl::set_params_0(data, params);
    static lcl::float4 tab[LENSED_EPL_TABLE_F4];
    build_table(params[4], params[3], tab);
    const lcl::float4 node = lcl::epl_pre_0(data, lcl::float2(dx, dy), false);
    for (int k = 0; k < w * h; ++k) {
        const float cx = 1.0f + (float)(k % w), cy = 1.0f + (float)(k / w);
        lcl::float4 b = lcl::epl_pre_0(data, lcl::float2(cx, cy), true);
        lcl::float4 pre(b.x + node.x, b.y + node.y, b.z + node.z, b.w + node.w);
        out[k] = lcl::compute_0(data, lcl::float2(cx + dx, cy + dy), tab, pre);
    }
}
'''


def _host_table_lib(source):
    lib = L.load()
    arr = (L.Object * 2)()
    for i, (oid, name) in enumerate([("lens", "epl_plus_shear"), ("src", source)]):
        assert lib.lensed_object_load(byref(arr[i]), oid.encode(), name.encode()) == 0
    src = L.take_string(lib.lensed_model_source_tabulated(2, arr))
    assert "lensed_frame_trace" in src  # the lens-frame form was generated
    series = open(os.path.join(ROOT, "lensed_amd", "kernels", "epl_series.h")).read().replace("#pragma once", "")
    harness = HARNESS.split("extern \"C\" void host_omega")[0]
    src += harness.replace("@SERIES@", series) + TABLE_HARNESS
    d = tempfile.mkdtemp(prefix="lensed_frame_")
    cpp, so = os.path.join(d, "t.cpp"), os.path.join(d, "t.so")
    open(cpp, "w").write(src)
    subprocess.check_call(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-w", "-o", so, cpp])
    h = ctypes.CDLL(so)
    fp = POINTER(c_float)
    h.host_rays.argtypes = [fp, c_int, c_int, c_float, c_float, fp]
    h.host_cull.argtypes = [fp, c_int, c_int, fp, c_int, c_float, POINTER(c_int)]
    return h


@pytest.mark.parametrize("source,sparams", [
    ("gauss", [52.0, 47.5, 2.5, 0.0, 0.8, 20.0]),
    ("sersic", [52.0, 47.5, 3.0, 0.0, 2.5, 0.7, 60.0]),
    ("exponential", [51.0, 48.0, 3.0, 0.0, 0.6, 100.0]),
])
def test_lens_frame_rays_match_oracle(source, sparams):
    """One ray per pixel at node offset (dx, dy): the generated table form
    (pixel and node parts, the source's frame, T bwd applied to the
    lens-frame deflection) against the FP64 oracle's ray-shoot."""
    h = _host_table_lib(source)
    objs = [("lens", "epl_plus_shear"), ("src", source)]
    params = np.array([50.3, 49.7, 12.0, 1.2, 0.65, 35.0, 0.03, -0.02] + sparams)
    fp = POINTER(c_float)
    for dx, dy in [(0.0, 0.0), (0.31, -0.22)]:
        out = np.zeros((100, 100), np.float32)
        h.host_rays(params.astype(np.float32).ctypes.data_as(fp), 100, 100, dx, dy, out.ctypes.data_as(fp))
        ref, _ = O.render(objs, params, (100, 100), [[dx, dy]], [[1.0, 0.0]])
        assert np.max(np.abs(out - ref)) < 2e-5 * np.max(ref), (dx, dy, np.max(np.abs(out - ref)) / np.max(ref))


@pytest.mark.parametrize("t", [0.7, 1.3])
def test_lens_frame_ray_through_the_lens_centre(t):
    """A ray exactly at the lens centre (u = 0): the deflection is 0 for
    t < 1 and not finite for t > 1; the guard must send the latter out of
    the source (brightness 0, as the reference's rejected ray), not return
    a NaN."""
    h = _host_table_lib("gauss")
    objs = [("lens", "epl_plus_shear"), ("src", "gauss")]
    params = np.array([50.0, 50.0, 12.0, t, 0.65, 35.0, 0.03, -0.02, 50.5, 50.2, 2.5, 0.0, 0.8, 20.0])
    fp = POINTER(c_float)
    out = np.zeros((100, 100), np.float32)
    h.host_rays(params.astype(np.float32).ctypes.data_as(fp), 100, 100, 0.0, 0.0, out.ctypes.data_as(fp))
    ref, _ = O.render(objs, params, (100, 100), [[0.0, 0.0]], [[1.0, 0.0]])
    assert np.all(np.isfinite(out))
    assert out[49, 49] == pytest.approx(ref[49, 49], abs=1e-6 * np.max(ref))  # pixel (50, 50)
    assert np.max(np.abs(out - ref)) < 2e-5 * np.max(ref)


@pytest.mark.parametrize("source", ["gauss", "sersic", "exponential"])
def test_cull_bound_never_skips_a_lit_ray(source):
    """The pixel cull (lensed_frame_cull) is a proof that every ray of a
    skipped pixel adds exactly 0: a brightness below FLT_MIN, which the
    kernels' FP32 arithmetic (subnormals flushed) takes as 0; brute force on the host over many
    random lenses and sources (including pixels next to the lens centre and
    slopes on both sides of t = 1): every ray of every skipped pixel traced,
    none may be lit; and some pixels must be skipped."""
    from lensed_amd import quadrature as Q
    h = _host_table_lib(source)
    qq, _ = Q.rule("g3k7")
    qn = np.ascontiguousarray(np.asarray(qq, np.float32).reshape(-1, 2))
    rho = float(np.max(np.hypot(qn[:, 0], qn[:, 1]))) * (1 + 1e-6)
    rng = np.random.default_rng(11)
    fp = POINTER(c_float)
    out = (c_int * 2)()
    skipped = 0
    for it in range(12):
        lens = [rng.uniform(45, 56), rng.uniform(45, 56), rng.uniform(5, 15), rng.uniform(0.3, 1.7),
                rng.uniform(0.3, 1.0), rng.uniform(0, 180), rng.uniform(-0.1, 0.1), rng.uniform(-0.1, 0.1)]
        if it % 4 == 0:
            lens[0], lens[1] = 50.0, 50.0  # the lens centre on a pixel centre
        sx, sy = rng.uniform(40, 60), rng.uniform(40, 60)
        src = {"gauss": [sx, sy, rng.uniform(0.5, 2.5), rng.uniform(-5, 5), rng.uniform(0.4, 1), rng.uniform(0, 180)],
               "sersic": [sx, sy, rng.uniform(0.5, 3), rng.uniform(-5, 5), rng.uniform(0.5, 4), rng.uniform(0.4, 1),
                          rng.uniform(0, 180)],
               "exponential": [sx, sy, rng.uniform(0.3, 1.5), rng.uniform(-5, 5), rng.uniform(0.4, 1),
                               rng.uniform(0, 180)]}[source]
        params = np.array(lens + src, np.float32)
        h.host_cull(params.ctypes.data_as(fp), 100, 100, qn.ctypes.data_as(fp), len(qn), rho, out)
        assert out[1] == 0, (it, params, out[0], out[1])
        skipped += out[0]
    assert skipped > 100, skipped  # not a vacuous check (the exponential's zero radius is ~111 scale lengths)
