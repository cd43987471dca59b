"""The C ABI (include/lensed.h): exports, priors, object metadata, program
generation and the plugin translation of unmodified reference objects.
No GPU is needed: hipRTC compiles for gfx950 without a device."""
import ctypes
import os
import re
from ctypes import POINTER, byref, c_char_p, c_size_t, c_void_p

import numpy as np
import pytest

from conftest import REFERENCE, ROOT
from lensed_amd import _lib as L
from lensed_amd.lensed import LensedObject, Prior
from oracle import oracle as O

OBJECTS = sorted(O.PARAMS)


def header_functions():
    text = open(os.path.join(ROOT, "include", "lensed.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"\btypedef\b[^;]*;", "", text)
    return sorted(set(re.findall(r"\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text)))


def test_library_exports_every_declared_function():
    lib = L.load()
    names = header_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
        assert n in L.SIGNATURES, n


def test_priors_match_oracle_and_closed_forms():
    u = np.linspace(0.001, 0.999, 97)
    p = Prior("norm 3 0.5")
    assert [p.apply(x) for x in u] == pytest.approx([3 + 0.5 * O.gauss_quantile(x) for x in u], abs=1e-15)
    assert (p.lower, p.upper) == (3 - 3.5, 3 + 3.5)
    assert not p.pseudo
    q = Prior("unif -2 6")
    assert [q.apply(x) for x in u] == pytest.approx(list(-2 + 8 * u), abs=1e-15)
    assert (q.lower, q.upper) == (-2.0, 6.0)
    d = Prior("1.25")
    assert d.pseudo and d.apply(0.3) == 1.25 and d.lower == d.upper == 1.25
    assert str(q) == "U(-2, 6)" and str(p) == "N(3, 0.5)"
    for bad in ("", "unif 1", "gauss 0 1", "norm a b"):
        with pytest.raises(Exception):
            Prior(bad)


@pytest.mark.parametrize("name", OBJECTS)
def test_object_metadata(name):
    o = LensedObject("obj", name)
    assert o.params == O.PARAMS[name]
    assert o.type == chr(O.lib().orc_type(O.lib().orc_kind(name.encode())))
    assert o.obj.size > 0
    defaults = [bool(o.obj.pars[i].pri) for i in range(o.obj.npars)]
    assert defaults == [(name, p) in O.DEFAULTS for p in O.PARAMS[name]]


def test_object_bounds_from_plugin_tables():
    epl = LensedObject("l", "epl_plus_shear")
    t = epl.obj.pars[epl.params.index("t")]
    assert (t.lower, t.upper) == (0.0, 2.0)
    old, new = LensedObject("s", "sersic-old"), LensedObject("s", "sersic")
    assert (old.obj.pars[4].lower, old.obj.pars[4].upper) == (0.5, 8.0)
    assert new.obj.pars[4].upper > 1e38


def test_missing_object_fails_cleanly():
    with pytest.raises(Exception, match="could not load object"):
        LensedObject("x", "no_such_object")


def test_kernel_options_and_name():
    lib = L.load()
    flags = (c_char_p * 3)(b"-cl-denorms-are-zero", b"-cl-fast-relaxed-math", None)
    s = L.take_string(lib.kernel_options(256, 200, 1, 21, 21, 49, flags))
    assert s.split() == ["-DIMAGE_SIZE=51200", "-DIMAGE_WIDTH=256", "-DIMAGE_HEIGHT=200", "-DPSF=1",
                         "-DPSF_WIDTH=21", "-DPSF_HEIGHT=21", "-DQUAD_POINTS=49",
                         "-cl-denorms-are-zero", "-cl-fast-relaxed-math"]
    assert L.take_string(lib.kernel_name(b"meta_", b"sie")) == "meta_sie"


def program_for(objects):
    lib = L.load()
    arr = (L.Object * len(objects))()
    for i, (oid, name) in enumerate(objects):
        assert lib.lensed_object_load(byref(arr[i]), oid.encode(), name.encode()) == 0, L.last_error()
    n = c_size_t()
    srcs = POINTER(c_char_p)()
    lib.main_program(len(objects), arr, byref(n), byref(srcs))
    assert n.value == 1, L.last_error()
    return srcs[0]


def compile_program(src):
    log = c_void_p()
    size = L.load().lensed_program_compile(src, b"gfx950", byref(log))
    text = L.take_string(log.value) if log.value else ""
    return size, text


@pytest.mark.parametrize("objects", [
    [("lens", "sie"), ("source", "sersic")],
    [("lens", "epl_plus_shear"), ("source", "gauss")],
    [("host", "devauc"), ("lens", "sie_plus_shear"), ("bulge", "sersic"), ("disk", "exponential"), ("sky", "sky")],
    [("lens", "nsie"), ("lens2", "point_mass"), ("src", "sersic-old"), ("fg", "sky")],
])
def test_main_program_compiles_for_gfx950(objects):
    src = program_for(objects)
    size, log = compile_program(src)
    assert size > 0, log
    assert b"lensed_render" in src and b"compute_0" in src


def test_object_program_has_metadata_symbols():
    lib = L.load()
    n = c_size_t()
    srcs = POINTER(c_char_p)()
    lib.object_program(b"sersic", byref(n), byref(srcs))
    assert n.value == 2
    assert b"lensed_meta_params" in srcs[1]


@pytest.mark.skipif(not os.path.isdir(os.path.join(REFERENCE, "objects")), reason="reference not present")
def test_reference_plugins_load_unmodified(monkeypatch):
    """Every objects/*.cl of the reference, unmodified, goes through the
    translator: metadata, a gfx950 build, and host renders that agree with
    this package's own objects and the FP64 oracle."""
    import host_model
    from lensed_amd import quadrature as Q

    monkeypatch.setenv("LENSED_OBJECTS_PATH", os.path.join(REFERENCE, "objects"))
    names = sorted(f[:-3] for f in os.listdir(os.path.join(REFERENCE, "objects")) if f.endswith(".cl"))
    assert len(names) == 15
    for name in names:
        o = LensedObject("obj", name)
        assert o.params == O.PARAMS[name], name
    qq, ww = Q.rule("g3k7")
    pts, wht, err = Q.table("g3k7")
    rng = np.random.default_rng(0)
    lenses = [n for n in names if O.lib().orc_type(O.lib().orc_kind(n.encode())) == ord("L")]
    sources = [n for n in names if O.lib().orc_type(O.lib().orc_kind(n.encode())) == ord("S")]
    cases = [[("l", ln), ("s", "gauss")] for ln in lenses] + [[("s", sn)] for sn in sources] + [[("f", "sky")]]
    values = {"x": 30.5, "y": 29.5, "r": 9.0, "rc": 1.5, "q": 0.7, "pa": 35.0, "g1": 0.04, "g2": -0.03,
              "t": 1.2, "mag": 0.0, "n": 2.0, "sigma": 3.0, "rs": 3.0, "bg": 1.0, "dx": 0.01, "dy": -0.02}
    for objs in cases:
        par = {f"{oid}.{p}": values[p] + (0.3 if oid == "s" and p in "xy" else 0.0)
               for oid, n in objs for p in O.PARAMS[n]}
        p = O.param_vector(objs, par)
        ref, _ = host_model.render(objs, p, (60, 60), qq, ww)
        exact, _ = O.render(objs, p, (60, 60), pts, np.stack([wht, err], 1))
        peak = np.abs(exact).max()
        assert np.max(np.abs(ref - exact)) <= 3e-5 * peak, objs
        monkeypatch.delenv("LENSED_OBJECTS_PATH")
        own, _ = host_model.render(objs, p, (60, 60), qq, ww)
        monkeypatch.setenv("LENSED_OBJECTS_PATH", os.path.join(REFERENCE, "objects"))
        assert np.max(np.abs(own - ref)) <= 3e-5 * peak, objs
    src = program_for([("lens", "epl_plus_shear"), ("src", "sersic-old"), ("fg", "sky")])
    assert compile_program(src)[0] > 0


def _tabulated_source(objects):
    lib = L.load()
    arr = (L.Object * len(objects))()
    for i, (oid, name) in enumerate(objects):
        assert lib.lensed_object_load(byref(arr[i]), oid.encode(), name.encode()) == 0, L.last_error()
    return L.take_string(lib.lensed_model_source_tabulated(len(objects), arr)), \
        L.take_string(lib.lensed_model_source(len(objects), arr))


def test_table_form_follows_the_data_layout(monkeypatch, tmp_path):
    """The tabulated EPL form reads the lens's data block by field name, so it
    is chosen by what the block declares, not by the plugin's name: this
    package's EPL objects get it; the same plugin with a renamed field, or
    under another name, gets (or keeps) the per-ray series."""
    objs = [("lens", "epl_plus_shear"), ("src", "gauss")]
    tab, plain = _tabulated_source(objs)
    assert "lensed_epl_table_deflection_pre" in tab or "lensed_epl_table_vector_pre" in tab
    assert "lensed_epl_pre_shear" in tab
    own = open(os.path.join(ROOT, "lensed_amd", "objects", "epl_plus_shear.cl")).read()
    # another name, same layout: tabulated
    (tmp_path / "my_power_law.cl").write_text(own)
    # the centre field renamed: the table form cannot read it
    (tmp_path / "epl_renamed.cl").write_text(own.replace("float2 c;", "float2 centre;")
                                             .replace("this->c ", "this->centre ").replace("this->c;", "this->centre;")
                                             .replace("this->c =", "this->centre ="))
    monkeypatch.setenv("LENSED_OBJECTS_PATH", str(tmp_path))
    tab2, _ = _tabulated_source([("lens", "my_power_law"), ("src", "gauss")])
    assert "lensed_epl_table_vector_pre" in tab2 and "lensed_epl_pre_shear" in tab2
    tab3, plain3 = _tabulated_source([("lens", "epl_renamed"), ("src", "gauss")])
    assert tab3 == plain3 and "lensed_epl_table" not in tab3.split("#line 1 \"generated\"")[1]


@pytest.mark.skipif(not os.path.isdir(os.path.join(REFERENCE, "objects")), reason="reference not present")
def test_reference_epl_plugin_keeps_the_series_under_table_mode(monkeypatch):
    """The reference's unmodified objects/epl_plus_shear.cl (fields x, m, w,
    g, t, f, n) asks for the table form under LENSED_TABLE: it gets the
    per-ray series, and the program builds for gfx950."""
    monkeypatch.setenv("LENSED_OBJECTS_PATH", os.path.join(REFERENCE, "objects"))
    objs = [("lens", "epl_plus_shear"), ("src", "gauss")]
    tab, plain = _tabulated_source(objs)
    assert tab == plain
    assert "lensed_epl_table" not in tab.split("#line 1 \"generated\"")[1]
    size, log = compile_program(program_for(objs))
    assert size > 0, log
