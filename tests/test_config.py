"""Configuration loading (CPU): INI files and --name=value options.

Behaviour follows src/input/ini.c:109-284 (groups, `=`/`:` assignment, `;`
and `#` comments, <object>.<param> names, one lens plane),
src/input/options.c (types, defaults, abbreviations resolved in table
order, boolean flags without a value, conditional requirements),
src/input/objects.c:328-359 (`wrap` and `image` keywords before a prior)
and src/input.c:214-239 (the final checks).  Objects are compiled from the
plugins with hipRTC (no GPU needed).
"""
import os

import pytest

from lensed_amd import config as C
from lensed_amd import data as D

HERE = os.path.dirname(os.path.abspath(__file__))

INI = """; a comment line
image  = ../img.fits     ; relative to this file
weight = 1000
root   = out/run-
# another comment

[objects]
lens   = sie
source : sersic

[priors]
lens.x     = unif 49 52
lens.y     = norm 50.5 0.5
lens.r     = unif 15 25
lens.q     = unif 0.5 0.95
lens.pa    = wrap unif 0 180
source.x   = image unif 48 53
source.y   = image unif 48 53
source.r   = unif 2 6
source.mag = unif -8 0
source.n   = unif 0.5 4
source.q   = unif 0.4 1
source.pa  = wrap unif 0 180

[labels]
lens.x  = x_L
lens.pa = \\theta_L
"""


@pytest.fixture
def ini(tmp_path):
    d = tmp_path / "cfg"
    d.mkdir()
    p = d / "run.ini"
    p.write_text(INI)
    return p


def pars(cfg):
    return {f"{o.id}.{o.obj.pars[j].name.decode()}": o.obj.pars[j] for o in cfg.objects for j in range(o.obj.npars)}


def test_reads_groups_objects_priors_labels(ini, tmp_path):
    cfg = C.Config(str(ini))
    cfg.check()
    assert [(o.id, o.name) for o in cfg.objects] == [("lens", "sie"), ("source", "sersic")]
    assert cfg.get("image") == str(tmp_path / "cfg") + "/../img.fits"
    assert cfg.get("weight") == "1000"  # a number stays a number
    assert cfg.get("root") == "out/run-"  # a string option, not a path
    p = pars(cfg)
    assert p["lens.pa"].wrap == 1 and p["lens.pa"].ipp == 0
    assert p["source.x"].ipp == 1 and p["source.x"].wrap == 0
    assert p["lens.x"].label == b"x_L" and p["lens.pa"].label == b"\\theta_L"
    from lensed_amd import _lib as L
    lib = L.load()
    assert lib.prior_lower(p["lens.r"].pri) == 15 and lib.prior_upper(p["lens.r"].pri) == 25
    assert lib.prior_apply(p["lens.y"].pri, 0.5) == pytest.approx(50.5)


def test_defaults_and_option_table():
    cfg = C.Config()
    expect = {"device": "auto", "output": "true", "rule": "g3k7", "nlive": "300", "ins": "true", "mmodal": "true",
              "ceff": "true", "acc": "0.05", "tol": "0.1", "shf": "0.8", "maxmodes": "100", "feedback": "false",
              "updint": "100", "seed": "-1", "resume": "false", "maxiter": "0", "offset": "0", "bscale": "0"}
    for k, v in expect.items():
        assert cfg.get(k) == v, k
    for k in ("root", "image", "gain", "weight", "xweight", "mask", "psf"):
        assert cfg.get(k) is None
    names = [r[0] for r in C.option_table()]
    assert names[:4] == ["device", "output", "root", "image"] and len(names) == 25


def test_command_line_options_override_and_abbreviate(ini):
    cfg = C.Config(str(ini), args=["--nl=500", "--rule=g5k11", "--resume", "--ceff=false", "--o=0"])
    assert cfg.get("nlive") == "500"       # prefix of the first matching option
    assert cfg.get("rule") == "g5k11"
    assert cfg.get("resume") == "true"     # a flag without a value
    assert cfg.get("ceff") == "false"
    assert cfg.get("output") == "false"    # 'o' selects 'output' (table order), not 'offset'
    with pytest.raises(C.ConfigError):
        cfg.arg("--nlive=many")
    with pytest.raises(C.ConfigError):
        cfg.arg("--nosuch=1")
    with pytest.raises(C.ConfigError):
        cfg.arg("--rule")                  # non-flags need a value


def test_requirements(tmp_path):
    objs = "[objects]\nsrc = gauss\n[priors]\nsrc.x = 1\nsrc.y = 1\nsrc.sigma = 1\nsrc.mag = 1\nsrc.q = 1\nsrc.pa = 0\n"
    def cfg_of(text):
        p = tmp_path / "r.ini"
        p.write_text(text + objs)
        return C.Config(str(p))
    with pytest.raises(C.ConfigError, match="missing required option: root"):
        cfg_of("image = a.fits\nweight = 1\n").check()          # output defaults to true
    with pytest.raises(C.ConfigError, match="missing required option: image"):
        cfg_of("output = false\nweight = 1\n").check()
    with pytest.raises(C.ConfigError, match="missing required option: gain"):
        cfg_of("output = false\nimage = a.fits\n").check()      # no weight: gain needed
    cfg_of("output = false\nimage = a.fits\ngain = 2\n").check()
    cfg_of("output = false\nimage = a.fits\nweight = 0\n").check()  # a weight of 0 is still given
    cfg_of("image = a.fits\nweight = 1\nroot = x\n").check()
    p = tmp_path / "np.ini"
    p.write_text("output = false\nimage = a.fits\nweight = 1\n[objects]\nsrc = gauss\n[priors]\nsrc.x = 1\n")
    with pytest.raises(C.ConfigError, match="missing prior: src.y"):
        C.Config(str(p)).check()
    p.write_text("output = false\nimage = a.fits\nweight = 1\n")
    with pytest.raises(C.ConfigError, match="no objects"):
        C.Config(str(p)).check()


@pytest.mark.parametrize("text,msg", [
    ("[nosuch]\n", "unknown group"),
    ("[objects\n", "missing closing"),
    ("[objects]\nlens = sie\nlens = sis\n", "duplicate object"),
    ("[objects]\nlens = sie\nsrc = gauss\nlens2 = sis\n", "multiple lensing planes"),
    ("[objects]\nlens = sie\n[priors]\nother.x = 1\n", "unknown object"),
    ("[objects]\nlens = sie\n[priors]\nlens.z = 1\n", "unknown parameter"),
    ("[objects]\nlens = sie\n[priors]\nlens = 1\n", "invalid parameter name"),
    ("[objects]\nlens = sie\n[priors]\nlens.x = unif 3\n", "invalid prior"),
    ("just words\n", "does not assign"),
    ("nlive = x\n", "invalid value"),
    ("[objects]\nlens = nosuchobject\n", "object lens"),
])
def test_errors_name_the_line(tmp_path, text, msg):
    p = tmp_path / "bad.ini"
    p.write_text(text)
    with pytest.raises(C.ConfigError, match=msg) as e:
        C.Config(str(p))
    assert "bad.ini:" in str(e.value)


def test_foreground_does_not_open_a_new_plane(tmp_path):
    p = tmp_path / "fg.ini"
    p.write_text("[objects]\nlens = sie\nsky = sky\nsrc = gauss\n")
    assert [o.id for o in C.Config(str(p)).objects] == ["lens", "sky", "src"]


def test_data_from_the_configuration(tmp_path):
    ref = os.path.join(HERE, "golden", "reference_sie.fits")
    p = tmp_path / "d.ini"
    p.write_text(f"output = false\nimage = {ref}[11:90,21:80]\ngain = 4\noffset = 0.5\n")
    cfg = C.Config(str(p))
    obs = cfg.observation()
    full, _ = D.read_fits(ref)
    assert obs.image.shape == (60, 80) and obs.pcs == (11.0, 21.0, 1.0, 1.0)
    assert (obs.image == full[20:80, 10:90]).all()
    assert abs(obs.weight - 4 / (obs.image + 0.5)).max() < 1e-6


def test_random_input_is_an_error_not_a_crash(tmp_path):
    from hypothesis import given, settings, strategies as st
    alphabet = st.sampled_from(list("abcdefghijklmnopqrstuvwxyz.=:;#[] \t0123456789-_\n"))
    p = tmp_path / "fuzz.ini"

    @settings(max_examples=200, deadline=None, derandomize=True)
    @given(st.lists(alphabet, max_size=300).map("".join), st.lists(st.text(max_size=30), max_size=3))
    def run(text, args):
        p.write_text("[objects]\nlens = sis\n" + text)
        try:
            cfg = C.Config(str(p))
            for a in args:
                try:
                    cfg.arg("--" + a)
                except C.ConfigError:
                    pass
            try:
                cfg.check()
            except C.ConfigError:
                pass
        except C.ConfigError:
            pass

    run()
