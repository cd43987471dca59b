"""A whole `lensed` run from an INI file on the GPU (lensed_amd.cli).

Input: the reference's own regression output tests/lens/sie.fits (its IMG
layer: an SIE + Sersic render, 100 x 100, weight 1000), copied as
tests/golden/reference_sie.fits.  The run reads the INI, prepares the data,
writes <root>.paramnames and <root>.ranges as src/lensed.c:1175-1227 does,
samples (one candidate per call through run(), and 512 per launch), lets
the dumper write <root>.fits, and the posterior must sit on the parameters
the reference rendered.
"""
import json
import os

import numpy as np
import pytest

from lensed_amd import cli
from lensed_amd import data as D

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
META = json.load(open(os.path.join(HERE, "golden", "reference_renders.json")))["lens/sie"]
ORDER = ["lens.x", "lens.y", "lens.r", "lens.q", "lens.pa", "source.x", "source.y", "source.r", "source.mag",
         "source.n", "source.q", "source.pa"]
WIDTH = {"lens.x": 0.5, "lens.y": 0.5, "lens.r": 1.0, "lens.q": 0.05, "lens.pa": 5.0, "source.x": 0.5,
         "source.y": 0.5, "source.r": 0.5, "source.mag": 0.5, "source.n": 0.2, "source.q": 0.1, "source.pa": 0.0}


def write_ini(tmp_path, root):
    t = META["params"]
    lines = [f"image  = {os.path.join(HERE, 'golden', 'reference_sie.fits')}", "weight = 1000",
             f"root   = {root}", "nlive  = 100", "seed   = 5", "", "[objects]", "lens   = sie",
             "source = sersic-old", "", "[priors]"]
    for k in ORDER:
        w = WIDTH[k]
        if w == 0:
            lines.append(f"{k} = {t[k]}")  # the round source's angle is fixed
        else:
            lo, hi = t[k] - w, t[k] + w
            if k == "source.q":
                hi = min(hi, 1.0)
            lines.append(f"{k} = {'wrap ' if k.endswith('.pa') else ''}unif {lo} {hi}")
    lines += ["", "[labels]", "lens.r = r_E"]
    p = tmp_path / "sie.ini"
    p.write_text("\n".join(lines) + "\n")
    return str(p)


@pytest.mark.parametrize("batch", [1, 512])
def test_cli_run_on_reference_data(tmp_path, capsys, batch):
    root = str(tmp_path / "out-")
    ini = write_ini(tmp_path, root)
    assert cli.main([ini, f"--batch={batch}"]) == 0
    out = capsys.readouterr().out
    print(out)
    assert "log-evidence" in out and "min chi²/dof" in out
    names = open(root + ".paramnames").read().splitlines()
    ranges = open(root + ".ranges").read().splitlines()
    # sampled parameters first, fixed ones after them (MultiNest's derived parameters)
    assert [n.split()[0] for n in names] == [k for k in ORDER if WIDTH[k]] + ["source.pa"]
    assert ranges[-1].split()[1:] == ["0.0000", "0.0000"]
    assert names[2].split()[1] == "r_E"
    lo, hi = map(float, ranges[2].split()[1:])
    assert (lo, hi) == (19.0, 21.0)
    layers = {n: D.read_fits(f"{root}.fits[{n}]")[0] for n in D.LAYERS}
    ref, _ = D.read_fits(os.path.join(HERE, "golden", "reference_sie.fits"))
    # the best sample of 100 live points lies within a few sigma of the truth, so
    # its model reproduces the reference render to well under 1 % of the peak
    assert np.max(np.abs(layers["RES"])) <= 1e-2 * np.max(ref)
    assert np.array_equal(layers["RES"], ref - layers["IMG"])
    assert np.all(layers["WHT"] == 1000)
    post = np.loadtxt(root + ".txt")
    assert post.shape[1] == 2 + 12
    # posterior table: every free parameter's mean within 3 sigma (or 1e-3 of its range) of the truth
    rows = {l.split()[0]: [float(v) for v in l.split()[1:]] for l in out.splitlines()
            if l.strip() and l.split()[0] in ORDER + ["r_E"] and len(l.split()) == 5}
    for k in ORDER:
        if not WIDTH[k]:
            continue
        mean, sig = rows["r_E" if k == "lens.r" else k][:2]
        assert abs(mean - META["params"][k]) <= max(3 * sig, 2e-3 * WIDTH[k]), (k, mean, sig)
