"""Quadrature rules regenerated from the mathematics (lensed_amd/csrc/gen_quad.py)
against their defining properties and, where the reference is present, its
own constant tables (src/quad/*.c)."""
import os
import re

import numpy as np
import pytest

from conftest import REFERENCE
from lensed_amd import quadrature as Q

RULES = ["point", "sub2", "sub4", "gm75", "g3k7", "g5k11", "g7k15"]


def test_rule_list_and_sizes():
    assert Q.names() == RULES
    sizes = {r: len(Q.table(r)[1]) for r in RULES}
    assert sizes == {"point": 1, "sub2": 4, "sub4": 16, "gm75": 17, "g3k7": 49, "g5k11": 121, "g7k15": 225}


@pytest.mark.parametrize("rule", RULES)
def test_weights_sum(rule):
    pts, w, e = Q.table(rule)
    assert w.sum() == pytest.approx(1.0, abs=1e-15)
    assert e.sum() == pytest.approx(0.0, abs=1e-15)
    assert np.all(np.abs(pts) <= 0.5)


def monomial_integral(a, b):
    # integral of x^a y^b over [-1/2, 1/2]^2
    f = lambda k: 0.0 if k % 2 else 2 * 0.5 ** (k + 1) / (k + 1)
    return f(a) * f(b)


@pytest.mark.parametrize("rule,n", [("g3k7", 3), ("g5k11", 5), ("g7k15", 7)])
def test_gauss_kronrod_exactness(rule, n):
    pts, w, e = Q.table(rule)
    x, y = pts[:, 0], pts[:, 1]
    for a in range(3 * n + 2):
        for b in range(3 * n + 2):
            exact = monomial_integral(a, b)
            got = np.sum(w * x ** a * y ** b)
            assert got == pytest.approx(exact, abs=1e-15), (a, b)
            # Kronrod minus Gauss vanishes where both are exact: the
            # product of 1-D differences is zero if either degree <= 2n-1
            if a <= 2 * n - 1 or b <= 2 * n - 1:
                assert abs(np.sum(e * x ** a * y ** b)) < 1e-15, (a, b)
    # ... and not beyond: the error weights see degree 2n in both variables
    assert abs(np.sum(e * x ** (2 * n) * y ** (2 * n))) > 1e-6 * monomial_integral(2 * n, 2 * n)


def test_genz_malik_degrees():
    pts, w, e = Q.table("gm75")
    x, y = pts[:, 0], pts[:, 1]
    for a in range(9):
        for b in range(9 - a):
            exact = monomial_integral(a, b)
            if a + b <= 7:
                assert np.sum(w * x ** a * y ** b) == pytest.approx(exact, abs=1e-15), (a, b)
            if a + b <= 5:  # the embedded degree-5 rule w - e is exact too
                assert np.sum((w - e) * x ** a * y ** b) == pytest.approx(exact, abs=1e-14), (a, b)


@pytest.mark.parametrize("rule,deg", [("point", 1), ("sub2", 1), ("sub4", 1)])
def test_subsampling_rules(rule, deg):
    pts, w, e = Q.table(rule)
    assert np.all(e == 0)
    assert np.allclose(w, 1.0 / len(w))
    for a in range(deg + 1):
        assert np.sum(w * pts[:, 0] ** a) == pytest.approx(monomial_integral(a, 0), abs=1e-15)


def test_scaled_rule_is_float_of_scaled_double():
    pts, w, e = Q.table("g3k7")
    qq, ww = Q.rule("g3k7", 0.5, 2.0)
    assert np.array_equal(qq[:, 0], (0.5 * pts[:, 0]).astype(np.float32))
    assert np.array_equal(qq[:, 1], (2.0 * pts[:, 1]).astype(np.float32))
    assert np.array_equal(ww[:, 0], w.astype(np.float32))
    assert np.array_equal(ww[:, 1], e.astype(np.float32))


def reference_table(rule):
    path = os.path.join(REFERENCE, "src", "quad", rule + ".c")
    text = open(path).read()
    blocks = {}
    for name in ("PTS", "WHT", "ERR"):
        m = re.search(r"QUAD_%s_%s\[\][^=]*=\s*\{(.*?)\};" % (rule.upper(), name), text, re.S)
        blocks[name] = np.array([float(v) for v in re.findall(r"[-+]?\d*\.?\d+(?:[eE][-+]?\d+)?", m.group(1))])
    return blocks["PTS"].reshape(-1, 2), blocks["WHT"], blocks["ERR"]


@pytest.mark.skipif(not os.path.isdir(os.path.join(REFERENCE, "src", "quad")), reason="reference not present")
@pytest.mark.parametrize("rule", RULES)
def test_matches_reference_constants(rule):
    pts, w, e = Q.table(rule)
    rp, rw, re_ = reference_table(rule)
    assert rp.shape == pts.shape
    assert np.max(np.abs(rp - pts)) < 2e-16
    assert np.max(np.abs(rw - w)) < 2e-16
    assert np.max(np.abs(re_ - e)) < 2e-16
