"""FITS input/output and data preparation (CPU).

The library's FITS reader is checked against an independent numpy parser
(tests/golden/make_fixtures.py) on the reference's own regression output
tests/lens/sie.fits (copied as tests/golden/reference_sie.fits: a CFITSIO-
written file with a primary image and three extensions, BITPIX -64), and on
files written here by a small numpy FITS writer covering every BITPIX with
BSCALE/BZERO.  Image sections follow CFITSIO's syntax and the pixel
coordinate system of src/data.c:236-278.  The writer's files are read back
by the numpy parser and checked against the FITS layout rules (80-byte
cards, 2880-byte records, END).  Data preparation follows
src/lensed.c:406-549; find_mode follows src/data.c:421-501 and is checked on
histograms whose mode and width are known in closed form.
"""
import os

import numpy as np
import pytest

from golden.make_fixtures import read_fits as np_read_fits
from lensed_amd import _lib as L
from lensed_amd import data as D

HERE = os.path.dirname(os.path.abspath(__file__))
REF_SIE = os.path.join(HERE, "golden", "reference_sie.fits")


def card(key, value=None, comment=""):
    if value is None:
        s = f"{key:<8}{comment}"
    elif isinstance(value, str):
        s = f"{key:<8}= '{value:<8}'" + (f" / {comment}" if comment else "")
    elif isinstance(value, bool):
        s = f"{key:<8}= {'T' if value else 'F':>20}"
    else:
        s = f"{key:<8}= {value:>20}" + (f" / {comment}" if comment else "")
    return s.ljust(80)[:80].encode()


def np_write_fits(path, hdus):
    """hdus: list of (array or None, bitpix, extra cards dict) -> FITS file."""
    out = b""
    for k, (arr, bitpix, extra) in enumerate(hdus):
        cards = [card("SIMPLE", True)] if k == 0 else [card("XTENSION", "IMAGE")]
        cards.append(card("BITPIX", bitpix))
        if arr is None:
            cards.append(card("NAXIS", 0))
        else:
            cards += [card("NAXIS", 2), card("NAXIS1", arr.shape[1]), card("NAXIS2", arr.shape[0])]
        if k == 0:
            cards.append(card("EXTEND", True))
        else:
            cards += [card("PCOUNT", 0), card("GCOUNT", 1)]
        for key, v in extra.items():
            cards.append(card(key, v))
        cards.append(card("END"))
        hdr = b"".join(cards)
        out += hdr + b" " * (-len(hdr) % 2880)
        if arr is not None:
            dt = {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}[bitpix]
            raw = np.ascontiguousarray(arr).astype(dt).tobytes()
            out += raw + b"\0" * (-len(raw) % 2880)
    with open(path, "wb") as f:
        f.write(out)


def test_reads_the_reference_output_file():
    hdus = [a for _, a in np_read_fits(REF_SIE) if a is not None]
    got, pcs = D.read_fits(REF_SIE)  # primary image first, as fits_open_image
    assert np.array_equal(got, hdus[0].astype(np.float32))
    assert pcs == (1.0, 1.0, 1.0, 1.0)
    for k in range(1, 4):
        layer, _ = D.read_fits(f"{REF_SIE}[{k}]")
        assert np.array_equal(layer, hdus[k].astype(np.float32))
    golden = np.load(os.path.join(HERE, "golden", "reference_renders.npz"))
    assert np.array_equal(got, golden["lens/sie/img"])
    assert np.array_equal(D.read_fits(f"{REF_SIE}[3]")[0], golden["lens/sie/err"])


@pytest.mark.parametrize("bitpix", [8, 16, 32, 64, -32, -64])
def test_every_bitpix_with_scaling(tmp_path, bitpix):
    rng = np.random.default_rng(abs(bitpix))
    raw = rng.integers(0, 200, size=(7, 5))
    if bitpix < 0:
        raw = rng.normal(size=(7, 5))
    scale, zero = (0.5, -3.0) if bitpix > 0 else (1.0, 0.0)
    path = str(tmp_path / "a.fits")
    np_write_fits(path, [(raw, bitpix, {"BSCALE": scale, "BZERO": zero})])
    got, _ = D.read_fits(path)
    assert got.shape == (7, 5)
    assert np.allclose(got, (zero + scale * raw).astype(np.float32), rtol=1e-7, atol=0)


def test_empty_primary_then_named_extensions(tmp_path):
    a = np.arange(20, dtype=np.float32).reshape(4, 5)
    path = str(tmp_path / "m.fits")
    np_write_fits(path, [(None, 16, {}), (a, -32, {"EXTNAME": "ONE"}), (2 * a, -32, {"EXTNAME": "two"})])
    assert np.array_equal(D.read_fits(path)[0], a)  # first HDU holding an image
    assert np.array_equal(D.read_fits(path + "[TWO]")[0], 2 * a)
    assert np.array_equal(D.read_fits(path + "[2]")[0], 2 * a)
    with pytest.raises(D.DataError):
        D.read_fits(path + "[THREE]")
    with pytest.raises(D.DataError):
        D.read_fits(path + "[0]")  # the primary HDU holds no image


def test_image_sections_and_pixel_coordinates(tmp_path):
    a = np.arange(6 * 8, dtype=np.float32).reshape(6, 8)  # NAXIS1 = 8 (x), NAXIS2 = 6 (y)
    path = str(tmp_path / "s.fits")
    np_write_fits(path, [(a, -32, {})])
    got, pcs = D.read_fits(path + "[3:7,2:5]")
    assert np.array_equal(got, a[1:5, 2:7]) and pcs == (3.0, 2.0, 1.0, 1.0)
    got, pcs = D.read_fits(path + "[7:3,*]")
    assert np.array_equal(got, a[:, 2:7][:, ::-1]) and pcs == (7.0, 1.0, -1.0, 1.0)
    got, pcs = D.read_fits(path + "[1:8:2,-*]")
    assert np.array_equal(got, a[::-1, ::2]) and pcs == (1.0, 6.0, 2.0, -1.0)
    p2 = (L.c_double * 4)()
    assert L.load().lensed_read_pcs((path + "[2:4,6:1:5]").encode(), p2) == 0
    assert tuple(p2) == (2.0, 6.0, 1.0, -5.0)
    for bad in ("[0:3,1:2]", "[1:9,1:2]", "[1:3]", "[a:b,1:2]"):
        with pytest.raises(D.DataError):
            D.read_fits(path + bad)


def test_written_layers_follow_the_fits_layout(tmp_path):
    rng = np.random.default_rng(1)
    layers = {n: rng.normal(size=(9, 13)).astype(np.float32) for n in D.LAYERS}
    path = str(tmp_path / "out.fits")
    D.write_fits("!" + path, layers)  # CFITSIO's overwrite prefix
    raw = open(path, "rb").read()
    assert len(raw) % 2880 == 0
    assert raw[:80].decode().startswith("SIMPLE  =                    T")
    hdus = np_read_fits(path)
    assert hdus[0][1] is None and hdus[0][0]["NAXIS"] == "0"
    assert [h["EXTNAME"] for h, _ in hdus[1:]] == list(D.LAYERS)
    for (h, a), n in zip(hdus[1:], D.LAYERS):
        assert h["XTENSION"] == "IMAGE" and h["BITPIX"] == "-32"
        assert np.array_equal(a, layers[n])
        assert np.array_equal(D.read_fits(f"{path}[{n}]")[0], layers[n])
    # every header card is 80 printable ASCII characters
    assert all(32 <= b < 127 for b in raw[:2880])


def test_find_mode_closed_forms():
    # 50 values at 0 and one at 10: bins of 0.1, mode at the first bin's middle,
    # the half maximum falls off one bin to the right, and the left edge is bin 0
    v = np.array([0.0] * 50 + [10.0], np.float32)
    mode, fwhm = D.find_mode(v)
    assert mode == pytest.approx(0.05) and fwhm == pytest.approx(0.2)
    # constant values: the value itself, zero width
    assert D.find_mode(np.full(10, 3.5)) == (3.5, 0.0)
    # masked pixels are ignored
    mode, _ = D.find_mode(np.array([0.0] * 50 + [10.0] * 60, np.float32), mask=np.array([1] * 50 + [0] * 60))
    assert mode < 1
    # a wide Gaussian: mode near the mean, FWHM near 2.355 sigma (histogram resolution)
    rng = np.random.default_rng(2)
    x = rng.normal(2.0, 1.0, 400000)
    x = x[np.abs(x - 2.0) < 5]
    mode, fwhm = D.find_mode(x)
    dx = (x.max() - x.min()) / 100
    assert abs(mode - 2.0) < 2 * dx
    assert abs(fwhm - 2.3548) < 3 * dx


def test_observation_weights_mask_psf(tmp_path):
    rng = np.random.default_rng(3)
    img = rng.uniform(1, 2, size=(10, 12)).astype(np.float32)
    gain = rng.uniform(2, 3, size=(10, 12)).astype(np.float32)
    xw = rng.uniform(0.5, 1.5, size=(10, 12)).astype(np.float32)
    mask = np.zeros((10, 12), np.int16)
    mask[2:4, 5:9] = 1
    psf = rng.uniform(0, 1, size=(5, 5))
    for name, a, bp in (("img", img, -32), ("gain", gain, -32), ("xw", xw, -32), ("mask", mask, 16),
                        ("psf", psf, -64)):
        np_write_fits(str(tmp_path / f"{name}.fits"), [(a, bp, {})])
    t = lambda n: str(tmp_path / f"{n}.fits")
    obs = D.Observation(t("img"), gain=t("gain"), offset=0.25, xweight=t("xw"), mask=t("mask"), psf=t("psf"))
    expect = gain / (img + np.float32(0.25)) * xw
    expect[mask != 0] = 0
    assert np.allclose(obs.weight, expect, rtol=1e-6)
    assert obs.masked == 8
    assert obs.psf.sum() == pytest.approx(1.0, rel=1e-6)
    assert np.allclose(obs.psf, psf / psf.sum(), rtol=1e-6)
    assert obs.pcs == (1.0, 1.0, 1.0, 1.0)
    # mode near 1.5 (not zero) with no sky object: the first warning
    assert obs.warnings & L.LENSED_WARN_SKY_NOT_ZERO
    # numeric options, bscale, and an image section
    obs2 = D.Observation(t("img") + "[2:11,3:8]", weight="4", bscale=2.0)
    assert np.array_equal(obs2.image, 2 * img[2:8, 1:11])
    assert np.all(obs2.weight == 4) and obs2.pcs == (2.0, 3.0, 1.0, 1.0)
    # background near zero, gain-based weights without an offset: the second warning
    np_write_fits(t("zero"), [(rng.normal(0, 0.01, size=(300, 300)), -32, {})])
    assert D.Observation(t("zero"), gain="1").warnings & L.LENSED_WARN_NO_OFFSET
    # negative weights
    assert D.Observation(t("img"), weight="-1").warnings & L.LENSED_WARN_NEGATIVE_WEIGHT
    with pytest.raises(D.DataError):
        D.Observation(t("img"), weight=t("psf"))  # wrong dimensions
    with pytest.raises(D.DataError):
        D.Observation(t("img"))  # neither weight nor gain


def test_malformed_files_are_errors_not_crashes(tmp_path):
    """Truncations, corrupted header values and random bytes: every outcome
    is an image or a DataError (the reader checks sizes against the file)."""
    from hypothesis import given, settings, strategies as st

    a = np.arange(12 * 9, dtype=np.float32).reshape(9, 12)
    good = str(tmp_path / "good.fits")
    np_write_fits(good, [(None, 16, {}), (a, -32, {"EXTNAME": "IMG"})])
    raw = open(good, "rb").read()
    path = str(tmp_path / "bad.fits")

    def attempt(data, spec=""):
        with open(path, "wb") as f:
            f.write(data)
        try:
            D.read_fits(path + spec)
        except D.DataError:
            pass

    @settings(max_examples=150, deadline=None, derandomize=True)
    @given(st.integers(0, len(raw)), st.lists(st.tuples(st.integers(0, len(raw) - 1), st.integers(0, 255)),
                                               max_size=8))
    def mutate(cut, flips):
        b = bytearray(raw)
        for i, v in flips:
            b[i] = v
        attempt(bytes(b[:cut]))
        attempt(bytes(b), "[1]")

    mutate()
    for key, val in (("NAXIS1", "99999999999"), ("NAXIS2", "-5"), ("NAXIS", "1000"), ("BITPIX", "13"),
                     ("GCOUNT", "4611686018427387904"), ("PCOUNT", "-1"), ("NAXIS1", "9223372036854775807")):
        text = raw.decode("latin-1")
        i = text.index(key.ljust(8) + "=", 2880)
        card = (key.ljust(8) + "= " + val.rjust(20)).ljust(80)
        bad = (text[:i] + card + text[i + 80:]).encode("latin-1")
        with open(path, "wb") as f:
            f.write(bad)
        with pytest.raises(D.DataError):
            D.read_fits(path + "[1]")
