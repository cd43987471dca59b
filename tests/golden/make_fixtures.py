"""Convert the reference's own regression fixtures into small committed vectors.

The reference's tests (tests/Makefile:1-28) each render one fixed model
(all priors are fixed values) at 100 x 100 px with the default g3k7 rule and
no PSF, and ship Lensed's own output image next to the .ini.  This script
reads those files from a reference checkout and writes

    reference_renders.npz   <test>/img (the IMG layer), <test>/err (the ERR layer,
                            quadrature error / value), float32 [100, 100]
    reference_renders.json  per test: objects [(id, name)], fixed parameter
                            values {"id.param": value}, weight, image size,
                            FITS ORIGIN if present

Run (in a container that has the reference at /root/reference):

    python tests/golden/make_fixtures.py [/root/reference]

Only data leave the reference: the rendered pixel values and the numbers of
the .ini files.  The output is what the tests read; the reference itself is
not needed afterwards.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

TESTS = [
    "foreground/sky", "lens/epl", "lens/epl-isothermal", "lens/epl_plus_shear",
    "lens/epl_plus_shear-isothermal", "lens/nsie", "lens/nsis", "lens/point_mass", "lens/sie",
    "lens/sie_plus_shear", "lens/sis", "lens/sis_plus_shear", "source/devauc",
    "source/exponential", "source/gauss", "source/sersic",
]


def read_fits(path):
    """Minimal FITS reader: list of (header dict, array or None) per HDU."""
    data = open(path, "rb").read()
    hdus, pos = [], 0
    while pos < len(data):
        hdr, end = {}, False
        while not end:
            block = data[pos:pos + 2880]
            pos += 2880
            for i in range(36):
                card = block[i * 80:(i + 1) * 80].decode("ascii")
                key = card[:8].strip()
                if key == "END":
                    end = True
                    break
                if card[8:10] == "= ":
                    hdr[key] = card[10:].split("/")[0].strip().strip("'").strip()
        bitpix, naxis = int(hdr["BITPIX"]), int(hdr.get("NAXIS", 0))
        shape = [int(hdr[f"NAXIS{i}"]) for i in range(naxis, 0, -1)]
        n = int(np.prod(shape)) if naxis else 0
        nbytes = n * abs(bitpix) // 8
        dt = {-64: ">f8", -32: ">f4", 8: "u1", 16: ">i2", 32: ">i4"}[bitpix]
        arr = np.frombuffer(data[pos:pos + nbytes], dtype=dt).reshape(shape) if n else None
        pos += (nbytes + 2879) // 2880 * 2880
        hdus.append((hdr, arr))
    return hdus


def read_ini(path):
    groups, cur = {}, None
    for line in open(path):
        line = line.split(";")[0].strip()
        if not line:
            continue
        if line.startswith("["):
            cur = line.strip("[]")
            groups[cur] = []
            continue
        k, v = [s.strip() for s in line.split("=", 1)]
        groups.setdefault(cur, []).append((k, v))
    return groups


def main(ref):
    arrays, meta = {}, {}
    for t in TESTS:
        ini = read_ini(os.path.join(ref, "tests", t + ".ini"))
        hdus = [h for h in read_fits(os.path.join(ref, "tests", t + ".fits")) if h[1] is not None]
        named = {h[0].get("EXTNAME"): h[1] for h in hdus if h[0].get("EXTNAME")}
        img = named.get("IMG", hdus[0][1])
        err = named.get("ERR", hdus[3][1])
        origin = next((h[0]["ORIGIN"] for h in read_fits(os.path.join(ref, "tests", t + ".fits"))
                       if "ORIGIN" in h[0]), None)
        arrays[t + "/img"] = np.asarray(img, dtype=np.float32)
        arrays[t + "/err"] = np.asarray(err, dtype=np.float32)
        opts = dict(ini.get(None, []))
        meta[t] = {
            "objects": ini["objects"],
            "params": {k: float(v) for k, v in ini["priors"]},
            "weight": float(opts.get("weight", "1")),
            "shape": list(img.shape),
            "origin": origin,
            "exact_f32": bool(np.all(np.asarray(img, np.float64) == np.asarray(img, np.float32))),
        }
    np.savez_compressed(os.path.join(HERE, "reference_renders.npz"), **arrays)
    with open(os.path.join(HERE, "reference_renders.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print(f"wrote {len(meta)} fixtures")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
