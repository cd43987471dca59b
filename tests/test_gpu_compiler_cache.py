"""The program cache against the compiler in use.

A process can run hipRTC from one of two ROCm copies: the system's, or the
one PyTorch bundles, which the library's hipRTC calls resolve to when torch
was imported first.  The two compile the same program to code objects that
round differently in places.  A cached code object must therefore only be
used by a process that runs the compiler that built it: each process here
gets, through a cache that the other kind of process filled first, exactly
the logL bits of its own uncached build.
"""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# a 64 x 64 SIE + Sersic system and five draws near its centre, whose logL
# differs in its last bits between the two compilers' builds
SCRIPT = r"""
import sys
if sys.argv[1] == "torch":
    import torch  # noqa: F401  (its ROCm libraries load before the library's)
sys.path.insert(0, sys.argv[2])
import numpy as np
from lensed_amd.lensed import Context, SystemSpec
from oracle import oracle as O
rng = np.random.default_rng(5)
n = 64
objs = [("lens", "sie"), ("src", "sersic")]
names = [f"{o}.{p}" for o, nm in objs for p in O.PARAMS[nm]]
base = {"lens.x": n / 2, "lens.y": n / 2, "lens.r": n / 6, "lens.q": 0.7, "lens.pa": 20.0,
        "src.x": n / 2 + 1, "src.y": n / 2, "src.r": 3.0, "src.mag": 0.0, "src.n": 2.0, "src.q": 0.8,
        "src.pa": 5.0}
pri = {k: f"unif {base[k] * 0.95 - 0.5} {base[k] * 1.05 + 0.5}" for k in names}
img = rng.normal(0, 0.05, (n, n)).astype(np.float32)
phys = np.array([[base[k] * (1 + 0.01 * rng.normal()) for k in names] for _ in range(5)])
with Context(SystemSpec(objs, pri, img, np.full_like(img, 400.0)), max_batch=8) as ctx:
    print(" ".join(v.hex() for v in ctx.loglike_phys(phys)))
"""


def run(kind, cache, fresh=False):
    env = dict(os.environ, LENSED_CACHE_DIR=str(cache))
    if fresh:
        env["LENSED_NO_CACHE"] = "1"
    print(f"compiler cache: {kind} process, {'uncached' if fresh else 'cached'} build", flush=True)
    r = subprocess.run([sys.executable, "-c", SCRIPT, kind, ROOT], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout.split()


def test_cached_program_comes_from_the_compiler_in_use(tmp_path):
    shared = tmp_path / "shared"
    own = {kind: run(kind, tmp_path / f"fresh_{kind}", fresh=True) for kind in ("torch", "plain")}
    # the torch process fills the shared cache; the plain one then finds it,
    # and after it the torch process again
    assert run("torch", shared) == own["torch"]
    assert run("plain", shared) == own["plain"]
    assert run("torch", shared) == own["torch"]
    print("programs in the shared cache:", len(list(shared.glob("*.co"))),
          "(same bits from both compilers)" if own["torch"] == own["plain"] else "(bits differ between compilers)")
