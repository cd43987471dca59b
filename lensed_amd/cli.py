"""Command line: a full Lensed run from INI files (src/lensed.c:main).

    python -m lensed_amd [-v|-q] [--batch=K] [--table] [--emulator=FILE] [--profile]
                         [--rules] file.ini ... [--name=value ...]

Reads the configuration (later files and options override earlier ones),
prepares the data, builds the likelihood context, writes
<root>.paramnames/.ranges, runs nested sampling -- one candidate per call
through run() as Lensed does, or `--batch=K` candidates per launch -- with
the dumper writing <root>.fits, and prints Lensed's summary: log-evidence,
maximum log-likelihood, minimum chi^2/dof and the parameter table.
"""
import ctypes
import sys
import time
from ctypes import byref, c_double

import numpy as np

from . import _lib as L
from . import nested
from .config import Config, ConfigError
from .data import DataError
from .lensed import Context, LensedError

EXTRA_FLAGS = ("batch", "table", "emulator", "profile", "rules")


def _log(level, quiet, *a):
    if level >= quiet:
        print(*a, flush=True)


def _results(ctx, n):
    mean, sigma, ml, mp = [(c_double * n)() for _ in range(4)]
    lz, lze, mx = c_double(), c_double(), c_double()
    L.load().lensed_results(ctx.ptr, mean, sigma, ml, mp, byref(lz), byref(lze), byref(mx))
    return list(mean), list(sigma), list(ml), list(mp), lz.value, lze.value, mx.value


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    verbose, quiet = False, 0
    extra = {"batch": "1", "table": None, "emulator": None, "profile": None, "rules": None}
    files, opts = [], []
    for a in argv:
        if a.startswith("--"):
            name = a[2:].split("=", 1)[0]
            if name in EXTRA_FLAGS:
                extra[name] = a.split("=", 1)[1] if "=" in a else "true"
            else:
                opts.append(a)
        elif a.startswith("-") and len(a) > 1:
            for ch in a[1:]:
                if ch == "v":
                    verbose = True
                elif ch == "q":
                    quiet = 2
                else:
                    print(__doc__)
                    return 2
        else:
            files.append(a)
    if extra["rules"]:
        lib = L.load()
        i = 0
        while lib.lensed_quad_name(i):
            print(lib.lensed_quad_name(i).decode())
            i += 1
        return 0
    if not files and not opts:
        print(__doc__)
        return 2
    try:
        cfg = Config(*files, args=opts)
        cfg.check()
        spec = cfg.system()
    except (ConfigError, DataError, LensedError) as e:
        print(f"error: {e}", file=sys.stderr)
        return 1
    obs = spec.observation
    for w in obs.warning_names:
        print(f"warning: {w}", file=sys.stderr)
    if verbose:
        print("options")
        for name in ("image", "weight", "gain", "offset", "psf", "mask", "rule", "nlive", "ceff", "acc", "tol",
                     "shf", "root", "output"):
            print(f"  {name} = {cfg.get(name)}")
        print(f"data\n  image size: {obs.image.shape[1]} x {obs.image.shape[0]}\n  masked pixels: {obs.masked}\n"
              f"  background: {obs.mode:f} +- {0.5 * obs.fwhm:f}")
    batch = int(extra["batch"])
    flags = L.LENSED_GRAPH if batch > 1 else 0
    if extra["table"]:
        flags |= L.LENSED_TABLE
    if extra["emulator"]:
        flags |= L.LENSED_EMULATE
    if extra["profile"]:
        flags |= L.LENSED_PROFILE
    try:
        ctx = Context(spec, flags=flags, max_batch=max(batch, 1), emulator=extra["emulator"])
    except LensedError as e:
        print(f"error: {e}", file=sys.stderr)
        return 1
    npar, ndim, pmap = ctx.npars(), ctx.ndims(), ctx.pmap()
    pars = [o.obj.pars[j] for o in spec.objects for j in range(o.obj.npars)]
    wrap = [int(pars[pmap[i]].wrap) for i in range(npar)]
    output = cfg.get_bool("output")
    root = cfg.get("root") if output else None
    if output:
        if L.load().lensed_write_paramfiles(ctx.ptr, 0, root.encode()) != 0:
            print(f"error: {L.last_error()}", file=sys.stderr)
            return 1
        ctx.set_output("!" + root + ".fits")
    ceff = cfg.get_bool("ceff")
    kw = dict(nlive=cfg.get_int("nlive"), tol=cfg.get_real("tol"),
              efr=cfg.get_real("acc") if ceff else cfg.get_real("shf"), ceff=ceff,
              seed=cfg.get_int("seed"), updint=cfg.get_int("updint"), maxiter=cfg.get_int("maxiter"),
              resume=cfg.get_bool("resume"), wrap=wrap, root=root)
    _log(1, quiet, "find posterior")
    t0 = time.time()
    if batch > 1:
        kw["seed"] = kw["seed"] if kw["seed"] >= 0 else int(time.time())
        kw["outfile"] = output
        kw["feedback"] = cfg.get_bool("feedback")
        nested.run_context(ctx, 0, batch=batch, **kw)
    else:
        lib = L.load()
        nested.run_multinest_abi(ctypes.cast(lib.loglike, L.LOGLIKE_FN), ndim, npar,
                                 ctypes.cast(lib.dumper, L.DUMPER_FN), outfile=output,
                                 feedback=cfg.get_bool("feedback"), context=ctx.ptr, **kw)
    dur = time.time() - t0
    _log(1, quiet, "  done in %02d:%02d:%02d" % (dur // 3600, dur % 3600 // 60, dur % 60))
    mean, sigma, ml, mp, logev, logev_err, maxl = _results(ctx, npar)
    chi2_dof = -2 * maxl / (obs.image.size - obs.masked - npar)
    _log(1, quiet, "summary\n")
    _log(1, quiet, f"  log-evidence: {logev:.4f} ± {logev_err:.4f}")
    _log(1, quiet, f"  max log-like: {maxl:.4f}")
    _log(1, quiet, f"  min chi²/dof: {chi2_dof:.4f}\n")
    _log(1, quiet, "parameters\n")
    _log(1, quiet, "  %-12s  %10s  %10s  %10s  %10s" % ("parameter", "mean", "sigma", "ML", "MAP"))
    _log(1, quiet, "  " + "-" * 60)
    for i, p in enumerate(pars):
        if p.defval:
            continue
        name = (p.label or p.id).decode()
        _log(1, quiet, "  %-12s  %10.4f  %10.4f  %10.4f  %10.4f" % (name, mean[i], sigma[i], ml[i], mp[i]))
    if extra["profile"]:
        calls, ms = ctx.profile()
        _log(1, quiet, "\nprofiler\n")
        for k, v in ms.items():
            _log(1, quiet, "  %-14s %12.3f ms  %10.4f ms/call" % (k, v, v / max(calls, 1)))
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
