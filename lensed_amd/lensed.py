"""Host-side mirror of Lensed's likelihood setup, over the C ABI.

A system is what one Lensed configuration describes (src/input/ini.c): an
ordered list of objects (`[objects]`), priors per parameter (`[priors]`,
docs/configuration.md:157-214), the observed image with its weights and an
optional PSF (src/lensed.c:406-511) and the quadrature rule.  `Context`
builds the device program for one or more systems and evaluates
log-likelihoods through the same entry points a sampler uses: the
MultiNest-style `loglike` callback (src/nested.c:17-131) one candidate at a
time, or the batch extension many candidates per launch.

All device work happens in liblensed_amd.so; nothing here computes a model.
"""
import ctypes
from ctypes import POINTER, byref, c_double, c_float, c_int, c_size_t

import numpy as np

from . import _lib as L

LOG_ZERO = -np.finfo(np.float64).max


class LensedError(RuntimeError):
    pass


def _check(rc, what):
    if rc != 0:
        raise LensedError(f"{what}: {L.last_error()}")


class Prior:
    """A prior in the reference syntax: a value, `unif a b` or `norm m s`."""

    def __init__(self, text):
        lib = L.load()
        self.text = str(text)
        self.ptr = lib.prior_read(self.text.encode())
        if not self.ptr:
            raise LensedError(f"invalid prior definition: {text}")

    def apply(self, u):
        return L.load().prior_apply(self.ptr, float(u))

    @property
    def lower(self):
        return L.load().prior_lower(self.ptr)

    @property
    def upper(self):
        return L.load().prior_upper(self.ptr)

    @property
    def pseudo(self):
        return bool(L.load().prior_pseudo(self.ptr))

    def __str__(self):
        buf = ctypes.create_string_buffer(64)
        L.load().prior_print(self.ptr, buf, 64)
        return buf.value.decode()

    def __del__(self):
        if getattr(self, "ptr", None):
            try:
                L.load().prior_free(self.ptr)
            except Exception:
                pass
            self.ptr = None


class LensedObject:
    """One object of the line of sight, loaded from objects/<name>.cl."""

    def __init__(self, oid, name):
        lib = L.load()
        self.obj = L.Object()
        _check(lib.lensed_object_load(byref(self.obj), oid.encode(), name.encode()), f"object {oid}")
        self.id, self.name = oid, name
        self._priors = []

    @property
    def type(self):
        return chr(self.obj.type)

    @property
    def params(self):
        return [self.obj.pars[i].name.decode() for i in range(self.obj.npars)]

    def set_prior(self, pname, text):
        """`[wrap] [image] <prior>`, as the [priors] section takes it."""
        words = str(text).split()
        wrap = image = False
        while words and words[0] in ("wrap", "image"):
            wrap |= words[0] == "wrap"
            image |= words[0] == "image"
            words = words[1:]
        names = self.params
        if pname not in names:
            raise LensedError(f"object {self.id} has no parameter {pname}")
        par = self.obj.pars[names.index(pname)]
        pri = Prior(" ".join(words))
        lib = L.load()
        if par.pri:
            lib.prior_free(par.pri)
        par.pri = lib.prior_read(pri.text.encode())  # owned by the object
        par.wrap = int(wrap)
        par.ipp = int(image)

    def __del__(self):
        if getattr(self, "obj", None) is not None:
            try:
                L.load().lensed_object_free(byref(self.obj))
            except Exception:
                pass
            self.obj = None


class SystemSpec:
    """Objects, priors and data of one observed system."""

    @classmethod
    def from_observation(cls, objects, priors, obs, rule="g3k7"):
        """A system from prepared data (lensed_amd.data.Observation)."""
        return cls(objects, priors, obs.image, obs.weight, psf=obs.psf, pcs=obs.pcs, rule=rule)

    def __init__(self, objects, priors, image, weight, psf=None, pcs=(1.0, 1.0, 1.0, 1.0), rule="g3k7"):
        self.objects = [o if hasattr(o, "obj") else LensedObject(*o) for o in objects]
        byid = {o.id: o for o in self.objects}
        for key, text in priors.items():
            oid, pname = key.split(".", 1)
            byid[oid].set_prior(pname, text)
        self.image = np.ascontiguousarray(image, dtype=np.float32)
        self.weight = np.ascontiguousarray(weight, dtype=np.float32)
        if self.image.shape != self.weight.shape or self.image.ndim != 2:
            raise LensedError("image and weight must be 2-D arrays of the same shape")
        self.psf = None if psf is None else np.ascontiguousarray(psf, dtype=np.float32)
        self.pcs = tuple(float(v) for v in pcs)
        self.rule = rule

    @property
    def param_ids(self):
        return [f"{o.id}.{p}" for o in self.objects for p in o.params]

    def _struct(self):
        s = L.System()
        h, w = self.image.shape
        s.width, s.height = w, h
        s.image = self.image.ctypes.data_as(POINTER(c_float))
        s.weight = self.weight.ctypes.data_as(POINTER(c_float))
        if self.psf is not None:
            s.psf = self.psf.ctypes.data_as(POINTER(c_float))
            s.psfh, s.psfw = self.psf.shape
        s.pcs = (c_double * 4)(*self.pcs)
        s.rule = self.rule.encode()
        self._objs = (L.Object * len(self.objects))(*[o.obj for o in self.objects])
        s.nobjs = len(self.objects)
        s.objs = self._objs
        return s


class Context:
    """The likelihood context for one or more systems (struct lensed)."""

    def __init__(self, systems, flags=0, max_batch=1, device=0, emulator=None, emulator_tol=0.0):
        if isinstance(systems, SystemSpec):
            systems = [systems]
        self.systems = list(systems)
        lib = L.load()
        structs = (L.System * len(self.systems))(*[s._struct() for s in self.systems])
        opts = L.Options(device=device, flags=flags, max_batch=max_batch,
                         emulator=None if emulator is None else str(emulator).encode(), emulator_tol=emulator_tol)
        self.ptr = lib.lensed_create(structs, len(self.systems), byref(opts))
        if not self.ptr:
            raise LensedError(L.last_error())
        self.max_batch = max_batch

    def close(self):
        if getattr(self, "ptr", None):
            L.load().lensed_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def npars(self, s=0):
        return L.load().lensed_npars(self.ptr, s)

    def ndims(self, s=0):
        return L.load().lensed_ndims(self.ptr, s)

    def pmap(self, s=0):
        out = (c_size_t * self.npars(s))()
        L.load().lensed_pmap(self.ptr, s, out)
        return list(out)

    def wrap(self, s=0):
        """Periodic flags of the parameters in sampler order (what Lensed
        passes to MultiNest's run(), src/lensed.c:1254-1259)."""
        flags = [o.obj.pars[j].wrap for o in self.systems[s].objects for j in range(o.obj.npars)]
        return [int(flags[k]) for k in self.pmap(s)]

    def loglike(self, cube):
        """One candidate through the MultiNest callback; returns (logL, physical)."""
        npar = self.npars(0)
        ndim = self.ndims(0)
        buf = (c_double * npar)(*list(np.asarray(cube, dtype=np.float64)[:ndim]), *([0.0] * (npar - ndim)))
        lnew = c_double()
        L.load().loglike(buf, byref(c_int(ndim)), byref(c_int(npar)), byref(lnew), self.ptr)
        return lnew.value, np.array(buf[:], dtype=np.float64)

    def loglike_batch(self, cubes, s=0):
        """ncand unit cubes -> (logL[ncand], physical[ncand, npar]) in one launch per max_batch."""
        cubes = np.ascontiguousarray(cubes, dtype=np.float64)
        n = cubes.shape[0]
        phys = np.empty((n, self.npars(s)))  # both filled by the call
        lnew = np.empty(n)
        _check(L.load().lensed_loglike_batch(self.ptr, s, cubes.ctypes.data_as(POINTER(c_double)), n,
                                             phys.ctypes.data_as(POINTER(c_double)),
                                             lnew.ctypes.data_as(POINTER(c_double))), "loglike_batch")
        return lnew, phys

    def loglike_phys(self, phys_per_system):
        """Physical parameters in object order, one [ncand, npar] array per system."""
        if not isinstance(phys_per_system, (list, tuple)):
            phys_per_system = [phys_per_system]
        ns = len(self.systems)
        arrs = [np.ascontiguousarray(np.atleast_2d(p), dtype=np.float64) for p in phys_per_system]
        arrs += [np.zeros((0, self.npars(s))) for s in range(len(arrs), ns)]
        outs = [np.zeros(a.shape[0]) for a in arrs]
        pp = (POINTER(c_double) * ns)(*[a.ctypes.data_as(POINTER(c_double)) for a in arrs])
        nc = (c_size_t * ns)(*[a.shape[0] for a in arrs])
        oo = (POINTER(c_double) * ns)(*[o.ctypes.data_as(POINTER(c_double)) for o in outs])
        _check(L.load().lensed_loglike_systems(self.ptr, pp, nc, oo), "loglike_systems")
        return outs if len(phys_per_system) > 1 else outs[0]

    def render(self, phys, s=0):
        """Dumper planes for one physical parameter vector (object order)."""
        spec = self.systems[s]
        h, w = spec.image.shape
        planes = {k: np.zeros((h, w), dtype=np.float32) for k in ("value", "error", "image", "chi2")}
        p = np.ascontiguousarray(phys, dtype=np.float64)
        fp = POINTER(c_float)
        _check(L.load().lensed_render(self.ptr, s, p.ctypes.data_as(POINTER(c_double)),
                                      planes["value"].ctypes.data_as(fp), planes["error"].ctypes.data_as(fp),
                                      planes["image"].ctypes.data_as(fp), planes["chi2"].ctypes.data_as(fp)),
               "render")
        return planes

    def set_output(self, path):
        """FITS file that every dumper() call writes the layers to (None: off)."""
        _check(L.load().lensed_set_output(self.ptr, None if path is None else str(path).encode()), "set_output")

    def write_layers(self, phys, path, s=0):
        """IMG, RES, RAW, ERR, WHT and PVL of one parameter vector (object order)."""
        p = np.ascontiguousarray(phys, dtype=np.float64)
        _check(L.load().lensed_write_layers(self.ptr, s, p.ctypes.data_as(POINTER(c_double)), str(path).encode()),
               "write_layers")

    def emulator_stats(self):
        """(candidates with exact-series tables, candidates with a table, largest
        held-out error of the certification) of the last emulated launch."""
        from ctypes import c_long
        nf, nc, mx = c_long(), c_long(), c_double()
        _check(L.load().lensed_emulator_stats(self.ptr, byref(nf), byref(nc), byref(mx)), "emulator_stats")
        return nf.value, nc.value, mx.value

    def emulator_cells(self):
        """(cells, certified cells, largest held-out error) of the emulator's
        certification at context creation."""
        from ctypes import c_long
        n, nc, mx = c_long(), c_long(), c_double()
        _check(L.load().lensed_emulator_cells(self.ptr, byref(n), byref(nc), byref(mx)), "emulator_cells")
        return n.value, nc.value, mx.value

    def emulator_rechecks(self):
        """(rechecked, changed, largest |delta logL|): exact re-evaluations of
        emulated candidates by the batched nested sampler."""
        from ctypes import c_long
        n, ch, mx = c_long(), c_long(), c_double()
        _check(L.load().lensed_emulator_rechecks(self.ptr, byref(n), byref(ch), byref(mx)), "emulator_rechecks")
        return n.value, ch.value, mx.value

    def render_stats(self):
        """(render items, items the pixel cull kept, 4 x 4 quads traced) of
        the last launch, and the cull pass's accumulated ms under
        LENSED_PROFILE."""
        from ctypes import c_long
        n, kept, quads, ms = c_long(), c_long(), c_long(), c_double()
        _check(L.load().lensed_render_stats(self.ptr, byref(n), byref(kept), byref(quads), byref(ms)),
               "render_stats")
        return n.value, kept.value, quads.value, ms.value

    def cull_ms(self):
        """The cull pass's accumulated ms under LENSED_PROFILE (without
        reading the render list back)."""
        ms = c_double()
        _check(L.load().lensed_render_stats(self.ptr, None, None, None, byref(ms)), "render_stats")
        return ms.value

    def set_tuning(self, **knobs):
        """Launch shapes for the launches from now on (lensed_tuning fields;
        unnamed ones keep their defaults).  Results do not depend on them."""
        t = L.Tuning(**knobs)
        _check(L.load().lensed_set_tuning(self.ptr, byref(t)), "set_tuning")

    def tuning(self):
        """The launch shapes of the last launch, resolved (or as requested
        before any launch), as a dict of the lensed_tuning fields."""
        t = L.Tuning()
        _check(L.load().lensed_get_tuning(self.ptr, byref(t)), "get_tuning")
        return {name: getattr(t, name) for name, _ in L.Tuning._fields_}

    def autotune_batch(self, s=0):
        """(best candidates per launch, {K: logL evals/s}) from timing
        lensed_loglike_batch() over K = 256 .. max_batch."""
        best = c_size_t()
        ks, rates, nk = (c_size_t * 32)(), (c_double * 32)(), c_int()
        _check(L.load().lensed_autotune_batch(self.ptr, s, byref(best), ks, rates, byref(nk)), "autotune_batch")
        return best.value, {ks[i]: rates[i] for i in range(nk.value)}

    def profile(self):
        """Accumulated stage times (ms) with the reference's stage names."""
        ms = (c_double * 8)()
        calls = L.load().lensed_profile(self.ptr, ms)
        names = ["+params", "-params", "set_params", "render", "convolve", "loglike", "+loglike_mem",
                 "-loglike_mem"]
        return calls, dict(zip(names, list(ms)))
