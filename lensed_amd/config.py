"""Lensed configurations (INI files and --name=value options) over the C ABI.

The parsing, the option table and the object loading are native
(csrc/config.cpp, after src/input.c, src/input/ini.c and
src/input/options.c); this module wraps them and turns a configuration into
the inputs of a likelihood context.
"""
from ctypes import byref, c_char_p

from . import _lib as L
from .data import DataError, Observation
from .lensed import SystemSpec


class ConfigError(RuntimeError):
    pass


def option_table():
    """[(name, type, help, default or None)] in the reference's order."""
    lib = L.load()
    out, i = [], 0
    while True:
        n, t, h, d = c_char_p(), c_char_p(), c_char_p(), c_char_p()
        if lib.lensed_config_option(i, byref(n), byref(t), byref(h), byref(d)) != 0:
            return out
        out.append((n.value.decode(), t.value.decode(), h.value.decode(), d.value.decode() if d.value else None))
        i += 1


class _ConfigObject:
    """An object owned by a configuration (the struct the context reads)."""

    def __init__(self, struct):
        self.obj = struct
        self.id = struct.id.decode()
        self.name = struct.name.decode()

    @property
    def type(self):
        return chr(self.obj.type)

    @property
    def params(self):
        return [self.obj.pars[i].name.decode() for i in range(self.obj.npars)]

    def label(self, pname):
        p = self.obj.pars[self.params.index(pname)]
        return p.label.decode() if p.label else None


class Config:
    """INI files in order, then command-line options, as `lensed` reads them."""

    def __init__(self, *ini_files, args=()):
        lib = L.load()
        self.ptr = lib.lensed_config_create()
        for f in ini_files:
            if lib.lensed_config_read_ini(self.ptr, str(f).encode()) != 0:
                raise ConfigError(L.last_error())
        for a in args:
            self.arg(a)

    def __del__(self):
        if getattr(self, "ptr", None):
            L.load().lensed_config_free(self.ptr)
            self.ptr = None

    def arg(self, text):
        if L.load().lensed_config_arg(self.ptr, str(text).encode()) != 0:
            raise ConfigError(L.last_error())

    def set(self, name, value):
        if L.load().lensed_config_set(self.ptr, name.encode(), str(value).encode()) != 0:
            raise ConfigError(L.last_error())

    def check(self):
        if L.load().lensed_config_check(self.ptr) != 0:
            raise ConfigError(L.last_error())

    def get(self, name):
        v = L.load().lensed_config_get(self.ptr, name.encode())
        return None if v is None else v.decode()

    def get_bool(self, name):
        return self.get(name) == "true"

    def get_int(self, name):
        return int(self.get(name))

    def get_real(self, name):
        return float(self.get(name))

    @property
    def objects(self):
        lib = L.load()
        n = lib.lensed_config_nobjects(self.ptr)
        arr = lib.lensed_config_objects(self.ptr)
        return [_ConfigObject(arr[i]) for i in range(n)]

    def observation(self):
        d = L.Data()
        lib = L.load()
        if lib.lensed_config_data(self.ptr, byref(d)) != 0:
            raise DataError(L.last_error())
        return Observation._from_struct(d)

    def system(self):
        """SystemSpec of this configuration (keeps the configuration alive)."""
        obs = self.observation()
        spec = SystemSpec(self.objects, {}, obs.image, obs.weight, psf=obs.psf, pcs=obs.pcs, rule=self.get("rule"))
        spec._config = self
        spec.observation = obs
        return spec
