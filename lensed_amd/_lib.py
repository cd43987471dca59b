"""ctypes binding of include/lensed.h (the C ABI of liblensed_amd.so)."""
import ctypes
import os
from ctypes import (POINTER, Structure, c_char_p, c_double, c_float, c_int, c_long, c_size_t,
                    c_void_p)

HERE = os.path.dirname(os.path.abspath(__file__))
# LENSED_LIB selects another build of the library (the sanitized host build
# of the CPU fuzz tests)
LIB_PATH = os.environ.get("LENSED_LIB") or os.path.join(HERE, "liblensed_amd.so")

OBJ_LENS, OBJ_SOURCE, OBJ_FOREGROUND = ord("L"), ord("S"), ord("F")
(PAR_PARAMETER, PAR_POSITION_X, PAR_POSITION_Y, PAR_RADIUS, PAR_MAGNITUDE, PAR_AXIS_RATIO,
 PAR_POS_ANGLE) = range(7)
LENSED_PROFILE, LENSED_EMULATE, LENSED_TABLE, LENSED_GRAPH, LENSED_NO_CULL, LENSED_AUTOTUNE = 1, 2, 4, 8, 16, 32


class Param(Structure):
    _fields_ = [("name", c_char_p), ("id", c_char_p), ("type", c_int), ("lower", c_double),
                ("upper", c_double), ("bounded", c_int), ("pri", c_void_p), ("wrap", c_int),
                ("ipp", c_int), ("derived", c_int), ("defval", c_int), ("label", c_char_p)]


class Object(Structure):
    _fields_ = [("type", c_int), ("size", c_size_t), ("id", c_char_p), ("name", c_char_p),
                ("npars", c_size_t), ("pars", POINTER(Param))]


class System(Structure):
    _fields_ = [("width", c_size_t), ("height", c_size_t), ("image", POINTER(c_float)),
                ("weight", POINTER(c_float)), ("psf", POINTER(c_float)), ("psfw", c_size_t),
                ("psfh", c_size_t), ("pcs", c_double * 4), ("rule", c_char_p), ("nobjs", c_size_t),
                ("objs", POINTER(Object))]


class Options(Structure):
    _fields_ = [("device", c_int), ("flags", c_int), ("max_batch", c_size_t), ("emulator", c_char_p),
                ("emulator_tol", c_double)]


class Tuning(Structure):
    _fields_ = [("list_fraction", c_double), ("render_grid", c_double), ("cull_grid", c_double),
                ("conv_grid", c_double), ("conv_rows", c_int), ("conv_order", c_int), ("default_ms", c_double),
                ("tuned_ms", c_double)]


class NsOptions(Structure):
    _fields_ = [("nlive", c_int), ("tol", c_double), ("efr", c_double), ("ceff", c_int), ("batch", c_int),
                ("seed", c_int), ("updint", c_int), ("maxiter", c_int), ("resume", c_int), ("outfile", c_int),
                ("wrap", POINTER(c_int)), ("root", c_char_p), ("feedback", c_int)]


class NsResult(Structure):
    _fields_ = [("logz", c_double), ("logz_err", c_double), ("info", c_double), ("max_loglike", c_double),
                ("nevals", c_long), ("niter", c_long)]


BATCH_FN = ctypes.CFUNCTYPE(None, POINTER(c_double), c_int, POINTER(c_double), POINTER(c_double), c_void_p)
LOGLIKE_FN = ctypes.CFUNCTYPE(None, POINTER(c_double), POINTER(c_int), POINTER(c_int), POINTER(c_double), c_void_p)
DUMPER_FN = ctypes.CFUNCTYPE(None, POINTER(c_int), POINTER(c_int), POINTER(c_int), POINTER(POINTER(c_double)),
                             POINTER(POINTER(c_double)), POINTER(POINTER(c_double)), POINTER(c_double),
                             POINTER(c_double), POINTER(c_double), POINTER(c_double), c_void_p)

class DataOptions(Structure):
    _fields_ = [("image", c_char_p), ("weight", c_char_p), ("gain", c_char_p), ("offset", c_double),
                ("bscale", c_double), ("xweight", c_char_p), ("mask", c_char_p), ("psf", c_char_p)]


class Data(Structure):
    _fields_ = [("width", c_size_t), ("height", c_size_t), ("image", POINTER(c_float)),
                ("weight", POINTER(c_float)), ("psf", POINTER(c_float)), ("psfw", c_size_t), ("psfh", c_size_t),
                ("pcs", c_double * 4), ("masked", c_size_t), ("mode", c_double), ("fwhm", c_double),
                ("warnings", c_int)]


LENSED_WARN_SKY_NOT_ZERO, LENSED_WARN_NO_OFFSET, LENSED_WARN_NEGATIVE_WEIGHT = 1, 2, 4

# every entry point declared in include/lensed.h: name -> (restype, argtypes)
SIGNATURES = {
    "object_program": (None, [c_char_p, POINTER(c_size_t), POINTER(POINTER(c_char_p))]),
    "main_program": (None, [c_size_t, POINTER(Object), POINTER(c_size_t), POINTER(POINTER(c_char_p))]),
    "kernel_options": (c_void_p, [c_size_t, c_size_t, c_int, c_size_t, c_size_t, c_size_t, POINTER(c_char_p)]),
    "kernel_name": (c_void_p, [c_char_p, c_char_p]),
    "lensed_program_compile": (c_long, [c_char_p, c_char_p, POINTER(c_void_p)]),
    "lensed_model_source": (c_void_p, [c_size_t, POINTER(Object)]),
    "lensed_model_source_tabulated": (c_void_p, [c_size_t, POINTER(Object)]),
    "prior_read": (c_void_p, [c_char_p]),
    "prior_default": (c_void_p, [c_double]),
    "prior_free": (None, [c_void_p]),
    "prior_print": (None, [c_void_p, c_char_p, c_size_t]),
    "prior_apply": (c_double, [c_void_p, c_double]),
    "prior_lower": (c_double, [c_void_p]),
    "prior_upper": (c_double, [c_void_p]),
    "prior_pseudo": (c_int, [c_void_p]),
    "lensed_object_load": (c_int, [POINTER(Object), c_char_p, c_char_p]),
    "lensed_object_free": (None, [POINTER(Object)]),
    "lensed_quad_size": (c_int, [c_char_p]),
    "lensed_quad_rule": (c_int, [c_char_p, POINTER(c_float), POINTER(c_float), c_double, c_double]),
    "lensed_quad_name": (c_char_p, [c_int]),
    "lensed_quad_table": (c_int, [c_char_p, POINTER(c_double), POINTER(c_double), POINTER(c_double)]),
    "lensed_create": (c_void_p, [POINTER(System), c_size_t, POINTER(Options)]),
    "lensed_destroy": (None, [c_void_p]),
    "lensed_last_error": (c_char_p, []),
    "lensed_npars": (c_size_t, [c_void_p, c_size_t]),
    "lensed_ndims": (c_size_t, [c_void_p, c_size_t]),
    "lensed_pmap": (c_int, [c_void_p, c_size_t, POINTER(c_size_t)]),
    "loglike": (None, [POINTER(c_double), POINTER(c_int), POINTER(c_int), POINTER(c_double), c_void_p]),
    "dumper": (None, [POINTER(c_int), POINTER(c_int), POINTER(c_int), POINTER(POINTER(c_double)),
                      POINTER(POINTER(c_double)), POINTER(POINTER(c_double)), POINTER(c_double),
                      POINTER(c_double), POINTER(c_double), POINTER(c_double), c_void_p]),
    "lensed_loglike_batch": (c_int, [c_void_p, c_size_t, POINTER(c_double), c_size_t, POINTER(c_double),
                                     POINTER(c_double)]),
    "lensed_loglike_batch_noalloc": (c_int, [c_void_p, c_size_t, POINTER(c_double), c_size_t, POINTER(c_double),
                                             POINTER(c_double)]),
    "lensed_loglike_systems": (c_int, [c_void_p, POINTER(POINTER(c_double)), POINTER(c_size_t),
                                       POINTER(POINTER(c_double))]),
    "lensed_render": (c_int, [c_void_p, c_size_t, POINTER(c_double), POINTER(c_float), POINTER(c_float),
                              POINTER(c_float), POINTER(c_float)]),
    "lensed_profile": (c_long, [c_void_p, POINTER(c_double)]),
    "lensed_emulator_train": (c_int, [c_char_p, ctypes.c_longlong, c_int, c_int, c_double, ctypes.c_ulonglong,
                                      POINTER(c_double)]),
    "lensed_emulator_read": (c_int, [c_char_p, POINTER(c_float)]),
    "lensed_emulator_loss_grad": (c_double, [POINTER(c_float), POINTER(c_float), POINTER(c_float), c_int,
                                             POINTER(c_float)]),
    "lensed_emulator_eval": (c_int, [POINTER(c_float), POINTER(c_float), POINTER(c_float), c_int, c_int,
                                     POINTER(c_float), POINTER(c_float)]),
    "lensed_emulator_stats": (c_int, [c_void_p, POINTER(c_long), POINTER(c_long), POINTER(c_double)]),
    "lensed_emulator_cells": (c_int, [c_void_p, POINTER(c_long), POINTER(c_long), POINTER(c_double)]),
    "lensed_emulator_rechecks": (c_int, [c_void_p, POINTER(c_long), POINTER(c_long), POINTER(c_double)]),
    "lensed_render_stats": (c_int, [c_void_p, POINTER(c_long), POINTER(c_long), POINTER(c_long), POINTER(c_double)]),
    "lensed_set_tuning": (c_int, [c_void_p, POINTER(Tuning)]),
    "lensed_autotune_batch": (c_int, [c_void_p, c_size_t, POINTER(c_size_t), POINTER(c_size_t), POINTER(c_double),
                                      POINTER(c_int)]),
    "lensed_get_tuning": (c_int, [c_void_p, POINTER(Tuning)]),
    "lensed_fits_read": (c_int, [c_char_p, POINTER(c_size_t), POINTER(c_size_t), POINTER(POINTER(c_float)),
                                 POINTER(c_double)]),
    "lensed_fits_free": (None, [c_void_p]),
    "lensed_read_pcs": (c_int, [c_char_p, POINTER(c_double)]),
    "lensed_fits_write": (c_int, [c_char_p, c_size_t, c_size_t, c_size_t, POINTER(POINTER(c_float)),
                                  POINTER(c_char_p)]),
    "lensed_find_mode": (None, [c_size_t, POINTER(c_float), POINTER(c_float), POINTER(c_double),
                                POINTER(c_double)]),
    "lensed_data_read": (c_int, [POINTER(DataOptions), POINTER(Data)]),
    "lensed_data_free": (None, [POINTER(Data)]),
    "lensed_set_output": (c_int, [c_void_p, c_char_p]),
    "lensed_write_layers": (c_int, [c_void_p, c_size_t, POINTER(c_double), c_char_p]),
    "lensed_config_create": (c_void_p, []),
    "lensed_config_free": (None, [c_void_p]),
    "lensed_config_read_ini": (c_int, [c_void_p, c_char_p]),
    "lensed_config_set": (c_int, [c_void_p, c_char_p, c_char_p]),
    "lensed_config_arg": (c_int, [c_void_p, c_char_p]),
    "lensed_config_check": (c_int, [c_void_p]),
    "lensed_config_get": (c_char_p, [c_void_p, c_char_p]),
    "lensed_config_option": (c_int, [c_int, POINTER(c_char_p), POINTER(c_char_p), POINTER(c_char_p),
                                     POINTER(c_char_p)]),
    "lensed_config_nobjects": (c_size_t, [c_void_p]),
    "lensed_config_objects": (POINTER(Object), [c_void_p]),
    "lensed_config_data": (c_int, [c_void_p, POINTER(Data)]),
    "lensed_write_paramfiles": (c_int, [c_void_p, c_size_t, c_char_p]),
    "run": (None, [c_int, c_int, c_int, c_int, c_double, c_double, c_int, c_int, c_int, c_int, c_int, c_double,
                   c_char_p, c_int, POINTER(c_int), c_int, c_int, c_int, c_int, c_double, c_int, LOGLIKE_FN,
                   DUMPER_FN, c_void_p]),
    "lensed_nested_sample": (c_int, [c_int, c_int, BATCH_FN, c_void_p, POINTER(NsOptions), POINTER(c_double),
                                     POINTER(c_double), POINTER(c_double), POINTER(NsResult)]),
    "lensed_nested_run": (c_int, [c_void_p, c_size_t, POINTER(NsOptions), POINTER(c_double), POINTER(c_double),
                                  POINTER(c_double), POINTER(NsResult)]),
    "lensed_results": (c_int, [c_void_p, POINTER(c_double), POINTER(c_double), POINTER(c_double),
                               POINTER(c_double), POINTER(c_double), POINTER(c_double), POINTER(c_double)]),
}

_lib = None


def load():
    """Load liblensed_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("liblensed_amd.so is not built; run python -m lensed_amd.build")
        lib = ctypes.CDLL(LIB_PATH)
        lib.free = ctypes.CDLL(None).free
        lib.free.argtypes = [c_void_p]
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def take_string(ptr):
    """Copy and free a malloc'd C string returned by the library."""
    lib = load()
    s = ctypes.cast(ptr, c_char_p).value.decode()
    lib.free(ptr)
    return s


def last_error():
    return (load().lensed_last_error() or b"").decode()
