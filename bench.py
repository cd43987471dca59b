"""Benchmark: logL evaluations per second of the Lensed likelihood on MI355X.

Workload (the configuration the metric is quoted on): 256 x 256 px, EPL +
external shear lens, Gaussian source (sigma 0.1" = 2 px), 21 x 21 Gaussian
PSF (FWHM 3 px), g3k7 quadrature (49 rays per pixel), mock data at the true
parameters plus seeded noise (sigma 0.05).  A step is one batched launch of
4096 candidate parameter vectors drawn uniformly from the priors, fresh
draws every launch as a sampler's launches are: parameters
in, one logL per candidate out (set_params -> render -> conv+chi^2 ->
finalize, replayed from a captured hipGraph).

Also reported: the plain GPU path (one candidate per call through the
MultiNest `loglike` callback, the reference's launch structure), the render
kernel's roofline position from hipEvent timings on its own stream, the
CPU restatement of the reference path timed on the host cores, and one full
nested-sampling run of the workload, batched + emulated against the plain
run() path (wall time and the difference of posterior means).

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--no-nested]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "logL evals/sec at 256² px, batched+emulated vs plain GPU, |ΔlogL|≤0.01"
PEAK_FP32_NONPACKED = 78.6  # TFLOP/s: one wave64 FMA per CU per cycle (packed FP32 is not used)
PEAK_HBM = 8000.0  # GB/s
# The render kernel is bound by VALU issue: one wave64 VALU instruction per
# CU per cycle, 256 CUs at 2.4 GHz
VALU_ISSUE_PEAK = 256 * 2.4e9  # wave instructions per second
# reference-formula work of this workload (add/mul/div 1, FMA 2,
# transcendental 1): 165 flop per ray, 49 rays per pixel
REF_FLOP_PER_RAY = 165
PROFILES = os.path.join(ROOT, "profiles")
RENDER_PROFILE = "lensed_render_c3_counters.md"
CONV_PROFILE = "lensed_conv_c3_counters.md"


def profile_figure(name, key):
    """A figure a profiles/ summary records as a line `key: value` (None if
    the summary or the line is missing)."""
    try:
        with open(os.path.join(PROFILES, name)) as f:
            for line in f:
                if line.strip().startswith(key + ":"):
                    return float(line.split(":", 1)[1].split()[0])
    except (OSError, ValueError):
        pass
    return None


def make_workload(K):
    from lensed_amd import workloads as W
    from lensed_amd.lensed import Context, LensedObject, SystemSpec

    wl = W.c3()
    objs = [LensedObject(oid, name) for oid, name in wl.objects]
    pids = [f"{o.id}.{p}" for o in objs for p in o.params]

    def render(objects, params, shape, psf):
        pri = {k: repr(float(v)) for k, v in zip(pids, params)}
        spec = SystemSpec(objects, pri, np.zeros(shape, np.float32), np.ones(shape, np.float32), psf=psf)
        with Context(spec, max_batch=1) as c:
            return c.render(params)["image"].astype(np.float64)

    wl.make_data(render, pids, mag_keys=("source.mag",))
    wl.ncand = K
    return wl, pids


EMULATOR = os.path.join(ROOT, "lensed_amd", "data", "epl_mlp.bin")


def draws(wl, ndim, K, n):
    """n batches of K unit-cube draws, each launch its own as a sampler's
    launches are (nothing a launch leaves behind matches the next one's
    candidates); the last is the workload's seeded set."""
    rng = np.random.default_rng(wl.cand_seed + 1)
    return [rng.random((K, ndim)) for _ in range(n - 1)] + [wl.candidates(ndim, K)]


def time_batched(wl, K, steps, warmup, flags):
    """Seconds for `steps` launches of K candidates (fresh draws every
    launch); logL of the last; the emulator's statistics (last launch, and
    the certification) when the emulated path runs; the launch shapes the
    autotuner chose (at the first warm-up launch)."""
    from lensed_amd._lib import LENSED_EMULATE
    from lensed_amd.lensed import Context
    emu = EMULATOR if flags & LENSED_EMULATE else None
    with Context(wl.spec(), max_batch=K, flags=flags, emulator=emu) as ctx:
        sets = draws(wl, ctx.ndims(), K, warmup + steps)
        for i in range(warmup):
            ctx.loglike_batch(sets[i])
        t0 = time.perf_counter()
        for i in range(warmup, warmup + steps):
            lnew, _ = ctx.loglike_batch(sets[i])
        dt = time.perf_counter() - t0
        stats = (ctx.emulator_stats(), ctx.emulator_cells()) if emu else None
        tuning = ctx.tuning()
    return dt, lnew, stats, tuning


def emulator_block(stats):
    (nf, nc, _), (ncells, ncert, cert_err) = stats
    return {"tables_from_exact_series": nf, "tables": nc, "certified_cells": ncert, "cells": ncells,
            "largest_heldout_abs_error_vs_fp64_series": cert_err, "certification_bound": 4e-6,
            "weights": "lensed_amd/data/epl_mlp.bin (2^22 seeded samples, 40 epochs, trained on the GPU by "
                       "lensed_amd.train_emulator)",
            "note": "the network builds a candidate's table only in a (q, t) cell certified at the bound "
                    "(certified once per weights when the context is created); the exact series elsewhere"}


def time_plain(wl, n):
    """The sampler's calling pattern: one C loglike() callback per candidate
    (ctypes arguments prebuilt, so the loop adds ~1 us per call)."""
    import ctypes
    from ctypes import byref, c_double, c_int
    from lensed_amd import _lib as L
    from lensed_amd.lensed import Context
    lib = L.load()
    with Context(wl.spec(), max_batch=1) as ctx:
        ndim, npar = ctx.ndims(), ctx.npars()
        cubes = wl.candidates(ndim, n)
        bufs = [(c_double * npar)(*cubes[i], *([0.0] * (npar - ndim))) for i in range(n)]
        warm = [(c_double * npar)(*cubes[i], *([0.0] * (npar - ndim))) for i in range(min(n, 20))]
        nd, npr, ptr = byref(c_int(ndim)), byref(c_int(npar)), ctypes.c_void_p(ctx.ptr)
        lnew = [c_double() for _ in range(n)]
        for b in warm:
            lib.loglike(b, nd, npr, byref(c_double()), ptr)
        f = lib.loglike
        t0 = time.perf_counter()
        for i in range(n):
            f(bufs[i], nd, npr, lnew[i], ptr)
        dt = time.perf_counter() - t0
    return n / dt, np.array([v.value for v in lnew])


TUNED = 0  # flags of the timed legs (LENSED_AUTOTUNE, set in main)


def kernel_times(wl, K, launches, flags=0, warm=30):
    """Stage times (ms per launch, hipEvents on the launch stream, fresh
    draws every launch), with the render stage split into the cull pass and
    the render kernel, and what the render traced in the last launch (the
    workload's seeded candidates).  `warm` launches first: after a pause
    the first launches run slower (C3's render 2.3 ms falling to 2.0 ms over
    about ten launches, as the GPU's clock comes up), and the timed loop runs
    warm."""
    from lensed_amd._lib import LENSED_EMULATE, LENSED_PROFILE
    from lensed_amd.lensed import Context
    emu = EMULATOR if flags & LENSED_EMULATE else None
    with Context(wl.spec(), max_batch=K, flags=LENSED_PROFILE | TUNED | flags, emulator=emu) as ctx:
        sets = draws(wl, ctx.ndims(), K, warm + launches)
        for i in range(warm):
            ctx.loglike_batch(sets[i])
        calls0, ms0 = ctx.profile()
        cull0 = ctx.cull_ms()
        for i in range(launches):
            ctx.loglike_batch(sets[warm + i])
        calls, ms = ctx.profile()
        items, kept, quads, cull1 = ctx.render_stats()
    n = calls - calls0
    out = {k: (ms[k] - ms0[k]) / n for k in ms}
    out["cull_pass"] = (cull1 - cull0) / n
    out["render_kernel"] = out["render"] - out["cull_pass"]
    return out, {"render_items": items, "items_kept_by_cull": kept, "quads_traced": quads,
                 "rays_traced_frac": quads / (items * 16)}


def render_roofline(kt, tr, K):
    """The render kernel against its bound, VALU issue.  Its VALU wave
    instructions per traced 4 x 4 quad (16 pixels, 49 rays each), less the
    transcendental ones (which issue to their own unit beside the VALU: with
    them the kernel counts 1.07 VALU instructions per CU per cycle, without
    them 0.95), come from the rocprofv3 SQ_INSTS_VALU and
    SQ_INSTS_VALU_TRANS_F32 counts of the same launch, recorded in profiles/;
    the quads traced and the kernel's time from this run."""
    quads = tr["quads_traced"]
    rays = quads * 16 * 49
    s = kt["render_kernel"] / 1e3
    per_all = profile_figure(RENDER_PROFILE, "valu_wave_instructions_per_traced_quad")
    per_trans = profile_figure(RENDER_PROFILE, "valu_transcendental_wave_instructions_per_traced_quad")
    per_q = per_all - per_trans if per_all and per_trans is not None else None
    flop_q = profile_figure(RENDER_PROFILE, "fp32_flop_per_traced_quad")
    per_cycle = profile_figure(RENDER_PROFILE, "valu_wave_instructions_per_cu_cycle")
    traffic = profile_figure(RENDER_PROFILE, "hbm_bytes_per_launch")
    achieved = per_q * quads / s / 1e9 if per_q else None
    peak = VALU_ISSUE_PEAK / 1e9
    return {"bound": "valu", "kernel": "lensed_render", "unit": "G wave-instructions/s",
            "achieved": None if achieved is None else round(achieved, 2), "peak": peak,
            "frac": None if achieved is None else round(achieved / peak, 4),
            "traffic": None if traffic is None else int(traffic * K / 4096),
            "traffic_unit": "HBM bytes per launch (rocprofv3 --pmc FETCH_SIZE + WRITE_SIZE, calibrated)",
            "avg_launch_ms": round(kt["render_kernel"], 4), "cull_pass_ms": round(kt["cull_pass"], 4),
            "valu_wave_instructions_per_traced_quad": per_all,
            "valu_transcendental_wave_instructions_per_traced_quad": per_trans, "quads_traced": quads,
            "rays_traced_per_launch": rays, "rays_traced_frac": round(tr["rays_traced_frac"], 4),
            "valu_wave_instructions_per_cu_cycle_from_profile": per_cycle,
            "fp32_vector": None if not flop_q else {
                "tflops": round(flop_q * quads / s / 1e12, 2),
                "frac_of_published_peak": round(flop_q * quads / s / 157.3e12, 4),
                "frac_of_nonpacked_peak": round(flop_q * quads / s / (PEAK_FP32_NONPACKED * 1e12), 4),
                "flop_per_ray": round(flop_q / (16 * 49), 1),
                "note": "FP32 flop the kernel executes, from the rocprofv3 SQ_INSTS_VALU_FMA_F32 (x2), ADD_F32, "
                        "MUL_F32 and TRANS_F32 counts of the seeded launch (x64 lanes) per traced quad, over "
                        "this run's quads and kernel time; the published 157.3 TFLOP/s counts packed FP32, "
                        "which the build disables, so one FP32 operation per lane per cycle (78.6 TFLOP/s) is "
                        "the ceiling here and the ray's non-FP32 work (LDS, conversions, copysigns) issues "
                        "between them"},
            "reference_flop_equivalent": {
                "tflops": round(REF_FLOP_PER_RAY * rays / s / 1e12, 2),
                "note": "the reference formulas' 165 flop per traced ray (its tenth-order series); the table "
                        "path computes the same ray in fewer instructions, so this rate is not a fraction of "
                        "any hardware peak"},
            "note": "frac = non-transcendental VALU wave instructions issued / (256 CUs x 2.4 GHz x kernel "
                    "time), the nominal issue peak of the VALU (one wave64 instruction per CU per cycle); at "
                    "the clock the kernel runs (GRBM_GUI_ACTIVE in the profile) it issues 0.95 of them per CU "
                    "per cycle, with the transcendentals on their own unit beside them"}


def conv_roofline(kt, K, conv_rows, npix=65536, taps=21):
    """The convolution + chi^2 kernel against its bounds: the VALU issue
    rate it reaches (its instruction count from the profile), the bytes it
    reads against HBM, and the dense convolution's FP32 work as an
    equivalent rate.  The kernel is the one the launch shapes chose: the
    column sweep (conv_rows 1) or the tile."""
    sweep = conv_rows == 1
    kernel = "lensed_conv_chi2_sweep" if sweep else "lensed_conv_chi2"
    s = kt["convolve"] / 1e3
    flop = (2 * 2 * taps + 4) * npix * K  # row pass, column pass, w (m - d)^2 of every pixel
    model_bytes = 4 * npix * K
    which = "sweep" if sweep else "tile"
    fetch = profile_figure(CONV_PROFILE, "hbm_bytes_per_launch_" + which)
    valu = profile_figure(CONV_PROFILE, "valu_wave_instructions_per_launch_" + which)
    hbm = None if fetch is None else fetch * K / 4096
    valu = None if valu is None else valu * K / 4096
    return {"kernel": kernel, "avg_launch_ms": round(kt["convolve"], 4),
            "valu_wave_instructions_per_launch": None if valu is None else int(valu),
            "valu_issue_frac": None if valu is None else round(valu / s / VALU_ISSUE_PEAK, 4),
            "dense_equivalent_tflops": round(flop / s / 1e12, 3),
            "dense_equivalent_note": "the separable passes and chi^2 term of every pixel; the sweep skips the "
                                     "row passes of dark rows, the column passes of dark windows and whole "
                                     "dark end blocks (whose chi^2 is the data's own), so this is not a "
                                     "fraction of an FP32 peak",
            "model_bytes_per_launch": model_bytes,
            "hbm_bytes_per_launch": None if hbm is None else int(hbm),
            "hbm_over_model_bytes": None if hbm is None else round(hbm / model_bytes, 3),
            "achieved_gbps": None if hbm is None else round(hbm / s / 1e9, 1),
            "frac_hbm": None if hbm is None else round(hbm / s / 1e9 / PEAK_HBM, 4)}


def near_truth(wl, pids, n, scale=2e-4, seed=7):
    """Parameter vectors scattered around the truth: the regime in which
    logL differences decide nested-sampling acceptance."""
    rng = np.random.default_rng(seed)
    truth = np.array([wl.truth[k] for k in pids])
    out = truth + scale * rng.normal(size=(n, len(pids))) * np.maximum(np.abs(truth), 1.0)
    qi = pids.index("lens.q")
    out[:, qi] = np.clip(out[:, qi], 0.3, 1.0)
    return out


def accuracy(wl, pids, K, flags):
    """|delta logL| of a path against the exact per-ray path on the same
    candidates: near the truth, and relative over the prior draws."""
    from lensed_amd._lib import LENSED_EMULATE
    from lensed_amd.lensed import Context
    emu = EMULATOR if flags & LENSED_EMULATE else None
    phys = near_truth(wl, pids, 512)
    with Context(wl.spec(), max_batch=K) as ex:
        cubes = wl.candidates(ex.ndims(), K)
        le = ex.loglike_phys(phys)
        lc, _ = ex.loglike_batch(cubes)
    with Context(wl.spec(), max_batch=K, flags=flags, emulator=emu) as c:
        lo = c.loglike_phys(phys)
        loc, _ = c.loglike_batch(cubes)
    return {"max_abs_dlogL_near_truth": float(np.max(np.abs(lo - le))),
            "near_truth_candidates": len(phys),
            "max_rel_dlogL_prior_draws": float(np.max(np.abs(loc - lc) / np.abs(lc)))}


def cpu_baseline(wl, pids, n=20, warm=3):
    """The CPU baseline: the oracle's fast FP32 build (render + convolve +
    chi^2, OpenMP over pixels, -O3 -ffast-math) timed per logL at C3 (this
    workload) and C1, median over n evaluations of prior draws after `warm`,
    on all host threads and on one."""
    from lensed_amd import quadrature as Q
    from lensed_amd import workloads as W
    from oracle import oracle as O
    qq, ww = Q.rule("g3k7")
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

    def draws_of(w, keys, count, seed):
        rng = np.random.default_rng(seed)
        rngs = [w.prior_range(k) for k in keys]
        return [np.array([a + x * (b - a) for x, (a, b) in zip(rng.random(len(keys)), rngs)], np.float32)
                for _ in range(count)]

    def latency(w, keys, nthreads):
        O.set_threads(nthreads)
        times = []
        for i, p in enumerate(draws_of(w, keys, warm + n, w.cand_seed)):
            t0 = time.perf_counter()
            v, _ = O.render(w.objects, p, w.shape, qq, ww, dtype=np.float32, fast=True)
            m = O.convolve(v, w.psf, dtype=np.float32, fast=True) if w.psf is not None else v
            O.chi2_f32terms(w.image, w.weight, m)
            if i >= warm:
                times.append(time.perf_counter() - t0)
        return float(np.median(times))

    c1 = W.c1()
    c1_ids = [f"{o}.{p}" for o, nm in c1.objects for p in O.PARAMS[nm]]
    pts, wht, err = Q.table("g3k7")
    c1.make_data(lambda objs, p, shape, psf: O.render(objs, p, shape, pts, np.stack([wht, err], 1))[0],
                 c1_ids, mag_keys=("source.mag",))
    t = {"c3_all": latency(wl, pids, threads), "c3_one": latency(wl, pids, 1),
         "c1_all": latency(c1, c1_ids, threads), "c1_one": latency(c1, c1_ids, 1)}
    O.set_threads(threads)
    return {"value": round(1.0 / t["c3_all"], 3), "unit": "logL evals/s", "cores": threads, "kind": "port",
           "sample": f"C3 (this workload) and C1, {n} evaluations of prior draws each after {warm} warm-up, "
                     "median latency; oracle/liboracle_fast.so FP32 (-O3 -ffast-math, OpenMP over pixels), "
                     "one candidate at a time",
           "median_ms_per_logL": {"C3_all_threads": round(t["c3_all"] * 1e3, 2),
                                  "C3_one_thread": round(t["c3_one"] * 1e3, 2),
                                  "C1_all_threads": round(t["c1_all"] * 1e3, 3),
                                  "C1_one_thread": round(t["c1_one"] * 1e3, 3)},
           "value_one_thread": round(1.0 / t["c3_one"], 3)}


def oracle_accuracy(wl, pids, K, flags):
    """The FP64 oracle (the checker) against the GPU path on the timed
    launch's own candidates: 8 seeded picks of its prior draws and the
    truth, model images and logL."""
    from lensed_amd import quadrature as Q
    from lensed_amd._lib import LENSED_EMULATE
    from lensed_amd.lensed import Context
    from oracle import oracle as O
    emu = EMULATOR if flags & LENSED_EMULATE else None
    pts, wht, err = Q.table("g3k7")
    with Context(wl.spec(), max_batch=K, flags=flags, emulator=emu) as ctx:
        cubes = wl.candidates(ctx.ndims(), K)
        lnew, phys_s = ctx.loglike_batch(cubes)
        phys = np.zeros((K, len(pids)))
        phys[:, ctx.pmap()] = phys_s
        picks = list(np.sort(np.random.default_rng(wl.cand_seed + 7).choice(K, 8, replace=False)))
        truth = wl.truth_vector(pids)
        lt = ctx.loglike_phys(truth[None])[0]
        rows = []
        for i, (p, lg) in enumerate([(phys[j], lnew[j]) for j in picks] + [(truth, lt)]):
            img = ctx.render(p)["image"]
            pcs = [1.0 + 0.5 * (wl.psf.shape[1] % 2 == 0), 1.0 + 0.5 * (wl.psf.shape[0] % 2 == 0), 1.0, 1.0]
            v, _ = O.render(wl.objects, p, wl.shape, pts, np.stack([wht, err], 1), pcs)
            m64 = O.convolve(v, wl.psf)
            l64 = -0.5 * O.chi2_d(wl.image, wl.weight, m64)
            peak = np.abs(m64).max()
            mask = np.abs(m64) > 1e-3 * peak
            rel = float(np.max(np.abs(img - m64)[mask] / np.abs(m64[mask]))) if mask.any() else 0.0
            rows.append((rel, abs(lg - l64), abs(lg - l64) / max(1.0, abs(l64)), i == 8))
    acc = {"candidates": f"8 seeded picks of the timed launch's {K} prior draws, and the truth",
           "model_max_rel_err": max(r[0] for r in rows),
           "max_abs_dlogL_prior_draws": max(r[1] for r in rows if not r[3]),
           "max_rel_dlogL_prior_draws": max(r[2] for r in rows if not r[3]),
           "abs_dlogL_truth": [r[1] for r in rows if r[3]][0],
           "bars": "model 1e-4 relative where above 1e-3 of its peak; |dlogL| 1e-2 near the truth, "
                   "1e-2 + 1e-5 |logL| for prior draws (FP32 model rounding at |logL| up to 1e7)"}
    return acc


def nested_compare(wl, pids, K, nlive=400):
    """A full nested-sampling run of the workload with Lensed's default
    sampler settings (constant efficiency, target acceptance 0.05), batched +
    emulated against the plain run() path: wall time, evaluations and the
    largest difference of posterior means in units of the plain run's sigma."""
    import ctypes
    from lensed_amd import _lib as L
    from lensed_amd import nested
    from lensed_amd._lib import LENSED_EMULATE, LENSED_GRAPH
    from lensed_amd.lensed import Context
    lib = L.load()
    kw = dict(nlive=nlive, tol=0.1, efr=0.05, ceff=True)
    with Context(wl.spec(), max_batch=K, flags=LENSED_GRAPH | LENSED_EMULATE, emulator=EMULATOR) as ctx:
        npar, pmap = ctx.npars(), ctx.pmap()
        t0 = time.perf_counter()
        rb = nested.run_context(ctx, 0, batch=K, seed=33, **kw)
        t_batch = time.perf_counter() - t0
    calls = [0]

    with Context(wl.spec(), max_batch=1) as ctx:
        base = ctypes.cast(lib.loglike, L.LOGLIKE_FN)

        def counted(cube, nd, np_, lnew, c):
            calls[0] += 1
            base(cube, nd, np_, lnew, c)
        t0 = time.perf_counter()
        nested.run_multinest_abi(L.LOGLIKE_FN(counted), ctx.ndims(), npar, ctypes.cast(lib.dumper, L.DUMPER_FN),
                                 updint=100000, seed=31, wrap=ctx.wrap(), context=ctx.ptr, **kw)
        t_plain = time.perf_counter() - t0
        n = npar
        mean, sigma, ml, mp = [(ctypes.c_double * n)() for _ in range(4)]
        lz, lze, mx = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        lib.lensed_results(ctx.ptr, mean, sigma, ml, mp, ctypes.byref(lz), ctypes.byref(lze), ctypes.byref(mx))
    mean_b = np.zeros(npar)
    mean_b[pmap] = rb.mean
    dev = np.abs(np.array(mean) - mean_b) / np.array(sigma)
    return {"settings": "nlive %d, ceff on, acc 0.05, tol 0.1 (Lensed defaults, src/input/options.c:150-227)" % nlive,
            "batched_emulated": {"seconds": round(t_batch, 3), "evaluations": rb.nevals, "iterations": rb.niter,
                                 "logz": round(rb.logz, 3), "logz_err": round(rb.logz_err, 3),
                                 "max_candidates_per_launch": K},
            "plain": {"seconds": round(t_plain, 3), "evaluations": calls[0], "logz": round(lz.value, 3),
                      "logz_err": round(lze.value, 3)},
            "wallclock_speedup": round(t_plain / t_batch, 3),
            "max_abs_mean_difference_in_sigma": round(float(dev.max()), 4),
            "logz_difference_in_sigma": round(abs(rb.logz - lz.value) / math.hypot(rb.logz_err, lze.value), 3)}


def c4_workload():
    """The uneven-batch workload's 64 systems with mock data, and 256 seeded
    prior draws per system (physical parameters)."""
    from lensed_amd import workloads as W
    from lensed_amd.lensed import Context, LensedObject, SystemSpec
    wls = W.c4()
    for wl in wls:
        objs = [LensedObject(oid, name) for oid, name in wl.objects]
        pids = [f"{o.id}.{p}" for o in objs for p in o.params]

        def render(objects, params, shape, psf, pids=pids):
            pri = {k: repr(float(v)) for k, v in zip(pids, params)}
            spec = SystemSpec(objects, pri, np.zeros(shape, np.float32), np.ones(shape, np.float32), psf=psf)
            with Context(spec, max_batch=1) as c:
                return c.render(params)["image"].astype(np.float64)
        wl.make_data(render, pids, mag_keys=("source.mag",))
        wl.pids = pids
    rng = np.random.default_rng(404)
    phys = [np.array([[rng.uniform(*wl.prior_range(k)) for k in wl.pids] for _ in range(256)]) for wl in wls]
    return wls, phys


def stage_ms(ctx, run, launches):
    """ms per launch of each profiled stage over `launches` runs after one."""
    run()
    c0, m0 = ctx.profile()
    for _ in range(launches):
        run()
    c1, m1 = ctx.profile()
    return {k: round((m1[k] - m0[k]) / (c1 - c0), 4) for k in m1}


def c4_launches(launches=3, flags_extra=0):
    """The uneven-batch workload (64 systems of 64..512 px, SIE or EPL +
    Sersic, 15 x 15 or 31 x 31 PSFs, 256 prior draws per system, one launch):
    ms per launch and per stage, on the table path with the render walking
    the pixel cull's list (list_fraction 2) and walking the items (-1), and
    without culling."""
    from lensed_amd._lib import LENSED_NO_CULL, LENSED_PROFILE, LENSED_TABLE
    from lensed_amd.lensed import Context
    wls, phys = c4_workload()
    out = {}
    for name, extra, frac in (("list", 0, 2.0), ("walk", 0, -1.0), ("no_cull", LENSED_NO_CULL, 0.0)):
        with Context([wl.spec() for wl in wls], max_batch=256,
                     flags=LENSED_TABLE | LENSED_PROFILE | extra | flags_extra) as ctx:
            ctx.set_tuning(list_fraction=frac)
            ctx.loglike_phys(phys)
            c0, m0 = ctx.profile()
            cull0 = ctx.cull_ms()
            for _ in range(launches):
                ctx.loglike_phys(phys)
            c1, m1 = ctx.profile()
            items, kept, quads, cull1 = ctx.render_stats()
        n = c1 - c0
        st = {k: round((m1[k] - m0[k]) / n, 3) for k in m1}
        st["cull_pass"] = round((cull1 - cull0) / n, 3)
        out[name] = {"ms_per_launch": round(sum(st[k] for k in m1), 3), "stages_ms": st,
                     "render_items": items, "items_kept": kept, "quads_traced": quads}
    npix = sum(wl.image.size for wl in wls)
    return {"workload": f"C4: 64 systems, {npix / 1e6:.2f} Mpx, 256 prior draws each, EPL table path",
            "launches": launches, "paths": out,
            "evals_per_s": {k: round(64 * 256 / (v["ms_per_launch"] / 1e3), 1) for k, v in out.items()}}


def conv_shapes(launches=5):
    """The convolution stage (ms per launch, hipEvents) under the 32-row tile
    (conv_rows 8) and the column sweep (conv_rows 1): C3 at 4096 prior draws
    per launch (culled) and C4 (64 systems, not culled; the sweep's rows all
    lit).  The tile's launches come first, then the sweep's."""
    from lensed_amd._lib import LENSED_PROFILE, LENSED_TABLE
    from lensed_amd.lensed import Context
    K = 4096
    wl, _ = make_workload(K)
    out = {"C3": {}, "C4": {}}
    with Context(wl.spec(), max_batch=K, flags=LENSED_TABLE | LENSED_PROFILE) as ctx:
        sets = draws(wl, ctx.ndims(), K, launches + 1)
        for rows in (8, 1):
            ctx.set_tuning(conv_rows=rows)
            it = iter(sets)
            ms = stage_ms(ctx, lambda: ctx.loglike_batch(next(it)), launches)["convolve"]
            out["C3"][f"conv_rows_{rows}"] = {"ms": ms, "resolved_conv_rows": ctx.tuning()["conv_rows"]}
    wls, phys = c4_workload()
    with Context([w.spec() for w in wls], max_batch=256, flags=LENSED_TABLE | LENSED_PROFILE) as ctx:
        for rows in (8, 1):
            ctx.set_tuning(conv_rows=rows)
            ms = stage_ms(ctx, lambda: ctx.loglike_phys(phys), 2)["convolve"]
            out["C4"][f"conv_rows_{rows}"] = {"ms": ms, "resolved_conv_rows": ctx.tuning()["conv_rows"]}
    return {"convolve_ms_per_launch": out,
            "note": "conv_rows 8: the 32-row tile; conv_rows 1: the column sweep (the default where every "
                    "system is culled; C4's 31 x 31 PSFs keep the tile)"}


def copy_probe(mib, reps=3):
    """Plain device-to-device copies of a known size: what the profiler's
    byte counters read for them calibrates the counters."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "copies": reps}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--plain", type=int, default=1000, help="candidates timed on the plain path")
    ap.add_argument("--no-nested", action="store_true", help="skip the end-to-end nested-sampling comparison")
    ap.add_argument("--value-only", action="store_true",
                    help="time only the batched + emulated launches and the roofline (for profiler runs)")
    ap.add_argument("--no-autotune", action="store_true", help="time the default launch shapes")
    ap.add_argument("--c4", action="store_true",
                    help="only time the uneven 64-system workload (render list, item walk, no cull) and exit")
    ap.add_argument("--conv-shapes", action="store_true",
                    help="only time the convolution under the tile and the column sweep (C3, C4) and exit")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (byte-counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: candidates are independent, so the path would shard by splitting a "
              "launch's candidates across GPUs; runs on several GPUs are not implemented.",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.c4:
        print(json.dumps(c4_launches()))
        return
    if args.conv_shapes:
        print(json.dumps(conv_shapes()))
        return
    from lensed_amd._lib import LENSED_AUTOTUNE, LENSED_EMULATE, LENSED_GRAPH, LENSED_TABLE
    global TUNED
    TUNED = 0 if args.no_autotune else LENSED_AUTOTUNE
    K = args.batch
    wl, pids = make_workload(K)
    if args.value_only:
        dt, _, emu_stats, tuning = time_batched(wl, K, args.steps, args.warmup, LENSED_GRAPH | LENSED_EMULATE | TUNED)
        kt, tr = kernel_times(wl, K, 3, LENSED_EMULATE)
        print(json.dumps({"metric": METRIC, "value": round(args.steps * K / dt, 2), "unit": "logL evals/s",
                          "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": round(dt / args.steps * 1e3, 4), "higher_is_better": True,
                          "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                          "config": {"workload": "C3 (see the full run)", "candidates_per_launch": K,
                                     "legs": "batched + emulated only"},
                          "roofline": render_roofline(kt, tr, K), "render_trace": tr,
                          "emulator": emulator_block(emu_stats), "launch_shapes": tuning,
                          "stages_ms_per_launch": {k: round(v, 4) for k, v in kt.items()},
                          "cpu_baseline": None}))
        return

    dt_exact, lnew, _, _ = time_batched(wl, K, args.steps, args.warmup, LENSED_GRAPH | TUNED)
    exact_rate = args.steps * K / dt_exact
    dt_table, _, _, _ = time_batched(wl, K, args.steps, args.warmup, LENSED_GRAPH | LENSED_TABLE | TUNED)
    table_rate = args.steps * K / dt_table
    dt, _, emu_stats, tuning = time_batched(wl, K, args.steps, args.warmup, LENSED_GRAPH | LENSED_EMULATE | TUNED)
    value = args.steps * K / dt
    acc = accuracy(wl, pids, K, LENSED_EMULATE)
    acc_table = accuracy(wl, pids, K, LENSED_TABLE)
    plain_rate, lplain = time_plain(wl, args.plain)
    # plain and batched use the same kernels: identical logL per candidate
    same = bool(np.array_equal(lplain, lnew[: len(lplain)]))
    kt, tr = kernel_times(wl, K, 3, LENSED_EMULATE)
    kt_exact, _ = kernel_times(wl, K, 2)
    nest = None if args.no_nested else nested_compare(wl, pids, K)
    # the CPU leg, last: the oracle as the checker (strict IEEE build), then
    # as the baseline (its fast-math build sets flush-to-zero in the process)
    acc64 = oracle_accuracy(wl, pids, K, LENSED_EMULATE)
    cpu = cpu_baseline(wl, pids)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "logL evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "C3: 256x256 px, epl_plus_shear + gauss (sigma 2 px), 21x21 Gaussian PSF "
                               "FWHM 3 px, g3k7 49 rays/px, sigma 0.05, candidates uniform from the priors",
                   "candidates_per_launch": K,
                   "path": "batched + emulated: EPL angular-factor tables from the BF16-MFMA emulator in the "
                           "(q, t) cells its certification passed at 4e-6 held-out error, from the exact series "
                           "elsewhere; pixel cull; launch shapes from the autotuner; hipGraph replay",
                   "seeds": {"noise": wl.seed, "candidates": wl.cand_seed}},
        "emulator": emulator_block(emu_stats),
        "launch_shapes": tuning,
        "batched_table": {"value": round(table_rate, 2), "unit": "logL evals/s",
                          "path": "batched, EPL angular factor from the per-candidate FP64-built table",
                          "accuracy_vs_exact": acc_table},
        "batched_exact": {"value": round(exact_rate, 2), "unit": "logL evals/s",
                          "path": "batched, per-ray tenth-order EPL series"},
        "accuracy_vs_exact": acc,
        "plain_gpu": {"value": round(plain_rate, 2), "unit": "logL evals/s",
                      "path": "one candidate per loglike() call, 4 kernels + 2 copies + sync",
                      "identical_logL_to_batched": same},
        "speedup_vs_plain": round(value / plain_rate, 3),
        "roofline": render_roofline(kt, tr, K),
        "render_trace": tr,
        "stages_ms_per_launch": {k: round(v, 4) for k, v in kt.items()},
        "stages_note": "set_params includes the emulator and exact-series table kernels; render is "
                       "cull_pass + render_kernel",
        "stages_ms_per_launch_exact": {k: round(v, 4) for k, v in kt_exact.items()},
        "conv_chi2": conv_roofline(kt, K, tuning.get("conv_rows")),
        "accuracy_vs_fp64_oracle": acc64,
        "cpu_baseline": cpu,
        "nested_sampling": nest,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
