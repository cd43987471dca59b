"""The EPL emulator on the GPU: training gradients and BF16-MFMA inference
against the CPU oracle (oracle/mlp_oracle.py), training, the certification
of the shipped network per (q, t) cell and what it gains over the exact
table, and the exact recheck of accepted points in nested sampling.

The emulated table costs an MLP evaluation (value and tangent) per table
node where the exact table costs a tenth-order series in FP64; the network
builds a candidate's table only in a cell certified at the bound that keeps
|delta logL| small, the exact series elsewhere, and the batched sampler
re-evaluates every candidate it may accept with the exact series."""
from ctypes import POINTER, c_double, c_float

import numpy as np
import pytest

import os
import time

from lensed_amd import _lib as L
from oracle import mlp_oracle as M

EMULATOR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lensed_amd", "data",
                        "epl_mlp.bin")

pytestmark = pytest.mark.gpu

FP = POINTER(c_float)


def rand_weights(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in M.SIZES:
        fan = shape[1] if len(shape) == 2 else 1
        p[name] = rng.uniform(-1, 1, shape) * (scale * np.sqrt(3.0 / fan) if len(shape) == 2 else 0.1)
    return M.pack(p).astype(np.float32)


def inputs(n, seed):
    rng = np.random.default_rng(seed)
    psi = rng.uniform(0, np.pi / 2, n)
    X = np.stack([rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), np.cos(psi), np.sin(psi)], 1).astype(np.float32)
    dX = np.stack([np.zeros(n), np.zeros(n), -np.sin(psi), np.cos(psi)], 1).astype(np.float32)
    return X, dX


def test_training_gradients_match_cpu_oracle():
    w = rand_weights(1)
    X, _ = inputs(300, 2)
    Y = np.random.default_rng(3).normal(0, 0.3, (300, 2)).astype(np.float32)
    g = np.zeros(M.NPARAMS, np.float32)
    loss = L.load().lensed_emulator_loss_grad(w.ctypes.data_as(FP), X.ctypes.data_as(FP), Y.ctypes.data_as(FP),
                                              300, g.ctypes.data_as(FP))
    lref, gref = M.loss_grad(w, X, Y)
    assert loss == pytest.approx(lref, rel=1e-5)
    scale = np.abs(gref).max()
    assert np.max(np.abs(g - gref)) <= 1e-4 * scale
    # and the oracle's gradient against a central difference of its own loss
    i = 5000
    e = np.zeros(M.NPARAMS)
    e[i] = 1e-6
    fd = (M.loss_grad(w + e, X, Y)[0] - M.loss_grad(w - e, X, Y)[0]) / 2e-6
    assert fd == pytest.approx(gref[i], rel=1e-4, abs=1e-9)


def test_bf16_mfma_inference_matches_cpu_oracle():
    w = rand_weights(4)
    X, dX = inputs(1000, 5)  # not a multiple of the 32-sample wave tile
    Y, dY = np.zeros((1000, 2), np.float32), np.zeros((1000, 2), np.float32)
    lib = L.load()
    assert lib.lensed_emulator_eval(w.ctypes.data_as(FP), X.ctypes.data_as(FP), dX.ctypes.data_as(FP), 1000, 1,
                                    Y.ctypes.data_as(FP), dY.ctypes.data_as(FP)) == 0
    yref, dyref = M.forward(w, X, dX)
    # bf16 weights and activations (8-bit mantissa), FP32 accumulation
    assert np.max(np.abs(Y - yref)) <= 3e-2 * np.abs(yref).max()
    assert np.max(np.abs(dY - dyref)) <= 3e-2 * np.abs(dyref).max()
    assert np.corrcoef(Y.ravel(), yref.ravel())[0, 1] > 0.999
    Y32 = np.zeros((1000, 2), np.float32)
    assert lib.lensed_emulator_eval(w.ctypes.data_as(FP), X.ctypes.data_as(FP), dX.ctypes.data_as(FP), 1000, 0,
                                    Y32.ctypes.data_as(FP), None) == 0
    assert np.max(np.abs(Y32 - yref)) <= 1e-5 * np.abs(yref).max()


@pytest.fixture(scope="module")
def trained(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("emu") / "epl_mlp.bin")
    stats = (c_double * 2)()
    rc = L.load().lensed_emulator_train(path.encode(), 1 << 18, 4, 2048, 2e-3, 303, stats)
    assert rc == 0, L.last_error()
    return path, stats[0], stats[1]


def test_training_learns_the_series(trained):
    path, mse, seconds = trained
    w = np.zeros(M.NPARAMS, np.float32)
    assert L.load().lensed_emulator_read(path.encode(), w.ctypes.data_as(FP)) == 0
    # held-out points against the oracle's series
    from oracle import oracle as O
    rng = np.random.default_rng(99)
    q, t, psi = rng.uniform(0.3, 1, 2000), rng.uniform(0.5, 1.5, 2000), rng.uniform(0, np.pi / 2, 2000)
    X = np.stack([(q - 0.65) / 0.35, (t - 1) / 0.5, np.cos(psi), np.sin(psi)], 1).astype(np.float32)
    pred, _ = M.forward(w, X)
    exact = O.epl_omega(q, t, psi) - np.stack([np.cos(psi), np.sin(psi)], 1)
    rms = np.sqrt(np.mean((pred - exact) ** 2))
    print(f"emulator: train mse {mse:.3e}, held-out rms {rms:.3e}, {seconds:.1f} s")
    # a short training (2^18 samples, 4 epochs): an order of magnitude below
    # the signal; the production network trains on 2^22 samples
    assert rms < 0.2 * np.sqrt(np.mean(exact ** 2))


def test_emulator_certification_and_what_it_gains():
    """What the trained network (the shipped weights) gains over the method
    without it, the exact-series table, measured on a 96 x 96 EPL + shear
    system with 512 prior draws per launch.

    At the production bound (4e-6 on the held-out angular factor, what
    keeps |delta logL| <= 3e-4 near the truth) the certification finds the
    cells where the network may build tables; candidates elsewhere take the
    exact table, so the emulated path equals the table path bit for bit on
    them, and with no certified cell it runs exactly the table path's
    kernels (its throughput within 5 % of it).  At a loose bound every cell
    is certified and every table comes from the network: the test records
    its throughput and |delta logL| against the exact table near the truth
    (the gain, or the loss, in numbers)."""
    from lensed_amd import workloads as W
    from lensed_amd._lib import LENSED_EMULATE, LENSED_GRAPH, LENSED_TABLE
    from lensed_amd.lensed import Context
    from test_gpu_parity import near_truth_phys, oracle_render_cb
    from oracle import oracle as O
    wl = W.c3()
    pids = [f"{o}.{p}" for o, n in wl.objects for p in O.PARAMS[n]]
    wl.shape = (96, 96)
    for k, v in (("lens.x", 48.5), ("lens.y", 48.5), ("source.x", 50.0), ("source.y", 47.5)):
        wl.truth[k] = v
    wl.priors.update({"lens.x": "unif 46 51", "lens.y": "unif 46 51", "source.x": "unif 45 55",
                      "source.y": "unif 42 52"})
    wl.make_data(oracle_render_cb, pids, mag_keys=("source.mag",))
    K = 512
    cubes = [np.random.default_rng(s).random((K, 14)) for s in range(6)]
    near = near_truth_phys(wl, pids, 64, 2e-4, seed=3)

    def rate(ctx):
        ctx.loglike_batch(cubes[0])
        ts = []
        for c in cubes[1:]:
            t0 = time.perf_counter()
            ctx.loglike_batch(c)
            ts.append(time.perf_counter() - t0)
        return K / float(np.median(ts))

    flags = LENSED_GRAPH
    with Context(wl.spec(), max_batch=K, flags=flags | LENSED_TABLE) as tb:
        lt = [tb.loglike_batch(c)[0] for c in cubes]
        ln_t = tb.loglike_phys(near)
        r_table = rate(tb)
    with Context(wl.spec(), max_batch=K, flags=flags | LENSED_EMULATE, emulator=EMULATOR) as em:
        ncells, ncert, cert_err = em.emulator_cells()
        le = [em.loglike_batch(c)[0] for c in cubes]
        nf, nc, _ = em.emulator_stats()
        r_emu = rate(em)
    print(f"production bound 4e-6: {ncert} of {ncells} cells certified (largest held-out error {cert_err:.2e}); "
          f"last launch {nc - nf} of {nc} tables from the network; evals/s table {r_table:.0f}, "
          f"emulated {r_emu:.0f}", flush=True)
    assert ncells == 1024 and 0 <= ncert <= ncells and cert_err > 0
    if ncert == 0:
        assert nf == nc == K
        for a, b in zip(le, lt):
            assert np.array_equal(a, b)
        assert r_emu >= 0.95 * r_table, (r_emu, r_table)
    with Context(wl.spec(), max_batch=K, flags=flags | LENSED_EMULATE, emulator=EMULATOR,
                 emulator_tol=1.0) as loose:
        assert loose.emulator_cells()[1] == 1024
        ll = [loose.loglike_batch(c)[0] for c in cubes]
        nf2, nc2, _ = loose.emulator_stats()
        ln_l = loose.loglike_phys(near)
        r_loose = rate(loose)
    d_near = float(np.max(np.abs(ln_l - ln_t)))
    d_prior = float(max(np.max(np.abs(a - b) / np.abs(b)) for a, b in zip(ll, lt)))
    print(f"every table from the network: evals/s {r_loose:.0f} ({r_loose / r_table:.2f} x the exact table); "
          f"max |dlogL| near the truth {d_near:.3e}, max relative dlogL over prior draws {d_prior:.2e}",
          flush=True)
    assert nf2 == 0 and nc2 == K
    assert all(np.all(np.isfinite(a)) for a in ll)



def test_recheck_gives_every_accepted_point_its_exact_logl(tmp_path):
    """The batched nested sampler with every EPL table from the network (a
    loose certification bound, so every cell is certified): candidates that
    may be accepted are evaluated again with exact-series tables, so every
    point in the posterior file (<root>.txt: weight, -2 logL, parameters)
    carries the exact logL of its parameters -- checked for a seeded sample
    of them against the exact table path."""
    from lensed_amd import nested
    from lensed_amd import workloads as W
    from lensed_amd._lib import LENSED_EMULATE, LENSED_GRAPH, LENSED_TABLE
    from lensed_amd.lensed import Context
    from test_gpu_parity import oracle_render_cb
    from oracle import oracle as O
    wl = W.c3()
    pids = [f"{o}.{p}" for o, n in wl.objects for p in O.PARAMS[n]]
    wl.shape = (64, 64)
    for k, v in (("lens.x", 32.5), ("lens.y", 32.5), ("source.x", 33.0), ("source.y", 31.5), ("lens.r", 10.0)):
        wl.truth[k] = v
    wl.priors.update({"lens.x": "unif 31 34", "lens.y": "unif 31 34", "source.x": "unif 30 36",
                      "source.y": "unif 28 34", "lens.r": "unif 8 12"})
    wl.make_data(oracle_render_cb, pids, mag_keys=("source.mag",))
    root = str(tmp_path / "emu")
    with Context(wl.spec(), max_batch=512, flags=LENSED_EMULATE | LENSED_GRAPH, emulator=EMULATOR,
                 emulator_tol=1.0) as ctx:
        pmap = ctx.pmap()
        rb = nested.run_context(ctx, 0, nlive=100, tol=0.5, efr=0.8, batch=512, seed=61, root=root, outfile=True)
        rechecked, changed, max_d = ctx.emulator_rechecks()
    print(f"emulated nested run: {rb}; rechecked {rechecked}, changed {changed}, largest change {max_d:.3e}",
          flush=True)
    assert rechecked > 0 and changed > 0
    rows = np.loadtxt(root + ".txt")
    assert rows.shape[1] == 2 + len(pids)
    pick = np.random.default_rng(62).choice(len(rows), min(300, len(rows)), replace=False)
    phys = np.zeros((len(pick), len(pids)))
    phys[:, pmap] = rows[pick, 2:]
    with Context(wl.spec(), max_batch=512, flags=LENSED_TABLE) as ex:
        exact = ex.loglike_phys(phys)
    assert np.allclose(-0.5 * rows[pick, 1], exact, rtol=1e-9, atol=1e-6), \
        np.max(np.abs(-0.5 * rows[pick, 1] - exact))
