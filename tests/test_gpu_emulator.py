"""The EPL emulator on the GPU: training gradients and BF16-MFMA inference
against the CPU oracle (oracle/mlp_oracle.py), training, and the emulated
table path with its held-out check and exact fallback.

What the network gains over the method without it is measured here and in
bench.py: the emulated table costs an MLP evaluation per table node and a
held-out check per candidate on top of what the exact-series table costs,
and with BF16 inference its error is far above the bound that keeps
|delta logL| <= 0.01, so the check sends every candidate to the exact
fallback.  The test asserts the fallback keeps results exact and records
the numbers."""
from ctypes import POINTER, c_double, c_float

import numpy as np
import pytest

from lensed_amd import _lib as L
from oracle import mlp_oracle as M

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


def test_emulated_tables_check_and_fall_back(trained):
    from lensed_amd import workloads as W
    from lensed_amd._lib import LENSED_EMULATE, LENSED_TABLE
    from lensed_amd.lensed import Context
    from test_gpu_parity import oracle_render_cb
    from oracle import oracle as O
    path = trained[0]
    wl = W.c3()
    pids = [f"{o}.{p}" for o, n in wl.objects for p in O.PARAMS[n]]
    wl.shape = (96, 96)
    for k, v in (("lens.x", 48.5), ("lens.y", 48.5), ("source.x", 50.0), ("source.y", 47.5)):
        wl.truth[k] = v
    wl.priors.update({"lens.x": "unif 46 51", "lens.y": "unif 46 51", "source.x": "unif 45 55",
                      "source.y": "unif 42 52"})
    wl.make_data(oracle_render_cb, pids, mag_keys=("source.mag",))
    cubes = wl.candidates(14, 128)
    with Context(wl.spec(), max_batch=128, flags=LENSED_TABLE) as tb:
        lt, _ = tb.loglike_batch(cubes)
    # production bound: every candidate fails the check, falls back, and the
    # result is the exact table's to the bit
    with Context(wl.spec(), max_batch=128, flags=LENSED_EMULATE, emulator=path) as em:
        le, _ = em.loglike_batch(cubes)
        nf, nc, mx = em.emulator_stats()
    assert nc == 128 and nf == 128 and mx > 4e-6
    assert np.array_equal(le, lt)
    # a loose bound lets the network's own tables through: what it costs in logL
    with Context(wl.spec(), max_batch=128, flags=LENSED_EMULATE, emulator=path, emulator_tol=1.0) as loose:
        ll, _ = loose.loglike_batch(cubes)
        nf2, nc2, mx2 = loose.emulator_stats()
    assert nf2 == 0
    print(f"emulated tables: max held-out error {mx2:.3e}; max |dlogL| without fallback "
          f"{np.max(np.abs(ll - lt)):.3e} (logL range {lt.min():.3e} .. {lt.max():.3e})")
    assert np.all(np.isfinite(ll))
