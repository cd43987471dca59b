"""Train the EPL emulator on the GPU with the library's own kernels.

    python -m lensed_amd.train_emulator [--out lensed_amd/data/epl_mlp.bin]
        [--log2-samples 22] [--epochs 40] [--batch 4096] [--lr 2e-3] [--seed 303]

Training data: 2^log2-samples seeded draws of (q, t, psi) with the exact
tenth-order series as target (generated on the device); FP32 forward and
backward kernels with Adam and a cosine learning-rate schedule.
"""
import argparse
import os
from ctypes import c_double

from . import _lib as L

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT = os.path.join(HERE, "data", "epl_mlp.bin")


def train(out=DEFAULT, log2_samples=22, epochs=40, batch=4096, lr=2e-3, seed=303):
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    stats = (c_double * 2)()
    rc = L.load().lensed_emulator_train(out.encode(), 1 << log2_samples, epochs, batch, lr, seed, stats)
    if rc != 0:
        raise RuntimeError(L.last_error())
    return {"path": out, "train_mse": stats[0], "seconds": stats[1], "samples": 1 << log2_samples,
            "epochs": epochs, "batch": batch, "lr": lr, "seed": seed}


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT)
    ap.add_argument("--log2-samples", type=int, default=22)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=303)
    a = ap.parse_args()
    print(train(a.out, a.log2_samples, a.epochs, a.batch, a.lr, a.seed), flush=True)
