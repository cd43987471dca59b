"""CPU oracle of the EPL emulator network, numpy FP64.

TEST INFRASTRUCTURE: the checker for the GPU training kernels (loss and
gradients) and the BF16-MFMA inference kernel.  Network and flat weight
layout as lensed_amd/csrc/emulator.h: 4 -> 128 -> 128 -> 128 -> 2, SiLU on
the hidden layers, weights row-major [out][in].
"""
import numpy as np

H, NIN, NOUT = 128, 4, 2
SIZES = [("W1", (H, NIN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
         ("W4", (NOUT, H)), ("b4", (NOUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(w):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(w[o:o + n], np.float64).reshape(shape)
        o += n
    return out


def pack(p):
    return np.concatenate([p[name].ravel() for name, _ in SIZES])


def silu(z):
    return z / (1 + np.exp(-z))


def dsilu(z):
    s = 1 / (1 + np.exp(-z))
    return s * (1 + z * (1 - s))


def forward(w, X, dX=None):
    """Outputs [n, 2] and, with input tangents dX, output tangents."""
    p = unpack(w)
    h, dh = np.asarray(X, np.float64), None if dX is None else np.asarray(dX, np.float64)
    for i in (1, 2, 3):
        z = h @ p[f"W{i}"].T + p[f"b{i}"]
        if dh is not None:
            dh = (dh @ p[f"W{i}"].T) * dsilu(z)
        h = silu(z)
    y = h @ p["W4"].T + p["b4"]
    return (y, None if dh is None else dh @ p["W4"].T)


def loss_grad(w, X, Y):
    """Mean squared error over all outputs and its gradient (flat layout)."""
    p = unpack(w)
    X, Y = np.asarray(X, np.float64), np.asarray(Y, np.float64)
    hs, zs = [X], []
    for i in (1, 2, 3):
        z = hs[-1] @ p[f"W{i}"].T + p[f"b{i}"]
        zs.append(z)
        hs.append(silu(z))
    y = hs[-1] @ p["W4"].T + p["b4"]
    d = y - Y
    loss = np.mean(d ** 2)
    g = {}
    dz = 2 * d / d.size
    g["W4"], g["b4"] = dz.T @ hs[3], dz.sum(0)
    dh = dz @ p["W4"]
    for i in (3, 2, 1):
        dz = dh * dsilu(zs[i - 1])
        g[f"W{i}"], g[f"b{i}"] = dz.T @ hs[i - 1], dz.sum(0)
        dh = dz @ p[f"W{i}"]
    return loss, pack(g)
