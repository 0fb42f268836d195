"""CPU oracle of the surrogate MLP -- TEST INFRASTRUCTURE ONLY.

numpy FP64 restatement of the network, loss and gradients of the engine's surrogate ensemble
(latticemontecarlo_amd/csrc/lmc_mlp.h, lmc_surrogate.hip). The surrogate has no counterpart in
the reference; it stands in for GetBarrierAndDiffFromLatticeIdPair
(lmc/pred/src/VacancyMigrationPredictorE0.cpp:153-159), and the reference's predictor (restated
in lmc_oracle.cpp) remains the reference its errors are measured against.

With rounding=True the bf16 rounding points of the kernels are reproduced (weights, h1, h2 and
the activation gradients fed to the matrix cores are rounded to bf16, round to nearest even);
everything else is FP64. With rounding=False the network is plain FP64, and its gradients are
pinned by central finite differences (tests/test_surrogate_oracle.py).
"""
import numpy as np

HIDDEN = 128
OUT = 2
SLOTS = 60


def bf16(x):
    """Round to the nearest bfloat16 (ties to even) via float32; returns float64."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def layout(kp):
    """Parameter offsets of one member: W1 [kp][128] | b1 | W2 [128][128] | b2 | W3 [128][2] | b3."""
    o = {"W1": 0, "b1": kp * HIDDEN}
    o["W2"] = o["b1"] + HIDDEN
    o["b2"] = o["W2"] + HIDDEN * HIDDEN
    o["W3"] = o["b2"] + HIDDEN
    o["b3"] = o["W3"] + HIDDEN * OUT
    o["P"] = (o["b3"] + OUT + 3) // 4 * 4
    return o


def unpack(p, kp):
    o = layout(kp)
    p = np.asarray(p, np.float64)
    return dict(W1=p[o["W1"]:o["b1"]].reshape(kp, HIDDEN), b1=p[o["b1"]:o["W2"]],
                W2=p[o["W2"]:o["b2"]].reshape(HIDDEN, HIDDEN), b2=p[o["b2"]:o["W3"]],
                W3=p[o["W3"]:o["b3"]].reshape(HIDDEN, OUT), b3=p[o["b3"]:o["b3"] + OUT])


def pack(d, kp):
    o = layout(kp)
    p = np.zeros(o["P"])
    for name in ("W1", "b1", "W2", "b2", "W3", "b3"):
        v = np.asarray(d[name]).reshape(-1)
        p[o[name]:o[name] + v.size] = v
    return p


def one_hot(env, channels, kp):
    """Compact one-hot input: x[r, j*C + c] = (env[r, j] == channels[c]), zero padded to kp."""
    env = np.asarray(env)
    C = len(channels)
    x = np.zeros((env.shape[0], kp))
    for c, code in enumerate(channels):
        x[:, c:SLOTS * C:C] = env[:, :SLOTS] == code
    return x


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * sigmoid(z)


def dsilu(z):
    s = sigmoid(z)
    return s * (1.0 + z * (1.0 - s))


def forward(p, x, rounding=True):
    """Returns the outputs y [n, 2] and the intermediates needed by backward."""
    kp = x.shape[1]
    w = unpack(p, kp)
    r = bf16 if rounding else (lambda v: np.asarray(v, np.float64))
    W1, W2, W3 = r(w["W1"]), r(w["W2"]), r(w["W3"])
    z1 = x @ W1 + w["b1"]
    h1 = r(silu(z1))
    z2 = h1 @ W2 + w["b2"]
    h2 = r(silu(z2))
    y = h2 @ W3 + w["b3"]
    return y, dict(x=x, z1=z1, h1=h1, z2=z2, h2=h2, W1=W1, W2=W2, W3=W3)


def loss(p, x, t, rounding=True):
    y, _ = forward(p, x, rounding)
    return float(np.sum((y - t) ** 2) / (2 * x.shape[0]))


def gradients(p, x, t, rounding=True):
    """Gradient of sum (y - t)^2 / (2n), packed like the parameters."""
    kp = x.shape[1]
    n = x.shape[0]
    r = bf16 if rounding else (lambda v: np.asarray(v, np.float64))
    y, c = forward(p, x, rounding)
    dy = (y - t) / n
    g = dict(W3=c["h2"].T @ dy, b3=dy.sum(0))
    dz2 = (dy @ c["W3"].T) * dsilu(c["z2"])
    g["b2"] = dz2.sum(0)
    dz2r = r(dz2)
    g["W2"] = c["h1"].T @ dz2r
    dz1 = (dz2r @ c["W2"].T) * dsilu(c["z1"])
    g["b1"] = dz1.sum(0)
    g["W1"] = c["x"].T @ r(dz1)
    return pack(g, kp)
