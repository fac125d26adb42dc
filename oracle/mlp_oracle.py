"""CPU oracle of the warm-start network (TEST INFRASTRUCTURE ONLY).

Forward and backward of 8 -> 64 SiLU -> 64 SiLU -> 1 in float64, with the BF16
roundings of the GPU kernels applied at the same points: inputs, weights, H1 and the
dZ fed to the matrix cores are rounded to BF16 (nearest even); biases, activations,
the output layer and all sums stay wide.  With `bf16=False` it is the plain float64
network, whose gradients tests check by central finite differences.

Parameter vector layout: W1[64][8], b1[64], W2[64][64], b2[64], W3[64], b3 (4801).
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID = 8, 64
SIZES = [("W1", (N_HID, N_IN)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)), ("b2", (N_HID,)),
         ("W3", (N_HID,)), ("b3", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + k], np.float64).reshape(shape)
        o += k
    return out


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SIZES])


def bf16(x) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(p, x, bf=True):
    w = unpack(p)
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    xb = r(x)
    z1 = xb @ r(w["W1"]).T + w["b1"]
    h1 = z1 * _sig(z1)
    z2 = r(h1) @ r(w["W2"]).T + w["b2"]
    h2 = z2 * _sig(z2)
    y = h2 @ r(w["W3"]) + w["b3"][0]
    return y, (xb, z1, h1, z2, h2)


def gradients(p, x, t, bf=True):
    """(grad[4801], loss) of 0.5 * sum (y - t)^2 / n."""
    w = unpack(p)
    r = bf16 if bf else (lambda a: np.asarray(a, np.float64))
    n = x.shape[0]
    y, (xb, z1, h1, z2, h2) = forward(p, x, bf)
    loss = 0.5 * np.sum((y - t) ** 2)
    dy = (y - t) / n
    s2 = _sig(z2)
    dz2 = dy[:, None] * r(w["W3"])[None, :] * (s2 * (1 + z2 * (1 - s2)))
    dz2b = r(dz2)
    s1 = _sig(z1)
    dh1 = dz2b @ r(w["W2"])
    dz1 = dh1 * (s1 * (1 + z1 * (1 - s1)))
    g = {"W3": h2.T @ dy, "b3": np.array([dy.sum()]), "W2": dz2b.T @ r(h1), "b2": dz2.sum(0),
         "W1": r(dz1).T @ xb, "b1": dz1.sum(0)}
    return pack(g), loss


def init_params(seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    d = {"W1": rng.normal(0, np.sqrt(2 / N_IN), (N_HID, N_IN)), "b1": np.zeros(N_HID),
         "W2": rng.normal(0, np.sqrt(2 / N_HID), (N_HID, N_HID)), "b2": np.zeros(N_HID),
         "W3": rng.normal(0, np.sqrt(1 / N_HID), N_HID), "b3": np.zeros(1)}
    return pack(d)
