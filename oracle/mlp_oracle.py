"""CPU oracle of the learned rate-and-state Newton guess (test infrastructure).

Only tests and bench.py's baseline leg use this module; the product path runs
the HIP kernels in eqdyna_amd/csrc/eqd_mlp.hip and never imports it.

The network is 6 -> 64 -> 64 -> 64 -> 1 with GELU (erf form) on the hidden
layers.  The GPU computes in mixed precision: activations and weights enter
the matrix products as bfloat16 (round to nearest even), products accumulate
in float32, biases, GELU and the loss are float32.  ``forward`` and
``backward`` here restate exactly that data flow in float64 with the same
bfloat16 rounding points, so the GPU's results should match them to float32
accumulation-order differences (plus rare one-ulp bfloat16 flips next to a
rounding tie).  With ``bf16=False`` they are the plain float64 network, which
the tests check against central finite differences.

There is no counterpart in the reference (src/faulting.f90:459-516 starts its
Newton iteration from the previous slip rate); the inputs follow the
reference's solveRSF quantities (src/faulting.f90:190-327):
  x0 trial shear traction |tau_trial| (Pa), x1 effective normal stress (Pa),
  x2 state variable (psi for the slip law, log theta for the ageing law),
  x3 a, x4 b, x5 ln of the plain guess |V_prev| (m/s);
output: r = ln(V / V_prev) of the converged slip rate V (the proposal is
V_prev * exp(r)).

Parameter vector (float32), the layout of include/eqdyna_amd.h:
  W1[64][6] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] w4[64] b4[1]
  in_shift[6] in_scale[6] out_shift out_scale
Normalised input xn = (x - in_shift) * in_scale; output r = y * out_scale + out_shift.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf

NIN, NH = 6, 64
SIZES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH, NH)), ("b3", (NH,)),
         ("w4", (NH,)), ("b4", (1,))]
NTRAIN = sum(int(np.prod(s)) for _, s in SIZES)  # 8833 trainable parameters
NPARAMS = NTRAIN + 2 * NIN + 2


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float values to bfloat16 (nearest, ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p: np.ndarray) -> dict:
    p = np.asarray(p, dtype=np.float64)
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    out["in_shift"] = p[o:o + NIN]
    out["in_scale"] = p[o + NIN:o + 2 * NIN]
    out["out_shift"] = p[o + 2 * NIN]
    out["out_scale"] = p[o + 2 * NIN + 1]
    return out


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def _r(x, on):
    return bf16(x) if on else np.asarray(x, dtype=np.float64)


def forward(p: np.ndarray, X: np.ndarray, bf16_points: bool = True, keep: bool = False):
    """Network output y (normalised) for raw inputs X[n, 6]; with keep, also the
    intermediate values backward() needs."""
    q = unpack(p)
    on = bf16_points
    xn = (np.asarray(X, dtype=np.float32).astype(np.float64) - q["in_shift"]) * q["in_scale"]
    if on:
        xn = xn.astype(np.float32).astype(np.float64)
    a0 = _r(xn, on)
    W1, W2, W3, w4 = _r(q["W1"], on), _r(q["W2"], on), _r(q["W3"], on), _r(q["w4"], on)
    z1 = a0 @ W1.T + q["b1"]
    a1 = _r(gelu(z1), on)
    z2 = a1 @ W2.T + q["b2"]
    a2 = _r(gelu(z2), on)
    z3 = a2 @ W3.T + q["b3"]
    a3 = _r(gelu(z3), on)
    y = a3 @ w4 + q["b4"][0]
    if keep:
        return y, dict(a0=a0, z1=z1, a1=a1, z2=z2, a2=a2, z3=z3, a3=a3, W2=W2, W3=W3, w4=w4)
    return y


def predict(p: np.ndarray, X: np.ndarray, bf16_points: bool = True) -> np.ndarray:
    q = unpack(p)
    return forward(p, X, bf16_points) * q["out_scale"] + q["out_shift"]


def loss_and_grad(p: np.ndarray, X: np.ndarray, target: np.ndarray, bf16_points: bool = True):
    """Mean squared error of the normalised output against the normalised
    target, and its gradient with respect to the trainable parameters
    (NTRAIN values in the parameter-vector order)."""
    q = unpack(p)
    on = bf16_points
    n = len(X)
    y, c = forward(p, X, on, keep=True)
    t = (np.asarray(target, dtype=np.float32).astype(np.float64) - q["out_shift"]) / q["out_scale"]
    r = y - t
    loss = float(np.mean(r * r))
    dy = 2.0 * r / n
    g = {}
    g["b4"] = np.array([dy.sum()])
    g["w4"] = c["a3"].T @ dy
    dz3 = np.outer(dy, c["w4"]) * gelu_grad(c["z3"])
    g["b3"] = dz3.sum(axis=0)
    dz3r = _r(dz3, on)
    g["W3"] = dz3r.T @ c["a2"]
    dz2 = (dz3r @ c["W3"]) * gelu_grad(c["z2"])
    g["b2"] = dz2.sum(axis=0)
    dz2r = _r(dz2, on)
    g["W2"] = dz2r.T @ c["a1"]
    dz1 = (dz2r @ c["W2"]) * gelu_grad(c["z1"])
    g["b1"] = dz1.sum(axis=0)
    g["W1"] = dz1.T @ c["a0"]
    return loss, np.concatenate([g[name].ravel() for name, _ in SIZES])


def init_params(X: np.ndarray, target: np.ndarray, seed: int) -> np.ndarray:
    """Seeded initial parameters: normalisation from the data (mean and
    inverse standard deviation of each input, mean and standard deviation of
    the target), weights uniform in +-sqrt(6/(fan_in+fan_out)), zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SIZES:
        if name.startswith("b"):
            parts.append(np.zeros(shape))
        else:
            fan_in = shape[1] if len(shape) == 2 else shape[0]
            fan_out = shape[0] if len(shape) == 2 else 1
            lim = math.sqrt(6.0 / (fan_in + fan_out))
            parts.append(rng.uniform(-lim, lim, size=shape))
    X = np.asarray(X, dtype=np.float64)
    mu = X.mean(axis=0)
    sd = X.std(axis=0)
    sd = np.where(sd > 0, sd, 1.0)
    lv = np.asarray(target, dtype=np.float64)
    parts += [mu, 1.0 / sd, np.array([lv.mean()]), np.array([max(lv.std(), 1e-6)])]
    return np.concatenate([np.ravel(x) for x in parts]).astype(np.float32)
