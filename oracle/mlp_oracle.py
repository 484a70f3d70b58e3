"""CPU restatement of the learned Newton seed network (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, and only as the checker.  There is no reference counterpart: the
network replaces the guess block of Utoprim_2d (utoprim_2d.c:268-307), so this
file is pinned by mathematics rather than by the reference (see
tests/test_mlp_oracle.py):
  - the backward pass against central finite differences of the forward pass;
  - Adam against its closed form after the first steps;
  - BF16 rounding against hand-computed round-to-nearest-even cases;
  - the forward pass against an explicit scalar loop on one row.

Parameter vector (float32, row-major W[out][in]):
  W1 64x8 | b1 64 | W2 64x64 | b2 64 | W3 64x64 | b3 64 | W4 2x64 | b4 2.
Loss on a batch of B rows: L = 1/(2B) sum_b mask_b |O_b - T_b|^2.
"""
from __future__ import annotations

import numpy as np

NIN, NH, NOUT = 8, 64, 2
SHAPES = [(NH, NIN), (NH,), (NH, NH), (NH,), (NH, NH), (NH,), (NOUT, NH), (NOUT,)]
NPARAMS = sum(int(np.prod(s)) for s in SHAPES)  # 9026


def split(params: np.ndarray):
    """Views W1, b1, W2, b2, W3, b3, W4, b4 of the flat parameter vector."""
    out, o = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(params[o:o + n].reshape(s))
        o += n
    return out


def forward(params: np.ndarray, X: np.ndarray):
    """FP64 forward pass; returns (O, [X, H1, H2, H3])."""
    W1, b1, W2, b2, W3, b3, W4, b4 = [np.asarray(a, np.float64) for a in split(params)]
    X = np.asarray(X, np.float64)
    H1 = np.tanh(X @ W1.T + b1)
    H2 = np.tanh(H1 @ W2.T + b2)
    H3 = np.tanh(H2 @ W3.T + b3)
    O = H3 @ W4.T + b4
    return O, [X, H1, H2, H3]


def loss_grad(params: np.ndarray, X: np.ndarray, T: np.ndarray, mask=None, out_w=(1.0, 1.0)):
    """Loss and gradient (FP64) of 1/(2B) sum mask sum_j out_w[j] (O_j - T_j)^2."""
    B = X.shape[0]
    w = np.ones(B) if mask is None else np.asarray(mask, np.float64)
    O, (X, H1, H2, H3) = forward(params, X)
    E = (O - np.asarray(T, np.float64)) * np.sqrt(np.asarray(out_w, np.float64))[None, :]
    loss = 0.5 * np.sum(w[:, None] * E * E) / B
    E = E * np.sqrt(np.asarray(out_w, np.float64))[None, :]
    W1, b1, W2, b2, W3, b3, W4, b4 = [np.asarray(a, np.float64) for a in split(params)]
    dO = w[:, None] * E / B
    gW4 = dO.T @ H3
    gb4 = dO.sum(0)
    dZ3 = (dO @ W4) * (1 - H3 * H3)
    gW3 = dZ3.T @ H2
    gb3 = dZ3.sum(0)
    dZ2 = (dZ3 @ W3) * (1 - H2 * H2)
    gW2 = dZ2.T @ H1
    gb2 = dZ2.sum(0)
    dZ1 = (dZ2 @ W2) * (1 - H1 * H1)
    gW1 = dZ1.T @ X
    gb1 = dZ1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3, gW4, gb4)])
    return loss, g


def adam(params, m, v, t, grad, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1); t = step number >= 1."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v


def bf16_round(x) -> np.ndarray:
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    x = np.asarray(x, np.float32)
    a = np.ascontiguousarray(x.reshape(-1))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).reshape(x.shape)


def forward_bf16(params: np.ndarray, X: np.ndarray) -> np.ndarray:
    """Emulation of the inference path: BF16 weights and activations, FP32
    bias and accumulation (accumulated here in FP64, then rounded to FP32)."""
    W1, b1, W2, b2, W3, b3, W4, b4 = split(np.asarray(params, np.float32))
    h = bf16_round(X).astype(np.float64)
    for W, b, act in ((W1, b1, True), (W2, b2, True), (W3, b3, True), (W4, b4, False)):
        z = (h @ bf16_round(W).astype(np.float64).T + b.astype(np.float64)).astype(np.float32)
        h = bf16_round(np.tanh(z)).astype(np.float64) if act else z.astype(np.float64)
    return h


def init_params(seed: int = 0) -> np.ndarray:
    """Uniform +-sqrt(3/fan_in) weights, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    parts = []
    for s in SHAPES:
        if len(s) == 2:
            a = np.sqrt(3.0 / s[1])
            parts.append(rng.uniform(-a, a, size=s).ravel())
        else:
            parts.append(np.zeros(s))
    return np.concatenate(parts).astype(np.float32)


def standardize(feat: np.ndarray, norm: np.ndarray) -> np.ndarray:
    """Raw features [n][8] -> network inputs with the 20-entry norm block
    (mean[8], 1/std[8], output mean[2], output std[2])."""
    return ((feat - norm[:8]) * norm[8:16]).astype(np.float32)
