"""NumPy oracle of the learned proposal network (TEST INFRASTRUCTURE ONLY).

No reference counterpart: the network is new in this library. Restates, in
plain NumPy, what egg_amd/csrc/mlp.hpp and mlp.hip compute:

* ``bf16``: round-to-nearest-even of float32 to bfloat16 (kept as float32);
* ``forward`` / ``loss_grad``: the BF16-operand, FP32-accumulate network
  (inputs, weights, hidden activations and back-propagated dZ's rounded to
  BF16 exactly where the matrix-core kernels round them); ``*_f64`` are the
  same maps in float64 without rounding, for finite-difference checks;
* ``frame_features``: the local frame and the input features of a node from
  its one-ring and previous displacement.

Parameters use the library's flat layout: W1[n_in][H], b1[H], W2[H][H],
b2[H], W3[H][n_out], b3[n_out].
"""

from __future__ import annotations

import numpy as np

DIMS = {2: dict(n_in=22, n_in_padded=32, hidden=64, n_out=2), 3: dict(n_in=85, n_in_padded=96, hidden=128, n_out=3)}


def bf16(x):
    """Round float32 values to the nearest bfloat16 (ties to even), as float32."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(p, dim):
    d = DIMS[dim]
    I, H, O = d["n_in"], d["hidden"], d["n_out"]
    p = np.asarray(p)
    o = 0
    W1 = p[o:o + I * H].reshape(I, H); o += I * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H * O].reshape(H, O); o += H * O
    b3 = p[o:o + O]; o += O
    assert o == p.size
    return W1, b1, W2, b2, W3, b3


def n_params(dim):
    d = DIMS[dim]
    I, H, O = d["n_in"], d["hidden"], d["n_out"]
    return I * H + H + H * H + H + H * O + O


C_GELU = np.float32(0.7978845608)


def gelu(x):
    x = np.asarray(x, np.float32)
    t = np.tanh(C_GELU * (x + np.float32(0.044715) * x * x * x))
    return np.float32(0.5) * x * (np.float32(1) + t)


def gelu_grad(x):
    x = np.asarray(x, np.float32)
    t = np.tanh(C_GELU * (x + np.float32(0.044715) * x * x * x))
    return (np.float32(0.5) * (1 + t) + np.float32(0.5) * x * (1 - t * t) * C_GELU
            * (1 + np.float32(0.134145) * x * x)).astype(np.float32)


def _mm(a, b):
    return (np.asarray(a, np.float64) @ np.asarray(b, np.float64)).astype(np.float32)


def forward(p, X, dim, keep=False):
    """Network output for features X[n][n_in_padded] as the device computes it."""
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, np.float32), dim)
    I = DIMS[dim]["n_in"]
    Xb = bf16(np.asarray(X, np.float32)[:, :I])
    Z1 = _mm(Xb, bf16(W1)) + b1
    H1 = bf16(gelu(Z1))
    Z2 = _mm(H1, bf16(W2)) + b2
    H2 = bf16(gelu(Z2))
    Y = _mm(H2, bf16(W3)) + b3
    if keep:
        return Y, (Xb, Z1, H1, Z2, H2)
    return Y


def loss_grad(p, X, T, dim):
    """(mean squared error, gradient) with the device's BF16 roundings."""
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, np.float32), dim)
    B = X.shape[0]
    Y, (Xb, Z1, H1, Z2, H2) = forward(p, X, dim, keep=True)
    R = Y - np.asarray(T, np.float32)
    loss = float(np.sum(R.astype(np.float64) ** 2) / B)
    dY = bf16(np.float32(2.0 / B) * R)
    dZ2 = (_mm(dY, bf16(W3).T) * gelu_grad(Z2)).astype(np.float32)
    dZ2b = bf16(dZ2)
    dZ1 = (_mm(dZ2b, bf16(W2).T) * gelu_grad(Z1)).astype(np.float32)
    dZ1b = bf16(dZ1)
    g = [_mm(Xb.T, dZ1b), dZ1b.astype(np.float64).sum(0), _mm(H1.T, dZ2b), dZ2b.astype(np.float64).sum(0),
         _mm(H2.T, dY), dY.astype(np.float64).sum(0)]
    return loss, np.concatenate([np.ravel(a) for a in g]).astype(np.float32)


def _gelu64(x):
    return 0.5 * x * (1 + np.tanh(0.7978845608 * (x + 0.044715 * x ** 3)))


def _gelu64_grad(x):
    t = np.tanh(0.7978845608 * (x + 0.044715 * x ** 3))
    return 0.5 * (1 + t) + 0.5 * x * (1 - t * t) * 0.7978845608 * (1 + 0.134145 * x * x)


def loss_grad_f64(p, X, T, dim):
    """Exact float64 loss and gradient (no rounding): the FD-checkable map."""
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, np.float64), dim)
    I = DIMS[dim]["n_in"]
    X = np.asarray(X, np.float64)[:, :I]
    B = X.shape[0]
    Z1 = X @ W1 + b1
    H1 = _gelu64(Z1)
    Z2 = H1 @ W2 + b2
    H2 = _gelu64(Z2)
    Y = H2 @ W3 + b3
    R = Y - T
    loss = float(np.sum(R ** 2) / B)
    dY = 2.0 * R / B
    dZ2 = (dY @ W3.T) * _gelu64_grad(Z2)
    dZ1 = (dZ2 @ W2.T) * _gelu64_grad(Z1)
    g = [X.T @ dZ1, dZ1.sum(0), H1.T @ dZ2, dZ2.sum(0), H2.T @ dY, dY.sum(0)]
    return loss, np.concatenate([np.ravel(a) for a in g])


def frame_features(ring, dprev, dim):
    """Local frame (rows), length scale h and the n_in_padded features of one
    node. ring[a] = coordinates of one-ring node a (row-major over offsets
    {-1,0,1}^dim, axis 0 most significant); dprev = previous displacement."""
    ring = np.asarray(ring, np.float64)
    C = (3 ** dim - 1) // 2
    S = [3 ** (dim - 1 - k) for k in range(dim)]
    a = [ring[C + S[k]] - ring[C - S[k]] for k in range(dim)]
    n = [np.linalg.norm(v) for v in a]
    if dim == 2:
        h = 0.25 * (n[0] + n[1])
        e0 = a[0] / n[0]
        E = np.array([e0, [-e0[1], e0[0]]])
    else:
        h = sum(n) / 6.0
        e0 = a[0] / n[0]
        v = a[1] - (a[1] @ e0) * e0
        e1 = v / np.linalg.norm(v)
        E = np.array([e0, e1, np.cross(e0, e1)])
    rel = (ring - ring[C]) @ E.T / h
    const = [1.0, 0.0] if dim == 2 else [1.0]
    f = np.concatenate([rel.ravel(), const, (E @ np.asarray(dprev, np.float64)) / h])
    out = np.zeros(DIMS[dim]["n_in_padded"])
    out[:f.size] = f
    return E, h, out
