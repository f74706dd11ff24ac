"""CPU restatement of the learned neighbour-radius network (numpy).  TEST
INFRASTRUCTURE ONLY: used by tests/ as the checker of the gpunb_mlp_* kernels.

The network has no counterpart in the reference (which sets RS only by the
Ahmad-Cohen volume rule, src/Main/regcor_gpu.F:146-253); what is restated here
is the arithmetic the library documents for it (matrix products on BF16
operands with FP32 accumulation, biases added in FP32 after the product):

  xh = bf16((x - xmean) / xstd)                       (FP32 ops, then BF16)
  h1 = bf16(silu(bf16(W1) xh + bf16(b1)))
  h2 = bf16(silu(bf16(W2) h1 + bf16(b2)))
  yhat = bf16(W3) h2 + bf16(b3);   y = yhat * ystd + ymean
  L = mean_s (yhat_s - (t_s - ymean) / ystd)^2

bf16() rounds FP32 to the nearest BF16 value, ties to even.  Backward: the
BF16 rounding is differentiated as the identity (straight-through); the
deltas d2 = d3 W3 silu'(z2) and d1 = (W2^T bf16(d2)) silu'(z1) enter the
matrix products (d1 and the weight gradients) as BF16, the bias and
output-layer gradients use them unrounded.

Modes:
  emulate=True   FP32 values as the kernels produce them: every matrix-vector
                 product is the exact sum of its (exact) BF16 x BF16 products
                 rounded once to FP32, then the FP32 bias is added.  The
                 matrix cores may round their partial sums differently (FP32
                 internal accumulation), and the kernels' exp() is hardware
                 exp2 based; both can move a bf16() result by one BF16 step on
                 rare elements, which the tests' tolerances allow for.
  emulate=False  FP64 with the same BF16 roundings.
  bf16_round=False  FP64 without any rounding (deltas unrounded too): a
                 smooth function of the parameters, used to pin backward() by
                 finite differences.

Parameter vector layout (4801 FP32): W1[64][8] b1[64] W2[64][64] b2[64]
W3[64] b3.
"""
from __future__ import annotations

import numpy as np

NI, NH = 8, 64
OW1 = 0
OB1 = OW1 + NH * NI
OW2 = OB1 + NH
OB2 = OW2 + NH * NH
OW3 = OB2 + NH
OB3 = OW3 + NH
NP = OB3 + 1


def bf16(a):
    """Round FP32 values to BF16 (nearest, ties to even); returns FP32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(p):
    p = np.asarray(p)
    return (p[OW1:OB1].reshape(NH, NI), p[OB1:OW2], p[OW2:OB2].reshape(NH, NH),
            p[OB2:OW3], p[OW3:OB3], p[OB3])


def _matvec_f32(W, b, h):
    """FP32(exact sum_a W[:, a] h[:, a]) + b in FP32 (W, h BF16 values)."""
    acc = (h.astype(np.float64) @ W.astype(np.float64).T).astype(np.float32)
    return (acc + b.astype(np.float32)).astype(np.float32)


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, norm, X, emulate=True, bf16_round=True):
    """Returns (y, cache).  norm = (xmean[8], xstd[8], ymean, ystd)."""
    xm, xs, ym, ys = norm
    X = np.asarray(X)
    if emulate:
        f32 = np.float32
        r = bf16 if bf16_round else (lambda a: a)
        xh = r((X.astype(f32) - np.asarray(xm, f32)) / np.asarray(xs, f32))
        W1, b1, W2, b2, W3, b3 = (r(np.asarray(t, f32)) for t in unpack(p))
        z1 = _matvec_f32(W1, b1, xh)
        h1 = r(_silu(z1.astype(np.float64)).astype(f32))
        z2 = _matvec_f32(W2, b2, h1)
        h2 = r(_silu(z2.astype(np.float64)).astype(f32))
        out = _matvec_f32(W3[None, :], np.atleast_1d(b3), h2)[:, 0]
        y = (out * f32(ys)).astype(f32) + f32(ym)
        cache = (xh.astype(np.float64), z1.astype(np.float64),
                 h1.astype(np.float64), z2.astype(np.float64),
                 h2.astype(np.float64), out.astype(np.float64))
        return y.astype(np.float64), cache
    r = (lambda a: bf16(a).astype(np.float64)) if bf16_round else (
        lambda a: np.asarray(a, np.float64))
    xh = r((X - np.asarray(xm)) / np.asarray(xs))
    W1, b1, W2, b2, W3, b3 = (r(t) for t in unpack(np.asarray(p, np.float64)))
    z1 = xh @ W1.T + b1
    h1 = r(_silu(z1))
    z2 = h1 @ W2.T + b2
    h2 = r(_silu(z2))
    out = h2 @ W3 + b3
    return out * ys + ym, (xh, z1, h1, z2, h2, out)


def loss_and_grad(p, norm, X, t, emulate=True, bf16_round=True):
    """(L, dL/dp) in FP64 from the forward cache; BF16 weights in the
    backward pass as in the forward pass, BF16 deltas in the matrix
    products (see the module docstring)."""
    xm, xs, ym, ys = norm
    _, (xh, z1, h1, z2, h2, out) = forward(p, norm, X, emulate, bf16_round)
    if bf16_round:
        r = lambda a: bf16(a).astype(np.float64)
        W1, b1, W2, b2, W3, b3 = (r(np.asarray(q, np.float32)) for q in unpack(p))
    else:
        r = lambda a: np.asarray(a, np.float64)
        W1, b1, W2, b2, W3, b3 = (r(q) for q in unpack(p))
    ns = xh.shape[0]
    e = out - (np.asarray(t, np.float64) - ym) / ys
    L = np.mean(e * e)
    d3 = 2.0 * e / ns                                   # (ns,)
    d2 = d3[:, None] * W3[None, :] * _dsilu(z2)         # (ns, NH)
    d2q = r(d2)
    d1 = (d2q @ W2) * _dsilu(z1)                        # (ns, NH)
    d1q = r(d1)
    g = np.zeros(NP)
    g[OW1:OB1] = (d1q.T @ xh).ravel()
    g[OB1:OW2] = d1.sum(0)
    g[OW2:OB2] = (d2q.T @ h1).ravel()
    g[OB2:OW3] = d2.sum(0)
    g[OW3:OB3] = d3 @ h2
    g[OB3] = d3.sum()
    return L, g


def adam_step(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1); returns (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
