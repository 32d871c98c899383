"""Numpy restatement of the learned-physics network of the fine transport.

TEST INFRASTRUCTURE ONLY (the checker of the library's BF16 matrix-core kernels): the
4-128-128-128-33 MLP with GELU, its loss (soft-target cross-entropy over the 4 x 8 width
classes, squared errors of the decoded widths and of the attenuation output) and the
gradients, in float64.
With `emulate=True` the operands are rounded to BF16 where the library rounds them (inputs,
weights, hidden activations, output gradients), so the library can be compared to it up to
the FP32 accumulation order.  The gradient formulas themselves are pinned by central finite
differences of the float64 loss (tests/test_mlp_oracle.py).
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

NIN, NH, NOUT, NCE, NCT = 4, 128, 33, 4, 8
NCLS = NCE * NCT
OFF_W0 = 0
OFF_B0 = OFF_W0 + NH * NIN
OFF_W1 = OFF_B0 + NH
OFF_B1 = OFF_W1 + NH * NH
OFF_W2 = OFF_B1 + NH
OFF_B2 = OFF_W2 + NH * NH
OFF_W3 = OFF_B2 + NH
OFF_B3 = OFF_W3 + NOUT * NH
NPARAM = OFF_B3 + NOUT

SHAPES = [("W0", OFF_W0, (NH, NIN)), ("b0", OFF_B0, (NH,)), ("W1", OFF_W1, (NH, NH)),
          ("b1", OFF_B1, (NH,)), ("W2", OFF_W2, (NH, NH)), ("b2", OFF_B2, (NH,)),
          ("W3", OFF_W3, (NOUT, NH)), ("b3", OFF_B3, (NOUT,))]


def bf16(x):
    """Round to BF16 (nearest even), returned as float32 (NaN kept quiet)."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, (u >> 16) | 0x40, r)
    return (r.astype(np.uint32) << 16).view(np.float32)


def unpack(p):
    p = np.asarray(p, np.float64)
    return {name: p[off:off + int(np.prod(shape))].reshape(shape) for name, off, shape in SHAPES}


def pack(d):
    p = np.zeros(NPARAM, np.float64)
    for name, off, shape in SHAPES:
        p[off:off + int(np.prod(shape))] = np.asarray(d[name]).ravel()
    return p


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def soft_targets(labels, dict4):
    """Joint soft assignment of (ln sigma_E, ln sigma_theta) onto the 4 x 8 classes."""
    lb = np.asarray(labels, np.float64)

    def assign(v, lo, hi, n):
        p = np.clip((v - lo) / (hi - lo) * (n - 1), 0.0, n - 1)
        i0 = np.minimum(p.astype(np.int64), n - 2)
        f = p - i0
        out = np.zeros((len(v), n))
        out[np.arange(len(v)), i0] = 1.0 - f
        out[np.arange(len(v)), i0 + 1] += f
        return out

    pe = assign(lb[:, 0], dict4[0], dict4[1], NCE)
    pt = assign(lb[:, 1], dict4[2], dict4[3], NCT)
    return (pe[:, :, None] * pt[:, None, :]).reshape(len(lb), NCLS)


def decode(z, dict4):
    """(ln sigma_E, ln sigma_theta, a) of raw outputs z[B][33] (log-moments of the classes)."""
    z = np.asarray(z, np.float64)
    lg = z[:, :NCLS] - z[:, :NCLS].max(axis=1, keepdims=True)
    q = np.exp(lg)
    q /= q.sum(axis=1, keepdims=True)
    ce = np.repeat(np.arange(NCE), NCT)
    ct = np.tile(np.arange(NCT), NCE)
    vE = dict4[0] + (dict4[1] - dict4[0]) * ce / (NCE - 1)
    vT = dict4[2] + (dict4[3] - dict4[2]) * ct / (NCT - 1)
    return np.stack([q @ vE, q @ vT, np.maximum(np.exp(z[:, NCLS]) - 1e-9, 0.0)], axis=1)


def forward(p, x, emulate=True):
    W = unpack(p)
    rnd = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    a0 = rnd(x).astype(np.float64)
    cache = {"a0": a0}
    a = a0
    for k in range(3):
        z = a @ rnd(W[f"W{k}"]).astype(np.float64).T + W[f"b{k}"]
        cache[f"z{k + 1}"] = z
        a = rnd(gelu(z)).astype(np.float64) if emulate else gelu(z)
        cache[f"a{k + 1}"] = a
    z4 = a @ rnd(W["W3"]).astype(np.float64).T + W["b3"]
    cache["z4"] = z4
    return z4, cache


def loss_and_grad(p, x, labels, dict4, w_att=1.0, w_mom=1.0, emulate=True):
    """Mean loss and its gradient (flat parameter layout) of a batch: soft-target cross-entropy
    + w_mom * squared errors of the decoded ln widths + w_att * attenuation output error^2."""
    W = unpack(p)
    rnd = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    z4, c = forward(p, x, emulate)
    B = len(x)
    lab = np.asarray(labels, np.float64)
    tgt = soft_targets(labels, dict4)
    lg = z4[:, :NCLS] - z4[:, :NCLS].max(axis=1, keepdims=True)
    lse = np.log(np.exp(lg).sum(axis=1, keepdims=True))
    logp = lg - lse
    q = np.exp(logp)
    ce = np.repeat(np.arange(NCE), NCT)
    ct = np.tile(np.arange(NCT), NCE)
    vE = dict4[0] + (dict4[1] - dict4[0]) * ce / (NCE - 1)
    vT = dict4[2] + (dict4[3] - dict4[2]) * ct / (NCT - 1)
    mE, mT = q @ vE, q @ vT
    eE, eT = mE - lab[:, 0], mT - lab[:, 1]
    d = z4[:, NCLS] - lab[:, 2]
    loss = (-(tgt * logp).sum() + w_mom * (eE * eE + eT * eT).sum() + w_att * (d * d).sum()) / B
    dz = np.zeros_like(z4)
    gm = 2.0 * w_mom * (eE[:, None] * q * (vE[None, :] - mE[:, None]) + eT[:, None] * q * (vT[None, :] - mT[:, None]))
    dz[:, :NCLS] = (q - tgt + gm) / B
    dz[:, NCLS] = 2.0 * w_att * d / B
    g = {}
    dzr = rnd(dz).astype(np.float64)
    g["W3"] = dzr.T @ c["a3"]
    g["b3"] = dz.sum(axis=0)
    da = dzr @ rnd(W["W3"]).astype(np.float64)
    for k in (2, 1, 0):
        dzk = da * gelu_grad(c[f"z{k + 1}"])
        dzkr = rnd(dzk).astype(np.float64)
        g[f"W{k}"] = dzkr.T @ c[f"a{k}"]
        g[f"b{k}"] = dzk.sum(axis=0)
        if k:
            da = dzkr @ rnd(W[f"W{k}"]).astype(np.float64)
    return loss, pack(g)
