"""numpy restatement of the learned-closure MLP (TEST INFRASTRUCTURE ONLY).

The closure has no counterpart in the reference: pyranda evaluates the LAD
chain with compact operators only (examples/RM3D.py:95-104). This module
restates, point by point on the CPU, what the BF16 matrix-core kernels
(pyr_mlp_infer, pyr_mlp_train_grad) compute, so that their outputs and
gradients are checked against a CPU computation on the same BF16-rounded
operands. Only tests/ import it.

Operands and rounding (the kernels' contract):
  * features x[p, f*9 + 3*(dy+1) + (dx+1)] = bf16((F_f(i+dx, j+dy) - shift_f) * scale_f)
    for f < 3 (27 values; entries 27..31 are zero), neighbours wrapped on
    periodic axes and clamped to the edge on bounded ones, point p = j*nx + i;
  * weights W1 (32x64), W2 (64x64), W3 (64x16) in BF16, biases in FP32;
  * h1 = bf16(relu(x W1 + b1)), h2 = bf16(relu(h1 W2 + b2)), o = h2 W3 + b3,
    products summed exactly here (FP32 on the matrix cores);
  * inference: out_j = max(o_j, 0) * oscale_j for j = 0, 1;
  * training (mean squared error on the scaled targets, no clamp):
    e_j = o_j - t_j / oscale_j, loss = sum_p sum_{j<2} e_j^2,
    dO = bf16(2 e / npoints) (outputs 2..15 zero),
    dZ2 = bf16([h2 > 0] * dO W3^T), dZ1 = bf16([h1 > 0] * dZ2 W2^T),
    gW3 = h2^T dO, gW2 = h1^T dZ2, gW1 = x^T dZ1, gb = column sums of dO, dZ2, dZ1.
"""
import numpy as np

NIN, NH, NO, NF = 32, 64, 16, 27


def bf16(a):
    """Round float32 values to the nearest BF16 (ties to even), as float32;
    NaN stays NaN."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).reshape(a.shape)
    return np.where(np.isnan(a), np.float32(np.nan), out)


def features(fields, shift, scale, periodic):
    """(npoints, 32) BF16-rounded inputs from three (nx, ny) host arrays."""
    nx, ny = fields[0].shape
    x = np.zeros((nx * ny, NIN), dtype=np.float32)
    i = np.arange(nx)
    j = np.arange(ny)
    for k in range(NF):
        f, r = divmod(k, 9)
        dx, dy = r % 3 - 1, r // 3 - 1
        ii = (i + dx) % nx if periodic[0] else np.clip(i + dx, 0, nx - 1)
        jj = (j + dy) % ny if periodic[1] else np.clip(j + dy, 0, ny - 1)
        v = np.asarray(fields[f], dtype=np.float64)[np.ix_(ii, jj)]
        v = ((v.astype(np.float32) - np.float32(shift[f])) * np.float32(scale[f])).astype(np.float32)
        x[:, k] = v.T.reshape(-1)          # point p = j * nx + i
    return bf16(x)


def _mm(a, b, exact=False):
    r = a.astype(np.float64) @ b.astype(np.float64)
    return r if exact else r.astype(np.float32)


def _rnd(a, exact):
    return np.asarray(a, dtype=np.float64) if exact else bf16(a)


def forward(x, params, exact=False):
    """exact: float64 throughout and no BF16 rounding (the mathematics the
    rounded computation approximates; used to pin the backward pass against
    finite differences)."""
    W1, W2, W3, b1, b2, b3 = params
    h1 = _rnd(np.maximum(_mm(x, _rnd(W1, exact), exact) + b1, 0.0), exact)
    h2 = _rnd(np.maximum(_mm(h1, _rnd(W2, exact), exact) + b2, 0.0), exact)
    o = _mm(h2, _rnd(W3, exact), exact) + b3
    return h1, h2, (o if exact else o.astype(np.float32))


def infer(x, params, oscale):
    """(out0, out1) per point."""
    o = forward(x, params)[2]
    return [np.maximum(o[:, j], 0.0) * np.float32(oscale[j]) for j in range(2)]


def loss_of(x, targets, params, oscale, exact=False):
    o = forward(x, params, exact)[2]
    return sum(float(np.sum((o[:, j].astype(np.float64) - np.asarray(targets[j], np.float64) / oscale[j]) ** 2))
               for j in range(2))


def gradient(x, targets, params, oscale, exact=False):
    """(gW1, gW2, gW3, gb1, gb2, gb3), loss for targets (t0, t1) per point;
    the gradient is that of loss / npoints."""
    W1, W2, W3, b1, b2, b3 = params
    n = x.shape[0]
    ft = np.float64 if exact else np.float32
    h1, h2, o = forward(x, params, exact)
    dO = np.zeros((n, NO), dtype=ft)
    loss = 0.0
    for j in range(2):
        e = (o[:, j] - np.asarray(targets[j], dtype=ft) / ft(oscale[j])).astype(ft)
        loss += float(np.sum(e.astype(np.float64) ** 2))
        dO[:, j] = e * ft(2.0 / n)
    dO = _rnd(dO, exact)
    dZ2 = _rnd(np.where(h2 > 0, _mm(dO, _rnd(W3, exact).T, exact), 0.0), exact)
    dZ1 = _rnd(np.where(h1 > 0, _mm(dZ2, _rnd(W2, exact).T, exact), 0.0), exact)
    grads = (_mm(x.T, dZ1, exact), _mm(h1.T, dZ2, exact), _mm(h2.T, dO, exact),
             dZ1.sum(0, dtype=np.float64).astype(ft), dZ2.sum(0, dtype=np.float64).astype(ft),
             dO.sum(0, dtype=np.float64).astype(ft))
    return grads, loss
