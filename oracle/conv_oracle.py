"""CPU restatement of the learned-seed network's arithmetic: 3-D convolution
(cross-correlation, 3x3x3 kernel, zero padding 1, stride 1 or 2), its two
backward products, ReLU and the mean-squared-error loss.  TEST
INFRASTRUCTURE ONLY: imported by tests/, never by the product path.

The network has no reference counterpart (it is the build's own extension
of the solver); these functions are pinned in tests/test_conv_oracle.py by a
brute-force 7-loop convolution, scipy.signal.correlate, finite differences
of a scalar loss, and the adjoint identity <conv(x), y> = <x, conv^T(y)>.

Layout (the device layout): activations [n][z][y][x][c] (channels last),
weights [o][tap][c] with tap = (dz*3 + dy)*3 + dx, dz, dy, dx in 0..2.
"""
import numpy as np


def out_size(n, stride):
    """Output extent of a 3-tap, pad-1 convolution: positions 0, s, 2s, ...
    up to n - 1."""
    return (n - 1) // stride + 1


def conv3d_forward(x, w, b, stride):
    """y[n, p, o] = b[o] + sum_{tap, c} w[o, tap, c] x[n, p*s + d - 1, c]."""
    N, D, H, W, C = x.shape
    O = w.shape[0]
    Do, Ho, Wo = (out_size(s, stride) for s in (D, H, W))
    xp = np.zeros((N, D + 2, H + 2, W + 2, C), dtype=np.float64)
    xp[:, 1:-1, 1:-1, 1:-1] = x
    y = np.zeros((N, Do, Ho, Wo, O), dtype=np.float64)
    y += np.asarray(b, np.float64)
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                tap = (dz * 3 + dy) * 3 + dx
                patch = xp[:, dz:dz + stride * (Do - 1) + 1:stride,
                           dy:dy + stride * (Ho - 1) + 1:stride,
                           dx:dx + stride * (Wo - 1) + 1:stride]
                y += patch @ w[:, tap, :].T.astype(np.float64)
    return y


def conv3d_backward(x, w, dy, stride):
    """(dx, dw, db) of conv3d_forward for upstream gradient dy."""
    N, D, H, W, C = x.shape
    O = w.shape[0]
    Do, Ho, Wo = dy.shape[1:4]
    xp = np.zeros((N, D + 2, H + 2, W + 2, C), dtype=np.float64)
    xp[:, 1:-1, 1:-1, 1:-1] = x
    dxp = np.zeros_like(xp)
    dw = np.zeros(w.shape, dtype=np.float64)
    dyf = dy.astype(np.float64)
    for dz in range(3):
        for dyy in range(3):
            for dx in range(3):
                tap = (dz * 3 + dyy) * 3 + dx
                sl = (slice(None), slice(dz, dz + stride * (Do - 1) + 1, stride),
                      slice(dyy, dyy + stride * (Ho - 1) + 1, stride),
                      slice(dx, dx + stride * (Wo - 1) + 1, stride))
                patch = xp[sl]
                dw[:, tap, :] += np.einsum('nzyxo,nzyxc->oc', dyf, patch)
                dxp[sl] += dyf @ w[:, tap, :].astype(np.float64)
    db = dyf.sum(axis=(0, 1, 2, 3))
    return dxp[:, 1:-1, 1:-1, 1:-1], dw, db


def relu(x):
    return np.maximum(x, 0.0)


def relu_backward(act, grad):
    """Gradient through a ReLU whose output is `act`."""
    return np.where(act > 0, grad, 0.0)


def mse(pred, label):
    """(loss, dloss/dpred) of the mean squared error over every entry."""
    d = pred.astype(np.float64) - label
    return float(np.mean(d * d)), 2.0 * d / d.size


def bf16_round(a):
    """Round-to-nearest-even to bfloat16, returned as float64 (the value a
    bfloat16 operand carries into the matrix core)."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def network_forward(x, params, strides, bf16=False):
    """The seed network: conv layers with ReLU after all but the last.
    Returns the list of activations (input first, output last).  With
    bf16=True every product operand is rounded to bfloat16 first, as the
    matrix cores see it."""
    q = bf16_round if bf16 else (lambda a: a)
    acts = [x]
    for li, ((w, b), s) in enumerate(zip(params, strides)):
        y = conv3d_forward(q(acts[-1]), q(w), b, s)
        if li < len(params) - 1:
            y = relu(y)
        acts.append(y)
    return acts


def network_backward(acts, params, strides, dout, bf16=False):
    """Gradients of every (w, b) for upstream gradient dout at the output
    (product operands rounded to bfloat16 when bf16=True; the bias gradient
    sums the unrounded upstream gradient)."""
    q = bf16_round if bf16 else (lambda a: a)
    grads = [None] * len(params)
    g = dout
    for li in range(len(params) - 1, -1, -1):
        w, _ = params[li]
        dx, _, _ = conv3d_backward(acts[li], q(w), q(g), strides[li])
        _, dw, _ = conv3d_backward(q(acts[li]), w, q(g), strides[li])
        db = np.asarray(g, np.float64).sum(axis=(0, 1, 2, 3))
        grads[li] = (dw, db)
        if li > 0:
            g = relu_backward(acts[li], dx)
    return grads
