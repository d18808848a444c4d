"""Pins the conv oracle (oracle/conv_oracle.py) the seed-network kernels are
checked against: brute-force loops, scipy.signal.correlate, the adjoint
identity, finite differences and the bfloat16 rounding of torch."""
import numpy as np
import pytest

from oracle import conv_oracle as CO


def brute_conv(x, w, b, s):
    N, D, H, W, C = x.shape
    O = w.shape[0]
    Do, Ho, Wo = (CO.out_size(q, s) for q in (D, H, W))
    y = np.zeros((N, Do, Ho, Wo, O))
    for n in range(N):
        for z in range(Do):
            for yy in range(Ho):
                for xx in range(Wo):
                    for o in range(O):
                        acc = b[o]
                        for dz in range(3):
                            for dy in range(3):
                                for dx in range(3):
                                    iz, iy, ix = z * s + dz - 1, yy * s + dy - 1, xx * s + dx - 1
                                    if 0 <= iz < D and 0 <= iy < H and 0 <= ix < W:
                                        tap = (dz * 3 + dy) * 3 + dx
                                        acc += w[o, tap] @ x[n, iz, iy, ix]
                        y[n, z, yy, xx, o] = acc
    return y


def rnd(seed, *shape):
    return np.random.default_rng(seed).normal(size=shape)


@pytest.mark.parametrize('s', [1, 2])
@pytest.mark.parametrize('dims', [(4, 5, 3), (5, 5, 5), (1, 2, 3)])
def test_forward_brute_force(s, dims):
    x = rnd(1, 2, *dims, 3)
    w = rnd(2, 4, 27, 3)
    b = rnd(3, 4)
    assert np.abs(CO.conv3d_forward(x, w, b, s) - brute_conv(x, w, b, s)).max() < 1e-12


def test_forward_scipy_correlate():
    from scipy.signal import correlate
    x = rnd(4, 1, 6, 5, 7, 2)
    w = rnd(5, 3, 27, 2)
    y = CO.conv3d_forward(x, w, np.zeros(3), 1)
    for o in range(3):
        ref = sum(correlate(x[0, ..., c], w[o, :, c].reshape(3, 3, 3), mode='same')
                  for c in range(2))
        assert np.abs(y[0, ..., o] - ref).max() < 1e-12


@pytest.mark.parametrize('s', [1, 2])
def test_backward_adjoint_and_bias(s):
    x = rnd(6, 2, 5, 4, 6, 3)
    w = rnd(7, 2, 27, 3)
    y = CO.conv3d_forward(x, w, np.zeros(2), s)
    r = rnd(8, *y.shape)
    dx, dw, db = CO.conv3d_backward(x, w, r, s)
    # <conv(x), r> = <x, conv^T(r)> = <w, dW(x, r)>
    lhs = float((y * r).sum())
    assert abs(lhs - float((x * dx).sum())) < 1e-9 * abs(lhs)
    assert abs(lhs - float((w * dw).sum())) < 1e-9 * abs(lhs)
    assert np.abs(db - r.sum(axis=(0, 1, 2, 3))).max() < 1e-12


def test_network_gradient_finite_differences():
    rng = np.random.default_rng(9)
    strides = [1, 2, 1]
    params = [(rng.normal(size=(4, 27, 2)) * 0.3, rng.normal(size=4) * 0.1),
              (rng.normal(size=(5, 27, 4)) * 0.3, rng.normal(size=5) * 0.1),
              (rng.normal(size=(3, 27, 5)) * 0.3, rng.normal(size=3) * 0.1)]
    x = rng.normal(size=(2, 5, 5, 5, 2))
    acts = CO.network_forward(x, params, strides)
    label = rng.normal(size=acts[-1].shape)
    loss, dout = CO.mse(acts[-1], label)
    grads = CO.network_backward(acts, params, strides, dout)
    eps = 1e-6
    for li, (w, b) in enumerate(params):
        for idx in [(0, 0, 0), (1, 13, 1), (w.shape[0] - 1, 26, w.shape[2] - 1)]:
            wp = [(p[0].copy(), p[1].copy()) for p in params]
            wp[li][0][idx] += eps
            lp, _ = CO.mse(CO.network_forward(x, wp, strides)[-1], label)
            wp[li][0][idx] -= 2 * eps
            lm, _ = CO.mse(CO.network_forward(x, wp, strides)[-1], label)
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - grads[li][0][idx]) < 1e-6 * max(1.0, abs(fd))
        wp = [(p[0].copy(), p[1].copy()) for p in params]
        wp[li][1][0] += eps
        lp, _ = CO.mse(CO.network_forward(x, wp, strides)[-1], label)
        wp[li][1][0] -= 2 * eps
        lm, _ = CO.mse(CO.network_forward(x, wp, strides)[-1], label)
        assert abs((lp - lm) / (2 * eps) - grads[li][1][0]) < 1e-6


def test_bf16_round_matches_torch():
    torch = pytest.importorskip('torch')
    a = np.concatenate([np.random.default_rng(10).normal(size=4096) * 10.0 ** np.arange(-4, 4).repeat(512),
                        [0.0, -0.0, 1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 65504.0]]).astype(np.float32)
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(CO.bf16_round(a), ref)
