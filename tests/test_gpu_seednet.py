"""GPU parity of the learned-seed network kernels against the conv oracle
(oracle/conv_oracle.py): every product with bfloat16-rounded operands in
float64 versus the matrix-core kernels' float32 accumulation."""
import numpy as np
import pytest

from oracle import conv_oracle as CO

pytestmark = pytest.mark.gpu

SHAPES = [(1, 5, 4, 6), (2, 3, 7, 5), (1, 9, 9, 9)]


def rel(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(1e-30, np.abs(b).max()))


def dev(a):
    from voxeldvc_amd import _lib as L
    return L.to_dev_flat(np.ascontiguousarray(a, dtype=np.float32))


@pytest.mark.parametrize('layer', [(2, 16, 1), (16, 32, 2), (32, 32, 1), (32, 3, 1), (3, 5, 2),
                                   (16, 5, 1), (32, 20, 2)])
@pytest.mark.parametrize('shape', SHAPES)
def test_conv_forward(gpu, layer, shape):
    from voxeldvc_amd import _lib as L
    C, O, s = layer
    N, D, H, W = shape
    print('conv fwd', layer, shape, flush=True)
    rng = np.random.default_rng(sum(layer) + sum(shape))
    x = rng.normal(size=(N, D, H, W, C)).astype(np.float32)
    w = (rng.normal(size=(O, 27, C)) * 0.2).astype(np.float32)
    b = rng.normal(size=O).astype(np.float32)
    ref = CO.conv3d_forward(CO.bf16_round(x), CO.bf16_round(w), b.astype(np.float64), s)
    y = L.empty(ref.size)
    for relu in (0, 1):
        L.call('vdvc_conv3d_forward', L.ctx(), N, D, H, W, C, O, s, dev(x), dev(w),
               dev(b), relu, L.ptr(y))
        r = CO.relu(ref) if relu else ref
        assert rel(L.to_host(y).reshape(ref.shape), r) < 1e-5


@pytest.mark.parametrize('layer', [(16, 32, 2), (32, 32, 1), (32, 3, 1)])
def test_conv_forward_many_blocks(gpu, layer):
    """A 17^3 (and 33^3 for the stride-2 layer) volume: tens of blocks of
    4 x 4 x 16 output positions, partial blocks on every axis."""
    from voxeldvc_amd import _lib as L
    C, O, s = layer
    n = 33 if s == 2 else 17
    print('conv fwd many blocks', layer, n, flush=True)
    rng = np.random.default_rng(7 + C + O)
    x = rng.normal(size=(1, n, n, n, C)).astype(np.float32)
    w = (rng.normal(size=(O, 27, C)) * 0.2).astype(np.float32)
    b = rng.normal(size=O).astype(np.float32)
    ref = CO.relu(CO.conv3d_forward(CO.bf16_round(x), CO.bf16_round(w), b.astype(np.float64), s))
    y = L.empty(ref.size)
    L.call('vdvc_conv3d_forward', L.ctx(), 1, n, n, n, C, O, s, dev(x), dev(w), dev(b), 1,
           L.ptr(y))
    assert rel(L.to_host(y).reshape(ref.shape), ref) < 1e-5


@pytest.mark.parametrize('layer', [(2, 16, 1), (16, 32, 2), (32, 32, 1), (32, 3, 1), (3, 5, 2)])
@pytest.mark.parametrize('shape', SHAPES)
def test_conv_backward(gpu, layer, shape):
    from voxeldvc_amd import _lib as L
    C, O, s = layer
    N, D, H, W = shape
    print('conv bwd', layer, shape, flush=True)
    rng = np.random.default_rng(100 + sum(layer) + sum(shape))
    x = rng.normal(size=(N, D, H, W, C)).astype(np.float32)
    act = np.where(rng.random(x.shape) < 0.3, 0.0, np.abs(x)).astype(np.float32)
    w = (rng.normal(size=(O, 27, C)) * 0.2).astype(np.float32)
    Do, Ho, Wo = (CO.out_size(q, s) for q in (D, H, W))
    dy = rng.normal(size=(N, Do, Ho, Wo, O)).astype(np.float32)
    dx_ref, _, _ = CO.conv3d_backward(x, CO.bf16_round(w), CO.bf16_round(dy), s)
    _, dw_ref, _ = CO.conv3d_backward(CO.bf16_round(x), w, CO.bf16_round(dy), s)
    db_ref = dy.astype(np.float64).sum(axis=(0, 1, 2, 3))
    dx = L.empty(x.size)
    L.call('vdvc_conv3d_backward_data', L.ctx(), N, D, H, W, C, O, s, dev(dy),
           dev(w), None, L.ptr(dx))
    assert rel(L.to_host(dx).reshape(x.shape), dx_ref) < 1e-5
    # fused ReLU derivative of the layer input
    L.call('vdvc_conv3d_backward_data', L.ctx(), N, D, H, W, C, O, s, dev(dy),
           dev(w), dev(act), L.ptr(dx))
    assert rel(L.to_host(dx).reshape(x.shape), CO.relu_backward(act, dx_ref)) < 1e-5
    dw = L.empty(w.size)
    db = L.empty(O)
    L.call('vdvc_conv3d_backward_weight', L.ctx(), N, D, H, W, C, O, s, dev(x),
           dev(dy), L.ptr(dw), L.ptr(db))
    assert rel(L.to_host(dw).reshape(w.shape), dw_ref) < 1e-5
    assert rel(L.to_host(db), db_ref) < 1e-5


def test_adam_mse_and_input(gpu):
    from voxeldvc_amd import _lib as L
    rng = np.random.default_rng(3)
    n = 1000
    p, g = rng.normal(size=n), rng.normal(size=n)
    m, v = rng.normal(size=n) * 0.1, rng.random(n) * 0.1
    dp, dm, dv = dev(p), dev(m), dev(v)
    L.call('vdvc_adam_step', L.ctx(), n, L.ptr(dp), dev(g), L.ptr(dm), L.ptr(dv), 1e-2,
           0.9, 0.999, 1e-8, 7)
    m1 = 0.9 * m + 0.1 * g
    v1 = 0.999 * v + 0.001 * g * g
    p1 = p - 1e-2 * (m1 / (1 - 0.9 ** 7)) / (np.sqrt(v1 / (1 - 0.999 ** 7)) + 1e-8)
    assert rel(L.to_host(dp), p1) < 1e-5 and rel(L.to_host(dm), m1) < 1e-6
    pred, lab = rng.normal(size=n), rng.normal(size=n)
    gr = L.empty(n)
    loss = (L._d * 1)()
    L.call('vdvc_mse', L.ctx(), n, dev(pred), dev(lab), L.ptr(gr), loss)
    rl, rg = CO.mse(pred.astype(np.float32).astype(np.float64),
                    lab.astype(np.float32).astype(np.float64))
    # pred - label is formed in float32, the squares summed in FP64
    assert abs(loss[0] - rl) < 1e-6 * rl and rel(L.to_host(gr), rg) < 1e-6
    # network input: stride-2 decimation, per-channel standardisation
    from voxeldvc_amd import seednet
    f = rng.uniform(0, 255, size=(11, 9, 7)).astype(np.float32)
    gi = rng.uniform(0, 255, size=(11, 9, 7)).astype(np.float32)
    fd, shp = L.to_dev_volume(f)
    gd, _ = L.to_dev_volume(gi)
    x, dims = seednet.network_input(fd, gd, shp, 4)
    assert dims == (1, 4, 5, 6)
    got = L.to_host(x).reshape(4, 5, 6, 2)
    for ch, img in enumerate((f, gi)):
        sub = img[::2, ::2, ::2].astype(np.float64)
        ref = ((sub - sub.mean()) / sub.std()).transpose(2, 1, 0)
        assert np.abs(got[..., ch] - ref).max() < 1e-5


def test_network_gradient_matches_oracle(gpu):
    """Full 5-layer forward/backward through the C ABI against the oracle
    network with the same bfloat16 operand rounding per product (float32
    against float64 accumulation flips an occasional intermediate rounding,
    hence bounds of a few 1e-3)."""
    from voxeldvc_amd import _lib as L
    from voxeldvc_amd import seednet
    net = seednet.SeedNet(seed=4)
    rng = np.random.default_rng(5)
    N, D = 2, 9
    x = rng.normal(size=(N, D, D, D, 2)).astype(np.float32)
    out, acts, shapes = net.forward(dev(x), (N, D, D, D), keep=True)
    params = [(L.to_host(w).astype(np.float64), L.to_host(b).astype(np.float64))
              for w, b in zip(net.w, net.b)]
    strides = [s for _, _, s in seednet.LAYERS]
    racts = CO.network_forward(x.astype(np.float64), params, strides, bf16=True)
    assert rel(L.to_host(out).reshape(racts[-1].shape), racts[-1]) < 5e-3
    lab = rng.normal(size=racts[-1].shape)
    _, dref = CO.mse(racts[-1], lab)
    rgrads = CO.network_backward(racts, params, strides, dref, bf16=True)
    grad = L.torch().empty_like(out)
    loss = (L._d * 1)()
    L.call('vdvc_mse', L.ctx(), out.numel(), L.ptr(out), dev(lab), L.ptr(grad), loss)
    g = net.backward(acts, shapes, grad)
    nl = len(seednet.LAYERS)
    for li in range(nl):
        gw = L.to_host(g[li]).reshape(params[li][0].shape)
        gb = L.to_host(g[nl + li])
        ew = np.linalg.norm(gw - rgrads[li][0]) / np.linalg.norm(rgrads[li][0])
        eb = np.linalg.norm(gb - rgrads[li][1]) / np.linalg.norm(rgrads[li][1])
        print('layer', li, 'grad rel err w %.2e b %.2e' % (ew, eb), flush=True)
        assert ew < 1e-2 and eb < 1e-2


def test_training_reduces_loss(gpu):
    """A few hundred Adam steps on a fixed batch fit it (the training loop
    runs entirely on the library kernels)."""
    from voxeldvc_amd import seednet
    rng = np.random.default_rng(6)
    N, D = 4, 9
    x = dev(rng.normal(size=(N, D, D, D, 2)))
    m = seednet.out_extent(D, 2)
    y = dev(rng.normal(size=(N, m, m, m, 3)) * 0.5)
    net = seednet.SeedNet(seed=7)
    l0 = net.train_step(x, (N, D, D, D), y, 0.0)
    for _ in range(300):
        loss = net.train_step(x, (N, D, D, D), y, 3e-3)
    print('loss', l0, '->', loss, flush=True)
    assert loss < 0.5 * l0
