"""The boundary's data-movement kernels (vdvc_convert, vdvc_copy, vdvc_fill,
vdvc_copy_volume, vdvc_axpby_x) against NumPy: conversions round to nearest
even like NumPy's astype, the volume copy is the reference's
ravel(order='F') into the padded level layout (engine/correlate_gpu.py:50-52)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NP = {0: np.float32, 2: np.float16, 3: np.float64, 4: np.uint8, 5: np.uint16, 6: np.int16}


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _bf16_round(x64):
    """float64 -> bfloat16 bits, nearest even, one rounding (exact rational
    comparison through float64 halves)."""
    x = np.asarray(x64, np.float64)
    out = np.empty(x.shape, np.uint16)
    for i, v in enumerate(x.ravel()):
        f = np.float32(v)
        b = int(np.array(f).view(np.uint32))
        lo_b = b & 0xFFFF0000
        lo = float(np.array(lo_b, np.uint32).view(np.float32))
        if abs(lo) > abs(v):            # the float rounding carried past v
            lo_b -= 0x10000
            lo = float(np.array(lo_b, np.uint32).view(np.float32))
        hi = float(np.array(lo_b + 0x10000, np.uint32).view(np.float32))
        dl, dh = abs(v - lo), abs(hi - v)
        r = lo_b + 0x10000 if (dh < dl or (dh == dl and (lo_b >> 16) & 1)) else lo_b
        out.ravel()[i] = r >> 16
    return out


@pytest.mark.parametrize('st,dt', [(3, 0), (0, 3), (3, 2), (0, 2), (2, 0), (4, 0), (5, 3),
                                   (6, 0), (0, 0)])
def test_convert_matches_numpy(gpu, st, dt):
    from voxeldvc_amd import _lib as L
    rng = np.random.default_rng(st * 10 + dt)
    n = 100_003
    if NP[st] in (np.uint8, np.uint16, np.int16):
        info = np.iinfo(NP[st])
        a = rng.integers(info.min, info.max, n, endpoint=True).astype(NP[st])
    else:
        a = (rng.standard_normal(n) * 300).astype(NP[st])
        a[:8] = [0.0, -0.0, 1.0, -1.5, 65504.0, 1e-8, 3.0, 0.1]
    d = _dev(a)
    out = L.empty(n, np.dtype(NP[dt]).name)
    L.call('vdvc_convert', L.ctx(), n, d, st, out, dt)
    got = out.cpu().numpy()
    want = a.astype(NP[dt])
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


def test_convert_bf16(gpu):
    import torch
    from voxeldvc_amd import _lib as L
    rng = np.random.default_rng(5)
    # float64 values near bfloat16 rounding ties, where rounding through
    # float32 first would round twice
    base = (rng.standard_normal(2000) * 100).astype(np.float32).astype(np.float64)
    tie = base + np.ldexp(1.0, np.frexp(base)[1] - 9) * 0.5 * (1 + 1e-12)
    a = np.concatenate([base, tie, [0.0, 1.0, -2.5, 255.0]])
    out = L.empty(a.size, 'bfloat16')
    L.call('vdvc_convert', L.ctx(), a.size, _dev(a), 3, out, 1)
    got = out.view(torch.int16).cpu().numpy().view(np.uint16)
    assert np.array_equal(got, _bf16_round(a))
    # float32 -> bfloat16 equals torch's rounding
    f = (rng.standard_normal(50_000) * 1000).astype(np.float32)
    out = L.empty(f.size, 'bfloat16')
    L.call('vdvc_convert', L.ctx(), f.size, _dev(f), 0, out, 1)
    assert torch.equal(out, _dev(f).to(torch.bfloat16))
    # and back exactly
    back = L.empty(f.size, 'float32')
    L.call('vdvc_convert', L.ctx(), f.size, out, 1, back, 0)
    assert torch.equal(back, out.float())


@pytest.mark.parametrize('nbytes', [0, 1, 15, 16, 4099, 1 << 20])
def test_copy_and_fill(gpu, nbytes):
    import torch
    from voxeldvc_amd import _lib as L
    a = torch.randint(0, 255, (nbytes,), dtype=torch.uint8, device='cuda')
    b = torch.zeros(nbytes + 1, dtype=torch.uint8, device='cuda')
    L.call('vdvc_copy', L.ctx(), nbytes, a, b)
    assert torch.equal(b[:nbytes], a) and int(b[nbytes]) == 0
    x = L.empty(nbytes + 3, 'float32')
    L.call('vdvc_fill', L.ctx(), nbytes + 3, 0, -2.5, x)
    assert bool((x == -2.5).all())


@pytest.mark.parametrize('shape', [(1, 1, 1), (5, 3, 2), (65, 40, 33), (130, 70, 66),
                                   (513, 3, 65)])
@pytest.mark.parametrize('src_order', ['C', 'F', 'strided'])
def test_copy_volume_layouts(gpu, shape, src_order):
    import torch
    from voxeldvc_amd import _lib as L
    nx, ny, nz = shape
    rng = np.random.default_rng(nx * 7 + nz)
    a = rng.integers(0, 255, shape).astype(np.float32)          # [x][y][z]
    if src_order == 'C':
        src = _dev(a)
    elif src_order == 'F':
        src = _dev(np.ascontiguousarray(a.transpose(2, 1, 0))).permute(2, 1, 0)
    else:
        big = np.zeros((nx, ny + 2, nz * 2), np.float32)
        big[:, :ny, ::2] = a
        src = _dev(big)[:, :ny, ::2]
    want = np.ravel(a, order='F')
    for pitched, dt, dcode in ((False, 'float32', 0), (True, 'float16', 2),
                               (True, 'bfloat16', 1), (True, 'float64', 3)):
        d, shp = L.to_dev_volume(src, dt, pitched=pitched)
        assert shp == shape
        p = L.pitch_of(nx) if pitched else nx
        got = d.float().cpu().numpy().reshape(nz, ny, p)
        assert np.array_equal(got[:, :, :nx].ravel(), want), (pitched, dt)
        assert not got[:, :, nx:].any()
        # and back to the dense order
        if pitched:
            dense = L.unpitch(d, shape)
            assert np.array_equal(dense.float().cpu().numpy(), want)
    # host C-order input: uploaded as it is and transposed on the device
    d, _ = L.to_dev_volume(a, 'float32')
    assert np.array_equal(d.cpu().numpy(), want)


@pytest.mark.parametrize('dtype', [np.uint8, np.uint16, np.int16, np.float64, np.bool_])
def test_image_and_flat_uploads_of_strided_tensors(gpu, dtype):
    """Device tensors of the post-solve images in any strides, and strided
    1-D / 2-D tensors, are laid out by the library's volume copy: the
    reference's ravel(order='F') of the image in its own dtype, and C-order
    flattening of the flat inputs."""
    import torch
    from voxeldvc_amd import _lib as L
    rng = np.random.default_rng(5)
    a = rng.integers(0, 200, (9, 7, 5))
    a = (a > 100) if dtype == np.bool_ else a.astype(dtype)
    host = a.view(np.uint8) if dtype == np.bool_ else a
    for src in (torch.from_numpy(np.ascontiguousarray(a)).cuda(),
                torch.from_numpy(np.ascontiguousarray(a.transpose(2, 0, 1))).cuda()
                .permute(1, 2, 0)):
        assert tuple(src.shape) == a.shape
        d, shp, code = L.to_dev_image(src)
        assert shp == a.shape
        got = d.cpu().numpy()
        want = np.ravel(host, order='F')
        assert np.array_equal(got.view(want.dtype) if got.dtype != want.dtype else got, want)
    # strided flat inputs, flattened in C order
    m = torch.from_numpy(rng.random((6, 10)).astype(np.float32)).cuda()
    np.testing.assert_array_equal(L.to_dev_flat(m[:, ::3].t()).cpu().numpy(),
                                  m.cpu().numpy()[:, ::3].T.ravel())
    v = torch.from_numpy(np.arange(40, dtype=np.int16)).cuda()
    d, code = L.to_dev_image_flat(v[::3])
    assert np.array_equal(d.cpu().numpy(), np.arange(40, dtype=np.int16)[::3])


def test_axpby_widen(gpu):
    from voxeldvc_amd import _lib as L
    rng = np.random.default_rng(3)
    n = 10_001
    x = rng.standard_normal(n).astype(np.float32)
    y = rng.standard_normal(n) * 1e3
    yd = _dev(y)
    L.call('vdvc_axpby_x', L.ctx(), n, 1.0, _dev(x), 0, 1.0, yd, 3)
    assert np.array_equal(yd.cpu().numpy(), y + x.astype(np.float64))
    y32 = _dev(y.astype(np.float32))
    L.call('vdvc_axpby_x', L.ctx(), n, 0.5, _dev(x), 0, -1.0, y32, 0)
    want = (0.5 * x.astype(np.float64) - y.astype(np.float32).astype(np.float64)).astype(np.float32)
    assert np.array_equal(y32.cpu().numpy(), want)


def test_stream_read_probe(gpu):
    """The streaming-read probe stores nothing unless a lane's xor fold of
    its 16-byte words equals the key; then it stores that fold in its block's
    word (lane q of a grid of G blocks reads rows q, q + 256 G, ...)."""
    import torch
    from voxeldvc_amd import _lib as L
    grid = 4
    stride = grid * 256
    n16 = 5 * stride + 37                  # the unrolled loop and its tail
    words = np.random.default_rng(9).integers(0, 2 ** 32, size=(n16, 4), dtype=np.uint32)
    src = _dev(words.view(np.int32))
    fold = np.zeros(stride, np.uint32)
    for q in range(stride):
        r = words[q::stride]
        fold[q] = np.bitwise_xor.reduce(r.ravel())
    tid = 2 * 256 + 77                     # a lane of block 2
    assert np.sum(fold == fold[tid]) == 1
    out = torch.full((grid,), -1, dtype=torch.int32, device='cuda')
    L.call('vdvc_stream_read', L.ctx(), n16 * 16, src, grid, int(fold[tid]), out)
    got = out.cpu().numpy().view(np.uint32)
    assert got[2] == fold[tid]
    assert np.all(got[[0, 1, 3]] == 0xFFFFFFFF)
    # a key no lane folds to: nothing is written
    miss = int(next(k for k in range(1, 1000) if k not in set(fold.tolist())))
    out.fill_(-1)
    L.call('vdvc_stream_read', L.ctx(), n16 * 16, src, grid, miss, out)
    assert np.all(out.cpu().numpy() == -1)
    with pytest.raises(ValueError):
        L.call('vdvc_stream_read', L.ctx(), 24, src, grid, 0, out)
