"""GPU parity of the post-solve voxel stages (libvdvc's vdvc_interp_field,
vdvc_active_mask, vdvc_push_forward, vdvc_zncc, vdvc_quality_cell) against
oracle/post_oracle.py on seeded inputs.  The interpolation, the active mask
and the pushes follow the oracle's operation order and are compared bit for
bit; the FP64 reductions within 1e-12 (ZNCC) and 1e-10 (per element)
relative, with identical NaN patterns."""
import numpy as np
import pytest

from oracle import post_oracle as P

pytestmark = pytest.mark.gpu

# h = 64 is the largest spacing whose per-plane weight factors come from the
# launch's table, h = 66 forms them in the kernel
SHAPES = [(2, 2, 2, 8), (3, 2, 4, 4), (1, 1, 1, 1), (5, 3, 2, 3), (17, 2, 2, 4), (2, 1, 3, 16),
          (70, 1, 2, 1), (1, 1, 1, 64), (1, 1, 1, 66)]


def _U(seed, Nx, Ny, Nz, scale, dtype):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(3 * (Nx + 1) * (Ny + 1) * (Nz + 1)) * scale).astype(dtype)


def _image(seed, shape, dtype):
    rng = np.random.default_rng(seed)
    if dtype == np.bool_:
        return rng.random(shape) < 0.5
    if np.dtype(dtype).kind in 'iu':
        lo = -300 if np.dtype(dtype).kind == 'i' else 0
        hi = 255 if dtype == np.uint8 else 300
        return rng.integers(lo, hi + 1, size=shape).astype(dtype)
    img = rng.random(shape) * 255
    img[img < 25] = 0.0
    return img.astype(dtype)


@pytest.mark.parametrize('dtype', [np.float32, np.float64])
@pytest.mark.parametrize('shape', SHAPES)
def test_interp_field_vox(gpu, shape, dtype):
    from voxeldvc_amd.engine import correlate_gpu as C
    Nx, Ny, Nz, h = shape
    print('interp_field_vox', shape, np.dtype(dtype).name, flush=True)
    U = _U(1, Nx, Ny, Nz, 2.0, dtype)
    want = P.interp_field_vox(U, h, Nx, Ny, Nz)
    got = C.interp_field_vox(U, C.build_N_stencil(h)[0], h, Nx, Ny, Nz, np)
    assert got.dtype == dtype and got.shape == want.shape
    np.testing.assert_array_equal(got, want)
    gotc = C.interp_field_vox(U, None, h, Nx, Ny, Nz, np, channels_first=True)
    np.testing.assert_array_equal(gotc, want.T)


def test_interp_rejects_other_stencils(gpu):
    from voxeldvc_amd.engine import correlate_gpu as C
    U = _U(1, 2, 2, 2, 1.0, np.float32)
    with pytest.raises(ValueError):
        C.interp_field_vox(U, C.build_N_stencil(8)[0] * 2, 8, 2, 2, 2, np)
    with pytest.raises(ValueError):
        C.interp_field_vox(U[:-3], None, 8, 2, 2, 2, np)


@pytest.mark.parametrize('udtype', [np.float32, np.float64])
@pytest.mark.parametrize('gdtype', [np.float32, np.float64, np.uint8, np.uint16, np.int16,
                                    np.bool_])
@pytest.mark.parametrize('shape', [(3, 2, 2, 4), (2, 3, 1, 8), (5, 3, 2, 3), (1, 1, 1, 66)])
def test_compute_active_mask(gpu, shape, gdtype, udtype):
    from voxeldvc_amd.engine import correlate_gpu as C
    Nx, Ny, Nz, h = shape
    print('compute_active_mask', shape, np.dtype(gdtype).name, np.dtype(udtype).name, flush=True)
    Nvx, Nvy, Nvz = Nx * h + 1, Ny * h + 1, Nz * h + 1
    g = _image(2, (Nvx + 3, max(2, Nvy - 1), Nvz + 2), gdtype)
    U = _U(3, Nx, Ny, Nz, 1.5, udtype)
    rng = np.random.default_rng(4)
    mask0 = rng.random(Nvx * Nvy * Nvz) < 0.9
    glt = 0.5 if gdtype == np.bool_ else 10.0
    for origin, m0 in (((2, -1, 1), mask0), ((0, 0, 0), None)):
        want = P.compute_active_mask(U, g, h, Nx, Ny, Nz, g_origin=origin, mask0_vox=m0, glt=glt)
        got = C.compute_active_mask(U, g, Nx, Ny, Nz, h, np, g_origin=origin, mask0_vox=m0,
                                    glt=glt)
        assert got.dtype == np.bool_
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize('order,mode,cval', [(1, 'nearest', 0.0), (0, 'nearest', 0.0),
                                             (0, 'constant', 1.0), (1, 'constant', -5.0)])
@pytest.mark.parametrize('fdtype', [np.float32, np.float64, np.uint8, np.uint16, np.int16,
                                    np.bool_])
def test_push_field_forward(gpu, fdtype, order, mode, cval):
    from voxeldvc_amd.engine import write_output as W
    if fdtype == np.bool_ and order == 1:
        pytest.skip('bool fields are pushed with order 0 (write_output.py:891-894)')
    if np.dtype(fdtype).kind == 'u' and cval < 0:
        pytest.skip('negative cval for an unsigned field is not a case the output stage uses')
    for shape, udt in (((3, 2, 2, 4), np.float32), ((2, 3, 2, 8), np.float64),
                       ((5, 2, 3, 3), np.float32)):
        Nx, Ny, Nz, h = shape
        print('push_field_forward', shape, np.dtype(fdtype).name, order, mode, flush=True)
        field = _image(5, (Nx * h + 1, Ny * h + 1, Nz * h + 1), fdtype)
        U = _U(6, Nx, Ny, Nz, 3.0, udt)
        want = P.push_field_forward(field, U, Nx, Ny, Nz, h, order=order, mode=mode, cval=cval)
        got = W.push_field_forward(field, U, Nx, Ny, Nz, h, order=order, mode=mode, cval=cval)
        assert got.dtype == field.dtype and got.shape == field.shape
        np.testing.assert_array_equal(got, want)


def test_push_reference_forward_device_arrays(gpu):
    """torch CUDA tensors in, xp=torch: the result stays on the device as
    an (Nvx, Nvy, Nvz) [x][y][z] array."""
    import torch
    from voxeldvc_amd.engine import write_output as W
    Nx, Ny, Nz, h = 3, 2, 2, 4
    ref = _image(7, (Nx * h + 1, Ny * h + 1, Nz * h + 1), np.float32)
    U = _U(8, Nx, Ny, Nz, 1.0, np.float32)
    want = P.push_field_forward(ref, U, Nx, Ny, Nz, h)
    got = W.push_reference_forward(torch.from_numpy(ref).cuda(), torch.from_numpy(U).cuda(),
                                   Nx, Ny, Nz, h, xp=torch)
    assert got.is_cuda and tuple(got.shape) == ref.shape
    np.testing.assert_array_equal(got.cpu().numpy(), want)


def test_zncc(gpu):
    from voxeldvc_amd.engine import write_output as W
    rng = np.random.default_rng(9)
    n = 3_000_001
    f = rng.integers(0, 256, n).astype(np.uint8)
    res = (rng.standard_normal(n) * 7).astype(np.float32)
    m = rng.random(n) < 0.8
    g = f.astype(np.float64) - res
    want = P.compute_zncc(f, g, m)
    assert W.compute_zncc(f.astype(np.float64), g, m) == pytest.approx(want, rel=1e-12)
    assert W.zncc_of_residual(f, res, m) == pytest.approx(want, rel=1e-12)
    assert W.compute_zncc(f, g, None) == pytest.approx(P.compute_zncc(f, g, np.ones(n, bool)),
                                                       rel=1e-12)
    assert np.isnan(W.compute_zncc(f, g, np.zeros(n, bool)))
    assert np.isnan(W.compute_zncc(np.full(n, 4.0), g, m))
    assert W.compute_zncc(f, 0.5 * f.astype(np.float64) + 3, m) == pytest.approx(1.0, abs=1e-13)


@pytest.mark.parametrize('fdtype', [np.float32, np.uint8])
@pytest.mark.parametrize('rdtype', [np.float32, np.float64])
@pytest.mark.parametrize('shape', [(3, 2, 2, 4), (2, 2, 3, 8), (2, 1, 1, 16), (4, 3, 2, 1),
                                   (70, 2, 1, 2)])
def test_quality_cell(gpu, shape, rdtype, fdtype):
    from voxeldvc_amd.engine import write_output as W
    Nx, Ny, Nz, h = shape
    print('quality_cell', shape, np.dtype(rdtype).name, np.dtype(fdtype).name, flush=True)
    Nvx, Nvy, Nvz = Nx * h + 1, Ny * h + 1, Nz * h + 1
    rng = np.random.default_rng(10)
    ref = _image(11, (Nvx, Nvy, Nvz), fdtype)
    res = (rng.standard_normal((Nvx, Nvy, Nvz)) * 6).astype(rdtype)
    res[rng.random(res.shape) < 0.3] = 0
    if h >= 4:
        res[:h + 1, :h + 1, :h + 1] = 0        # element (0,0,0): too few active voxels
        res[0, 0, 0] = 1
        if Nx > 1:
            ref[h:2 * h + 1, :h + 1, :h + 1] = 50   # element (1,0,0): constant f
    res_flat = res.reshape(-1, order='F')
    ws, wz = P.compute_correlation_quality_cell(res_flat, ref, h, Nx, Ny, Nz)
    gs, gz = W.compute_correlation_quality_cell(res_flat, ref, h, Nx, Ny, Nz)
    assert gs.shape == (Nx, Ny, Nz) and gs.dtype == np.float64
    for got, want in ((gs, ws), (gz, wz)):
        np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
        ok = ~np.isnan(want)
        np.testing.assert_allclose(got[ok], want[ok], rtol=1e-10, atol=1e-12)
    if h >= 4:
        assert np.isnan(gs[0, 0, 0])
    # a (Nvx, Nvy, Nvz) residual grid is accepted as the reference accepts it
    gs2, _ = W.compute_correlation_quality_cell(res, ref, h, Nx, Ny, Nz)
    np.testing.assert_array_equal(np.isnan(gs2), np.isnan(gs))
    np.testing.assert_array_equal(gs2[~np.isnan(gs2)], gs[~np.isnan(gs)])


def test_post_stages_full_size_properties(gpu):
    """257^3 voxels (32^3 elements, h = 8): properties that hold at any size.
    An affine nodal field with dyadic coefficients interpolates exactly; a
    zero displacement pushes every field onto itself and an integer
    translation shifts it; the active mask of g = f at U = 0 is mask0 & f >
    glt; ZNCC of f against a*f + b is 1."""
    from voxeldvc_amd.engine import correlate_gpu as C
    from voxeldvc_amd.engine import write_output as W
    Nx = Ny = Nz = 32
    h = 8
    n = Nx * h + 1
    print('post stages 257^3', flush=True)
    k = np.arange(Nx + 1, dtype=np.float64) * h
    kx, ky, kz = np.meshgrid(k, k, k, indexing='ij')
    A = np.array([[0.25, -0.125, 0.0], [0.125, 0.0, 0.5], [-0.25, 0.125, 0.25]])
    t = np.array([1.5, -0.75, 2.0])
    X = np.stack([kx.ravel(order='F'), ky.ravel(order='F'), kz.ravel(order='F')], axis=1)
    U = (X @ A.T + t).reshape(-1).astype(np.float32)
    u = C.interp_field_vox(U, None, h, Nx, Ny, Nz, np, channels_first=True)
    vox = np.arange(n, dtype=np.float64)
    for c in range(3):
        uc = u[c].reshape(n, n, n, order='F')
        want = (A[c, 0] * vox[:, None, None] + A[c, 1] * vox[None, :, None]
                + A[c, 2] * vox[None, None, :] + t[c])
        np.testing.assert_array_equal(uc, want.astype(np.float32))
    rng = np.random.default_rng(12)
    f = rng.integers(0, 256, size=(n, n, n)).astype(np.uint8)
    Z = np.zeros(3 * (Nx + 1) ** 3, np.float32)
    np.testing.assert_array_equal(W.push_reference_forward(f, Z, Nx, Ny, Nz, h), f)
    T = np.tile(np.array([2.0, -3.0, 1.0], np.float32), (Nx + 1) ** 3)
    sh = W.push_field_forward(f, T, Nx, Ny, Nz, h, order=0, mode='constant', cval=7)
    assert np.array_equal(sh[2:, :n - 3, 1:], f[:n - 2, 3:, :n - 1])
    assert np.all(sh[:2] == 7) and np.all(sh[:, n - 3:] == 7) and np.all(sh[:, :, :1] == 7)
    m0 = rng.random(n ** 3) < 0.95
    act = C.compute_active_mask(Z, f, Nx, Ny, Nz, h, np, mask0_vox=m0, glt=20.0)
    np.testing.assert_array_equal(act, m0 & (f.reshape(-1, order='F') > 20))
    ff = f.reshape(-1, order='F')
    assert W.compute_zncc(ff, 3.0 * ff.astype(np.float64) - 8, m0) == pytest.approx(1.0, abs=1e-12)
