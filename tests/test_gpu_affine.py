"""GPU parity of the preprocess affine stage (libvdvc's vdvc_gaussian_filter,
vdvc_smooth_masked, vdvc_gradient3, vdvc_affine_normal, vdvc_affine_rms,
vdvc_warp_affine, vdvc_overlap_bounds) against oracle/affine_oracle.py:
smoothing, gradients and the warp bit for bit; the float64 normal-equation
and RMS sums within 1e-11 relative; the whole refinement within 1e-9 of the
oracle's A, t with the same iteration counts."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import affine_oracle as AO

pytestmark = pytest.mark.gpu


def _vol(seed, shape, lo=0.0, hi=255.0):
    return (np.random.default_rng(seed).random(shape) * (hi - lo) + lo).astype(np.float32)


def _dev(a):
    from voxeldvc_amd import _lib as L
    d, shp = L.to_dev_volume(a, 'float32')
    return d, shp


def _host(d, shp):
    from voxeldvc_amd import _lib as L
    return L.to_host(d).reshape(shp, order='F')


# (300, 70, 130): several 256-voxel row tiles along x and 64-position tiles
# along y and z, with partial last tiles; sigma 9 (radius 36) takes the
# per-voxel pass instead of the LDS tiles
@pytest.mark.parametrize('shape', [(40, 9, 7), (33, 20, 17), (300, 70, 130)])
@pytest.mark.parametrize('sigma', [4.0, 2.0, 1.0, 9.0])
@pytest.mark.parametrize('mode', ['nearest', 'constant'])
def test_gaussian_filter(gpu, shape, sigma, mode):
    from voxeldvc_amd.engine import correlate_gpu as C
    print('gaussian_filter', shape, sigma, mode, flush=True)
    x = _vol(1, shape)
    d, shp = _dev(x)
    got = _host(C._gauss_dev(d, shp, sigma, mode), shp)
    np.testing.assert_array_equal(got, ndimage.gaussian_filter(x, sigma, mode=mode, cval=0.0))


@pytest.mark.parametrize('cval', [2.5, 0.1])
def test_gaussian_filter_constant_cval(gpu, cval):
    """'constant' extension with a cval float32 holds exactly (2.5: the LDS
    tiles store it as float32) and one it does not (0.1: SciPy extends the
    float64 line with the float64 cval, so the per-voxel pass runs)."""
    from voxeldvc_amd import _lib as L
    from voxeldvc_amd.engine import correlate_gpu as C
    shape = (70, 66, 67)
    print('gaussian_filter cval', cval, flush=True)
    x = _vol(3, shape)
    d, shp = _dev(x)
    w, r = C.gaussian_weights(3.0)
    out = L.torch().empty_like(d)
    L.call('vdvc_gaussian_filter', L.ctx(), shp[0], shp[1], shp[2], d, out,
           w.ctypes.data_as(L._dp), r, 1, cval)
    np.testing.assert_array_equal(_host(out, shp),
                                  ndimage.gaussian_filter(x, 3.0, mode='constant', cval=cval))


@pytest.mark.parametrize('sigma', [0.0, 2.0, 4.0])
def test_smooth_masked_and_valid(gpu, sigma):
    from voxeldvc_amd import _lib as L
    from voxeldvc_amd.engine import correlate_gpu as C
    g = _vol(2, (30, 21, 14))
    g[:, :5] = 0.0
    g[g < 20] = 0.0
    d, shp = _dev(g)
    valid = L.empty(d.numel(), 'uint8')
    out = _host(C._smooth_masked_dev(d, shp, sigma, 0.0, valid), shp)
    want = AO.smooth_masked(g, (g != 0).astype(np.float32), sigma)
    np.testing.assert_array_equal(out, want)
    np.testing.assert_array_equal(_host(valid, shp).astype(bool), g != 0)


def test_gradient(gpu):
    from voxeldvc_amd import _lib as L
    x = _vol(3, (17, 12, 9))
    d, shp = _dev(x)
    gs = [L.empty(d.numel()) for _ in range(3)]
    L.call('vdvc_gradient3', L.ctx(), *shp, d, *gs)
    for got, want in zip(gs, np.gradient(x)):
        np.testing.assert_array_equal(_host(got, shp), want)


def _affine_case(n=33, seed=4):
    from voxeldvc_amd import synth
    f = synth.speckle_numpy((n, n, n), 800, seed)
    c = np.full(3, (n - 1) / 2.0)
    M = np.array([[0.015, -0.005, 0.002], [0.003, -0.01, 0.004], [-0.002, 0.006, 0.02]])
    tr = np.array([0.7, -0.3, 0.45])

    def u(X, Y, Z):
        q = [X - c[0], Y - c[1], Z - c[2]]
        return tuple(M[i, 0] * q[0] + M[i, 1] * q[1] + M[i, 2] * q[2] + tr[i] for i in range(3))
    g = synth.warp_backward(f, u, 0.5, seed + 1)
    g[:, :, :3] = 0.0        # a background fill the fit mask excludes
    return f.astype(np.float32), g.astype(np.float32), M, tr, c


def test_normal_equations_and_rms(gpu):
    from voxeldvc_amd import _lib as L
    from voxeldvc_amd.engine import correlate_gpu as C
    f, g, M, tr, c = _affine_case()
    f_s = ndimage.gaussian_filter(f, 2.0, mode='nearest')
    g_s = AO.smooth_masked(g, (g != 0).astype(np.float32), 2.0)
    gg = AO.gradient(f_s)
    valid = g != 0
    Mt = M * 0.5
    want_H, want_b = AO.normal_equations(f_s, g_s, gg, valid, Mt, tr * 0.5, c)
    d = {k: _dev(v)[0] for k, v in (('f', f_s), ('g', g_s), ('gx', gg[0]), ('gy', gg[1]),
                                     ('gz', gg[2]), ('f0', f), ('g0', g))}
    shp = f.shape
    vd = L.to_dev_flat(valid.reshape(-1, order='F'), 'uint8')
    H = np.zeros((12, 12))
    b = np.zeros(12)
    dp = lambda a: np.ascontiguousarray(a, np.float64).ctypes.data_as(L._dp)
    Mc, trc, cc = np.ascontiguousarray(Mt), np.ascontiguousarray(tr * 0.5), np.ascontiguousarray(c)
    L.call('vdvc_affine_normal', L.ctx(), *shp, d['f'], d['g'], d['gx'], d['gy'], d['gz'], vd,
           dp(Mc), dp(trc), dp(cc), 0.0, H.ctypes.data_as(L._dp), b.ctypes.data_as(L._dp))
    np.testing.assert_allclose(H, want_H, rtol=1e-11, atol=1e-11 * np.abs(want_H).max())
    np.testing.assert_allclose(b, want_b, rtol=1e-11, atol=1e-11 * np.abs(want_b).max())
    out = (L._d * 2)()
    L.call('vdvc_affine_rms', L.ctx(), *shp, d['f0'], d['g0'], vd, dp(Mc), dp(trc), dp(cc), 0.0,
           out)
    assert np.sqrt(out[0] / out[1]) == pytest.approx(
        AO.masked_rms(f, g, valid, Mt, tr * 0.5, c), rel=1e-11)
    del C


def test_affine_refine_matches_oracle(gpu):
    from voxeldvc_amd.engine import correlate_gpu as C
    f, g, M, tr, c = _affine_case()
    print('affine_refine_gn 33^3', flush=True)
    A_w, t_w, info_w = AO.affine_refine_gn(f, g, np.eye(3), np.zeros(3), c)
    A_g, t_g, info_g = C.affine_refine_gn(f, g, np, np.eye(3), np.zeros(3), c)
    print('  iters', info_g['iters'], info_w['iters'], 'rms', info_g['rms_final'],
          info_w['rms_final'], flush=True)
    assert info_g['iters'] == info_w['iters']
    np.testing.assert_allclose(A_g, A_w, rtol=0, atol=1e-9)
    np.testing.assert_allclose(t_g, t_w, rtol=0, atol=1e-8)
    assert info_g['rms_final'] == pytest.approx(info_w['rms_final'], rel=1e-9)
    A_true = np.linalg.inv(np.eye(3) - M)
    assert np.abs(A_g - A_true).max() < 2e-3


def test_warp_onto_ref_grid(gpu):
    from voxeldvc_amd import preprocess as P
    g = _vol(5, (26, 23, 19))
    Aff = np.array([[1.02, 0.01, -0.02], [0.0, 0.97, 0.015], [0.01, -0.01, 1.01]])
    t = np.array([1.3, -0.7, 0.4])
    c = np.array([12.5, 11.0, 9.0])
    Nf = (28, 20, 17)
    got = P.warp_onto_ref_grid(Aff, t, c, Nf, g, np)
    np.testing.assert_array_equal(got, AO.warp_onto_ref_grid(Aff, t, c, Nf, g))


@pytest.mark.parametrize('fdtype', [np.float32, np.uint8, np.uint16])
def test_compute_crop(gpu, fdtype):
    from voxeldvc_amd import preprocess as P
    rng = np.random.default_rng(6)
    f = np.zeros((40, 35, 30), fdtype)
    g = np.zeros((40, 35, 30), np.float32)
    f[3:37, 2:33, 5:28] = (rng.random((34, 31, 23)) * 100 + 1).astype(fdtype)
    g[6:39, 0:30, 1:25] = rng.random((33, 30, 24)) * 100 + 1
    for h in (4, 8):
        lo, hi = P.compute_crop(f, g, h, glt=0.0)
        wlo, whi = AO.compute_crop(f, g, h, glt=0.0)
        np.testing.assert_array_equal(lo, wlo)
        np.testing.assert_array_equal(hi, whi)
    with pytest.raises(RuntimeError):
        P.compute_crop(f, np.zeros_like(g), 8)


def test_affine_refine_full_size(gpu):
    """257^3 speckle (GPU generator) pulled by a known affine: the refinement
    recovers it (closed form: A = (I - M)^-1, t = A t_r)."""
    import time
    import torch
    from voxeldvc_amd import synth
    from voxeldvc_amd.engine import correlate_gpu as C
    n = 257
    f = synth.speckle_torch((n, n, n), synth.blobs_for((n, n, n)), seed=7)
    c = np.full(3, (n - 1) / 2.0)
    M = np.array([[0.02, -0.004, 0.003], [0.002, -0.012, 0.005], [-0.003, 0.004, 0.03]])
    tr = np.array([1.2, -0.8, 0.6])
    Aff = np.eye(3) + M
    g = synth.warp_backward_torch(f, (n, n, n), None, 0.5, seed=8, quantize=False,
                                  affine=(Aff, c), extra_u=lambda X, Y, Z: (tr[0] + 0 * X,
                                                                           tr[1] + 0 * Y,
                                                                           tr[2] + 0 * Z))
    fv = f.view(n, n, n).permute(2, 1, 0)
    gv = g.view(n, n, n).permute(2, 1, 0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    A_g, t_g, info = C.affine_refine_gn(fv, gv, torch, np.eye(3), np.zeros(3), c)
    ms = (time.perf_counter() - t0) * 1e3
    A_true = np.linalg.inv(np.eye(3) - M)
    print('  257^3 refinement %.0f ms, iters %s, rms %.3f -> %.3f, |A - A_true| %.2e, '
          '|t - t_true| %.2e' % (ms, info['iters'], info['rms_initial'], info['rms_final'],
                                 np.abs(A_g - A_true).max(), np.abs(t_g - A_true @ tr).max()),
          flush=True)
    assert np.abs(A_g - A_true).max() < 1e-3
    assert np.abs(t_g - A_true @ tr).max() < 0.02
