"""Pins of the preprocess affine-stage oracle (oracle/affine_oracle.py):
SciPy's gaussian_filter and map_coordinates, np.gradient, the reference's
own affine_refine_gn result (golden), a known affine, brute force.  CPU."""
import hashlib
import os

import numpy as np
import pytest
from scipy import ndimage

from conftest import GOLDEN, golden
from oracle import affine_oracle as A


@pytest.mark.parametrize('shape', [(40, 7, 5), (9, 33, 12)])
@pytest.mark.parametrize('sigma', [4.0, 2.0, 1.0, 0.7])
@pytest.mark.parametrize('mode', ['nearest', 'constant'])
def test_gaussian_filter_matches_scipy_bitwise(shape, sigma, mode):
    x = (np.random.default_rng(1).random(shape) * 255).astype(np.float32)
    np.testing.assert_array_equal(A.gaussian_filter(x, sigma, mode=mode),
                                  ndimage.gaussian_filter(x, sigma, mode=mode, cval=0.0))


def test_gradient_matches_numpy():
    x = (np.random.default_rng(2).random((6, 5, 4)) * 255).astype(np.float32)
    for got, want in zip(A.gradient(x), np.gradient(x)):
        assert got.dtype == np.float32
        np.testing.assert_array_equal(got, want)


def test_smooth_masked_keeps_background_out():
    """Normalised convolution: a constant material region next to a zero
    fill is not darkened by the fill (it extends into the fill as far as
    the kernel reaches, radius 8 at sigma 2), and the fill beyond the
    kernel's reach keeps its value."""
    g = np.zeros((30, 6, 6), np.float32)
    g[:10] = 100.0
    v = (g != 0).astype(np.float32)
    s = A.smooth_masked(g, v, 2.0)
    np.testing.assert_allclose(s[:18], 100.0, rtol=1e-6)
    np.testing.assert_array_equal(s[19:], 0.0)


def _affine49():
    d = golden('affine_refine49')
    shas = dict(ln.split() for ln in open(os.path.join(GOLDEN, 'INPUT_SHA256.txt'))
                .read().splitlines())
    return d, shas


def test_affine_refine_matches_reference_golden():
    from voxeldvc_amd import synth
    d, shas = _affine49()
    n = 49
    f = synth.speckle_numpy((n, n, n), 1500, 31)
    c = d['c']
    M, tr = d['M_true'], d['tr_true']

    def u(X, Y, Z):
        q = [X - c[0], Y - c[1], Z - c[2]]
        return tuple(M[i, 0] * q[0] + M[i, 1] * q[1] + M[i, 2] * q[2] + tr[i] for i in range(3))
    g = synth.warp_backward(f, u, 0.5, 32)
    f, g = f.astype(np.float32), g.astype(np.float32)
    assert hashlib.sha256(f.tobytes()).hexdigest() == shas['affine49_f']
    assert hashlib.sha256(g.tobytes()).hexdigest() == shas['affine49_g']
    A_, t_, info = A.affine_refine_gn(f, g, np.eye(3), np.zeros(3), c)
    np.testing.assert_allclose(A_, d['A'], rtol=0, atol=1e-9)
    np.testing.assert_allclose(t_, d['t'], rtol=0, atol=1e-8)
    assert info['iters'] == list(d['iters'])
    assert info['rms_final'] == pytest.approx(float(d['rms_final']), rel=1e-9)
    assert info['rms_initial'] == pytest.approx(float(d['rms_initial']), rel=1e-9)
    # the known affine: A = (I - M)^-1, t = A t_r
    A_true = np.linalg.inv(np.eye(3) - M)
    assert np.abs(A_ - A_true).max() < 2e-3
    assert np.abs(t_ - A_true @ tr).max() < 0.05


def test_warp_onto_ref_grid_matches_scipy():
    rng = np.random.default_rng(3)
    g = (rng.random((20, 18, 16)) * 255).astype(np.float32)
    Aff = np.array([[1.02, 0.01, -0.02], [0.0, 0.97, 0.015], [0.01, -0.01, 1.01]])
    t = np.array([1.3, -0.7, 0.4])
    c = np.array([9.5, 8.5, 7.5])
    Nf = (22, 17, 15)
    got = A.warp_onto_ref_grid(Aff, t, c, Nf, g)
    X = np.stack(np.meshgrid(*[np.arange(n, dtype=np.float64) for n in Nf], indexing='ij'))
    pts = np.einsum('ij,jxyz->ixyz', Aff, X - c[:, None, None, None]) + (c + t)[:, None, None, None]
    want = ndimage.map_coordinates(g, pts, order=1, mode='constant', cval=0.0)
    assert got.shape == Nf and got.dtype == np.float32
    np.testing.assert_allclose(got, want, rtol=0, atol=0.05)
    assert np.count_nonzero(got == 0) == pytest.approx(np.count_nonzero(want == 0), abs=40)


def test_compute_crop_brute_force():
    rng = np.random.default_rng(4)
    f = np.zeros((40, 35, 30), np.float32)
    g = np.zeros_like(f)
    f[3:37, 2:33, 5:28] = rng.random((34, 31, 23)) * 100 + 1
    g[6:39, 0:30, 1:25] = rng.random((33, 30, 24)) * 100 + 1
    h = 8
    lo, hi = A.compute_crop(f, g, h)
    valid = (f > 0) & (g > 0)
    ix = np.argwhere(valid)
    blo, bhi = ix.min(0), ix.max(0)
    for a in range(3):
        size = hi[a] - lo[a]
        assert (size - 1) % h == 0 and size <= bhi[a] - blo[a] + 1 < size + h
        assert 0 <= lo[a] and hi[a] <= f.shape[a]
        assert abs((lo[a] + hi[a] - 1) / 2 - (blo[a] + bhi[a]) / 2) <= 1.0
    with pytest.raises(RuntimeError):
        A.compute_crop(f, np.zeros_like(g), h)
