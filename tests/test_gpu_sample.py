"""GPU parity of the deformed-image sampler (vdvc_sample) across image types,
masks, row counts and strip splits, and of the plane-wave calibration mode
(vdvc_calibration_mode)."""
import numpy as np
import pytest

from oracle import dvc_oracle as O

pytestmark = pytest.mark.gpu

# (Nx, Ny, Nz, h): several 64-voxel strips, odd and even rows per element,
# one-voxel elements, the last element row's extra row; Nx h = 64 k (65 and
# 129 voxel columns), where the last column is folded into the last strip
SHAPES = [(10, 3, 2, 8), (5, 3, 2, 3), (70, 1, 2, 1), (4, 2, 3, 5), (2, 2, 2, 16),
          (8, 3, 2, 8), (32, 1, 2, 4)]


def _run_sample(U, g, gcode, shp_g, Nx, Ny, Nz, h, go, m0, bm, glt, f32w=False):
    from voxeldvc_amd import _lib as L
    from voxeldvc_amd.engine import kernels_dvc as K
    shp = K.level_shape(h, Nx, Ny, Nz)
    Np = L.pitch_of(shp[0]) * shp[1] * shp[2]
    dU = L.to_dev_flat(U, 'float32')
    d0 = None if m0 is None else L.to_dev_level(m0, shp, 'uint8')
    db = None if bm is None else L.to_dev_level(bm, shp, 'uint8')
    gv = L.zeros(Np, 'float32')
    mk = L.zeros(Np, 'uint8')
    s = (L._d * 3)()
    L.call('vdvc_sample', L.ctx(), Nx, Ny, Nz, h, L.ptr(dU), L.ptr(g), gcode, shp_g[0], shp_g[1],
           shp_g[2], float(go[0]), float(go[1]), float(go[2]), float(glt), L.ptr(d0), L.ptr(db),
           L.ptr(gv), L.ptr(mk), s, int(f32w))
    return (L.to_host(L.unpitch(gv, shp)), L.to_host(L.unpitch(mk, shp)).astype(bool),
            np.array(list(s)))


def _problem(seed, Nx, Ny, Nz, h, scale=2.0):
    rng = np.random.default_rng(seed)
    gimg = rng.integers(0, 256, size=(Nx * h + 3, Ny * h + 2, Nz * h + 1)).astype(np.float32)
    U = rng.normal(scale=scale, size=3 * (Nx + 1) * (Ny + 1) * (Nz + 1)).astype(np.float32)
    nv = (Nx * h + 1) * (Ny * h + 1) * (Nz * h + 1)
    m0 = rng.random(nv) < 0.85
    bm = rng.random(nv) < 0.6
    return gimg, U, m0, bm


def _stats_np(gv, bm, Nx, Ny, Nz, h):
    """(sum mlt, sum mlt*g, sum mlt*g^2) over the b-mask, mlt = product of
    the per-axis voxel multiplicities (2 on an interior element face)."""
    def mult(n_el):
        t = np.arange(n_el * h + 1)
        return np.where((t > 0) & (t < n_el * h) & (t % h == 0), 2.0, 1.0)
    mx, my, mz = mult(Nx), mult(Ny), mult(Nz)
    mlt = (mz[:, None, None] * my[None, :, None] * mx[None, None, :]).ravel()
    g = gv.astype(np.float64)
    w = mlt * bm
    return np.array([w.sum(), (w * g).sum(), (w * g * g).sum()])


@pytest.mark.parametrize('shape', SHAPES)
def test_sample_half_and_float_images_agree(gpu, shape):
    """Integer intensities are exact in float16: sampling the float16 and the
    float32 copy of one image gives identical values, masks and statistics."""
    from voxeldvc_amd import _lib as L
    Nx, Ny, Nz, h = shape
    print('sample f16/f32', shape, flush=True)
    gimg, U, m0, bm = _problem(21, Nx, Ny, Nz, h)
    g32, gs = L.to_dev_volume(gimg, 'float32')
    g16, _ = L.to_dev_volume(gimg, 'float16')
    for masks in [(m0, bm), (None, None), (m0, None)]:
        a = _run_sample(U, g32, L.F32, gs, Nx, Ny, Nz, h, (1, -1, 0), *masks, glt=20.0)
        b = _run_sample(U, g16, L.F16, gs, Nx, Ny, Nz, h, (1, -1, 0), *masks, glt=20.0)
        assert np.array_equal(a[0], b[0])
        assert np.array_equal(a[1], b[1])
        assert np.array_equal(a[2], b[2])


@pytest.mark.parametrize('gnx', [1, 2, 3])
def test_sample_narrow_images(gpu, gnx):
    """Images one, two and three voxels wide in x: a float32 image takes both
    x-neighbours of a tap from one 8-byte load clamped to the row (a
    one-voxel row reads its high word past the image), a float16 image loads
    each tap alone.  Values, masks and statistics are identical, and match
    the oracle."""
    from voxeldvc_amd import _lib as L
    Nx, Ny, Nz, h = 3, 2, 2, 4
    print('sample narrow image', gnx, flush=True)
    rng = np.random.default_rng(40 + gnx)
    gimg = rng.integers(0, 256, size=(gnx, Ny * h + 2, Nz * h + 1)).astype(np.float32)
    U = rng.normal(scale=1.5, size=3 * (Nx + 1) * (Ny + 1) * (Nz + 1)).astype(np.float32)
    g32, gs = L.to_dev_volume(gimg, 'float32')
    g16, _ = L.to_dev_volume(gimg, 'float16')
    go = (-0.25, -1, 0)
    for f32w in (False, True):
        a = _run_sample(U, g32, L.F32, gs, Nx, Ny, Nz, h, go, None, None, -1.0, f32w)
        b = _run_sample(U, g16, L.F16, gs, Nx, Ny, Nz, h, go, None, None, -1.0, f32w)
        assert all(np.array_equal(x, y) for x, y in zip(a, b))
    # clamped samples carry values outside the box too; the mask is the
    # in-bounds test there
    a = _run_sample(U, g32, L.F32, gs, Nx, Ny, Nz, h, go, None, None, -1.0)
    rv, rm = O.sample_deformed_field(U, gimg, h, Nx, Ny, Nz, go, None, glt=-1.0)
    assert (a[1] != rm).mean() < 2e-3
    assert np.abs(a[0] - rv).max() < 2e-3


@pytest.mark.parametrize('shape', SHAPES)
def test_sample_matches_oracle(gpu, shape):
    from voxeldvc_amd import _lib as L
    Nx, Ny, Nz, h = shape
    print('sample oracle', shape, flush=True)
    gimg, U, m0, bm = _problem(22, Nx, Ny, Nz, h)
    g32, gs = L.to_dev_volume(gimg, 'float32')
    go = (2, 0, -1)
    gv, mk, st = _run_sample(U, g32, L.F32, gs, Nx, Ny, Nz, h, go, m0, bm, glt=30.0)
    rv, rm = O.sample_deformed_field(U, gimg, h, Nx, Ny, Nz, go, m0, glt=30.0)
    # u at a voxel is summed in a different float32 order; allow a few
    # in-bounds / threshold flips at the boundary of the box
    assert (mk != rm).mean() < 2e-3
    ok = mk == rm
    assert np.abs(gv[ok] - rv[ok]).max() < 2e-3
    # statistics over the given b-mask, from the kernel's own values
    ref = _stats_np(gv, bm, Nx, Ny, Nz, h)
    assert np.allclose(st, ref, rtol=1e-12, atol=0)
    # without a b-mask the statistics run over the sampled mask
    _, mk2, st2 = _run_sample(U, g32, L.F32, gs, Nx, Ny, Nz, h, go, m0, None, glt=30.0)
    assert np.array_equal(mk2, mk)
    assert np.allclose(st2, _stats_np(gv, mk, Nx, Ny, Nz, h), rtol=1e-12, atol=0)


@pytest.mark.parametrize('shape', SHAPES)
def test_sample_f32_weights(gpu, shape):
    """The mixed-precision form's float32 interpolation weights against the
    float64 (SciPy-order) weights on the same coordinates: values within
    float32 rounding of the 0..255 range, the mask equal wherever the
    value is not within that rounding of the threshold, statistics of the
    kernel's own values; float16 and float32 images identical."""
    from voxeldvc_amd import _lib as L
    Nx, Ny, Nz, h = shape
    print('sample f32 weights', shape, flush=True)
    gimg, U, m0, bm = _problem(24, Nx, Ny, Nz, h)
    g32, gs = L.to_dev_volume(gimg, 'float32')
    g16, _ = L.to_dev_volume(gimg, 'float16')
    go = (1, 0, -1)
    a = _run_sample(U, g32, L.F32, gs, Nx, Ny, Nz, h, go, m0, bm, 30.0)
    b = _run_sample(U, g32, L.F32, gs, Nx, Ny, Nz, h, go, m0, bm, 30.0, f32w=True)
    c = _run_sample(U, g16, L.F16, gs, Nx, Ny, Nz, h, go, m0, bm, 30.0, f32w=True)
    assert np.abs(a[0] - b[0]).max() <= 256 * 2.0 ** -22
    near = np.abs(a[0] - 30.0) <= 256 * 2.0 ** -22
    assert np.array_equal(a[1][~near], b[1][~near])
    assert np.allclose(b[2], _stats_np(b[0], bm, Nx, Ny, Nz, h), rtol=1e-12, atol=0)
    assert all(np.array_equal(x, y) for x, y in zip(b, c))


def test_sample_nan_displacement(gpu):
    """A NaN nodal displacement gives a false mask at every voxel whose
    displacement is interpolated from that node (the kernel's interpolation
    carries the NaN even where the node's weight is zero); every other voxel
    matches the oracle."""
    from voxeldvc_amd import _lib as L
    Nx, Ny, Nz, h = 9, 2, 2, 8
    gimg, U, m0, bm = _problem(23, Nx, Ny, Nz, h)
    U = U.copy()
    U[3 * 5] = np.nan                 # x component of node (5, 0, 0)
    g32, gs = L.to_dev_volume(gimg, 'float32')
    gv, mk, _ = _run_sample(U, g32, L.F32, gs, Nx, Ny, Nz, h, (0, 0, 0), None, None, glt=0.0)
    vol = lambda a: a.reshape(Nz * h + 1, Ny * h + 1, Nx * h + 1)
    # voxels interpolated from node (5, 0, 0): x in [4h, 6h), y and z in [0, h)
    box = np.zeros((Nz * h + 1, Ny * h + 1, Nx * h + 1), bool)
    box[:h, :h, 4 * h:6 * h] = True
    assert not vol(mk)[box].any()
    Uo = U.copy()
    Uo[3 * 5] = 0.0
    rv, rm = O.sample_deformed_field(Uo, gimg, h, Nx, Ny, Nz, (0, 0, 0), None, glt=0.0)
    out = ~box.ravel()
    assert (mk[out] == rm[out]).mean() > 0.998
    ok = out & mk & rm
    assert np.abs(gv[ok] - rv[ok]).max() < 2e-3


@pytest.mark.parametrize('dtype', ['float32', 'float64'])
@pytest.mark.parametrize('shape', [(4, 6, 3, 8), (1, 1, 1, 2), (16, 33, 2, 4)])
def test_calibration_mode(gpu, dtype, shape):
    """cos(j*h/T*2*pi) on the x components (T = 10h), zero elsewhere, one
    rounded operation at a time in the working type."""
    from voxeldvc_amd import _lib as L
    Nx, Ny, Nz, h = shape
    nn = (Nx + 1) * (Ny + 1) * (Nz + 1)
    out = L.empty(3 * nn, dtype)
    L.call('vdvc_calibration_mode', L.ctx(), Nx, Ny, Nz, h, L.rtype_of(out), L.ptr(out))
    got = L.to_host(out).reshape(-1, 3)
    dt = np.dtype(dtype).type
    jj = (np.arange(nn) // (Nx + 1)) % (Ny + 1)
    v = jj.astype(dt) * dt(h)
    v = v / dt(10 * h)
    v = v * dt(2)
    v = v * dt(np.pi)
    assert v.dtype == np.dtype(dtype)
    ref = np.cos(v)
    tol = 2e-7 if dtype == 'float32' else 1e-15
    assert np.abs(got[:, 0] - ref).max() <= tol
    assert not got[:, 1:].any()
