"""Pins of the post-solve oracle (oracle/post_oracle.py) against SciPy, closed
forms and brute force.  CPU only."""
import numpy as np
import pytest
from scipy import ndimage

from oracle import dvc_oracle as O
from oracle import post_oracle as P


def _img(dtype, shape, seed):
    rng = np.random.default_rng(seed)
    if dtype == np.bool_:
        return rng.random(shape) < 0.4
    if np.dtype(dtype).kind in 'iu':
        info = np.iinfo(dtype)
        lo, hi = max(info.min, -300), min(info.max, 300 if info.max > 300 else info.max)
        return rng.integers(lo, hi + 1, size=shape).astype(dtype)
    return (rng.standard_normal(shape) * 40 + 100).astype(dtype)


def _coords(shape, n, seed):
    """Random coordinates with out-of-range values, exact ties (k + 0.5),
    exact integers and the box faces +-1e-9."""
    rng = np.random.default_rng(seed)
    cs = []
    for s in shape:
        c = rng.uniform(-1.5, s + 0.5, n)
        k = rng.integers(0, s, n)
        c[: n // 6] = k[: n // 6] + 0.5
        c[n // 6: n // 3] = k[n // 6: n // 3]
        c[n // 3: n // 3 + 4] = [0.0, s - 1.0, -1e-9, s - 1 + 1e-9]
        cs.append(c)
    return np.stack(cs)


@pytest.mark.parametrize('dtype', [np.float32, np.float64, np.uint8, np.uint16, np.int16,
                                   np.bool_])
@pytest.mark.parametrize('order', [0, 1])
@pytest.mark.parametrize('mode,cval', [('nearest', 0.0), ('constant', 1.0), ('constant', -3.0)])
def test_map_coordinates_01_matches_scipy_bitwise(dtype, order, mode, cval):
    if dtype == np.bool_ and order == 1:
        pytest.skip('bool fields are pushed with order 0 only')
    if np.dtype(dtype).kind == 'u' and cval < 0:
        pytest.skip('negative cval for unsigned output is not a case the stages produce')
    img = _img(dtype, (7, 5, 6), 3)
    c = _coords(img.shape, 3000, 4)
    want = ndimage.map_coordinates(img, c, order=order, mode=mode, cval=cval)
    got = P.map_coordinates_01(img, c, order=order, mode=mode, cval=cval)
    assert got.dtype == want.dtype
    np.testing.assert_array_equal(got, want)


def test_map_coordinates_01_float32_coords_like_the_stages():
    img = _img(np.float32, (9, 8, 7), 5)
    c = _coords(img.shape, 2000, 6).astype(np.float32)
    np.testing.assert_array_equal(P.map_coordinates_01(img, c),
                                  ndimage.map_coordinates(img, c, order=1, mode='nearest'))


def _affine_U(A, t, h, Nx, Ny, Nz, dtype):
    n = np.arange((Nx + 1) * (Ny + 1) * (Nz + 1))
    X = np.stack([n % (Nx + 1), (n // (Nx + 1)) % (Ny + 1), n // ((Nx + 1) * (Ny + 1))],
                 axis=1) * h
    u = X @ np.asarray(A).T + np.asarray(t)
    return u.reshape(-1).astype(dtype)


@pytest.mark.parametrize('h,N', [(4, (3, 2, 2)), (8, (2, 3, 1)), (3, (2, 2, 3))])
def test_interp_affine_field_exact(h, N):
    """Trilinear interpolation reproduces an affine nodal field exactly; with
    dyadic coefficients every operation is exact in float64."""
    Nx, Ny, Nz = N
    A = np.array([[0.25, -0.125, 0.5], [0.0, 0.375, -0.25], [0.125, 0.0, -0.5]])
    t = np.array([1.5, -2.25, 0.75])
    U = _affine_U(A, t, h, Nx, Ny, Nz, np.float64)
    u = P.interp_field_vox(U, h, Nx, Ny, Nz)
    Nvx, Nvy, Nvz = Nx * h + 1, Ny * h + 1, Nz * h + 1
    n = np.arange(Nvx * Nvy * Nvz)
    X = np.stack([n % Nvx, (n // Nvx) % Nvy, n // (Nvx * Nvy)], axis=1).astype(np.float64)
    want = X @ A.T + t
    if h == 3:
        np.testing.assert_allclose(u, want, rtol=0, atol=1e-12)
    else:
        np.testing.assert_array_equal(u, want)
    uc = P.interp_field_vox(U, h, Nx, Ny, Nz, channels_first=True)
    np.testing.assert_array_equal(uc, u.T)


def test_interp_nodal_voxels_equal_nodal_values():
    h, Nx, Ny, Nz = 4, 3, 2, 2
    U = np.random.default_rng(1).standard_normal(3 * (Nx + 1) * (Ny + 1) * (Nz + 1)).astype(
        np.float32)
    u = P.interp_field_vox(U, h, Nx, Ny, Nz)
    Nvx, Nvy = Nx * h + 1, Ny * h + 1
    for k in range(Nz + 1):
        for j in range(Ny + 1):
            for i in range(Nx + 1):
                node = i + (Nx + 1) * (j + (Ny + 1) * k)
                v = i * h + Nvx * (j * h + Nvy * k * h)
                np.testing.assert_array_equal(u[v], U[3 * node:3 * node + 3])


@pytest.mark.parametrize('mode,cval', [('nearest', 0.0), ('constant', 7.0)])
@pytest.mark.parametrize('dtype', [np.float32, np.uint8])
def test_push_integer_translation_closed_form(mode, cval, dtype):
    """U = integer translation t: pushed[x] = field[x - t] with edge clamp
    ('nearest') or cval outside ('constant'), for both orders."""
    h, Nx, Ny, Nz = 4, 3, 2, 2
    shp = (Nx * h + 1, Ny * h + 1, Nz * h + 1)
    field = _img(dtype, shp, 9)
    t = np.array([2.0, -1.0, 3.0])
    U = _affine_U(np.zeros((3, 3)), t, h, Nx, Ny, Nz, np.float32)
    ix = [np.arange(s)[:, None, None] if a == 0 else np.arange(s)[None, :, None] if a == 1
          else np.arange(s)[None, None, :] for a, s in enumerate(shp)]
    src = [ix[a] - int(t[a]) for a in range(3)]
    inside = np.ones(shp, dtype=bool)
    for a in range(3):
        inside &= (src[a] >= 0) & (src[a] <= shp[a] - 1)
    clamped = [np.clip(src[a], 0, shp[a] - 1) for a in range(3)]
    want = field[clamped[0], clamped[1], clamped[2]]
    if mode == 'constant':
        want = np.where(inside, want, np.asarray(cval).astype(dtype))
    for order in (0, 1):
        got = P.push_field_forward(field, U, Nx, Ny, Nz, h, order=order, mode=mode, cval=cval)
        np.testing.assert_array_equal(got, want)


def test_push_linear_ramp_inside_domain():
    """Order-1 interpolation of a linear field is exact where x - u stays in
    the domain: pushed(x) = a . (x - u(x)) + b."""
    h, Nx, Ny, Nz = 4, 3, 3, 2
    shp = (Nx * h + 1, Ny * h + 1, Nz * h + 1)
    a = np.array([0.5, -0.25, 1.0])
    b = 3.0
    X, Y, Z = np.meshgrid(*[np.arange(s) for s in shp], indexing='ij')
    field = a[0] * X + a[1] * Y + a[2] * Z + b
    A = np.array([[0.02, 0.01, 0.0], [0.0, -0.03, 0.01], [0.01, 0.0, 0.02]])
    U = _affine_U(A, [0.3, -0.2, 0.1], h, Nx, Ny, Nz, np.float64)
    got = P.push_field_forward(field, U, Nx, Ny, Nz, h)
    u = P.interp_field_vox(U, h, Nx, Ny, Nz)
    Nvox = np.prod(shp)
    n = np.arange(Nvox)
    P_ = np.stack([n % shp[0], (n // shp[0]) % shp[1], n // (shp[0] * shp[1])], axis=1) - u
    ok = np.all((P_ >= 0) & (P_ <= np.array(shp) - 1), axis=1)
    want = (P_ @ a + b).reshape(shp, order='F')
    okg = ok.reshape(shp, order='F')
    np.testing.assert_allclose(got[okg], want[okg], rtol=0, atol=1e-12)
    assert okg.mean() > 0.5


@pytest.mark.parametrize('origin', [(0, 0, 0), (2, -1, 1)])
def test_active_mask_float_image_equals_solver_sampler(origin):
    """For a float32 image the active mask is the GN loop's dynamic mask
    (dvc_oracle.sample_deformed_field, pinned by the reference's solves)."""
    h, Nx, Ny, Nz = 4, 3, 3, 2
    rng = np.random.default_rng(11)
    g = (rng.random((Nx * h + 3, Ny * h + 1, Nz * h + 2)) * 200).astype(np.float32)
    g[g < 30] = 0.0
    U = (rng.standard_normal(3 * (Nx + 1) * (Ny + 1) * (Nz + 1)) * 1.5).astype(np.float32)
    Nvox = (Nx * h + 1) * (Ny * h + 1) * (Nz * h + 1)
    mask0 = rng.random(Nvox) < 0.9
    want = O.sample_deformed_field(U, g, h, Nx, Ny, Nz, g_origin=origin, mask0_vox=mask0,
                                   glt=10.0)[1]
    got = P.compute_active_mask(U, g, h, Nx, Ny, Nz, g_origin=origin, mask0_vox=mask0, glt=10.0)
    np.testing.assert_array_equal(got, want)


def test_active_mask_integer_image_rounds_before_threshold():
    """A uint8 image is sampled in uint8 (rounded half away from zero) before
    the > glt test, as map_coordinates returns the input dtype."""
    h, Nx, Ny, Nz = 4, 2, 2, 2
    rng = np.random.default_rng(12)
    g = rng.integers(0, 3, size=(Nx * h + 1, Ny * h + 1, Nz * h + 1)).astype(np.uint8)
    U = (rng.standard_normal(3 * (Nx + 1) * (Ny + 1) * (Nz + 1)) * 0.7).astype(np.float32)
    got = P.compute_active_mask(U, g, h, Nx, Ny, Nz, glt=0.0)
    u = P.interp_field_vox(U, h, Nx, Ny, Nz)
    Nvx, Nvy, Nvz = g.shape
    n = np.arange(Nvx * Nvy * Nvz)
    c = np.stack([(n % Nvx).astype(np.float32) + u[:, 0],
                  ((n // Nvx) % Nvy).astype(np.float32) + u[:, 1],
                  (n // (Nvx * Nvy)).astype(np.float32) + u[:, 2]])
    inb = np.all((c >= 0) & (c <= np.array(g.shape)[:, None] - 1), axis=0)
    c = np.stack([np.clip(c[a], 0, g.shape[a] - 1) for a in range(3)])
    gv = ndimage.map_coordinates(g, c, order=1, mode='nearest')
    np.testing.assert_array_equal(got, inb & (gv > 0.0))
    # the float sample would differ somewhere (values in (0, 0.5) round to 0)
    gf = ndimage.map_coordinates(g.astype(np.float64), c, order=1, mode='nearest')
    assert np.any((gf > 0) != (gv > 0))


def test_zncc_invariants():
    rng = np.random.default_rng(13)
    f = rng.random(5000) * 255
    m = rng.random(5000) < 0.7
    g = 0.5 * f + 20.0
    assert P.compute_zncc(f, g, m) == pytest.approx(1.0, abs=1e-14)
    assert P.compute_zncc(f, -3.0 * f + 1.0, m) == pytest.approx(-1.0, abs=1e-14)
    g2 = f + rng.standard_normal(5000) * 30
    assert P.compute_zncc(f, g2, m) == pytest.approx(np.corrcoef(f[m], g2[m])[0, 1], rel=1e-12)
    assert np.isnan(P.compute_zncc(f, g2, np.zeros(5000, bool)))
    assert np.isnan(P.compute_zncc(np.full(5000, 3.0), g2, m))


def test_quality_cell_brute_force():
    h, Nx, Ny, Nz = 4, 3, 2, 2
    Nvx, Nvy, Nvz = Nx * h + 1, Ny * h + 1, Nz * h + 1
    rng = np.random.default_rng(14)
    ref = (rng.random((Nvx, Nvy, Nvz)) * 255).astype(np.float32)
    res = (rng.standard_normal(Nvx * Nvy * Nvz) * 5).astype(np.float32)
    res[rng.random(res.size) < 0.3] = 0.0
    # element (0,0,0): fewer than 8 active voxels; element (1,0,0): constant f
    res3 = res.reshape(Nvx, Nvy, Nvz, order='F')
    res3[0:h + 1, 0:h + 1, 0:h + 1] = 0.0
    res3[0, 0, 0] = 1.0
    res3[1, 1, 1] = 2.0
    ref[h:2 * h + 1, 0:h + 1, 0:h + 1] = 50.0
    res = res3.reshape(-1, order='F')
    std, zn = P.compute_correlation_quality_cell(res, ref, h, Nx, Ny, Nz)
    rf = ref.reshape(-1, order='F').astype(np.float64)
    for k in range(Nz):
        for j in range(Ny):
            for i in range(Nx):
                sl = (slice(i * h, i * h + h + 1), slice(j * h, j * h + h + 1),
                      slice(k * h, k * h + h + 1))
                r = res3[sl].astype(np.float64).ravel()
                f = ref[sl].astype(np.float64).ravel()
                a = r != 0
                if a.sum() < 8:
                    assert np.isnan(std[i, j, k]) and np.isnan(zn[i, j, k])
                    continue
                assert std[i, j, k] == pytest.approx(np.std(r[a]), rel=1e-9)
                fa, ga = f[a], f[a] - r[a]
                if np.all(fa == fa[0]):
                    assert np.isnan(zn[i, j, k])
                else:
                    assert zn[i, j, k] == pytest.approx(np.corrcoef(fa, ga)[0, 1], rel=1e-9)
    assert np.isnan(std[0, 0, 0]) and np.isnan(zn[1, 0, 0]) and not np.isnan(std[1, 0, 0])
    assert rf.size == res.size
