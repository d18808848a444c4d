"""Pins of the element-scale output oracle (oracle/cells_oracle.py) with
known answers: the boundary shell alone, one inactive voxel inside an
element and on a corner shared by eight, the overlap-loss threshold and the
material floor, linear displacement fields (which Hex8 reproduces exactly,
so the strains are the eigenvalues of the imposed symmetric gradient),
scalar-multiple convergence ratios, and the identity and one-element-shift
nearest remaps; the SciPy-based remap against an exhaustive search."""
import numpy as np
import pytest

from oracle import cells_oracle as CO


def shell_set(Nx, Ny, Nz):
    return {(i, j, k) for k in range(Nz) for j in range(Ny) for i in range(Nx)
            if i in (0, Nx - 1) or j in (0, Ny - 1) or k in (0, Nz - 1)}


def elem_set(mask, Nx, Ny, Nz):
    m = np.asarray(mask).reshape(-1)
    return {(e % Nx, (e // Nx) % Ny, e // (Nx * Ny)) for e in np.nonzero(m)[0]}


def box(lo, hi):
    return {(i, j, k) for i in range(lo[0], hi[0] + 1) for j in range(lo[1], hi[1] + 1)
            for k in range(lo[2], hi[2] + 1)}


def vid(x, y, z, Nx, Ny, h):
    return x + (Nx * h + 1) * (y + (Ny * h + 1) * z)


@pytest.mark.parametrize('dims,h', [((6, 5, 7), 2), ((7, 7, 6), 3)])
def test_all_active_is_shell_only(dims, h):
    Nx, Ny, Nz = dims
    nv = (Nx * h + 1) * (Ny * h + 1) * (Nz * h + 1)
    act = np.ones(nv, bool)
    ue, un = CO.unsafe_elements(act, h, Nx, Ny, Nz)
    assert elem_set(ue, Nx, Ny, Nz) == shell_set(Nx, Ny, Nz)
    ol = CO.overlap_lost_element_mask(act, np.ones(nv, bool), h, Nx, Ny, Nz)
    assert elem_set(ol, Nx, Ny, Nz) == shell_set(Nx, Ny, Nz)
    # shell nodes: nodes within one element of the outer surface
    nn = (Nx + 1, Ny + 1, Nz + 1)
    idx = np.arange(un.size)
    ni, nj, nk = idx % nn[0], (idx // nn[0]) % nn[1], idx // (nn[0] * nn[1])
    near = ((ni <= 1) | (ni >= Nx - 1) | (nj <= 1) | (nj >= Ny - 1) | (nk <= 1) | (nk >= Nz - 1))
    np.testing.assert_array_equal(un, near)
    # unsafe voxels: exactly the voxels of shell elements, i.e. within h of
    # the surface; safe cells: elements at distance >= 2 from the surface
    uv = CO.unsafe_voxel_mask(act, h, Nx, Ny, Nz).reshape(Nx * h + 1, Ny * h + 1, Nz * h + 1,
                                                          order='F')
    x = np.arange(Nx * h + 1)[:, None, None]
    y = np.arange(Ny * h + 1)[None, :, None]
    z = np.arange(Nz * h + 1)[None, None, :]
    want = ~((x > h) & (x < (Nx - 1) * h) & (y > h) & (y < (Ny - 1) * h) & (z > h) &
             (z < (Nz - 1) * h))
    np.testing.assert_array_equal(uv, want)
    safe = CO.safe_cell_mask(uv, h, Nx, Ny, Nz)
    assert elem_set(safe.reshape(-1, order='F'), Nx, Ny, Nz) == box((2, 2, 2), (Nx - 3, Ny - 3, Nz - 3))


def test_inactive_voxel_inside_element_dilates_27():
    Nx, Ny, Nz, h = 9, 8, 9, 4
    nv = (Nx * h + 1) * (Ny * h + 1) * (Nz * h + 1)
    act = np.ones(nv, bool)
    act[vid(4 * h + 1, 3 * h + 2, 4 * h + 3, Nx, Ny, h)] = False   # inside element (4, 3, 4)
    ue, _ = CO.unsafe_elements(act, h, Nx, Ny, Nz)
    assert elem_set(ue, Nx, Ny, Nz) == shell_set(Nx, Ny, Nz) | box((3, 2, 3), (5, 4, 5))


def test_inactive_corner_voxel_dilates_64():
    Nx, Ny, Nz, h = 10, 10, 10, 3
    nv = (Nx * h + 1) * (Ny * h + 1) * (Nz * h + 1)
    act = np.ones(nv, bool)
    act[vid(5 * h, 5 * h, 5 * h, Nx, Ny, h)] = False    # node (5, 5, 5): shared by 8 elements
    ue, _ = CO.unsafe_elements(act, h, Nx, Ny, Nz)
    assert elem_set(ue, Nx, Ny, Nz) == shell_set(Nx, Ny, Nz) | box((3, 3, 3), (6, 6, 6))


def test_overlap_loss_threshold_and_material_floor():
    Nx, Ny, Nz, h = 8, 8, 8, 2
    nv = (Nx * h + 1) * (Ny * h + 1) * (Nz * h + 1)
    act = np.ones(nv, bool)
    mat = np.ones(nv, bool)
    act[vid(4 * h + 1, 4 * h + 1, 4 * h + 1, Nx, Ny, h)] = False     # 1 of 27 lost in (4,4,4)
    mat[vid(3 * h + 1, 3 * h + 1, 3 * h + 1, Nx, Ny, h)] = False     # a pore: not a loss
    ol = CO.overlap_lost_element_mask(act, mat, h, Nx, Ny, Nz)
    assert elem_set(ol, Nx, Ny, Nz) == shell_set(Nx, Ny, Nz) | box((3, 3, 3), (5, 5, 5))
    ol2 = CO.overlap_lost_element_mask(act, mat, h, Nx, Ny, Nz, frac_threshold=1.0 / 27)
    assert elem_set(ol2, Nx, Ny, Nz) == shell_set(Nx, Ny, Nz)
    # material floor: element (2, 5, 2) keeps 13 of 27 material voxels (< 0.5), not dilated
    x0, y0, z0 = 2 * h, 5 * h, 2 * h
    cnt = 0
    for dz in range(h + 1):
        for dy in range(h + 1):
            for dx in range(h + 1):
                if cnt < 14:
                    mat[vid(x0 + dx, y0 + dy, z0 + dz, Nx, Ny, h)] = False
                    cnt += 1
    ol3 = CO.overlap_lost_element_mask(act, mat, h, Nx, Ny, Nz, frac_threshold=1.0 / 27)
    assert (2, 5, 2) in elem_set(ol3, Nx, Ny, Nz)
    assert (3, 5, 2) not in elem_set(ol3, Nx, Ny, Nz)


def rot(a, b, c):
    ca, sa, cb, sb, cc, sc = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(c), np.sin(c)
    Rz = np.array([[ca, -sa, 0], [sa, ca, 0], [0, 0, 1]])
    Ry = np.array([[cb, 0, sb], [0, 1, 0], [-sb, 0, cb]])
    Rx = np.array([[1, 0, 0], [0, cc, -sc], [0, sc, cc]])
    return Rz @ Ry @ Rx


def linear_grid(G, t, Nx, Ny, Nz, h):
    x = np.arange(Nx + 1)[:, None, None] * h
    y = np.arange(Ny + 1)[None, :, None] * h
    z = np.arange(Nz + 1)[None, None, :] * h
    X = np.stack(np.broadcast_arrays(x, y, z), axis=-1).astype(np.float64)
    return X @ np.asarray(G).T + np.asarray(t)


def test_strains_of_linear_field():
    R = rot(0.3, -0.7, 1.1)
    lam = np.array([0.012, -0.004, 0.0065])
    S = R @ np.diag(lam) @ R.T
    W = np.array([[0, 0.003, -0.001], [-0.003, 0, 0.002], [0.001, -0.002, 0]])  # rotation part
    U = linear_grid(S + W, [1.5, -2.0, 0.25], 5, 4, 6, 8)
    p = CO.principal_strains_cell(U, 8)
    want = np.sort(lam)[::-1]
    assert np.abs(p - want).max() < 1e-15
    e = CO.equivalent_strain(p)
    e1, e2, e3 = want
    assert np.abs(e - np.sqrt(2) / 3 * np.sqrt((e1 - e2) ** 2 + (e2 - e3) ** 2 + (e3 - e1) ** 2)).max() < 1e-16
    assert abs(CO.equivalent_strain(np.array([1.0, 0.0, -1.0])) - 2 / np.sqrt(3)) < 1e-15


def test_gn_convergence_scalar_multiple_and_nan():
    Nx, Ny, Nz = 4, 3, 5
    rng = np.random.default_rng(3)
    U = rng.standard_normal(3 * (Nx + 1) * (Ny + 1) * (Nz + 1))
    r = CO.gn_convergence_cell(-0.25 * U, U, Nx, Ny, Nz)
    assert np.allclose(r, 0.25, rtol=1e-15, atol=0)
    U0 = np.zeros_like(U)
    r0 = CO.gn_convergence_cell(U, U0, Nx, Ny, Nz)
    assert np.isnan(r0).all()


def test_remap_identity_shift_and_brute_force():
    Nx, Ny, Nz, h = 6, 5, 4, 4
    rng = np.random.default_rng(5)
    m = rng.random((Nx, Ny, Nz)) < 0.4
    Z = np.zeros((Nx + 1, Ny + 1, Nz + 1, 3))
    np.testing.assert_array_equal(CO.remap_mask_nearest(m, Z, h), m)
    Zs = Z.copy()
    Zs[..., 0] = h     # every centre moves one element along +x
    want = np.concatenate([m[:1], m[:-1]], axis=0)
    np.testing.assert_array_equal(CO.remap_mask_nearest(m, Zs, h), want)
    np.testing.assert_array_equal(CO.remap_mask_nearest_brute(m, Zs, h), want)
    U = rng.normal(0, 1.3, (Nx + 1, Ny + 1, Nz + 1, 3))
    np.testing.assert_array_equal(CO.remap_mask_nearest(m, U, h, (1, -2, 3)),
                                  CO.remap_mask_nearest_brute(m, U, h, (1, -2, 3)))


# ------------------------------------------ deformed-configuration field remap
def _smooth_U(Nx, Ny, Nz, h, amp=0.6, k=17.0):
    n = np.stack(np.meshgrid(np.arange(Nx + 1), np.arange(Ny + 1), np.arange(Nz + 1),
                             indexing='ij'), -1).astype(float) * h
    return amp * np.sin(n / k) + 0.3


def _second_diff_max(F):
    """Largest |second difference| of a lattice field along and across axes."""
    m = 0.0
    for a in range(3):
        m = max(m, np.abs(np.diff(F, 2, axis=a)).max())
        for b in range(a + 1, 3):
            m = max(m, np.abs(np.diff(np.diff(F, axis=a), axis=b)).max())
    return m


@pytest.mark.parametrize('crop', [(0, 0, 0), (3, -2, 5)])
def test_remap_field_linear_fields_are_exact(crop):
    """Any piecewise-linear interpolant reproduces a field that is affine in
    the moved positions: griddata (the reference) and the Kuhn split agree
    with it, and with each other, inside the moved lattice."""
    Nx, Ny, Nz, h = 6, 5, 4, 8
    U = _smooth_U(Nx, Ny, Nz, h)
    x_ref, x_def = CO.element_centres_deformed(U, h)
    rng = np.random.default_rng(1)
    A, b = rng.normal(size=(3, 3)), rng.normal(size=3)
    F = x_def @ A.T + b
    q = x_ref + np.asarray(crop, float)
    want = q @ A.T + b
    g = CO.remap_field_linear(F, U, h, crop)
    k = CO.remap_field_kuhn_brute(F, U, h, crop)
    # queries at least one cell inside the moved lattice
    inner = np.zeros((Nx, Ny, Nz), bool)
    inner[1:-1, 1:-1, 1:-1] = True
    inner &= np.all((q > x_def[0, 0, 0] + h) & (q < x_def[-1, -1, -1] - h), axis=-1)
    assert inner.sum() > 0
    assert np.abs(g - want)[inner].max() < 1e-11
    assert np.abs(k - want)[inner].max() < 1e-11


def test_remap_field_identity_and_nearest_fill():
    Nx, Ny, Nz, h = 5, 4, 6, 4
    rng = np.random.default_rng(2)
    F = rng.normal(size=(Nx, Ny, Nz, 3))
    Z = np.zeros((Nx + 1, Ny + 1, Nz + 1, 3))
    # no displacement: every query is a lattice point
    assert np.array_equal(CO.remap_field_kuhn_brute(F, Z, h), F)
    assert np.allclose(CO.remap_field_linear(F, Z, h), F, atol=1e-12)
    # queries shifted far outside the lattice take the nearest value
    far = (100.0 * h, 0.0, 0.0)
    k = CO.remap_field_kuhn_brute(F[..., 0], Z, h, far)
    g = CO.remap_field_linear(F[..., 0], Z, h, far)
    assert np.array_equal(k, g) and np.array_equal(k, np.broadcast_to(F[-1:, :, :, 0], k.shape))


def test_remap_field_smooth_agrees_with_griddata():
    """On a smooth field the two triangulations differ by at most a fraction
    of the field's second differences; a wrong vertex or weight would be
    off by first differences."""
    Nx, Ny, Nz, h = 7, 6, 5, 8
    U = _smooth_U(Nx, Ny, Nz, h)
    _, x_def = CO.element_centres_deformed(U, h)
    F = np.sin(x_def[..., 0] / 20) * np.cos(x_def[..., 1] / 15) + x_def[..., 2] / 50
    g = CO.remap_field_linear(F, U, h)
    k = CO.remap_field_kuhn_brute(F, U, h)
    inner = np.zeros(F.shape, bool)
    inner[1:-1, 1:-1, 1:-1] = True
    d = np.abs(g - k)[inner].max()
    assert d <= 0.5 * _second_diff_max(F)
    assert d < 0.1 * np.abs(np.diff(F, axis=0)).max()


@pytest.mark.parametrize('dtype', [np.uint8, np.uint16, np.float32])
def test_material_cell_mask(dtype):
    Nx, Ny, Nz, h = 4, 3, 5, 4
    rng = np.random.default_rng(3)
    shape = (Nx * h + 1, Ny * h + 1, Nz * h + 1)
    ref = rng.integers(0, 200, shape).astype(dtype)
    ref[: 2 * h] = 0
    glt = 90.0
    m = CO.material_cell_mask(ref, h, Nx, Ny, Nz, glt)
    for i in range(Nx):
        for j in range(Ny):
            for k in range(Nz):
                blk = ref[i * h:i * h + h + 1, j * h:j * h + h + 1, k * h:k * h + h + 1]
                assert m[i, j, k] == (blk.astype(np.float64).mean() > glt)
    assert not m[:2].any() and m.any()


def test_material_cell_mask_eroded():
    rng = np.random.default_rng(4)
    m = rng.random((6, 5, 7)) < 0.8
    e = CO.material_cell_mask_eroded(m)
    for i in range(6):
        for j in range(5):
            for k in range(7):
                nb = m[max(i - 1, 0):i + 2, max(j - 1, 0):j + 2, max(k - 1, 0):k + 2]
                assert e[i, j, k] == nb.all()
