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
