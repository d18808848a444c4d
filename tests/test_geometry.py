"""CPU pins of the product's geometry module (voxeldvc_amd/engine/geometry_dvc.py,
the constants the run stage and the solver feed the kernels): closed forms,
textbook element properties, brute-force indexing and agreement with the
oracle's independent restatement (reference engine/geometry_dvc.py:26-239)."""
import numpy as np
import pytest

from oracle import dvc_oracle as O
from voxeldvc_amd.engine import geometry_dvc as G

CORNERS = np.array(G.CORNER_OFFSETS)


def test_corner_order_is_the_reference_order():
    # engine/geometry_dvc.py:14-17: 0=(0,0,0) 1=(1,0,0) 2=(1,1,0) 3=(0,1,0),
    # then the same four at z = 1
    assert CORNERS.tolist() == [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                                [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]]
    assert np.array_equal(CORNERS, O.CORNERS)


@pytest.mark.parametrize('h', [1, 2, 3, 4, 8, 16])
def test_voxel_center_stencil(h):
    xg, yg, zg, wg = G.VoxelCenterStencil(h)
    n = (h + 1) ** 3
    assert xg.shape == yg.shape == zg.shape == wg.shape == (n,)
    xi = 2.0 * np.arange(h + 1) / h - 1.0
    # z fastest within the (h+1)^3 points (indexing 'ij' then ravel)
    q = np.arange(n)
    assert np.allclose(xg, xi[q // (h + 1) ** 2])
    assert np.allclose(yg, xi[(q // (h + 1)) % (h + 1)])
    assert np.allclose(zg, xi[q % (h + 1)])
    # trapezoid rule on [0, h]^3: total weight h^3, 1/8 at the 8 corners
    assert np.isclose(wg.sum(), h ** 3)
    corner = (np.abs(xg) == 1) & (np.abs(yg) == 1) & (np.abs(zg) == 1)
    assert np.allclose(wg[corner], 0.125) and corner.sum() == 8
    # the rule integrates trilinear polynomials exactly: int_0^h x y z = h^6 / 8
    X, Y, Z = (xg + 1) * h / 2, (yg + 1) * h / 2, (zg + 1) * h / 2
    assert np.isclose((wg * X * Y * Z).sum(), h ** 6 / 8.0)
    ref = O.voxel_center_stencil(h)
    for a, b in zip((xg, yg, zg, wg), ref):
        assert np.array_equal(a, b)


@pytest.mark.parametrize('h', [1, 2, 3, 4, 8])
def test_N_stencil(h):
    N, w = G.build_N_stencil(h)
    assert N.shape == ((h + 1) ** 3, 8) and w.shape == ((h + 1) ** 3,)
    assert np.allclose(N.sum(1), 1.0)                       # partition of unity
    xg, yg, zg, _ = G.VoxelCenterStencil(h)
    # isoparametric: sum_a N_a x_a reproduces the point (corners at +-1)
    for d, coord in enumerate((xg, yg, zg)):
        assert np.allclose(N @ (2.0 * CORNERS[:, d] - 1.0), coord)
    # Kronecker property at the element corners
    qi = lambda a, b, c: (a * (h + 1) + b) * (h + 1) + c
    for n, (a, b, c) in enumerate(CORNERS):
        row = N[qi(a * h, b * h, c * h)]
        assert np.isclose(row[n], 1.0) and np.isclose(np.abs(row).sum(), 1.0)
    Nr, wr = O.build_N_stencil(h)
    assert np.allclose(N, Nr, rtol=0, atol=1e-15) and np.array_equal(w, wr)


def test_laplacian_closed_form():
    K = G.build_K_ref_laplacian()
    L8 = K[0::3, 0::3]
    # unit-cube Q1 Laplacian: diagonal 1/3, edge neighbours 0, face and body
    # diagonals -1/12
    exp = np.zeros((8, 8))
    for a, pa in enumerate(CORNERS):
        for b, pb in enumerate(CORNERS):
            exp[a, b] = {0: 1 / 3, 1: 0.0, 2: -1 / 12, 3: -1 / 12}[int(np.abs(pa - pb).sum())]
    assert np.allclose(L8, exp, rtol=0, atol=1e-15)
    for c in range(3):
        for d in range(3):
            blk = K[c::3, d::3]
            assert np.allclose(blk, L8 if c == d else 0.0, rtol=0, atol=1e-15)
    assert np.allclose(K, K.T) and np.allclose(K.sum(1), 0.0)   # constants in the kernel
    assert np.allclose(K, O.build_K_ref_laplacian(), rtol=0, atol=1e-15)


@pytest.mark.parametrize('nu', [0.0, 0.25, 0.3, 0.45])
def test_elastic_stiffness(nu):
    K = G.build_K_ref(nu)
    assert K.shape == (24, 24) and np.allclose(K, K.T)
    ev = np.linalg.eigvalsh(K)
    # exactly the 6 rigid-body modes in the kernel, positive otherwise
    assert np.sum(np.abs(ev) < 1e-10) == 6 and ev.min() > -1e-10
    X = CORNERS.astype(float)
    c1 = (1 - nu) / ((1 + nu) * (1 - 2 * nu))
    c2 = nu / ((1 + nu) * (1 - 2 * nu))
    c3 = 1.0 / (2 * (1 + nu))
    # rigid rotation about z: u = (-y, x, 0) costs nothing
    u = np.zeros(24)
    u[0::3], u[1::3] = -X[:, 1], X[:, 0]
    assert abs(u @ K @ u) < 1e-12
    # uniaxial strain e_xx = 1 (unit volume): energy c1 / 2
    u = np.zeros(24)
    u[0::3] = X[:, 0]
    assert np.isclose(0.5 * u @ K @ u, 0.5 * c1)
    # equal biaxial strain e_xx = e_yy = 1: energy (2 c1 + 2 c2) / 2
    u[1::3] = X[:, 1]
    assert np.isclose(0.5 * u @ K @ u, c1 + c2)
    # engineering shear gamma_xy = 1 (u_x = y): energy c3 / 2
    u = np.zeros(24)
    u[0::3] = X[:, 1]
    assert np.isclose(0.5 * u @ K @ u, 0.5 * c3)
    assert np.allclose(K, O.build_K_ref(nu), rtol=0, atol=1e-14)


@pytest.mark.parametrize('shape', [(1, 1, 1), (3, 4, 2), (5, 2, 3)])
def test_node_index_and_element_nodes(shape):
    Nx, Ny, Nz = shape
    idx = G.build_node_index(Nx, Ny, Nz)
    assert idx.shape == (Nx + 1, Ny + 1, Nz + 1)
    for i in range(Nx + 1):
        for j in range(Ny + 1):
            for k in range(Nz + 1):
                assert idx[i, j, k] == i + j * (Nx + 1) + k * (Nx + 1) * (Ny + 1)
    conn = G.element_node_ids(idx, Nx, Ny, Nz)
    assert conn.shape == (Nx * Ny * Nz, 8)
    for k in range(Nz):
        for j in range(Ny):
            for i in range(Nx):
                e = i + j * Nx + k * Nx * Ny
                want = [idx[i + a, j + b, k + c] for (a, b, c) in CORNERS]
                assert conn[e].tolist() == want
    # the oracle's element DOF table names the same nodes
    dofs = O.elem_dofs(np.arange(Nx * Ny * Nz), Nx, Ny)
    assert np.array_equal(dofs[:, 0::3] // 3, conn)


@pytest.mark.parametrize('shape', [(1, 1, 1), (2, 2, 2), (3, 4, 2), (5, 3, 4)])
def test_interior_dof_mask(shape):
    Nx, Ny, Nz = shape
    m = G.build_interior_dof_mask(Nx, Ny, Nz)
    assert m.shape == (3 * (Nx + 1) * (Ny + 1) * (Nz + 1),) and m.dtype == bool
    n = np.arange((Nx + 1) * (Ny + 1) * (Nz + 1))
    i, j, k = n % (Nx + 1), (n // (Nx + 1)) % (Ny + 1), n // ((Nx + 1) * (Ny + 1))
    inner = (i > 0) & (i < Nx) & (j > 0) & (j < Ny) & (k > 0) & (k < Nz)
    assert np.array_equal(m, np.repeat(inner, 3))
    assert m.sum() == 3 * max(Nx - 1, 0) * max(Ny - 1, 0) * max(Nz - 1, 0)
    assert np.array_equal(m, O.build_interior_dof_mask(Nx, Ny, Nz))
