"""Host-side geometry constants of the structured Hex8 mesh: the voxel-centre
trapezoid stencil, Hex8 shape functions at its points, the unit-cube vector
Laplacian and elastic stiffness, and implicit node / DOF indexing.

Same functions, arguments and return conventions as the reference's
src/voxeldvc/engine/geometry_dvc.py (cited per function).  These are tiny
host arrays (at most 24x24 or (h+1)^3 x 8) consumed by the device kernels.
"""
import numpy as np

# Hex8 local corner order (geometry_dvc.py:14-17)
CORNER_OFFSETS = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                  (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))


def VoxelCenterStencil(h):
    """(xg, yg, zg, wg): the (h+1)^3 voxel-centre points xi = 2k/h - 1 and
    their trapezoid weights (1/2 per axis at k in {0, h}); geometry_dvc.py:26-51."""
    xi = 2.0 * np.arange(h + 1) / h - 1.0
    w1 = np.where((np.arange(h + 1) == 0) | (np.arange(h + 1) == h), 0.5, 1.0)
    g = np.meshgrid(xi, xi, xi, indexing='ij')
    wg = np.einsum('i,j,k->ijk', w1, w1, w1)
    return g[0].ravel(), g[1].ravel(), g[2].ravel(), wg.ravel()


def _hex8_N(x, y, z):
    """Hex8 shape functions at natural coordinates in [-1,1]^3, (npts, 8);
    geometry_dvc.py:155-168."""
    out = np.empty(np.shape(x) + (8,))
    for a, (i, j, k) in enumerate(CORNER_OFFSETS):
        out[..., a] = 0.125 * (1 + (2 * i - 1) * x) * (1 + (2 * j - 1) * y) * (1 + (2 * k - 1) * z)
    return out


def build_N_stencil(h):
    """(N_stencil (nq, 8), w_stencil (nq,)); geometry_dvc.py:171-184."""
    xg, yg, zg, wg = VoxelCenterStencil(h)
    return _hex8_N(xg, yg, zg), wg


def _unit_cube_grads(p):
    """(3, 8) gradients of the unit-cube trilinear basis at p in [0,1]^3."""
    xi, eta, zeta = p
    d = np.empty((3, 8))
    for a, (i, j, k) in enumerate(CORNER_OFFSETS):
        fx = (xi if i else 1.0 - xi)
        fy = (eta if j else 1.0 - eta)
        fz = (zeta if k else 1.0 - zeta)
        d[0, a] = (1.0 if i else -1.0) * fy * fz
        d[1, a] = fx * (1.0 if j else -1.0) * fz
        d[2, a] = fx * fy * (1.0 if k else -1.0)
    return d


def _gauss_points_unit():
    g = 0.5 / np.sqrt(3.0)
    pts = (0.5 - g, 0.5 + g)
    return [(a, b, c) for a in pts for b in pts for c in pts]


def build_K_ref_laplacian():
    """24x24 unit-cube vector Laplacian, 2x2x2 Gauss (weight 1/8 per point),
    entry [3a+c, 3b+c] = L8[a, b]; geometry_dvc.py:88-110."""
    L8 = sum(0.125 * (lambda d: d.T @ d)(_unit_cube_grads(p)) for p in _gauss_points_unit())
    K = np.zeros((24, 24))
    for c in range(3):
        K[c::3, c::3] = L8
    return K


def build_K_ref(nu):
    """24x24 unit-cube isotropic elastic stiffness (E = 1), 2x2x2 Gauss,
    node-major DOFs, engineering shear rows (xy, yz, zx); geometry_dvc.py:58-148."""
    c1 = (1.0 - nu) / ((1.0 + nu) * (1.0 - 2.0 * nu))
    c2 = nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    c3 = 1.0 / (2.0 * (1.0 + nu))
    D = np.diag([c1, c1, c1, c3, c3, c3])
    D[0, 1] = D[0, 2] = D[1, 0] = D[1, 2] = D[2, 0] = D[2, 1] = c2
    K = np.zeros((24, 24))
    for p in _gauss_points_unit():
        d = _unit_cube_grads(p)
        B = np.zeros((6, 24))
        B[0, 0::3] = d[0]
        B[1, 1::3] = d[1]
        B[2, 2::3] = d[2]
        B[3, 0::3] = d[1]
        B[3, 1::3] = d[0]
        B[4, 1::3] = d[2]
        B[4, 2::3] = d[1]
        B[5, 0::3] = d[2]
        B[5, 2::3] = d[0]
        K += 0.125 * B.T @ D @ B
    return K


def build_node_index(Nx, Ny, Nz):
    """(Nx+1, Ny+1, Nz+1) node ids i + j(Nx+1) + k(Nx+1)(Ny+1);
    geometry_dvc.py:191-198."""
    i, j, k = np.meshgrid(np.arange(Nx + 1), np.arange(Ny + 1), np.arange(Nz + 1), indexing='ij')
    return i + (Nx + 1) * (j + (Ny + 1) * k)


def build_interior_dof_mask(Nx, Ny, Nz):
    """DOF mask of nodes strictly inside the bounding box;
    geometry_dvc.py:201-217."""
    idx = build_node_index(Nx, Ny, Nz)
    inner = np.zeros(idx.shape, dtype=bool)
    inner[1:Nx, 1:Ny, 1:Nz] = True
    m = np.zeros(3 * idx.size, dtype=bool)
    for c in range(3):
        m[3 * idx.ravel() + c] = inner.ravel()
    return m


def element_node_ids(node_idx, Nx, Ny, Nz):
    """(Nx*Ny*Nz, 8) corner node ids per element, element id i + j Nx + k Nx Ny;
    geometry_dvc.py:220-239."""
    cols = [node_idx[i:i + Nx, j:j + Ny, k:k + Nz] for (i, j, k) in CORNER_OFFSETS]
    return np.stack(cols, axis=-1).transpose(2, 1, 0, 3).reshape(-1, 8)
