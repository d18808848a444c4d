"""Element-scale write-output fields restated in NumPy / SciPy.

Test infrastructure only: this module checks the HIP path and is imported by
tests, smoke() and the benchmark's CPU baseline, never by the package.

It follows the reference algorithm literally -- per-element voxel index
arrays over the (h+1)^3 voxels [i*h, i*h + h] of each axis, node dilation
through the 8 Hex8 corner nodes -- where the GPU path uses per-element
counts and 27-neighbourhood tests, so the two formulations check each other:

  * elem_voxels              engine/kernels_dvc.py:64-89
  * boundary_elements        engine/kernels_dvc.py:92-115
  * unsafe_elements          engine/kernels_dvc.py:118-163 (_unsafe_elements)
  * unsafe_voxel_mask        engine/kernels_dvc.py:166-233 (and the element /
                             node forms)
  * overlap_lost_element_mask engine/kernels_dvc.py:236-340
  * safe_cell_mask           engine/write_output.py:352-377
  * principal_strains_cell   engine/write_output.py:256-281
  * equivalent_strain        engine/write_output.py:284-306
  * gn_convergence_cell      engine/write_output.py:450-491
  * remap_mask_nearest       engine/write_output.py:494-530
  * remap_field_linear       engine/write_output.py:533-568
  * material_cell_mask       engine/write_output.py:309-338
  * material_cell_mask_eroded engine/write_output.py:341-349

Element order everywhere is x fastest (e = i + Nx*(j + Ny*k)); node order
n = i + (Nx+1)*(j + (Ny+1)*k); voxel order x fastest on the
(Nx*h+1, Ny*h+1, Nz*h+1) grid.
"""
import numpy as np

# Hex8 local corners (geometry_dvc.py element_node_ids order)
CORNERS = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
           (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))


def _elem_ijk(Nx, Ny, Nz):
    e = np.arange(Nx * Ny * Nz)
    return e % Nx, (e // Nx) % Ny, e // (Nx * Ny)


def elem_voxels(h, Nx, Ny, Nz):
    """(Ne, (h+1)^3) flat voxel indices of each element's voxel block."""
    Nvx, Nvy = Nx * h + 1, Ny * h + 1
    q = np.arange(h + 1)
    qz, qy, qx = np.meshgrid(q, q, q, indexing='ij')
    off = (qx + Nvx * (qy + Nvy * qz)).ravel()
    i, j, k = _elem_ijk(Nx, Ny, Nz)
    origin = i * h + Nvx * (j * h + Nvy * (k * h))
    return origin[:, None] + off[None, :]


def elem_nodes(Nx, Ny, Nz):
    """(Ne, 8) node ids of each element in Hex8 corner order."""
    i, j, k = _elem_ijk(Nx, Ny, Nz)
    cols = [(i + di) + (Nx + 1) * ((j + dj) + (Ny + 1) * (k + dk)) for di, dj, dk in CORNERS]
    return np.stack(cols, axis=1)


def boundary_elements(Nx, Ny, Nz):
    """(Ne,) True on the outer one-element shell."""
    i, j, k = _elem_ijk(Nx, Ny, Nz)
    return ((i == 0) | (i == Nx - 1) | (j == 0) | (j == Ny - 1) | (k == 0) | (k == Nz - 1))


def unsafe_elements(active, h, Nx, Ny, Nz):
    """(elem_unsafe (Ne,), unsafe_node (Nn,)): nodes of elements holding an
    inactive voxel are unsafe; an element touching an unsafe node is unsafe;
    the boundary shell is OR-ed in after the dilation (its nodes too)."""
    active = np.asarray(active, dtype=bool).ravel()
    ev = elem_voxels(h, Nx, Ny, Nz)
    en = elem_nodes(Nx, Ny, Nz)
    has_inactive = (~active[ev]).any(axis=1)
    unsafe_node = np.zeros((Nx + 1) * (Ny + 1) * (Nz + 1), dtype=bool)
    unsafe_node[en[has_inactive].ravel()] = True
    elem_unsafe = unsafe_node[en].any(axis=1)
    bnd = boundary_elements(Nx, Ny, Nz)
    elem_unsafe = elem_unsafe | bnd
    unsafe_node[en[bnd].ravel()] = True
    return elem_unsafe, unsafe_node


def unsafe_voxel_mask(active, h, Nx, Ny, Nz):
    """(Nvox,) True for every voxel of an unsafe element."""
    elem_unsafe, _ = unsafe_elements(active, h, Nx, Ny, Nz)
    ev = elem_voxels(h, Nx, Ny, Nz)
    out = np.zeros(np.asarray(active).size, dtype=bool)
    out[ev[elem_unsafe].ravel()] = True
    return out


def overlap_lost_element_mask(active, mask0, h, Nx, Ny, Nz, frac_threshold=0.0,
                              min_material_frac=0.5):
    """(Ne,) overlap loss (dilated through shared nodes) | material
    fraction below the floor | boundary shell."""
    active = np.asarray(active, dtype=bool).ravel()
    mask0 = np.asarray(mask0, dtype=bool).ravel()
    ev = elem_voxels(h, Nx, Ny, Nz)
    en = elem_nodes(Nx, Ny, Nz)
    n_mat = mask0[ev].sum(axis=1)
    n_lost = (mask0 & ~active)[ev].sum(axis=1)
    with np.errstate(invalid='ignore', divide='ignore'):
        frac = np.where(n_mat > 0, n_lost / n_mat, 0.0)
    lost = (n_mat > 0) & (frac > frac_threshold)
    node_lost = np.zeros((Nx + 1) * (Ny + 1) * (Nz + 1), dtype=bool)
    node_lost[en[lost].ravel()] = True
    dilated = node_lost[en].any(axis=1)
    sparse = n_mat / ev.shape[1] < min_material_frac
    return dilated | sparse | boundary_elements(Nx, Ny, Nz)


def safe_cell_mask(unsafe_vox, h, Nx, Ny, Nz):
    """(Nx, Ny, Nz) True where none of the element's voxels is unsafe."""
    u = np.asarray(unsafe_vox, dtype=bool).reshape(-1, order='F')
    anyu = u[elem_voxels(h, Nx, Ny, Nz)].any(axis=1)
    return (~anyu).reshape(Nx, Ny, Nz, order='F')


def shape_grads_centroid(h):
    """(3, 8) Hex8 shape-function gradients at the element centre, / h."""
    g = np.empty((3, 8))
    for a, (cx, cy, cz) in enumerate(CORNERS):
        sx, sy, sz = (1.0 if cx else -1.0), (1.0 if cy else -1.0), (1.0 if cz else -1.0)
        g[0, a] = sx * 0.25
        g[1, a] = sy * 0.25
        g[2, a] = sz * 0.25
    return g / h


def principal_strains_cell(U_grid, h):
    """(Nx, Ny, Nz, 3) eigenvalues of sym(grad u) at element centres,
    descending."""
    U_grid = np.asarray(U_grid, dtype=np.float64)
    Nx, Ny, Nz = (n - 1 for n in U_grid.shape[:3])
    dN = shape_grads_centroid(h)
    grad = np.zeros((Nx, Ny, Nz, 3, 3))
    for a, (di, dj, dk) in enumerate(CORNERS):
        blk = U_grid[di:di + Nx, dj:dj + Ny, dk:dk + Nz, :]
        grad += blk[..., :, None] * dN[None, None, None, :, a][..., None, :]
    eps = 0.5 * (grad + np.swapaxes(grad, -1, -2))
    return np.linalg.eigvalsh(eps)[..., ::-1]


def equivalent_strain(p):
    """von Mises equivalent strain of principal strains (..., 3)."""
    p = np.asarray(p, dtype=np.float64)
    e1, e2, e3 = p[..., 0], p[..., 1], p[..., 2]
    return (np.sqrt(2.0) / 3.0) * np.sqrt((e1 - e2) ** 2 + (e2 - e3) ** 2 + (e3 - e1) ** 2)


def _grid3(vec, Nx, Ny, Nz):
    v = np.asarray(vec, dtype=np.float64)
    return np.stack([v[c::3].reshape(Nx + 1, Ny + 1, Nz + 1, order='F') for c in range(3)],
                    axis=-1)


def gn_convergence_cell(dU, U, Nx, Ny, Nz):
    """(Nx, Ny, Nz) ||dU|| / ||U|| over each element's 24 corner DOFs, NaN
    where ||U|| is 0."""
    def sq(g):
        acc = np.zeros((Nx, Ny, Nz))
        for di, dj, dk in CORNERS:
            b = g[di:di + Nx, dj:dj + Ny, dk:dk + Nz, :]
            acc += (b * b).sum(axis=-1)
        return acc
    num = np.sqrt(sq(_grid3(dU, Nx, Ny, Nz)))
    den = np.sqrt(sq(_grid3(U, Nx, Ny, Nz)))
    with np.errstate(invalid='ignore', divide='ignore'):
        r = num / den
    r[den <= 0.0] = np.nan
    return r


def element_centres_deformed(U_grid, h):
    """(Nx, Ny, Nz, 3) reference centres (i + 0.5) h and the same moved by
    the mean of the 8 corner displacements."""
    U_grid = np.asarray(U_grid, dtype=np.float64)
    Nx, Ny, Nz = (n - 1 for n in U_grid.shape[:3])
    ii, jj, kk = np.meshgrid(np.arange(Nx), np.arange(Ny), np.arange(Nz), indexing='ij')
    x_ref = (np.stack([ii, jj, kk], axis=-1).astype(np.float64) + 0.5) * h
    u = np.zeros(x_ref.shape)
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                u += U_grid[ii + di, jj + dj, kk + dk]
    u /= 8.0
    return x_ref, x_ref + u


def remap_mask_nearest(elem_mask, U_grid, h, crop_lo=(0, 0, 0)):
    """(Nx, Ny, Nz) element mask carried to the deformed configuration: the
    value of the nearest deformed element centre at every regular centre
    (+ crop_lo)."""
    from scipy.interpolate import griddata
    x_ref, x_def = element_centres_deformed(U_grid, h)
    q = (x_ref + np.asarray(crop_lo, dtype=np.float64)).reshape(-1, 3)
    v = np.asarray(elem_mask).reshape(-1).astype(np.float64)
    out = griddata(x_def.reshape(-1, 3), v, q, method='nearest')
    return out.reshape(np.asarray(elem_mask).shape).astype(bool)


def remap_mask_nearest_brute(elem_mask, U_grid, h, crop_lo=(0, 0, 0)):
    """The same nearest remap by exhaustive search (ties: lowest index in
    the x-fastest element order) -- for small meshes."""
    elem_mask = np.asarray(elem_mask, dtype=bool)
    x_ref, x_def = element_centres_deformed(U_grid, h)
    Nx, Ny, Nz = elem_mask.shape
    P = x_def.transpose(2, 1, 0, 3).reshape(-1, 3)            # x-fastest points
    Q = (x_ref + np.asarray(crop_lo, dtype=np.float64)).transpose(2, 1, 0, 3).reshape(-1, 3)
    vals = elem_mask.transpose(2, 1, 0).ravel()
    out = np.empty(Q.shape[0], dtype=bool)
    for s in range(0, Q.shape[0], 512):
        d = Q[s:s + 512, None, :] - P[None, :, :]
        d2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
        out[s:s + 512] = vals[np.argmin(d2, axis=1)]
    return out.reshape(Nz, Ny, Nx).transpose(2, 1, 0)


def remap_field_linear(field, U_grid, h, crop_lo=(0, 0, 0)):
    """(Nx, Ny, Nz[, C]) per-element field carried to the deformed
    configuration the reference's way: scipy griddata 'linear' over the
    moved element centres (a Qhull Delaunay triangulation), the 'nearest'
    values where it gives NaN (outside the convex hull)."""
    from scipy.interpolate import griddata
    field = np.asarray(field, dtype=np.float64)
    x_ref, x_def = element_centres_deformed(U_grid, h)
    pts = x_def.reshape(-1, 3)
    q = (x_ref + np.asarray(crop_lo, dtype=np.float64)).reshape(-1, 3)
    vals = field.reshape(-1, field.shape[-1]) if field.ndim == 4 else field.reshape(-1, 1)
    lin = griddata(pts, vals, q, method='linear')
    near = griddata(pts, vals, q, method='nearest')
    bad = np.isnan(lin)
    lin[bad] = near[bad]
    return lin.reshape(field.shape)


def material_cell_mask(ref, h, Nx, Ny, Nz, glt=0.0):
    """(Nx, Ny, Nz) True where the mean of ref over the element's (h+1)^3
    voxel block exceeds glt (ref indexed [x][y][z])."""
    flat = np.asarray(ref).reshape(-1, order='F')
    mean = flat[elem_voxels(h, Nx, Ny, Nz)].mean(axis=1)
    return (mean > glt).reshape(Nx, Ny, Nz, order='F')


def material_cell_mask_eroded(mask):
    """One-cell erosion over the 26-neighbourhood, the border counting as
    True (scipy binary_erosion, 3x3x3 structure, border_value=1)."""
    from scipy.ndimage import binary_erosion
    return binary_erosion(np.asarray(mask, dtype=bool), structure=np.ones((3, 3, 3), bool),
                          border_value=1)


_KUHN = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))


def remap_field_kuhn_brute(field, U_grid, h, crop_lo=(0, 0, 0)):
    """The device path's interpolant by exhaustive search (small meshes):
    every cell of the moved-centre lattice split into the 6 tetrahedra
    around its (0,0,0)-(1,1,1) diagonal; a query takes the barycentric
    interpolation over the first tetrahedron (cells x fastest, tetrahedra
    in _KUHN order) whose coordinates are all >= -1e-10, else the value of
    the nearest moved centre (lowest index on ties)."""
    field = np.asarray(field, dtype=np.float64)
    x_ref, x_def = element_centres_deformed(U_grid, h)
    Nx, Ny, Nz = x_def.shape[:3]
    C = field.shape[3] if field.ndim == 4 else 1
    F = field.reshape(Nx, Ny, Nz, C)
    Q = (x_ref + np.asarray(crop_lo, dtype=np.float64)).reshape(-1, 3)
    P = x_def.transpose(2, 1, 0, 3).reshape(-1, 3)
    Fx = F.transpose(2, 1, 0, 3).reshape(-1, C)
    # tetrahedra: (4 vertex ids) over all cells
    tets = []
    for c in range(Nz - 1):
        for b in range(Ny - 1):
            for a in range(Nx - 1):
                vid = lambda o: (a + o[0]) + Nx * ((b + o[1]) + Ny * (c + o[2]))
                for p in _KUHN:
                    e0 = [0, 0, 0]
                    e1 = list(e0)
                    e1[p[0]] = 1
                    e2 = list(e1)
                    e2[p[1]] = 1
                    tets.append((vid(e0), vid(e1), vid(e2), vid((1, 1, 1))))
    tets = np.array(tets, dtype=np.int64).reshape(-1, 4)
    out = np.empty((Q.shape[0], C))
    if len(tets):
        V0 = P[tets[:, 0]]
        M = np.stack([P[tets[:, 1]] - V0, P[tets[:, 2]] - V0, P[tets[:, 3]] - V0], axis=2)
        Minv = np.linalg.inv(M)
    for qi, q in enumerate(Q):
        hit = None
        if len(tets):
            l123 = np.einsum('tij,tj->ti', Minv, q[None, :] - V0)
            lam = np.concatenate([1.0 - l123.sum(1, keepdims=True), l123], axis=1)
            ok = np.nonzero(lam.min(1) >= -1e-10)[0]
            if ok.size:
                hit = ok[0]
        if hit is not None:
            out[qi] = lam[hit] @ Fx[tets[hit]]
        else:
            d = q[None, :] - P
            d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
            out[qi] = Fx[np.argmin(d2)]
    # Q is in the reference centres' (i, j, k) order with k fastest
    o = out.reshape(Nx, Ny, Nz, C)
    return o if field.ndim == 4 else o[..., 0]
