"""GPU parity of the element-scale output fields (libvdvc's vdvc_cell_masks,
vdvc_safe_cells, vdvc_principal_strains, vdvc_equivalent_strain,
vdvc_gn_convergence, vdvc_remap_mask_deformed) against
oracle/cells_oracle.py on seeded inputs.  The masks and the remap are
compared exactly (the GPU's 27-neighbourhood and per-voxel element tests
against the oracle's element-voxel arrays and node dilation); the strains
within 1e-13 of the largest |strain| (float64 Jacobi against LAPACK); the
equivalent strain of given principal strains and the convergence ratio
within 1 ulp (NumPy's operation order, float64)."""
import numpy as np
import pytest

from oracle import cells_oracle as CO

pytestmark = pytest.mark.gpu

# (Nx > 64: element columns beyond one wavefront's lanes in the row-wise count)
MESHES = [(6, 5, 7, 2), (7, 7, 6, 3), (1, 1, 1, 4), (2, 3, 1, 5), (9, 4, 5, 8), (3, 3, 3, 1),
          (12, 11, 10, 4), (70, 2, 3, 2), (130, 1, 2, 3)]


def _fields(seed, Nx, Ny, Nz, h):
    """Active / material voxel masks with scattered pores, lost voxels and a
    lost slab."""
    rng = np.random.default_rng(seed)
    nv = (Nx * h + 1) * (Ny * h + 1) * (Nz * h + 1)
    mat = rng.random(nv) > 0.05
    act = mat & (rng.random(nv) > 0.002)
    g = act.reshape(Nx * h + 1, Ny * h + 1, Nz * h + 1, order='F')
    g[:, :, : max(1, h // 2)] = False            # a lost slab at z = 0
    return act, mat


@pytest.mark.parametrize('mesh', MESHES)
def test_cell_masks(gpu, mesh):
    from voxeldvc_amd.engine import kernels_dvc as K
    Nx, Ny, Nz, h = mesh
    print('cell masks', mesh, flush=True)
    act, mat = _fields(11, Nx, Ny, Nz, h)
    ue, un = CO.unsafe_elements(act, h, Nx, Ny, Nz)
    np.testing.assert_array_equal(K.compute_unsafe_element_mask(act, h, Nx, Ny, Nz), ue)
    np.testing.assert_array_equal(K.compute_unsafe_node_mask(act, h, Nx, Ny, Nz), un)
    uv = K.compute_unsafe_voxel_mask(act, h, Nx, Ny, Nz)
    np.testing.assert_array_equal(uv, CO.unsafe_voxel_mask(act, h, Nx, Ny, Nz))
    for ft, mf in ((0.0, 0.5), (0.01, 0.9), (0.2, 0.0)):
        np.testing.assert_array_equal(
            K.compute_overlap_lost_element_mask(act, mat, h, Nx, Ny, Nz, ft, mf),
            CO.overlap_lost_element_mask(act, mat, h, Nx, Ny, Nz, ft, mf))
    np.testing.assert_array_equal(K.boundary_element_mask(Nx, Ny, Nz),
                                  CO.boundary_elements(Nx, Ny, Nz))


@pytest.mark.parametrize('mesh', MESHES)
def test_safe_cells(gpu, mesh):
    from voxeldvc_amd.engine import write_output as W
    Nx, Ny, Nz, h = mesh
    print('safe cells', mesh, flush=True)
    act, _ = _fields(12, Nx, Ny, Nz, h)
    uv = CO.unsafe_voxel_mask(act, h, Nx, Ny, Nz)
    grid = uv.reshape(Nx * h + 1, Ny * h + 1, Nz * h + 1, order='F')
    want = CO.safe_cell_mask(grid, h, Nx, Ny, Nz)
    np.testing.assert_array_equal(W.compute_safe_cell_mask(grid, h, Nx, Ny, Nz), want)
    rng = np.random.default_rng(4)
    sparse = rng.random(uv.size) < 0.001            # an arbitrary voxel mask, flat
    np.testing.assert_array_equal(W.compute_safe_cell_mask(sparse, h, Nx, Ny, Nz),
                                  CO.safe_cell_mask(sparse, h, Nx, Ny, Nz))


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('mesh', MESHES)
def test_strains(gpu, mesh, dtype):
    from voxeldvc_amd.engine import write_output as W
    Nx, Ny, Nz, h = mesh
    print('strains', mesh, np.dtype(dtype).name, flush=True)
    rng = np.random.default_rng(21)
    U = (rng.standard_normal((Nx + 1, Ny + 1, Nz + 1, 3)) * 0.3).astype(dtype)
    U[0, 0, 0] = 0.0                               # a degenerate-free but mixed field
    want = CO.principal_strains_cell(U, h)
    got = W.compute_principal_strains_cell(U, h)
    assert got.shape == (Nx, Ny, Nz, 3) and got.dtype == np.float64
    scale = np.abs(want).max()
    assert np.abs(got - want).max() <= 1e-13 * scale
    assert (got[..., 0] >= got[..., 1]).all() and (got[..., 1] >= got[..., 2]).all()
    eq = W.compute_equivalent_strain(want)
    np.testing.assert_allclose(eq, CO.equivalent_strain(want), rtol=2.3e-16, atol=0)


def test_strains_degenerate_and_linear(gpu):
    from voxeldvc_amd.engine import write_output as W
    Nx, Ny, Nz, h = 4, 3, 5, 8
    # uniform dilation (triple eigenvalue), uniaxial (double), pure rotation (zero)
    x = np.arange(Nx + 1)[:, None, None] * h
    y = np.arange(Ny + 1)[None, :, None] * h
    z = np.arange(Nz + 1)[None, None, :] * h
    X = np.stack(np.broadcast_arrays(x, y, z), axis=-1).astype(np.float64)
    for G, lam in ((0.01 * np.eye(3), [0.01, 0.01, 0.01]),
                   (np.diag([0.02, -0.005, -0.005]), [0.02, -0.005, -0.005]),
                   (np.array([[0, 0.01, 0], [-0.01, 0, 0], [0, 0, 0]]), [0.0, 0.0, 0.0])):
        got = W.compute_principal_strains_cell(X @ G.T + 3.0, h)
        assert np.abs(got - np.array(lam)).max() < 1e-15


@pytest.mark.parametrize('dtypes', [(np.float64, np.float64), (np.float32, np.float64),
                                    (np.float32, np.float32)])
@pytest.mark.parametrize('mesh', MESHES)
def test_gn_convergence(gpu, mesh, dtypes):
    from voxeldvc_amd.engine import write_output as W
    Nx, Ny, Nz, _ = mesh
    print('gn convergence', mesh, dtypes, flush=True)
    rng = np.random.default_rng(31)
    n = 3 * (Nx + 1) * (Ny + 1) * (Nz + 1)
    dU = (rng.standard_normal(n) * 1e-3).astype(dtypes[0])
    U = rng.standard_normal(n).astype(dtypes[1])
    U[: 3 * (Nx + 1) * (Ny + 1)] = 0.0             # a zero plane -> NaN cells where all 8 are 0
    want = CO.gn_convergence_cell(dU, U, Nx, Ny, Nz)
    got = W.compute_gn_convergence_cell(dU, U, Nx, Ny, Nz)
    assert got.shape == (Nx, Ny, Nz)
    np.testing.assert_array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    np.testing.assert_allclose(got[ok], want[ok], rtol=2.3e-16, atol=0)


@pytest.mark.parametrize('mesh', MESHES)
def test_remap_mask(gpu, mesh):
    from voxeldvc_amd.engine import write_output as W
    Nx, Ny, Nz, h = mesh
    print('remap mask', mesh, flush=True)
    rng = np.random.default_rng(41)
    m = rng.random((Nx, Ny, Nz)) < 0.4
    for U, crop in ((np.zeros((Nx + 1, Ny + 1, Nz + 1, 3)), (0, 0, 0)),
                    (rng.normal(0, 0.7 * h, (Nx + 1, Ny + 1, Nz + 1, 3)), (2, -1, 3)),
                    (rng.normal(3 * h, 0.2 * h, (Nx + 1, Ny + 1, Nz + 1, 3)), (0, 0, 0))):
        got = W.remap_element_mask_to_deformed(m, U, h, crop)
        np.testing.assert_array_equal(got, CO.remap_mask_nearest_brute(m, U, h, crop))
        np.testing.assert_array_equal(got, CO.remap_mask_nearest(m, U, h, crop))


def test_cells_device_tensors(gpu):
    """Torch CUDA inputs stay on the device and give the same values."""
    from voxeldvc_amd.engine import kernels_dvc as K
    from voxeldvc_amd.engine import write_output as W
    torch = gpu
    Nx, Ny, Nz, h = 5, 4, 6, 4
    act, mat = _fields(13, Nx, Ny, Nz, h)
    ta = torch.from_numpy(act).cuda()
    got = K.compute_unsafe_voxel_mask(ta, h, Nx, Ny, Nz)
    assert got.is_cuda and got.dtype == torch.bool
    np.testing.assert_array_equal(got.cpu().numpy(), CO.unsafe_voxel_mask(act, h, Nx, Ny, Nz))
    rng = np.random.default_rng(5)
    U = rng.standard_normal((Nx + 1, Ny + 1, Nz + 1, 3))
    p = W.compute_principal_strains_cell(torch.from_numpy(U).cuda(), h)
    assert p.is_cuda and tuple(p.shape) == (Nx, Ny, Nz, 3)
    assert np.abs(p.cpu().numpy() - CO.principal_strains_cell(U, h)).max() < 1e-13 * np.abs(U).max()


def test_cells_full_size_properties(gpu):
    """A 64^3-element mesh at h = 8 (513^3 voxels): the unsafe voxel mask is
    the voxel expansion of the unsafe element mask, and the safe cells are
    the elements with no unsafe element within +-1."""
    from voxeldvc_amd.engine import kernels_dvc as K
    torch = gpu
    Nx = Ny = Nz = 64
    h = 8
    n = Nx * h + 1
    print('cells full size', n, flush=True)
    g = torch.Generator(device='cuda').manual_seed(3)
    act = torch.rand(n ** 3, device='cuda', generator=g) > 1e-6
    ue = K.compute_unsafe_element_mask(act, h, Nx, Ny, Nz).view(Nz, Ny, Nx)
    uv = K.compute_unsafe_voxel_mask(act, h, Nx, Ny, Nz).view(n, n, n)
    # voxel (x, y, z) with x not on an element face lies only in element x // h
    inner = torch.arange(n, device='cuda') % h != 0
    sel = uv[inner][:, inner][:, :, inner]
    idx = (torch.arange(n, device='cuda') // h).clamp(max=Nx - 1)[inner]
    exp = ue[idx][:, idx][:, :, idx]
    assert torch.equal(sel, exp)
    # inactive count check: every element holding an inactive voxel is unsafe
    ina = (~act).nonzero().view(-1)
    z, y, x = ina // (n * n), (ina // n) % n, ina % n
    assert bool(ue[(z // h).clamp(max=Nz - 1), (y // h).clamp(max=Ny - 1),
                   (x // h).clamp(max=Nx - 1)].all())
    from voxeldvc_amd.engine import write_output as W
    safe = W.compute_safe_cell_mask(uv.permute(2, 1, 0), h, Nx, Ny, Nz)   # (Nx, Ny, Nz)
    u = torch.nn.functional.max_pool3d(ue.float()[None, None], 3, stride=1, padding=1)[0, 0]
    assert torch.equal(safe.permute(2, 1, 0), u == 0)


def test_remap_mask_spread_and_errors(gpu):
    """A far outlier (the cell grid coarsens until it has at most ~4 cells
    per point), queries well outside the moved cloud, and a non-finite
    displacement (rejected)."""
    from voxeldvc_amd.engine import write_output as W
    Nx, Ny, Nz, h = 9, 7, 8, 4
    rng = np.random.default_rng(43)
    m = rng.random((Nx, Ny, Nz)) < 0.5
    U = rng.normal(0, 0.5 * h, (Nx + 1, Ny + 1, Nz + 1, 3))
    U[3, 2, 4] = (4000.0, -2500.0, 800.0)
    for crop in ((0, 0, 0), (60, -45, 30)):
        got = W.remap_element_mask_to_deformed(m, U, h, crop)
        np.testing.assert_array_equal(got, CO.remap_mask_nearest_brute(m, U, h, crop))
    U[1, 1, 1, 2] = np.nan
    with pytest.raises(ValueError):
        W.remap_element_mask_to_deformed(m, U, h)


def test_remap_mask_lattice_ties(gpu):
    """Whole-element shifts put queries exactly half way between moved
    centres: ties go to the lowest element index (x fastest), as in the
    exhaustive search."""
    from voxeldvc_amd.engine import write_output as W
    Nx, Ny, Nz, h = 6, 5, 4, 4
    rng = np.random.default_rng(44)
    m = rng.random((Nx, Ny, Nz)) < 0.5
    U = np.zeros((Nx + 1, Ny + 1, Nz + 1, 3))
    U[..., 0] = 0.5 * h          # every query midway between two x neighbours
    U[..., 2] = -1.5 * h
    got = W.remap_element_mask_to_deformed(m, U, h)
    np.testing.assert_array_equal(got, CO.remap_mask_nearest_brute(m, U, h))


def test_remap_mask_outlier_large_mesh(gpu):
    """One diverged element far from the rest on a 40^3-element mesh: the
    cell grid is sized around the queries (the outlier is filed in a
    boundary cell), and the result still equals the nearest-point remap."""
    from voxeldvc_amd.engine import write_output as W
    Nx = Ny = Nz = 40
    h = 8
    rng = np.random.default_rng(45)
    m = rng.random((Nx, Ny, Nz)) < 0.5
    U = rng.normal(0, 0.3 * h, (Nx + 1, Ny + 1, Nz + 1, 3))
    U[20, 20, 20] = (1e6, -3e5, 7e5)
    got = W.remap_element_mask_to_deformed(m, U, h)
    np.testing.assert_array_equal(got, CO.remap_mask_nearest(m, U, h))


FIELD_MESHES = [(6, 5, 4, 8), (7, 6, 5, 4), (2, 2, 2, 2), (9, 3, 4, 2), (1, 3, 2, 4)]


@pytest.mark.parametrize('mesh', FIELD_MESHES)
def test_remap_field_matches_kuhn_oracle(gpu, mesh):
    """vdvc_remap_field_deformed against the oracle's exhaustive-search
    restatement of the same interpolant (every tetrahedron of every cell
    tried): equal to rounding wherever the walk finds the tetrahedron, and
    the nearest fill elsewhere; smooth and random displacements, crops,
    1 and 3 channels."""
    from voxeldvc_amd.engine import write_output as W
    Nx, Ny, Nz, h = mesh
    print('remap field', mesh, flush=True)
    rng = np.random.default_rng(51)
    n = np.stack(np.meshgrid(np.arange(Nx + 1), np.arange(Ny + 1), np.arange(Nz + 1),
                             indexing='ij'), -1).astype(float) * h
    cases = [(0.6 * np.sin(n / 17.0) + 0.3, (0, 0, 0)),
             (rng.normal(0, 0.15 * h, n.shape), (1.5, -2.0, 0.5)),
             (np.zeros(n.shape), (0, 0, 0)),
             (0.6 * np.sin(n / 11.0) - 1.0, (3 * h, 0, -h))]
    for U, crop in cases:
        for F in (rng.normal(size=(Nx, Ny, Nz)), rng.normal(size=(Nx, Ny, Nz, 3))):
            got = W.remap_cell_field_to_deformed(F, U, h, crop)
            want = CO.remap_field_kuhn_brute(F, U, h, crop)
            assert got.shape == F.shape
            np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * (1 + np.abs(F).max()))


def test_remap_field_against_griddata(gpu):
    """Against the reference's scipy griddata ('linear' + 'nearest' fill):
    a field affine in the moved positions is reproduced exactly by both
    inside the moved lattice; a smooth field agrees to a fraction of its
    second differences."""
    from voxeldvc_amd.engine import write_output as W
    Nx, Ny, Nz, h = 12, 10, 9, 8
    n = np.stack(np.meshgrid(np.arange(Nx + 1), np.arange(Ny + 1), np.arange(Nz + 1),
                             indexing='ij'), -1).astype(float) * h
    U = 0.8 * np.sin(n / 23.0) + 0.4
    x_ref, x_def = CO.element_centres_deformed(U, h)
    rng = np.random.default_rng(52)
    A, b = rng.normal(size=(3, 3)), rng.normal(size=3)
    F = x_def @ A.T + b
    inner = np.zeros((Nx, Ny, Nz), bool)
    inner[1:-1, 1:-1, 1:-1] = True
    got = W.remap_cell_field_to_deformed(F, U, h)
    ref = CO.remap_field_linear(F, U, h)
    want = x_ref @ A.T + b
    assert np.abs(got - want)[inner].max() < 1e-11
    assert np.abs(ref - want)[inner].max() < 1e-11
    S = np.sin(x_def[..., 0] / 20) * np.cos(x_def[..., 1] / 15) + x_def[..., 2] / 50
    got = W.remap_cell_field_to_deformed(S, U, h)
    ref = CO.remap_field_linear(S, U, h)
    d2 = max(np.abs(np.diff(S, 2, axis=a)).max() for a in range(3))
    assert np.abs(got - ref)[inner].max() <= 0.5 * d2


@pytest.mark.parametrize('dtype', ['uint8', 'uint16', 'float32', 'float64', 'bool'])
@pytest.mark.parametrize('mesh', [(6, 5, 7, 2), (4, 3, 5, 4), (3, 2, 2, 8), (70, 2, 3, 2)])
def test_material_cells(gpu, dtype, mesh):
    from voxeldvc_amd.engine import write_output as W
    Nx, Ny, Nz, h = mesh
    rng = np.random.default_rng(61)
    shape = (Nx * h + 1, Ny * h + 1, Nz * h + 1)
    if dtype == 'bool':
        ref = rng.random(shape) < 0.6
        glt = 0.5
    else:
        ref = rng.integers(0, 200, shape).astype(dtype)
        ref[: h + 1] = 0
        glt = 99.5
    got = W.compute_material_cell_mask(ref, h, Nx, Ny, Nz, glt=glt)
    want = CO.material_cell_mask(ref.astype(np.float64) if dtype == 'bool' else ref, h, Nx, Ny,
                                 Nz, glt)
    assert got.dtype == bool and got.shape == (Nx, Ny, Nz)
    np.testing.assert_array_equal(got, want)
    ero = W.compute_material_cell_mask_eroded(got)
    np.testing.assert_array_equal(ero, CO.material_cell_mask_eroded(want))
