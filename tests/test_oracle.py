"""CPU pins of the oracle.  Each oracle function is checked against something
that fixes it independently of its own formula: closed forms, brute-force
dense assembly on tiny meshes, library routines (np.gradient,
scipy.ndimage, np.linalg.solve, np.fft), the reference's own tests, and the
golden vectors written from the reference engine (tests/golden/)."""
import numpy as np
import pytest
from scipy.ndimage import map_coordinates, minimum_filter

from conftest import golden, problem17
from oracle import dvc_oracle as O


# ---------------------------------------------------------------- geometry
@pytest.mark.parametrize('h', [1, 2, 3, 4, 8])
def test_stencil_closed_forms(h):
    N, w = O.build_N_stencil(h)
    assert N.shape == ((h + 1) ** 3, 8)
    assert np.allclose(N.sum(1), 1.0)                       # partition of unity
    assert np.isclose(w.sum(), h ** 3)                       # trapezoid rule volume
    # N at the element corners is the identity (node order of the reference)
    k = np.arange(h + 1)
    qi = lambda a, b, c: (a * (h + 1) + b) * (h + 1) + c
    for n, (a, b, c) in enumerate(O.CORNERS):
        row = N[qi(a * h, b * h, c * h)]
        assert np.isclose(row[n], 1.0) and np.isclose(np.abs(row).sum(), 1.0)


def test_laplacian_K_closed_form():
    K = O.build_K_ref_laplacian()
    L8 = K[0::3, 0::3]
    # unit-cube Q1 Laplacian: diag 1/3, edge neighbours 0, face/body diagonals -1/12
    exp = np.zeros((8, 8))
    for a, pa in enumerate(O.CORNERS):
        for b, pb in enumerate(O.CORNERS):
            d = int(np.abs(pa - pb).sum())
            exp[a, b] = {0: 1 / 3, 1: 0.0, 2: -1 / 12, 3: -1 / 12}[d]
    assert np.allclose(L8, exp, atol=1e-14)
    assert np.allclose(K[0::3, 1::3], 0) and np.allclose(K, K.T)


def test_elastic_K_properties():
    K = O.build_K_ref(0.3)
    assert np.allclose(K, K.T)
    ev = np.linalg.eigvalsh(K)
    assert np.sum(np.abs(ev) < 1e-10) == 6 and ev.min() > -1e-10   # 6 rigid modes
    # uniaxial strain e_xx = 1 on the unit cube: strain energy = c1 / 2
    nu = 0.3
    u = np.zeros(24)
    u[0::3] = O.CORNERS[:, 0]
    c1 = (1 - nu) / ((1 + nu) * (1 - 2 * nu))
    assert np.isclose(0.5 * u @ K @ u, 0.5 * c1)


def test_interior_mask():
    m = O.build_interior_dof_mask(3, 4, 2).reshape(3, 5, 4, 3)   # [k][j][i][c]
    assert m.sum() == 3 * 2 * 3 * 1 and m[1, 1:4, 1:3].all()


# ------------------------------------------------- dense brute-force pins
def dense_H(g, m, h, Nx, Ny, Nz, squared=False):
    """Assemble H = sum_e sum_q w m v_eq v_eq^T element by element."""
    Ndof = 3 * (Nx + 1) * (Ny + 1) * (Nz + 1)
    N, w = O.build_N_stencil(h)
    offs = O.voxel_offsets(h, Nx, Ny)
    H = np.zeros((Ndof, Ndof))
    for e in range(Nx * Ny * Nz):
        dofs = O.elem_dofs(np.array([e]), Nx, Ny)[0]
        vox = O.elem_voxel_origin(np.array([e]), h, Nx, Ny)[0] + offs
        for q in range(offs.size):
            veq = (N[q][:, None] * g[vox[q]][None, :]).ravel()
            H[np.ix_(dofs, dofs)] += w[q] * m[vox[q]] * np.outer(veq, veq)
    return H


@pytest.mark.parametrize('shape', [(1, 1, 1, 1), (2, 1, 1, 2), (1, 2, 2, 3), (2, 2, 1, 2)])
def test_H_dense(shape):
    Nx, Ny, Nz, h = shape
    rng = np.random.default_rng(1)
    Nvox = (Nx * h + 1) * (Ny * h + 1) * (Nz * h + 1)
    g = rng.normal(size=(Nvox, 3))
    m = rng.random(Nvox) < 0.7
    H = dense_H(g, m, h, Nx, Ny, Nz)
    v = rng.normal(size=H.shape[0])
    assert np.allclose(O.matvec_H(v, g, m, h, Nx, Ny, Nz), H @ v)
    assert np.allclose(O.build_diagonal_H(g, m, h, Nx, Ny, Nz), np.diag(H))
    assert np.allclose(H, H.T) and np.linalg.eigvalsh(H).min() > -1e-9
    res = rng.normal(size=Nvox)
    # b = sum w m res v_eq: the gradient of 1/2 sum w m (res - v_eq.u)^2 at u = 0
    b = O.compute_b(res, m, g, h, Nx, Ny, Nz)
    Ndof = H.shape[0]
    J = np.zeros((Ndof,))
    eps = 1e-6
    for i in range(Ndof):
        e = np.zeros(Ndof)
        e[i] = eps
        # b_i = -(d/du_i) phi(u) with phi = 1/2 sum w m (res - g.u(x))^2
        J[i] = (phi(res, m, g, h, Nx, Ny, Nz, -e) - phi(res, m, g, h, Nx, Ny, Nz, e)) / (2 * eps)
    assert np.allclose(b, J, atol=1e-5)


def phi(res, m, g, h, Nx, Ny, Nz, u):
    N, w = O.build_N_stencil(h)
    offs = O.voxel_offsets(h, Nx, Ny)
    s = 0.0
    for e in range(Nx * Ny * Nz):
        dofs = O.elem_dofs(np.array([e]), Nx, Ny)[0]
        vox = O.elem_voxel_origin(np.array([e]), h, Nx, Ny)[0] + offs
        ul = N @ u[dofs].reshape(8, 3)
        r = res[vox] - (g[vox] * ul).sum(1)
        s += 0.5 * (w * m[vox] * r * r).sum()
    return s


def test_unique_voxel_weight_rule():
    """Summed trapezoid weight per voxel = prod(1/2 on the outer planes)."""
    Nx, Ny, Nz, h = 3, 2, 2, 4
    Nvx, Nvy, Nvz = Nx * h + 1, Ny * h + 1, Nz * h + 1
    _, w = O.build_N_stencil(h)
    acc = np.zeros(Nvx * Nvy * Nvz)
    offs = O.voxel_offsets(h, Nx, Ny)
    for e in range(Nx * Ny * Nz):
        np.add.at(acc, O.elem_voxel_origin(np.array([e]), h, Nx, Ny)[0] + offs, w)
    W1 = lambda n: np.where((np.arange(n) == 0) | (np.arange(n) == n - 1), 0.5, 1.0)
    exp = np.einsum('i,j,k->ijk', W1(Nvx), W1(Nvy), W1(Nvz)).ravel(order='F')
    assert np.allclose(acc, exp)


def test_K_matvec_dense():
    Nx, Ny, Nz = 2, 3, 2
    for Kr in (O.build_K_ref_laplacian(), O.build_K_ref(0.25)):
        Ndof = 3 * (Nx + 1) * (Ny + 1) * (Nz + 1)
        KK = np.zeros((Ndof, Ndof))
        E = np.arange(1, Nx * Ny * Nz + 1, dtype=float)
        for e in range(Nx * Ny * Nz):
            d = O.elem_dofs(np.array([e]), Nx, Ny)[0]
            KK[np.ix_(d, d)] += E[e] * Kr
        v = np.random.default_rng(2).normal(size=Ndof)
        assert np.allclose(O.matvec_K_ref(v, Kr, E, Nx, Ny, Nz), KK @ v)
        assert np.allclose(O.build_diagonal_K_ref(Kr, E, Nx, Ny, Nz), np.diag(KK))
        P = np.diag(O.build_interior_dof_mask(Nx, Ny, Nz).astype(float))
        assert np.allclose(O.matvec_equilibrium_gap(v, Kr, np.diag(P), E, Nx, Ny, Nz),
                           KK @ P @ KK @ v)
    # constant fields are in the Laplacian's kernel; linear fields give zero
    # at interior rows
    Kl = O.build_K_ref_laplacian()
    n = np.arange((Nx + 1) * (Ny + 1) * (Nz + 1))
    ii = n % (Nx + 1)
    v = np.zeros(3 * n.size)
    v[0::3] = 1.0
    assert np.allclose(O.matvec_K_ref(v, Kl, 1.0, Nx, Ny, Nz), 0)
    v[0::3] = ii
    r = O.matvec_K_ref(v, Kl, 1.0, Nx, Ny, Nz)
    assert np.allclose(r[O.build_interior_dof_mask(Nx, Ny, Nz)], 0)


def test_pcg_exact_solve():
    rng = np.random.default_rng(3)
    A = rng.normal(size=(40, 40))
    A = A @ A.T + 40 * np.eye(40)
    b = rng.normal(size=40)
    x, it = O.pcg(lambda v: A @ v, b, 1.0 / np.diag(A), max_iter=200, tol=1e-12)
    assert np.allclose(x, np.linalg.solve(A, b), atol=1e-9)
    # the stopping rule: one iteration with max_iter=1
    x1, it1 = O.pcg(lambda v: A @ v, b, 1.0 / np.diag(A), max_iter=1, tol=1e-12)
    assert it1 == 1 and not np.allclose(x1, x)


# --------------------------------------------------------- voxel stages
def test_gradient_and_erosion():
    rng = np.random.default_rng(4)
    f = rng.integers(0, 256, size=(9, 7, 5)).astype(np.float32)
    fv, g = O.image_grad_and_values(f)
    for c in range(3):
        assert np.array_equal(g[:, c], np.ravel(np.gradient(f, axis=c), order='F'))
    m = rng.random(f.size) < 0.8
    fp = np.zeros((3, 3, 3), bool)
    fp[1, 1, :] = fp[1, :, 1] = fp[:, 1, 1] = True
    exp = np.ravel(minimum_filter(m.reshape(9, 7, 5, order='F'), footprint=fp, mode='nearest'),
                   order='F')
    assert np.array_equal(O.erode_mask_6(m, 9, 7, 5), exp)


def test_masked_stats_bruteforce():
    Nx, Ny, Nz, h = 2, 2, 1, 3
    rng = np.random.default_rng(5)
    Nvox = (Nx * h + 1) * (Ny * h + 1) * (Nz * h + 1)
    fv = rng.normal(size=Nvox)
    m = rng.random(Nvox) < 0.6
    offs = O.voxel_offsets(h, Nx, Ny)
    ids = (O.elem_voxel_origin(np.arange(Nx * Ny * Nz), h, Nx, Ny)[:, None] + offs).ravel()
    sel = fv[ids][m[ids]]
    mean, std = O.masked_quadpoint_mean_std(fv, m, h, Nx, Ny, Nz)
    assert np.isclose(mean, sel.mean()) and np.isclose(std, sel.std())
    assert O.masked_quadpoint_mean_std(fv, np.zeros(Nvox, bool), h, Nx, Ny, Nz) == (0.0, 0.0)


def test_sampling_affine_exact_and_scipy():
    rng = np.random.default_rng(6)
    img = rng.normal(size=(11, 9, 8))
    cx, cy, cz = (rng.uniform(0, n - 1, 500) for n in img.shape)
    ref = map_coordinates(img, np.stack([cx, cy, cz]), order=1, mode='nearest')
    assert np.allclose(O.trilinear_nearest(img, cx, cy, cz), ref)
    # a linear image is reproduced exactly by trilinear sampling
    X, Y, Z = np.meshgrid(*(np.arange(n, dtype=float) for n in img.shape), indexing='ij')
    lin = 2 * X - 3 * Y + 0.5 * Z + 1
    assert np.allclose(O.trilinear_nearest(lin, cx, cy, cz), 2 * cx - 3 * cy + 0.5 * cz + 1)
    # an affine nodal U samples g at exactly x + u(x)
    Nx, Ny, Nz, h = 2, 2, 2, 3
    n = np.arange(27)
    Xn = np.stack([n % 3, (n // 3) % 3, n // 9], 1) * h
    U = np.zeros(81)
    U[0::3] = 0.1 * Xn[:, 0] + 0.3
    U[1::3] = -0.2
    U[2::3] = 0.05 * Xn[:, 2]
    g = 1.5 * X[:7, :7, :7] + Y[:7, :7, :7] - Z[:7, :7, :7]
    gv, mk = O.sample_deformed_field(U, g, h, Nx, Ny, Nz)
    vx, vy, vz = (a.ravel(order='F') for a in np.meshgrid(np.arange(7.), np.arange(7.),
                                                            np.arange(7.), indexing='ij'))
    x = vx + 0.1 * vx + 0.3
    y = vy - 0.2
    z = vz + 0.05 * vz
    inb = (x <= 6) & (y >= 0) & (z <= 6)
    val = 1.5 * x + y - z
    assert np.array_equal(mk, inb & (gv > 0))        # in-bounds & mask0 & g > glt
    assert np.allclose(gv[inb], val[inb])


def test_fft_shift_reference_test():
    """tests/test_dvc_gpu_multiscale.py:75-85 of the reference."""
    rng = np.random.default_rng(42)
    f = rng.uniform(0, 1, size=(32, 32, 32))
    g = np.roll(f, (3, -2, 1), axis=(0, 1, 2))
    assert np.allclose(O.fft_rigid_shift(f, g), (3, -2, 1))


# ------------------------------------------------------ reference tests
def test_reference_single_scale_matches_direct():
    f, g, Nx, Ny, Nz, h = problem17()
    kw = dict(K_ref_laplacian=O.build_K_ref_laplacian(), l0=4.5 * h, maxiter=20, eps=1e-3)
    U1, r1, _ = O.correlate(f, g, Nx, Ny, Nz, h, **kw)
    U2, r2, _ = O.multiscale_correlate(f, g, Nx, Ny, Nz, h, scales=(0,), fft_prealign=False, **kw)
    assert np.allclose(U1, U2, atol=1e-12) and np.allclose(r1, r2, atol=1e-12)


def test_reference_rejects_indivisible():
    f, g, Nx, Ny, Nz, h = problem17()
    K = O.build_K_ref_laplacian()
    with pytest.raises(ValueError):
        O.multiscale_correlate(f, g, Nx, Ny, Nz, h, scales=(4, 0), K_ref_laplacian=K, l0=36.)
    with pytest.raises(ValueError):
        O.multiscale_correlate(f, g, Nx, Ny, Nz, h, scales=(1, 0), g_origin=(1, 0, 0),
                               K_ref_laplacian=K, l0=36.)
    with pytest.raises(ValueError):
        O.correlate(f, g, Nx, Ny, Nz, h, reg_type='tv')
    with pytest.raises(ValueError):
        O.correlate(f, g, 2 ** 20, 2 ** 20, 2, h)          # int32 guard before any work


# ------------------------------------------------------- golden vectors
@pytest.mark.parametrize('name,dt,tol', [('corr17_f64', np.float64, 1e-12),
                                         ('corr17_f32', np.float32, 1e-6)])
def test_golden_correlate_17(name, dt, tol):
    z = golden(name)
    U, res, dU = O.correlate(z['f'], z['g'], 2, 2, 2, 8, K_ref_laplacian=O.build_K_ref_laplacian(),
                             l0=36.0, maxiter=20, dtype=dt)
    assert np.abs(U - z['U']).max() < tol
    assert np.abs(dU - z['dU']).max() < tol
    assert np.abs(res - z['res']).max() < 1e3 * tol


def test_golden_multiscale():
    z = golden('multi17_fft')
    U, res, _ = O.multiscale_correlate(z['f'], z['g'], 2, 2, 2, 8, scales=(2, 1, 0), l0=36.0,
                                       K_ref_laplacian=O.build_K_ref_laplacian(), maxiter=30)
    assert np.abs(U - z['U']).max() < 1e-6
    z = golden('multi33_dyn')
    zr = golden('multi33_dyn_f32in')
    U, res, _ = O.multiscale_correlate(z['f'], z['g'], 4, 4, 4, 8, scales=(2, 1, 0),
                                       K_ref_laplacian=O.build_K_ref_laplacian(), l0=36.0,
                                       dynamic_mask=True, freeze_mask_after=4)
    assert np.abs(U - zr['U']).max() < 1e-6


def test_golden_eqgap():
    z = golden('multi33_dyn')
    zr = golden('eqgap33')
    U, _, _ = O.correlate(z['f'], z['g'], 4, 4, 4, 8, reg_type='equilibrium_gap',
                          K_ref_elastic=O.build_K_ref(0.3),
                          K_ref_laplacian=O.build_K_ref_laplacian(), l0=36.0, eps_reg=1e-3,
                          maxiter=10)
    assert np.abs(U - zr['U']).max() < 1e-5
    d = O.build_diagonal_equilibrium_gap(O.build_K_ref(0.3),
                                         O.build_interior_dof_mask(3, 3, 3).astype(np.float32),
                                         8, 3, 3, 3)
    assert np.array_equal(d, golden('eqgap_diag3')['diag'])


def test_golden_affine():
    z = golden('multi33_dyn')
    za = golden('affine33')
    A, t, c, info = O.affine_prealign(z['f'], z['g'], decimate=2)
    assert abs(info['cost_final'] - float(za['cost_final'])) < 1e-5
    assert np.abs(A - za['A']).max() < 1e-3 and np.abs(t - za['t']).max() < 1e-2
    assert np.allclose(O.fft_rigid_shift(z['f'], z['g']), golden('fft33')['d'])
