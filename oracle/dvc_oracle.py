"""CPU oracle for the global Gauss-Newton DVC solve (NumPy restatement).

TEST INFRASTRUCTURE ONLY.  This module is the checker the HIP library is
compared against, and the host-core baseline that bench.py times.  Nothing in
the product package (voxeldvc_amd) imports it; only tests/, the smoke check in
__graft_entry__.py and bench.py's CPU-baseline leg do.

It restates, in plain NumPy (float32 or float64, chosen by the caller), the
algorithm of ggalu/voxelDVC's engine with the same element-chunked structure:
gather per element, separable einsum over the (h+1)^3 voxel-centre stencil,
np.add.at scatter.  Every function cites the reference lines it follows
(paths relative to the reference's src/voxeldvc/).

Conventions (reference engine/kernels_dvc.py:9-21):
  * image shape (Nx*h+1, Ny*h+1, Nz*h+1); voxel flat index
    vidx = vx + vy*Nvx + vz*Nvx*Nvy (x fastest, i.e. ravel(order='F'));
  * node n = i + j*(Nx+1) + k*(Nx+1)*(Ny+1), DOF 3n+c;
  * element e = i + j*Nx + k*Nx*Ny;
  * local corner order 0=(0,0,0) 1=(1,0,0) 2=(1,1,0) 3=(0,1,0)
    4=(0,0,1) 5=(1,0,1) 6=(1,1,1) 7=(0,1,1).

Parity pins (what each function is checked against) are listed in
tests/test_oracle.py; functions pinned only by golden vectors written from the
reference are named there too.
"""
import numpy as np

# local corner offsets (di, dj, dk) in the reference's Hex8 order
# (engine/geometry_dvc.py:14-17)
CORNERS = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                    (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)], dtype=np.int64)


# ---------------------------------------------------------------------------
# geometry (engine/geometry_dvc.py)
# ---------------------------------------------------------------------------

def voxel_center_stencil(h):
    """Trapezoid voxel-centre rule on one element (geometry_dvc.py:26-51):
    (h+1)^3 points xi = 2k/h - 1, weight 1/2 per axis at k in {0, h}; point
    order is meshgrid 'ij' (first axis slowest)."""
    k = np.arange(h + 1)
    xi = 2.0 * k / h - 1.0
    w = np.ones(h + 1)
    w[[0, -1]] = 0.5
    X, Y, Z = np.meshgrid(xi, xi, xi, indexing='ij')
    WX, WY, WZ = np.meshgrid(w, w, w, indexing='ij')
    return X.ravel(), Y.ravel(), Z.ravel(), (WX * WY * WZ).ravel()


def hex8_shape(x, y, z):
    """Trilinear Hex8 shape functions on [-1,1]^3 (geometry_dvc.py:155-168)."""
    cols = []
    for (a, b, c) in CORNERS:
        sx = 2 * a - 1
        sy = 2 * b - 1
        sz = 2 * c - 1
        cols.append(0.125 * (1 + sx * x) * (1 + sy * y) * (1 + sz * z))
    return np.stack(cols, axis=-1)


def build_N_stencil(h):
    """(N_stencil (nq,8), w_stencil (nq,)) -- geometry_dvc.py:171-184."""
    x, y, z, w = voxel_center_stencil(h)
    return hex8_shape(x, y, z), w


def _shape_grads_unit(xi, eta, zeta):
    """d N_a / d(x,y,z) on the unit cube at (xi,eta,zeta) in [0,1]^3, (3,8)."""
    out = np.zeros((3, 8))
    for a, (i, j, k) in enumerate(CORNERS):
        fx = xi if i else 1 - xi
        fy = eta if j else 1 - eta
        fz = zeta if k else 1 - zeta
        dx = 1.0 if i else -1.0
        dy = 1.0 if j else -1.0
        dz = 1.0 if k else -1.0
        out[0, a] = dx * fy * fz
        out[1, a] = fx * dy * fz
        out[2, a] = fx * fy * dz
    return out


def _gauss2():
    g = 1.0 / np.sqrt(3.0)
    return [0.5 * (1 - g), 0.5 * (1 + g)]


def build_K_ref_laplacian():
    """24x24 vector Laplacian of the unit cube, 2x2x2 Gauss, block per
    component (geometry_dvc.py:88-110).  The reference integrates with
    weight 1 and detJ = 1/8 on the unit cube (it maps Gauss points of
    [-1,1]^3 into [0,1]^3 but keeps weight 1 per point)."""
    L8 = np.zeros((8, 8))
    for xi in _gauss2():
        for eta in _gauss2():
            for zeta in _gauss2():
                dN = _shape_grads_unit(xi, eta, zeta)
                L8 += 0.125 * dN.T @ dN
    K = np.zeros((24, 24))
    for c in range(3):
        K[c::3, c::3] = L8
    return K


def build_K_ref(nu):
    """24x24 isotropic elastic Hex8 stiffness on the unit cube, E=1, 2x2x2
    Gauss, node-major DOFs (geometry_dvc.py:58-85, 128-148)."""
    lam_c1 = (1.0 - nu) / ((1.0 + nu) * (1.0 - 2.0 * nu))
    lam_c2 = nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu_c3 = 1.0 / (2.0 * (1.0 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam_c2
    np.fill_diagonal(D[:3, :3], lam_c1)
    D[3, 3] = D[4, 4] = D[5, 5] = mu_c3
    K = np.zeros((24, 24))
    for xi in _gauss2():
        for eta in _gauss2():
            for zeta in _gauss2():
                dN = _shape_grads_unit(xi, eta, zeta)
                B = np.zeros((6, 24))
                for a in range(8):
                    B[0, 3 * a] = dN[0, a]
                    B[1, 3 * a + 1] = dN[1, a]
                    B[2, 3 * a + 2] = dN[2, a]
                    # engineering shear strains in the reference's row order
                    B[3, 3 * a] = dN[1, a]
                    B[3, 3 * a + 1] = dN[0, a]
                    B[4, 3 * a + 1] = dN[2, a]
                    B[4, 3 * a + 2] = dN[1, a]
                    B[5, 3 * a] = dN[2, a]
                    B[5, 3 * a + 2] = dN[0, a]
                K += 0.125 * B.T @ D @ B
    return K


def build_interior_dof_mask(Nx, Ny, Nz):
    """DOFs of nodes strictly inside the mesh bounding box
    (geometry_dvc.py:201-217)."""
    i = np.arange(Nx + 1)
    j = np.arange(Ny + 1)
    k = np.arange(Nz + 1)
    ii = (i > 0) & (i < Nx)
    jj = (j > 0) & (j < Ny)
    kk = (k > 0) & (k < Nz)
    node = (kk[:, None, None] & jj[None, :, None] & ii[None, None, :]).ravel()
    return np.repeat(node, 3)


# ---------------------------------------------------------------------------
# implicit connectivity (engine/kernels_dvc.py:32-89)
# ---------------------------------------------------------------------------

def elem_dofs(elem_ids, Nx, Ny):
    """(chunk, 24) DOFs of each element, node-major (kernels_dvc.py:32-61)."""
    i = elem_ids % Nx
    j = (elem_ids // Nx) % Ny
    k = elem_ids // (Nx * Ny)
    sj = Nx + 1
    sk = (Nx + 1) * (Ny + 1)
    base = i + j * sj + k * sk
    off = CORNERS[:, 0] + CORNERS[:, 1] * sj + CORNERS[:, 2] * sk
    nodes = base[:, None] + off[None, :]
    return (3 * nodes[:, :, None] + np.arange(3)[None, None, :]).reshape(-1, 24)


def voxel_offsets(h, Nx, Ny):
    """Flat offsets of the (h+1)^3 stencil voxels from an element's origin
    voxel, stencil order (kernels_dvc.py:64-78)."""
    Nvx = Nx * h + 1
    Nvy = Ny * h + 1
    k = np.arange(h + 1)
    qx, qy, qz = np.meshgrid(k, k, k, indexing='ij')
    return (qx + qy * Nvx + qz * Nvx * Nvy).ravel()


def elem_voxel_origin(elem_ids, h, Nx, Ny):
    """Flat index of each element's origin voxel (kernels_dvc.py:81-89)."""
    Nvx = Nx * h + 1
    Nvy = Ny * h + 1
    i = elem_ids % Nx
    j = (elem_ids // Nx) % Ny
    k = elem_ids // (Nx * Ny)
    return h * (i + j * Nvx + k * Nvx * Nvy)


def auto_chunk_size(nq, max_tensor_elems=20_000_000):
    """kernels_dvc.py:444-449."""
    return max(1, max_tensor_elems // (nq * 24))


def _chunks(Ne, chunk):
    for s in range(0, Ne, chunk):
        yield np.arange(s, min(s + chunk, Ne), dtype=np.int64)


# ---------------------------------------------------------------------------
# correlation Hessian, its diagonal and the RHS (kernels_dvc.py:452-543,
# correlate_gpu.py:341-378)
# ---------------------------------------------------------------------------

def matvec_H(v, grad_f_vox, mask_vox, h, Nx, Ny, Nz, chunk_size=None):
    """Hv = sum_e sum_q w_q m_q (v_eq . v_loc) v_eq, separable contraction
    (kernels_dvc.py:452-506).  Arithmetic in v.dtype."""
    dt = v.dtype
    Nst, wst = build_N_stencil(h)
    Nst = Nst.astype(dt)
    wst = wst.astype(dt)
    offs = voxel_offsets(h, Nx, Ny)
    chunk = chunk_size or auto_chunk_size(offs.size)
    out = np.zeros(v.shape[0], dtype=dt)
    for e in _chunks(Nx * Ny * Nz, chunk):
        dofs = elem_dofs(e, Nx, Ny)
        vox = elem_voxel_origin(e, h, Nx, Ny)[:, None] + offs[None, :]
        g = grad_f_vox[vox]
        m = mask_vox[vox]
        vl = v[dofs].reshape(-1, 8, 3)
        u = np.einsum('qa,cak->cqk', Nst, vl)
        s = wst[None, :] * m * np.einsum('cqk,cqk->cq', g, u)
        loc = np.einsum('qa,cqk->cak', Nst, s[:, :, None] * g).reshape(-1, 24)
        np.add.at(out, dofs.ravel(), loc.ravel())
    return out


def build_diagonal_H(grad_f_vox, mask_vox, h, Nx, Ny, Nz, chunk_size=None):
    """diag(H) = sum w m N^2 g^2 (kernels_dvc.py:509-543)."""
    dt = grad_f_vox.dtype
    Nst, wst = build_N_stencil(h)
    N2 = (Nst.astype(dt)) ** 2
    wst = wst.astype(dt)
    offs = voxel_offsets(h, Nx, Ny)
    chunk = chunk_size or auto_chunk_size(offs.size)
    Ndof = 3 * (Nx + 1) * (Ny + 1) * (Nz + 1)
    out = np.zeros(Ndof, dtype=dt)
    for e in _chunks(Nx * Ny * Nz, chunk):
        dofs = elem_dofs(e, Nx, Ny)
        vox = elem_voxel_origin(e, h, Nx, Ny)[:, None] + offs[None, :]
        g = grad_f_vox[vox]
        m = mask_vox[vox]
        t = (wst[None, :] * m)[:, :, None] * g ** 2
        loc = np.einsum('qa,cqk->cak', N2, t).reshape(-1, 24)
        np.add.at(out, dofs.ravel(), loc.ravel())
    return out


def compute_b(res_vox, mask_vox, grad_f_vox, h, Nx, Ny, Nz, chunk_size=None):
    """b = sum w m res v_eq (correlate_gpu.py:341-378)."""
    dt = grad_f_vox.dtype
    Nst, wst = build_N_stencil(h)
    Nst = Nst.astype(dt)
    wst = wst.astype(dt)
    offs = voxel_offsets(h, Nx, Ny)
    chunk = chunk_size or auto_chunk_size(offs.size)
    Ndof = 3 * (Nx + 1) * (Ny + 1) * (Nz + 1)
    out = np.zeros(Ndof, dtype=dt)
    for e in _chunks(Nx * Ny * Nz, chunk):
        dofs = elem_dofs(e, Nx, Ny)
        vox = elem_voxel_origin(e, h, Nx, Ny)[:, None] + offs[None, :]
        g = grad_f_vox[vox]
        s = wst[None, :] * mask_vox[vox] * res_vox[vox]
        loc = np.einsum('qa,cqk->cak', Nst, s[:, :, None] * g).reshape(-1, 24)
        np.add.at(out, dofs.ravel(), loc.ravel())
    return out


# ---------------------------------------------------------------------------
# regularizers (kernels_dvc.py:349-437)
# ---------------------------------------------------------------------------

def matvec_K_ref(v, K_ref, E_elem, Nx, Ny, Nz, chunk_size=500_000):
    """K v = sum_e E_e P_e^T K_ref P_e v (kernels_dvc.py:349-374)."""
    dt = v.dtype
    Ne = Nx * Ny * Nz
    K = np.asarray(K_ref, dtype=dt)
    E = np.broadcast_to(np.asarray(E_elem, dtype=dt), (Ne,))
    out = np.zeros(v.shape[0], dtype=dt)
    for e in _chunks(Ne, chunk_size):
        dofs = elem_dofs(e, Nx, Ny)
        loc = np.einsum('ij,kj->ki', K, v[dofs]) * E[e][:, None]
        np.add.at(out, dofs.ravel(), loc.ravel())
    return out


def build_diagonal_K_ref(K_ref, E_elem, Nx, Ny, Nz):
    """diag of the assembled K (kernels_dvc.py:377-392)."""
    d = np.diag(np.asarray(K_ref))
    Ne = Nx * Ny * Nz
    E = np.broadcast_to(np.asarray(E_elem), (Ne,))
    out = np.zeros(3 * (Nx + 1) * (Ny + 1) * (Nz + 1), dtype=d.dtype)
    for e in _chunks(Ne, 500_000):
        dofs = elem_dofs(e, Nx, Ny)
        np.add.at(out, dofs.ravel(), (E[e][:, None] * d[None, :]).ravel())
    return out


def matvec_equilibrium_gap(v, K_el, interior_mask, E_elem, Nx, Ny, Nz):
    """R_m v = K (P_i (K v)) (kernels_dvc.py:402-417)."""
    w = matvec_K_ref(v, K_el, E_elem, Nx, Ny, Nz)
    w = interior_mask * w
    return matvec_K_ref(w, K_el, E_elem, Nx, Ny, Nz)


def rademacher_probes(Ndof, n_probes=32, seed=0):
    """The reference's Hutchinson probe stream: rng.choice([-1,1], Ndof) per
    probe from default_rng(seed) (kernels_dvc.py:430-433)."""
    rng = np.random.default_rng(seed)
    return [rng.choice([-1.0, 1.0], size=Ndof) for _ in range(n_probes)]


def build_diagonal_equilibrium_gap(K_el, interior_mask, E_elem, Nx, Ny, Nz,
                                   n_probes=32, seed=0):
    """Hutchinson estimate of diag(R_m), clipped at 1e-12*max|.|
    (kernels_dvc.py:420-437).  Probes are float64, as in the reference."""
    Ndof = 3 * (Nx + 1) * (Ny + 1) * (Nz + 1)
    acc = np.zeros(Ndof, dtype=np.float64)
    for z in rademacher_probes(Ndof, n_probes, seed):
        acc = acc + z * matvec_equilibrium_gap(z, K_el, interior_mask, E_elem, Nx, Ny, Nz)
    d = acc / n_probes
    return np.clip(d, 1e-12 * float(np.abs(d).max()), None)


# ---------------------------------------------------------------------------
# PCG (kernels_dvc.py:550-583)
# ---------------------------------------------------------------------------

def pcg(matvec, b, M_inv, max_iter=1000, tol=1e-8):
    """Jacobi PCG from x0 = 0 with stop sqrt(r.r / (b.b + 1e-30)) < tol,
    checked after each update (kernels_dvc.py:550-583).  Returns (x, iters).
    Dtypes follow NumPy promotion exactly as in the reference: a float64
    M_inv makes z, p, x float64 even for float32 b."""
    x = np.zeros_like(b)
    r = b - matvec(x)
    z = M_inv * r
    p = z.copy()
    rz = np.dot(r, z)
    r0 = np.dot(b, b) + 1e-30
    it = 0
    for it in range(max_iter):
        Ap = matvec(p)
        alpha = rz / (np.dot(p, Ap) + 1e-30)
        x = x + alpha * p
        r = r - alpha * Ap
        if float(np.sqrt(np.dot(r, r) / r0)) < tol:
            return x, it + 1
        z = M_inv * r
        rz_new = np.dot(r, z)
        beta = rz_new / (rz + 1e-30)
        p = z + beta * p
        rz = rz_new
    return x, it + 1


# ---------------------------------------------------------------------------
# image setup and per-iteration voxel stages (correlate_gpu.py:49-378)
# ---------------------------------------------------------------------------

def image_grad_and_values(pix, dtype=np.float32):
    """F-ordered values and np.gradient (edge_order=1) per axis
    (correlate_gpu.py:49-72)."""
    a = np.asarray(pix, dtype=dtype)
    f_vox = np.ravel(a, order='F')
    g = np.empty((a.size, 3), dtype=dtype)
    for c in range(3):
        g[:, c] = np.ravel(np.gradient(a, axis=c), order='F')
    return f_vox.copy(), g


def erode_mask_6(mask_vox, Nvx, Nvy, Nvz):
    """6-neighbour erosion with edge replication ('nearest'), i.e. the
    reference's minimum_filter over the 3-D cross (correlate_gpu.py:75-94)."""
    m = mask_vox.reshape(Nvx, Nvy, Nvz, order='F')
    p = np.pad(m, 1, mode='edge')
    c = p[1:-1, 1:-1, 1:-1]
    out = (c & p[:-2, 1:-1, 1:-1] & p[2:, 1:-1, 1:-1] & p[1:-1, :-2, 1:-1]
           & p[1:-1, 2:, 1:-1] & p[1:-1, 1:-1, :-2] & p[1:-1, 1:-1, 2:])
    return np.ravel(out, order='F')


def masked_quadpoint_mean_std(field_vox, mask_vox, h, Nx, Ny, Nz, chunk_size=None):
    """Mean / population std over the multiplicity-weighted quadrature-point
    multiset restricted to the mask, float64 sums (correlate_gpu.py:113-157)."""
    offs = voxel_offsets(h, Nx, Ny)
    chunk = chunk_size or auto_chunk_size(offs.size)
    cnt = s1 = s2 = 0.0
    for e in _chunks(Nx * Ny * Nz, chunk):
        vox = elem_voxel_origin(e, h, Nx, Ny)[:, None] + offs[None, :]
        mf = mask_vox[vox].astype(np.float64)
        val = field_vox[vox].astype(np.float64) * mf
        cnt += mf.sum()
        s1 += val.sum()
        s2 += (val * val).sum()
    if cnt < 1:
        return 0.0, 0.0
    mean = s1 / cnt
    var = s2 / cnt - mean * mean
    return mean, (var if var > 0.0 else 0.0) ** 0.5


def trilinear_nearest(img, cx, cy, cz):
    """Order-1 interpolation of img[(x,y,z)] at coordinates already clipped to
    [0, n-1], clamping the upper neighbour ('nearest'), float64 weights --
    what map_coordinates(order=1, mode='nearest') computes for in-range
    coordinates."""
    nx, ny, nz = img.shape
    out_dt = img.dtype

    def split(c, n):
        c = c.astype(np.float64)
        i0 = np.floor(c).astype(np.int64)
        i0 = np.clip(i0, 0, n - 1)
        i1 = np.minimum(i0 + 1, n - 1)
        return i0, i1, c - i0

    x0, x1, fx = split(cx, nx)
    y0, y1, fy = split(cy, ny)
    z0, z1, fz = split(cz, nz)
    I = img.astype(np.float64)
    v = np.zeros(cx.shape, dtype=np.float64)
    for xi, wx in ((x0, 1 - fx), (x1, fx)):
        for yi, wy in ((y0, 1 - fy), (y1, fy)):
            for zi, wz in ((z0, 1 - fz), (z1, fz)):
                v += wx * wy * wz * I[xi, yi, zi]
    return v.astype(out_dt)


def sample_deformed_field(U, g_pix, h, Nx, Ny, Nz, g_origin=(0, 0, 0),
                          mask0_vox=None, glt=0.0, chunk_size=None):
    """g(x+u(x)) at every voxel plus mask = in_bounds & mask0 & (g > glt);
    in-bounds tested on the closed box [0, n-1] before clipping; later
    elements overwrite shared voxels (correlate_gpu.py:201-281)."""
    dt = U.dtype
    Nvx, Nvy, Nvz = Nx * h + 1, Ny * h + 1, Nz * h + 1
    Nvox = Nvx * Nvy * Nvz
    gnx, gny, gnz = g_pix.shape
    if mask0_vox is None:
        mask0_vox = np.ones(Nvox, dtype=bool)
    Nst, _ = build_N_stencil(h)
    Nst = Nst.astype(dt)
    offs = voxel_offsets(h, Nx, Ny)
    chunk = chunk_size or max(1, 4_000_000 // offs.size)
    g_vals = np.zeros(Nvox, dtype=dt)
    mask = np.zeros(Nvox, dtype=bool)
    for e in _chunks(Nx * Ny * Nz, chunk):
        dofs = elem_dofs(e, Nx, Ny)
        vox = elem_voxel_origin(e, h, Nx, Ny)[:, None] + offs[None, :]
        u = np.einsum('qa,kac->kqc', Nst, U[dofs].reshape(-1, 8, 3))
        vx = (vox % Nvx).astype(dt)
        vy = ((vox // Nvx) % Nvy).astype(dt)
        vz = (vox // (Nvx * Nvy)).astype(dt)
        cx = u[:, :, 0] + vx - dt.type(g_origin[0])
        cy = u[:, :, 1] + vy - dt.type(g_origin[1])
        cz = u[:, :, 2] + vz - dt.type(g_origin[2])
        inb = ((cx >= 0) & (cx <= gnx - 1) & (cy >= 0) & (cy <= gny - 1)
               & (cz >= 0) & (cz <= gnz - 1))
        cx = np.clip(cx, 0, gnx - 1)
        cy = np.clip(cy, 0, gny - 1)
        cz = np.clip(cz, 0, gnz - 1)
        gv = trilinear_nearest(g_pix.astype(dt, copy=False), cx, cy, cz)
        m = inb & mask0_vox[vox] & (gv > glt)
        g_vals[vox.ravel()] = gv.ravel()
        mask[vox.ravel()] = m.ravel()
    return g_vals, mask


# ---------------------------------------------------------------------------
# prealignment (correlate_gpu.py:381-567)
# ---------------------------------------------------------------------------

def fft_rigid_shift(f_pix, g_pix):
    """Integer shift d with g(x) ~ f(x - d) by phase correlation on the
    common corner crop (correlate_gpu.py:381-414)."""
    f = np.asarray(f_pix)
    g = np.asarray(g_pix)
    shape = tuple(min(a, b) for a, b in zip(f.shape, g.shape))
    sl = tuple(slice(0, n) for n in shape)
    F = np.fft.fftn(f[sl])
    G = np.fft.fftn(g[sl])
    R = F * np.conj(G)
    R = R / (np.abs(R) + 1e-12)
    r = np.fft.ifftn(R).real
    peak = np.unravel_index(np.argmax(r), r.shape)
    d = []
    for p, n in zip(peak, shape):
        p = int(p)
        if p > n // 2:
            p -= n
        d.append(-float(p))
    return np.array(d, dtype=np.float64)


def build_affine_matrix(p):
    """A = Rz Ry Rx diag(s) (correlate_gpu.py:436-448)."""
    rx, ry, rz, sx, sy, sz = p[:6]
    cx, snx = np.cos(rx), np.sin(rx)
    cy, sny = np.cos(ry), np.sin(ry)
    cz, snz = np.cos(rz), np.sin(rz)
    Rx = np.array([[1, 0, 0], [0, cx, -snx], [0, snx, cx]])
    Ry = np.array([[cy, 0, sny], [0, 1, 0], [-sny, 0, cy]])
    Rz = np.array([[cz, -snz, 0], [snz, cz, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx @ np.diag([sx, sy, sz])


def affine_cost(f_d, g_d, A, t, c):
    """1 - ZNCC of f against g sampled at A(x-c)+c+t over in-bounds samples
    (>= 1000 valid, else 2.0); correlate_gpu.py:417-433, 523-536."""
    nx, ny, nz = f_d.shape
    gnx, gny, gnz = g_d.shape
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing='ij')
    X = np.stack([ix.ravel(), iy.ravel(), iz.ravel()]).astype(np.float64)
    cc = np.asarray(c, dtype=np.float64).reshape(3, 1)
    Xp = np.asarray(A) @ (X - cc) + cc + np.asarray(t, dtype=np.float64).reshape(3, 1)
    inb = ((Xp[0] >= 0) & (Xp[0] <= gnx - 1) & (Xp[1] >= 0) & (Xp[1] <= gny - 1)
           & (Xp[2] >= 0) & (Xp[2] <= gnz - 1))
    if int(inb.sum()) < 1000:
        return 2.0
    gs = trilinear_nearest(g_d.astype(np.float64),
                           np.clip(Xp[0], 0, gnx - 1), np.clip(Xp[1], 0, gny - 1),
                           np.clip(Xp[2], 0, gnz - 1))
    fv = f_d.ravel().astype(np.float64)[inb]
    gv = gs[inb]
    fv = fv - fv.mean()
    gv = gv - gv.mean()
    den = np.sqrt((fv * fv).sum() * (gv * gv).sum())
    if den < 1e-8:
        return 2.0
    return 1.0 - float((fv * gv).sum() / den)


def affine_prealign(f_pix, g_pix, decimate=4, t0=None, angle_bound=0.3,
                    scale_bound=0.15, maxiter=200):
    """9-DOF Powell search of 1-ZNCC on stride-decimated volumes
    (correlate_gpu.py:451-567).  Returns (A, t_full, c_full, info)."""
    from scipy.optimize import minimize
    f_d = np.asarray(f_pix, dtype=np.float32)[::decimate, ::decimate, ::decimate]
    g_d = np.asarray(g_pix, dtype=np.float32)[::decimate, ::decimate, ::decimate]
    c = np.array([(n - 1) / 2.0 for n in f_d.shape])
    t0d = fft_rigid_shift(f_d, g_d) if t0 is None else np.asarray(t0, float) / decimate

    def cost(p):
        return affine_cost(f_d, g_d, build_affine_matrix(p), p[6:9], c)

    p_id = np.array([0., 0., 0., 1., 1., 1., 0., 0., 0.])
    p_tr = p_id.copy()
    p_tr[6:9] = t0d
    bounds = ([(-angle_bound, angle_bound)] * 3
              + [(1 - scale_bound, 1 + scale_bound)] * 3 + [(None, None)] * 3)
    res = minimize(cost, p_tr, method='Powell', bounds=bounds,
                   options={'maxiter': maxiter})
    A = build_affine_matrix(res.x)
    info = {'cost_identity': cost(p_id), 'cost_translation_only': cost(p_tr),
            'cost_final': float(res.fun), 'nfev': res.nfev, 'p': res.x}
    return A, res.x[6:9] * decimate, c * decimate, info


# ---------------------------------------------------------------------------
# Gauss-Newton drivers (correlate_gpu.py:828-1305)
# ---------------------------------------------------------------------------

def correlate(f_pix, g_pix, Nx, Ny, Nz, h, g_origin=(0, 0, 0), glt=0.0, l0=None,
              K_ref_laplacian=None, reg_exponent=2, c_reg=1.0, reg_type='laplacian',
              K_ref_elastic=None, eps_reg=1e-3, dynamic_mask=False,
              freeze_mask_after=None, U0=None, maxiter=30, eps=1e-3, pcg_tol=1e-6,
              pcg_max_iter=2000, disp=False, dtype=np.float32, log=None):
    """One-level GN solve (correlate_gpu.py:828-1140).  Returns (U, res, dU).
    `log`, when a list, receives one dict per GN iteration."""
    dt = np.dtype(dtype)
    Nvx, Nvy, Nvz = Nx * h + 1, Ny * h + 1, Nz * h + 1
    Ndof = 3 * (Nx + 1) * (Ny + 1) * (Nz + 1)
    Nvox = Nvx * Nvy * Nvz
    if Nvox > np.iinfo(np.int32).max or Ndof > np.iinfo(np.int32).max:
        raise ValueError("Nvox=%d / Ndof=%d exceed int32 range; matrix-free index "
                         "arrays assume int32 indices" % (Nvox, Ndof))
    f_x = np.asarray(f_pix, dtype=dt)
    g_x = np.asarray(g_pix, dtype=dt)
    f_vox, grad = image_grad_and_values(f_x, dt)
    mask0 = f_vox > glt
    grad[~erode_mask_6(mask0, Nvx, Nvy, Nvz)] = 0
    mean0, std0 = masked_quadpoint_mean_std(f_vox, mask0, h, Nx, Ny, Nz)
    f_vox -= float(mean0)

    if reg_type not in ('laplacian', 'equilibrium_gap'):
        raise ValueError("reg_type must be 'laplacian' or 'equilibrium_gap'")
    interior = None
    if reg_type == 'equilibrium_gap':
        interior = build_interior_dof_mask(Nx, Ny, Nz).astype(dt)

    def matvec_reg(v):
        if reg_type == 'laplacian':
            return matvec_K_ref(v, K_ref_laplacian, h, Nx, Ny, Nz)
        Rm = matvec_equilibrium_gap(v, K_ref_elastic, interior, h, Nx, Ny, Nz)
        return Rm + eps_reg * matvec_K_ref(v, K_ref_laplacian, h, Nx, Ny, Nz)

    diag_R_cache = []

    def diag_reg():
        if not diag_R_cache:
            if reg_type == 'laplacian':
                d = build_diagonal_K_ref(K_ref_laplacian, h, Nx, Ny, Nz)
            else:
                d = (build_diagonal_equilibrium_gap(K_ref_elastic, interior, h, Nx, Ny, Nz)
                     + eps_reg * build_diagonal_K_ref(K_ref_laplacian, h, Nx, Ny, Nz))
            diag_R_cache.append(d)
        return diag_R_cache[0]

    ll = 0.0
    if l0 is not None:
        if K_ref_laplacian is None:
            raise ValueError("K_ref_laplacian is required when l0 is given")
        if reg_type == 'equilibrium_gap' and K_ref_elastic is None:
            raise ValueError("K_ref_elastic is required when l0 is given and "
                             "reg_type='equilibrium_gap'")
        T = 10 * h
        if reg_type == 'equilibrium_gap':
            H0 = float(build_diagonal_H(grad, mask0, h, Nx, Ny, Nz).sum())
            L0 = float(diag_reg().sum())
        else:
            nodes = np.arange((Nx + 1) * (Ny + 1) * (Nz + 1))
            jj = (nodes // (Nx + 1)) % (Ny + 1)
            V = np.zeros(Ndof, dtype=dt)
            V[0::3] = np.cos(jj.astype(dt) * h / T * 2 * np.pi)
            H0 = float(np.dot(V, matvec_H(V, grad, mask0, h, Nx, Ny, Nz)))
            L0 = float(np.dot(V, matvec_reg(V)))
        ll = c_reg * (l0 / T) ** reg_exponent * H0 / L0

    U = np.zeros(Ndof, dtype=dt) if U0 is None else np.asarray(U0, dtype=dt).copy()
    res = np.zeros(Nvox, dtype=dt)
    dU = np.zeros(Ndof, dtype=dt)

    # No casts below: the reference's regularizer diagonal is float64 (its
    # K_ref is float64), so M_inv, the PCG vectors, dU and U are promoted to
    # float64 by NumPy even in the float32 solve (correlate_gpu.py:1042-1047,
    # 1113-1129); the restatement keeps that promotion.
    def jacobi(mask):
        d = build_diagonal_H(grad, mask, h, Nx, Ny, Nz)
        if l0 is not None:
            d = d + ll * diag_reg()
        return 1.0 / (d + 1e-5 * float(d.min()))

    def make_A(mask):
        def A(v):
            out = matvec_H(v, grad, mask, h, Nx, Ny, Nz)
            if l0 is not None:
                out = out + ll * matvec_reg(v)
            return out
        return A

    if not dynamic_mask:
        M_inv = jacobi(mask0)
        A = make_A(mask0)
    frozen = None
    for ik in range(maxiter):
        g_vals, mask_vox = sample_deformed_field(U, g_x, h, Nx, Ny, Nz, g_origin,
                                                 mask0, glt)
        if freeze_mask_after is not None and frozen is None and ik >= freeze_mask_after:
            frozen = mask_vox.copy()
        b_mask = frozen if frozen is not None else mask_vox
        g_mean, std1 = masked_quadpoint_mean_std(g_vals, b_mask, h, Nx, Ny, Nz)
        # python floats are weak scalars: they take the array's dtype
        g_vals -= float(g_mean)
        g_vals *= float(-(std0 / std1))
        g_vals += f_vox
        res = g_vals
        res[~b_mask] = 0
        b = compute_b(res, b_mask, grad, h, Nx, Ny, Nz)
        if l0 is not None:
            b = b - ll * matvec_reg(U)
        if dynamic_mask:
            M_inv = jacobi(b_mask)
            A = make_A(b_mask)
        dU, n_pcg = pcg(A, b, M_inv, max_iter=pcg_max_iter, tol=pcg_tol)
        U = U + dU
        err = float(np.linalg.norm(dU) / np.linalg.norm(U))
        if log is not None or disp:
            _, stdr = masked_quadpoint_mean_std(res, mask_vox, h, Nx, Ny, Nz)
            n_act = int(mask_vox.sum())
            if log is not None:
                log.append(dict(iter=ik + 1, std_res=stdr, err=err, active=n_act,
                                pcg_iters=n_pcg, ll=ll))
            if disp:
                print("Iter # %2d | std(res)=%2.2f gl | dU/U=%1.2e | active pts=%d/%d (%.1f%%)"
                      % (ik + 1, stdr, err, n_act, Nvox, 100 * n_act / Nvox))
        if err < eps:
            break
    return U, res, dU


def affine_nodal_U(A, t, c, s, h_i, Nx, Ny, Nz, dtype):
    """Uniform-affine nodal U0 at one pyramid level (correlate_gpu.py:1264-1281)."""
    c_i = np.asarray(c, float) / s
    t_i = np.asarray(t, float) / s
    n = np.arange((Nx + 1) * (Ny + 1) * (Nz + 1))
    X = np.stack([n % (Nx + 1), (n // (Nx + 1)) % (Ny + 1),
                  n // ((Nx + 1) * (Ny + 1))], axis=1).astype(np.float64) * h_i
    u = (X - c_i) @ np.asarray(A, float).T + c_i + t_i - X
    U = np.zeros(3 * n.size, dtype=dtype)
    for k in range(3):
        U[k::3] = u[:, k].astype(dtype)
    return U


def multiscale_correlate(f_pix, g_pix, Nx, Ny, Nz, h, scales=(2, 1, 0),
                         g_origin=(0, 0, 0), l0=None, fft_prealign=True,
                         prealign_affine=False, ext_affine=None, **kw):
    """Coarse-to-fine stride-decimation pyramid (correlate_gpu.py:1143-1305)."""
    scales = sorted(set(scales), reverse=True)
    dt = np.dtype(kw.get('dtype', np.float32))
    A = t = c = None
    if ext_affine is not None:
        A, t, c = (np.asarray(x, dtype=np.float64) for x in ext_affine)
    elif prealign_affine:
        A, t, c, _ = affine_prealign(f_pix, g_pix)
    U = None
    prev = None
    res = dU = None
    for isc in scales:
        s = 2 ** isc
        if h % s != 0:
            raise ValueError("h=%d not divisible by 2**%d=%d" % (h, isc, s))
        if any(o % s != 0 for o in g_origin):
            raise ValueError("g_origin=%r not divisible by 2**%d=%d for all "
                             "requested scales" % (g_origin, isc, s))
        h_i = h // s
        f_i = np.asarray(f_pix)[::s, ::s, ::s]
        g_i = np.asarray(g_pix)[::s, ::s, ::s]
        go_i = tuple(o // s for o in g_origin)
        l0_i = None if l0 is None else l0 / s
        if U is not None:
            U = U * (prev / s)
        elif A is not None:
            U = affine_nodal_U(A, t, c, s, h_i, Nx, Ny, Nz, dt)
        elif fft_prealign:
            d = fft_rigid_shift(f_i, g_i)
            U = np.zeros(3 * (Nx + 1) * (Ny + 1) * (Nz + 1), dtype=dt)
            for k in range(3):
                U[k::3] = d[k]
        U, res, dU = correlate(f_i, g_i, Nx, Ny, Nz, h_i, g_origin=go_i, l0=l0_i,
                               U0=U, **kw)
        prev = s
    return U, res, dU
