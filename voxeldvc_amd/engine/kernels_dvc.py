"""Matrix-free operators and PCG: drop-in equivalents of the reference's
src/voxeldvc/engine/kernels_dvc.py, computed by libvdvc.so.

Arrays may be NumPy (copied to the GPU and the result copied back) or torch
tensors on the GPU (used in place, result on the GPU).  `xp` is accepted for
signature compatibility; the computation always runs on the GPU.
"""
import numpy as np

from .. import _lib as L
from .geometry_dvc import CORNER_OFFSETS


# ---------------------------------------------------------------------------
# implicit connectivity (kernels_dvc.py:32-89) -- host index helpers
# ---------------------------------------------------------------------------

def get_elem_dofs(elem_ids, Nx, Ny, xp=np):
    """(chunk, 24) node-major DOFs of each element (kernels_dvc.py:32-61)."""
    e = np.asarray(elem_ids)
    i, j, k = e % Nx, (e // Nx) % Ny, e // (Nx * Ny)
    sj, sk = Nx + 1, (Nx + 1) * (Ny + 1)
    off = np.array([a + b * sj + c * sk for (a, b, c) in CORNER_OFFSETS], dtype=e.dtype)
    nodes = (i + j * sj + k * sk)[:, None] + off[None, :]
    return (3 * nodes[:, :, None] + np.arange(3, dtype=e.dtype)).reshape(-1, 24)


def get_voxel_offsets(h, Nx, Ny, xp=np):
    """Flat voxel offsets of the (h+1)^3 stencil points (kernels_dvc.py:64-78)."""
    Nvx, Nvy = Nx * h + 1, Ny * h + 1
    k = np.arange(h + 1, dtype=np.int32)
    a, b, c = np.meshgrid(k, k, k, indexing='ij')
    return (a + b * Nvx + c * Nvx * Nvy).ravel()


def get_elem_voxel_origins(elem_ids, h, Nx, Ny, xp=np):
    """Flat origin voxel of each element (kernels_dvc.py:81-89)."""
    e = np.asarray(elem_ids)
    Nvx, Nvy = Nx * h + 1, Ny * h + 1
    return h * (e % Nx + ((e // Nx) % Ny) * Nvx + (e // (Nx * Ny)) * Nvx * Nvy)


def auto_chunk_size(nq, max_tensor_elems=20_000_000):
    """Kept for signature compatibility (kernels_dvc.py:444-449); the device
    kernels stream voxels and need no element chunks."""
    return max(1, max_tensor_elems // (nq * 24))


# ---------------------------------------------------------------------------
# argument plumbing
# ---------------------------------------------------------------------------

def _is_f64(a):
    if a is None or type(a) in (float, int):   # Python scalars are weak
        return False
    if L.is_tensor(a):
        return a.dtype == L.torch().float64
    return np.asarray(a).dtype == np.float64


def work_dtype(*arrays):
    """The working type of an operator call, NumPy-style: float64 if any
    array (or NumPy scalar) input is float64 (the reference's dtype=float64
    mode), else float32; Python scalars do not promote.  Element matrices are
    applied in the working type of the vectors.  Returns (torch dtype name,
    library type code)."""
    return ('float64', L.F64) if any(_is_f64(a) for a in arrays) else ('float32', L.F32)


class _Out:
    """Remembers whether the caller passed host or device arrays."""

    def __init__(self, *arrays):
        self.device = any(L.is_tensor(a) for a in arrays if a is not None)

    def __call__(self, t):
        return t if self.device else L.to_host(t)


def level_shape(h, Nx, Ny, Nz):
    return (Nx * h + 1, Ny * h + 1, Nz * h + 1)


def grad_planes(grad_f_vox, shape, dtype='float32'):
    """(Nvox, 3) gradient (reference layout) -> three pitched device planes of
    `dtype`."""
    g = L.to_dev_flat(grad_f_vox, dtype).view(-1, 3)
    nx, ny, nz = shape
    p = L.pitch_of(nx)
    out = []
    for c in range(3):
        o = L.empty(p * ny * nz, dtype)
        # column c of the (Nvox, 3) array: element stride 3
        L.call('vdvc_copy_volume', L.ctx(), nx, ny, nz, g[:, c], L.type_code(g.dtype), 3,
               3 * nx, 3 * nx * ny, o, L.type_code(g.dtype), p)
        out.append(o)
    return out


def _mask_u8(mask_vox, shape):
    return None if mask_vox is None else L.to_dev_level(mask_vox, shape, 'uint8')


def _check_stencil(N_stencil, h):
    if N_stencil is not None and np.shape(N_stencil) != ((h + 1) ** 3, 8):
        raise ValueError('N_stencil shape %r does not match h=%d' % (np.shape(N_stencil), h))


def _kptr(K):
    a = np.ascontiguousarray(np.asarray(K, dtype=np.float64))
    if a.shape != (24, 24):
        raise ValueError('K_ref must be 24x24, got %r' % (a.shape,))
    return a


def _E(E_elem, Nx, Ny, Nz, dtype='float32'):
    """Scalar E, or a device per-element array of `dtype`."""
    a = np.asarray(E_elem) if not L.is_tensor(E_elem) else None
    if a is not None and a.ndim == 0:
        return float(a), None
    d = L.to_dev_flat(E_elem, dtype)
    if d.numel() == 1:
        return float(d.item()), None
    if d.numel() != Nx * Ny * Nz:
        raise ValueError('E_elem must be a scalar or have Nx*Ny*Nz entries')
    return 0.0, d


# ---------------------------------------------------------------------------
# correlation Hessian (kernels_dvc.py:452-543)
# ---------------------------------------------------------------------------

def matvec_H(v, grad_f_vox, mask0_vox, N_stencil, w_stencil, h, Nx, Ny, Nz, xp=np,
             chunk_size=None):
    """f = H @ v (kernels_dvc.py:452-506)."""
    _check_stencil(N_stencil, h)
    out = _Out(v, grad_f_vox)
    dt, code = work_dtype(v, grad_f_vox)
    shp = level_shape(h, Nx, Ny, Nz)
    gx, gy, gz = grad_planes(grad_f_vox, shp, dt)
    dv = L.to_dev_flat(v, dt)
    res = L.empty(dv.numel(), dt)
    m = _mask_u8(mask0_vox, shp)
    L.call('vdvc_matvec_H', L.ctx(), Nx, Ny, Nz, h, L.ptr(gx), L.ptr(gy), L.ptr(gz), code,
           L.ptr(m), L.ptr(dv), L.ptr(res))
    return out(res)


def build_diagonal_H(grad_f_vox, mask0_vox, N_stencil, w_stencil, h, Nx, Ny, Nz, Ndof, xp=np,
                     chunk_size=None):
    """diag(H) (kernels_dvc.py:509-543)."""
    _check_stencil(N_stencil, h)
    out = _Out(grad_f_vox)
    dt, code = work_dtype(grad_f_vox)
    shp = level_shape(h, Nx, Ny, Nz)
    gx, gy, gz = grad_planes(grad_f_vox, shp, dt)
    res = L.empty(Ndof, dt)
    m = _mask_u8(mask0_vox, shp)
    L.call('vdvc_diag_H', L.ctx(), Nx, Ny, Nz, h, L.ptr(gx), L.ptr(gy), L.ptr(gz), code,
           L.ptr(m), L.ptr(res))
    return out(res)


# ---------------------------------------------------------------------------
# regularizers (kernels_dvc.py:349-437)
# ---------------------------------------------------------------------------

def matvec_K_ref(v, K_ref, E_elem, Nx, Ny, Nz, xp=np, chunk_size=500_000):
    """f = sum_e E_e P_e^T K_ref P_e v (kernels_dvc.py:349-374)."""
    out = _Out(v, E_elem)
    dt, code = work_dtype(v, E_elem)
    K = _kptr(K_ref)
    E, dE = _E(E_elem, Nx, Ny, Nz, dt)
    dv = L.to_dev_flat(v, dt)
    res = L.empty(dv.numel(), dt)
    L.call('vdvc_matvec_K', L.ctx(), Nx, Ny, Nz, code, K.ctypes.data_as(L._dp), E, L.ptr(dE),
           L.ptr(dv), L.ptr(res))
    return out(res)


def build_diagonal_K_ref(K_ref, E_elem, Nx, Ny, Nz, Ndof, xp=np):
    """Diagonal of the assembled K (kernels_dvc.py:377-392)."""
    out = _Out(E_elem)
    dt, code = work_dtype(E_elem)
    K = _kptr(K_ref)
    E, dE = _E(E_elem, Nx, Ny, Nz, dt)
    res = L.empty(Ndof, dt)
    L.call('vdvc_diag_K', L.ctx(), Nx, Ny, Nz, code, K.ctypes.data_as(L._dp), E, L.ptr(dE),
           L.ptr(res))
    return out(res)


def matvec_equilibrium_gap(v, K_ref_elastic, interior_mask, E_elem, Nx, Ny, Nz, xp=np,
                           chunk_size=500_000):
    """f = K (P_i (K v)) (kernels_dvc.py:402-417).  `interior_mask` must be the
    interior-node DOF mask of the mesh (geometry_dvc.build_interior_dof_mask);
    the kernel evaluates it from node indices."""
    out = _Out(v, E_elem)
    dt, code = work_dtype(v, E_elem)
    _check_interior(interior_mask, Nx, Ny, Nz)
    K = _kptr(K_ref_elastic)
    E, dE = _E(E_elem, Nx, Ny, Nz, dt)
    dv = L.to_dev_flat(v, dt)
    res = L.empty(dv.numel(), dt)
    L.call('vdvc_matvec_eqgap', L.ctx(), Nx, Ny, Nz, code, K.ctypes.data_as(L._dp), E,
           L.ptr(dE), L.ptr(dv), L.ptr(res))
    return out(res)


def _check_interior(interior_mask, Nx, Ny, Nz):
    if interior_mask is None:
        return
    m = np.asarray(L.to_host(interior_mask) if L.is_tensor(interior_mask) else interior_mask)
    from .geometry_dvc import build_interior_dof_mask
    if m.shape != (3 * (Nx + 1) * (Ny + 1) * (Nz + 1),) or not np.array_equal(
            m.astype(bool), build_interior_dof_mask(Nx, Ny, Nz)):
        raise ValueError('interior_mask must be build_interior_dof_mask(Nx, Ny, Nz)')


def rademacher_probes(Ndof, n_probes=32, seed=0):
    """The reference's probe stream: rng.choice([-1, 1], Ndof) per probe from
    default_rng(seed) (kernels_dvc.py:430-433), as int8 (n_probes, Ndof)."""
    rng = np.random.default_rng(seed)
    return np.stack([rng.choice([-1.0, 1.0], size=Ndof) for _ in range(n_probes)]).astype(np.int8)


def build_diagonal_equilibrium_gap(K_ref_elastic, interior_mask, E_elem, Nx, Ny, Nz, Ndof,
                                   xp=np, n_probes=32, seed=0):
    """Hutchinson estimate of diag(K P_i K), clipped at 1e-12 max
    (kernels_dvc.py:420-437); probes from the reference's RNG stream."""
    out = _Out(E_elem)
    dt, code = work_dtype(E_elem)
    _check_interior(interior_mask, Nx, Ny, Nz)
    K = _kptr(K_ref_elastic)
    E, dE = _E(E_elem, Nx, Ny, Nz, dt)
    if dE is not None:
        raise ValueError('per-element E_elem is not supported by the Hutchinson diagonal')
    pr = L.to_dev_flat(rademacher_probes(Ndof, n_probes, seed).ravel(), 'int8')
    res = L.empty(Ndof, dt)
    L.call('vdvc_hutchinson_eqgap', L.ctx(), Nx, Ny, Nz, code, K.ctypes.data_as(L._dp), E,
           L.ptr(pr), n_probes, L.ptr(res))
    return out(res)


# ---------------------------------------------------------------------------
# PCG (kernels_dvc.py:550-583)
# ---------------------------------------------------------------------------

def dot(a, b):
    """FP64 dot product of two device vectors of one working type."""
    if a.dtype != b.dtype:
        raise TypeError('dot of %s and %s vectors' % (a.dtype, b.dtype))
    r = (L._d * 1)()
    L.call('vdvc_dot', L.ctx(), int(a.numel()), L.rtype_of(a), L.ptr(a), L.ptr(b), r)
    return float(r[0])


def axpby(a, x, b, y):
    """y = a x + b y in place (device vectors of one working type)."""
    if x.dtype != y.dtype:
        raise TypeError('axpby of %s and %s vectors' % (x.dtype, y.dtype))
    L.call('vdvc_axpby', L.ctx(), int(y.numel()), L.rtype_of(y), float(a), L.ptr(x), float(b),
           L.ptr(y))
    return y


def pcg(matvec, b, M_inv, xp=np, x0=None, max_iter=1000, tol=1e-8, verbose=False):
    """Jacobi PCG with a caller-supplied matvec (kernels_dvc.py:550-583).
    Vector updates and FP64 dots run on the GPU in the working type of b and
    M_inv; `matvec` receives and must return device tensors of that type."""
    out = _Out(b, M_inv)
    t = L.require_gpu()
    dt, code = work_dtype(b, M_inv)
    bd = L.to_dev_flat(b, dt)
    Md = L.to_dev_flat(M_inv, dt)
    x = L.zeros(bd.numel(), dt) if x0 is None else L.clone(L.to_dev_flat(x0, dt))
    r = L.clone(bd)
    axpby(-1.0, L.to_dev_flat(matvec(x), dt), 1.0, r)
    z = t.empty_like(r)
    L.call('vdvc_mul', L.ctx(), int(r.numel()), code, L.ptr(Md), L.ptr(r), L.ptr(z))
    p = L.clone(z)
    rz = dot(r, z)
    r0 = dot(bd, bd) + 1e-30
    for it in range(max_iter):
        Ap = L.to_dev_flat(matvec(p), dt)
        alpha = rz / (dot(p, Ap) + 1e-30)
        axpby(alpha, p, 1.0, x)
        axpby(-alpha, Ap, 1.0, r)
        res = (dot(r, r) / r0) ** 0.5
        if verbose and (it % 10 == 0 or it < 5):
            print(f"  PCG iter {it:4d}  rel.residual = {res:.3e}")
        if res < tol:
            if verbose:
                print(f"  Converged at iter {it}, residual = {res:.3e}")
            break
        L.call('vdvc_mul', L.ctx(), int(r.numel()), code, L.ptr(Md), L.ptr(r), L.ptr(z))
        rz_new = dot(r, z)
        beta = rz_new / (rz + 1e-30)
        axpby(1.0, z, beta, p)
        rz = rz_new
    return out(x)


# ---------------------------------------------------------------------------
# unsafe / overlap-lost masks of the output stage (kernels_dvc.py:92-340)
# ---------------------------------------------------------------------------

def boundary_element_mask(Nx, Ny, Nz):
    """(Nx*Ny*Nz,) bool, elements x fastest: True on the outer one-element
    shell (kernels_dvc.py:92-115)."""
    m = np.zeros((Nz, Ny, Nx), dtype=bool)
    m[[0, -1], :, :] = True
    m[:, [0, -1], :] = True
    m[:, :, [0, -1]] = True
    return m.ravel()


def _cell_masks(active_vox, h, Nx, Ny, Nz, mask0_vox=None, frac_threshold=0.0,
                min_material_frac=0.5, elem=False, node=False, vox=False, lost=False):
    """One vdvc_cell_masks call; returns the requested outputs (device
    uint8 tensors) in the order element, node, voxel, overlap-lost."""
    t = L.require_gpu()
    Nvox = (Nx * h + 1) * (Ny * h + 1) * (Nz * h + 1)
    a = L.to_dev_flat(active_vox, 'uint8')
    if a.numel() != Nvox:
        raise ValueError('active_vox has %d entries, expected Nvox = %d' % (a.numel(), Nvox))
    m0 = None
    if lost:
        m0 = L.to_dev_flat(mask0_vox, 'uint8')
        if m0.numel() != Nvox:
            raise ValueError('mask0_vox has %d entries, expected Nvox = %d' % (m0.numel(), Nvox))
    Ne, Nn = Nx * Ny * Nz, (Nx + 1) * (Ny + 1) * (Nz + 1)
    mk = lambda on, n: t.empty(n, dtype=t.uint8, device='cuda') if on else None
    oe, on_, ov, ol = mk(elem, Ne), mk(node, Nn), mk(vox, Nvox), mk(lost, Ne)
    L.call('vdvc_cell_masks', L.ctx(), Nx, Ny, Nz, h, a, m0, float(frac_threshold),
           float(min_material_frac), oe, on_, ov, ol)
    return [x for x in (oe, on_, ov, ol) if x is not None]


def _bool_out(d, like):
    return d.bool() if L.is_tensor(like) else L.to_host(d).astype(bool)


def compute_unsafe_voxel_mask(active_vox, h, Nx, Ny, Nz):
    """(Nvox,) bool: every voxel of an element that shares a node with an
    element holding an inactive voxel, plus the outer element shell's voxels
    (kernels_dvc.py:166-193)."""
    (v,) = _cell_masks(active_vox, h, Nx, Ny, Nz, vox=True)
    return _bool_out(v, active_vox)


def compute_unsafe_element_mask(active_vox, h, Nx, Ny, Nz):
    """(Nx*Ny*Nz,) bool, elements x fastest (kernels_dvc.py:196-213)."""
    (e,) = _cell_masks(active_vox, h, Nx, Ny, Nz, elem=True)
    return _bool_out(e, active_vox)


def compute_unsafe_node_mask(active_vox, h, Nx, Ny, Nz):
    """((Nx+1)*(Ny+1)*(Nz+1),) bool: nodes of elements holding an inactive
    voxel and of the shell elements (kernels_dvc.py:216-233)."""
    (n,) = _cell_masks(active_vox, h, Nx, Ny, Nz, node=True)
    return _bool_out(n, active_vox)


def compute_overlap_lost_element_mask(active_vox, mask0_vox, h, Nx, Ny, Nz,
                                      frac_threshold=0.0, min_material_frac=0.5):
    """(Nx*Ny*Nz,) bool, elements x fastest: shares a node with an element
    whose material lost overlap (lost / material > frac_threshold), or has
    a material count fraction below min_material_frac, or lies on the outer
    shell (kernels_dvc.py:236-340)."""
    (o,) = _cell_masks(active_vox, h, Nx, Ny, Nz, mask0_vox, frac_threshold, min_material_frac,
                       lost=True)
    return _bool_out(o, active_vox)
