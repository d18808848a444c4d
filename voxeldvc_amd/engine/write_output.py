"""Voxel-scale post-solve stages of a correlation run: drop-in equivalents of
the reference's src/voxeldvc/engine/write_output.py functions of the same
names (arguments, return values and layouts), computed by libvdvc.so.

  push_field_forward / push_reference_forward   write_output.py:571-622
  compute_zncc                                  write_output.py:658-690
  compute_correlation_quality_cell              write_output.py:380-447
  mesh_dims                                     write_output.py:228-230
  compute_principal_strains_cell                write_output.py:256-281
  compute_equivalent_strain                     write_output.py:284-306
  compute_safe_cell_mask                        write_output.py:352-377
  compute_gn_convergence_cell                   write_output.py:450-491
  remap_element_mask_to_deformed                write_output.py:494-530
  remap_cell_field_to_deformed                  write_output.py:533-568
  compute_material_cell_mask                    write_output.py:309-338
  compute_material_cell_mask_eroded             write_output.py:341-349

(the unsafe / overlap-lost masks are in engine/kernels_dvc.py, as in the
reference).  Not provided: the .vti export.

Images are read in their own dtype (float32, float64, uint8, uint16, int16,
bool), as scipy.ndimage.map_coordinates reads them; pushes return the
field's dtype.  `xp=np` returns NumPy arrays; torch CUDA tensors in and
`xp=torch` keep the result on the device.
"""
import numpy as np

from .. import _lib as L
from .correlate_gpu import _nodal

_MODES = {'nearest': 0, 'constant': 1}


def mesh_dims(shape, h):
    """Largest N with N*h <= n-1 per axis (write_output.py:228-230)."""
    return tuple((int(s) - 1) // h for s in shape)


def _np_dtype(a):
    if L.is_tensor(a):
        return np.dtype(str(a.dtype).replace('torch.', ''))
    return np.asarray(a).dtype


def _grid_view(d, shape):
    """Dense x-fastest device array -> (Nvx, Nvy, Nvz) view [x][y][z]."""
    nx, ny, nz = shape
    return d.view(nz, ny, nx).permute(2, 1, 0)


def push_field_forward(field, U, Nx_e, Ny_e, Nz_e, h, order=1, mode='nearest', cval=0.0,
                       xp=np):
    """field_pushed(x) = field(x - u(x)) on the voxel grid, u the Hex8
    interpolation of U at voxel x, sampled like map_coordinates(order, mode,
    cval); (Nvx, Nvy, Nvz) in field's dtype (write_output.py:571-608).
    Spline orders 0 and 1 and modes 'nearest' and 'constant' (the ones the
    output stage uses) are supported."""
    if order not in (0, 1):
        raise ValueError('order must be 0 or 1, got %r' % (order,))
    if mode not in _MODES:
        raise ValueError("mode must be 'nearest' or 'constant', got %r" % (mode,))
    shape = tuple(int(s) for s in field.shape)
    want = (Nx_e * h + 1, Ny_e * h + 1, Nz_e * h + 1)
    if shape != want:
        raise ValueError('field shape %r does not match the voxel grid %r of the mesh'
                         % (shape, want))
    dU, _, code = _nodal(U, Nx_e, Ny_e, Nz_e)
    f, _, fcode = L.to_dev_image(field)
    out = L.torch().empty_like(f)
    L.call('vdvc_push_forward', L.ctx(), Nx_e, Ny_e, Nz_e, h, dU, code, f, fcode, int(order),
           _MODES[mode], float(cval), out)
    fdt = _np_dtype(field)
    if xp is not None and getattr(xp, '__name__', '') == 'torch':
        t = L.torch()
        if fdt == np.bool_:
            out = out.bool()
        elif fdt == np.uint16:
            out = out.view(t.uint16)
        return _grid_view(out, shape)
    return L.image_to_host(out, fdt).reshape(shape, order='F')


def push_reference_forward(ref, U, Nx_e, Ny_e, Nz_e, h, xp=np):
    """ref(x - u(x)), trilinear, mode 'nearest' (write_output.py:611-622)."""
    return push_field_forward(ref, U, Nx_e, Ny_e, Nz_e, h, order=1, xp=xp)


def compute_zncc(f_flat, g_warped_flat, mask_vox, xp=np):
    """ZNCC of f and g over mask_vox, float64, two-pass (mean removed
    first); nan for an empty or constant selection (write_output.py:658-690)."""
    f, fcode = L.to_dev_image_flat(f_flat)
    g, gcode = L.to_dev_image_flat(g_warped_flat)
    if f.numel() != g.numel():
        raise ValueError('f and g have different sizes')
    m = None if mask_vox is None else L.to_dev_flat(mask_vox, 'uint8')
    if m is not None and m.numel() != f.numel():
        raise ValueError('mask_vox size does not match f')
    out = (L._d * 1)()
    L.call('vdvc_zncc', L.ctx(), f.numel(), f, fcode, g, gcode, 0, m, out)
    return float(out[0])


def zncc_of_residual(f_flat, res, mask_vox):
    """compute_zncc(f, f - res, mask) without forming f - res: the global
    correlation figure write_outputs prints (write_output.py:784-789)."""
    f, fcode = L.to_dev_image_flat(f_flat)
    r, rcode = L.to_dev_image_flat(res)
    if f.numel() != r.numel():
        raise ValueError('f and res have different sizes')
    m = None if mask_vox is None else L.to_dev_flat(mask_vox, 'uint8')
    out = (L._d * 1)()
    L.call('vdvc_zncc', L.ctx(), f.numel(), f, fcode, r, rcode, 1, m, out)
    return float(out[0])


def _flat_f(a):
    """Voxel array in the reference's order: (Nvx, Nvy, Nvz) grids are
    raveled x fastest, flat arrays pass through (write_output.py:408-409)."""
    if L.is_tensor(a):
        return a.permute(2, 1, 0).reshape(-1) if a.dim() == 3 else a.reshape(-1)
    a = np.asarray(a)
    return a.reshape(-1, order='F') if a.ndim == 3 else a.reshape(-1)


def compute_correlation_quality_cell(res, ref, h, Nx, Ny, Nz):
    """(residual_std_cell, zncc_cell), each (Nx, Ny, Nz) float64: the
    population std of res and the ZNCC of (ref, ref - res) over each
    element's (h+1)^3 voxels with res != 0; nan below 8 active voxels (and
    ZNCC for zero variance) (write_output.py:380-447)."""
    t = L.require_gpu()
    rf = _flat_f(res)
    wd = 'float32' if _np_dtype(rf) == np.float32 else 'float64'  # others convert exactly
    r = L.to_dev_flat(rf, wd)
    f, fcode = L.to_dev_image_flat(_flat_f(ref))
    Nvox = (Nx * h + 1) * (Ny * h + 1) * (Nz * h + 1)
    if r.numel() != Nvox or f.numel() != Nvox:
        raise ValueError('res / ref must have Nvox = %d voxels' % Nvox)
    Ne = Nx * Ny * Nz
    std = t.empty(Ne, dtype=t.float64, device='cuda')
    zn = t.empty(Ne, dtype=t.float64, device='cuda')
    L.call('vdvc_quality_cell', L.ctx(), Nx, Ny, Nz, h, r, L.rtype_of(r), f, fcode, std, zn)
    if L.is_tensor(res):
        return _grid_view(std, (Nx, Ny, Nz)), _grid_view(zn, (Nx, Ny, Nz))
    return (L.to_host(std).reshape(Nx, Ny, Nz, order='F'),
            L.to_host(zn).reshape(Nx, Ny, Nz, order='F'))


# ---------------------------------------------------------------------------
# element-scale fields (write_output.py:252-306, 352-377, 450-568)
# ---------------------------------------------------------------------------

def _nodal_from_grid(U_grid):
    """(Nx+1, Ny+1, Nz+1, 3) nodal grid -> flat node-major device vector
    (DOF 3*(i + (Nx+1)*(j + (Ny+1)*k)) + c) and its type code."""
    t = L.require_gpu()
    if L.is_tensor(U_grid):
        g = U_grid.detach().to('cuda')
        if g.dtype not in (t.float32, t.float64):
            g = g.to(t.float64)
        flat = g.permute(2, 1, 0, 3).contiguous().view(-1)
    else:
        g = np.asarray(U_grid)
        if g.dtype not in (np.float32, np.float64):
            g = g.astype(np.float64)
        flat = t.from_numpy(np.ascontiguousarray(g.transpose(2, 1, 0, 3)).reshape(-1)).to('cuda')
    return flat, L.rtype_of(flat)


def _grid_dims(U_grid):
    s = tuple(int(x) for x in U_grid.shape)
    if len(s) != 4 or s[3] != 3 or min(s[:3]) < 2:
        raise ValueError('U_grid must have shape (Nx+1, Ny+1, Nz+1, 3), got %r' % (s,))
    return s[0] - 1, s[1] - 1, s[2] - 1


def _cell_out(d, like, Nx, Ny, Nz, tail=()):
    """Element-ordered (x fastest) device array -> (Nx, Ny, Nz, *tail)."""
    if L.is_tensor(like):
        return d.view(Nz, Ny, Nx, *tail).permute(2, 1, 0, *range(3, 3 + len(tail)))
    a = L.to_host(d).reshape((Nz, Ny, Nx) + tuple(tail))
    return np.ascontiguousarray(a.transpose((2, 1, 0) + tuple(range(3, 3 + len(tail)))))


def compute_principal_strains_cell(U_grid, h):
    """(Nx, Ny, Nz, 3) float64 principal small strains e1 >= e2 >= e3 at the
    element centres: Hex8 gradients at the centre, eps = (grad u + grad
    u^T) / 2, eigenvalues in float64 (write_output.py:256-281)."""
    t = L.require_gpu()
    Nx, Ny, Nz = _grid_dims(U_grid)
    u, code = _nodal_from_grid(U_grid)
    p = t.empty(Nx * Ny * Nz * 3, dtype=t.float64, device='cuda')
    L.call('vdvc_principal_strains', L.ctx(), Nx, Ny, Nz, int(h), u, code, p, None)
    return _cell_out(p, U_grid, Nx, Ny, Nz, (3,))


def compute_equivalent_strain(principal_strains):
    """sqrt(2)/3 * sqrt((e1-e2)^2 + (e2-e3)^2 + (e3-e1)^2) over the last
    axis (write_output.py:284-306)."""
    t = L.require_gpu()
    shp = tuple(principal_strains.shape)
    if not shp or shp[-1] != 3:
        raise ValueError('principal_strains must have a last axis of 3')
    if L.is_tensor(principal_strains):
        p = principal_strains.detach().to(device='cuda', dtype=t.float64).contiguous().view(-1)
    else:
        p = t.from_numpy(np.ascontiguousarray(principal_strains, dtype=np.float64).reshape(-1)).to('cuda')
    n = p.numel() // 3
    out = t.empty(n, dtype=t.float64, device='cuda')
    L.call('vdvc_equivalent_strain', L.ctx(), n, p, out)
    out = out.view(shp[:-1])
    return out if L.is_tensor(principal_strains) else L.to_host(out)


def compute_safe_cell_mask(unsafe_vox, h, Nx, Ny, Nz):
    """(Nx, Ny, Nz) bool: True where none of the element's (h+1)^3 voxels
    is unsafe (write_output.py:352-377)."""
    t = L.require_gpu()
    u = L.to_dev_flat(_flat_f(unsafe_vox), 'uint8')
    Nvox = (Nx * h + 1) * (Ny * h + 1) * (Nz * h + 1)
    if u.numel() != Nvox:
        raise ValueError('unsafe_vox has %d entries, expected Nvox = %d' % (u.numel(), Nvox))
    s = t.empty(Nx * Ny * Nz, dtype=t.uint8, device='cuda')
    L.call('vdvc_safe_cells', L.ctx(), Nx, Ny, Nz, int(h), u, s)
    return _cell_out(s, unsafe_vox, Nx, Ny, Nz).astype(bool) if not L.is_tensor(unsafe_vox) \
        else _cell_out(s, unsafe_vox, Nx, Ny, Nz).bool()


def compute_gn_convergence_cell(dU, U, Nx, Ny, Nz):
    """(Nx, Ny, Nz) float64 ||dU|| / ||U|| over each element's 24 corner
    DOFs, NaN where ||U|| = 0 (write_output.py:450-491)."""
    t = L.require_gpu()
    Ndof = 3 * (Nx + 1) * (Ny + 1) * (Nz + 1)

    def dev(a):
        if L.is_tensor(a):
            d = a.detach().to('cuda').reshape(-1)
            d = d if d.dtype in (t.float32, t.float64) else d.to(t.float64)
        else:
            arr = np.asarray(a).reshape(-1)
            arr = arr if arr.dtype in (np.float32, np.float64) else arr.astype(np.float64)
            d = t.from_numpy(np.ascontiguousarray(arr)).to('cuda')
        if d.numel() != Ndof:
            raise ValueError('expected %d DOFs, got %d' % (Ndof, d.numel()))
        return d.contiguous()
    dd, uu = dev(dU), dev(U)
    out = t.empty(Nx * Ny * Nz, dtype=t.float64, device='cuda')
    L.call('vdvc_gn_convergence', L.ctx(), Nx, Ny, Nz, dd, L.rtype_of(dd), uu, L.rtype_of(uu), out)
    return _cell_out(out, U, Nx, Ny, Nz)


def remap_element_mask_to_deformed(elem_mask, U_grid, h, crop_lo=(0, 0, 0)):
    """(Nx, Ny, Nz) bool element mask carried to the deformed configuration:
    centres (i + 0.5) h moved by the mean of their 8 corner displacements,
    and each regular centre + crop_lo takes the value of the nearest moved
    centre (exact float64 search through a uniform grid of cells; ties:
    lowest element index, x
    fastest) (write_output.py:494-530, scipy griddata 'nearest')."""
    t = L.require_gpu()
    Nx, Ny, Nz = _grid_dims(U_grid)
    if tuple(int(x) for x in elem_mask.shape) != (Nx, Ny, Nz):
        raise ValueError('elem_mask shape %r does not match U_grid' % (tuple(elem_mask.shape),))
    u, code = _nodal_from_grid(U_grid)
    if L.is_tensor(elem_mask):
        m = elem_mask.detach().to('cuda').permute(2, 1, 0).reshape(-1).to(t.uint8).contiguous()
    else:
        m = t.from_numpy(np.ascontiguousarray(np.asarray(elem_mask, dtype=bool).transpose(2, 1, 0))
                         .reshape(-1).astype(np.uint8)).to('cuda')
    out = t.empty(Nx * Ny * Nz, dtype=t.uint8, device='cuda')
    crop = (L._d * 3)(*[float(c) for c in crop_lo])
    L.call('vdvc_remap_mask_deformed', L.ctx(), Nx, Ny, Nz, int(h), u, code, crop, m, out)
    r = _cell_out(out, elem_mask, Nx, Ny, Nz)
    return r.bool() if L.is_tensor(elem_mask) else r.astype(bool)


def remap_cell_field_to_deformed(field, U_grid, h, crop_lo=(0, 0, 0)):
    """(Nx, Ny, Nz[, C]) float64 per-element field carried to the deformed
    configuration (write_output.py:533-568): element centres moved by the
    mean of their 8 corner displacements; each regular centre + crop_lo
    takes the linear interpolation of the field over the moved-centre
    lattice, or the value of the nearest moved centre outside it.  The
    reference interpolates over scipy's Delaunay tetrahedra of the moved
    centres; here each lattice cell is split into the 6 tetrahedra around
    its main diagonal (vdvc_remap_field_deformed).  Both interpolants are
    piecewise linear through the same values at the same points: they agree
    exactly on linear fields and to the field's second differences
    otherwise."""
    t = L.require_gpu()
    Nx, Ny, Nz = _grid_dims(U_grid)
    shp = tuple(int(x) for x in field.shape)
    if shp[:3] != (Nx, Ny, Nz) or len(shp) not in (3, 4):
        raise ValueError('field shape %r does not match U_grid (%d, %d, %d)' % (shp, Nx, Ny, Nz))
    C = shp[3] if len(shp) == 4 else 1
    u, code = _nodal_from_grid(U_grid)
    if L.is_tensor(field):
        f = field.detach().to(device='cuda', dtype=t.float64)
        f = (f.permute(2, 1, 0, 3) if len(shp) == 4 else f.permute(2, 1, 0)).contiguous().view(-1)
    else:
        a = np.asarray(field, dtype=np.float64)
        a = a.transpose(2, 1, 0, 3) if len(shp) == 4 else a.transpose(2, 1, 0)
        f = t.from_numpy(np.ascontiguousarray(a).reshape(-1)).to('cuda')
    out = t.empty(Nx * Ny * Nz * C, dtype=t.float64, device='cuda')
    crop = (L._d * 3)(*[float(c) for c in crop_lo])
    L.call('vdvc_remap_field_deformed', L.ctx(), Nx, Ny, Nz, int(h), u, code, crop, f, C, out)
    o = out.view(Nz, Ny, Nx, C).permute(2, 1, 0, 3)
    if len(shp) == 3:
        o = o[..., 0]
    return o if L.is_tensor(field) else L.to_host(o.contiguous())


def compute_material_cell_mask(ref, h, Nx, Ny, Nz, glt=0.0):
    """(Nx, Ny, Nz) bool: True where the mean reference gray level over the
    element's (h+1)^3 voxels exceeds glt (write_output.py:309-338).  ref is
    the (Nvx, Nvy, Nvz) reference volume in its own dtype; float64 block
    sums (exact for integer images)."""
    t = L.require_gpu()
    d, shp, code = L.to_dev_image(ref)
    want = (Nx * h + 1, Ny * h + 1, Nz * h + 1)
    if tuple(shp) != want:
        raise ValueError('ref shape %r does not match the voxel grid %r of the mesh'
                         % (tuple(shp), want))
    out = t.empty(Nx * Ny * Nz, dtype=t.uint8, device='cuda')
    L.call('vdvc_material_cells', L.ctx(), Nx, Ny, Nz, int(h), d, code, float(glt), out)
    r = _cell_out(out, ref, Nx, Ny, Nz)
    return r.bool() if L.is_tensor(ref) else r.astype(bool)


def compute_material_cell_mask_eroded(material_cell_mask):
    """One-cell erosion of an (Nx, Ny, Nz) element mask over the
    26-neighbourhood, the mesh border counting as material
    (write_output.py:341-349)."""
    t = L.require_gpu()
    Nx, Ny, Nz = (int(x) for x in material_cell_mask.shape)
    if L.is_tensor(material_cell_mask):
        m = material_cell_mask.detach().to('cuda').permute(2, 1, 0).reshape(-1).to(t.uint8).contiguous()
    else:
        m = t.from_numpy(np.ascontiguousarray(np.asarray(material_cell_mask, dtype=bool)
                                              .transpose(2, 1, 0)).reshape(-1).astype(np.uint8)).to('cuda')
    out = t.empty(Nx * Ny * Nz, dtype=t.uint8, device='cuda')
    L.call('vdvc_erode_cells', L.ctx(), Nx, Ny, Nz, m, out)
    r = _cell_out(out, material_cell_mask, Nx, Ny, Nz)
    return r.bool() if L.is_tensor(material_cell_mask) else r.astype(bool)
