"""Voxel-scale post-solve stages of a correlation run: drop-in equivalents of
the reference's src/voxeldvc/engine/write_output.py functions of the same
names (arguments, return values and layouts), computed by libvdvc.so.

  push_field_forward / push_reference_forward   write_output.py:571-622
  compute_zncc                                  write_output.py:658-690
  compute_correlation_quality_cell              write_output.py:380-447
  mesh_dims                                     write_output.py:228-230

The element-scale stages of write_outputs (principal strains, griddata
remaps, unsafe / overlap-lost element masks, .vti export) are not part of
this module.

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
