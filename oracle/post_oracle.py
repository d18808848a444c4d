"""CPU oracle for the post-solve voxel stages of a correlation run (NumPy).

TEST INFRASTRUCTURE ONLY.  The HIP kernels of these stages are compared
against this module; nothing in the product package imports it.

Restated from ggalu/voxelDVC (paths relative to its src/voxeldvc/):
  * interp_field_vox          engine/correlate_gpu.py:160-198
  * compute_active_mask       engine/correlate_gpu.py:284-338
  * push_field_forward        engine/write_output.py:571-608
  * compute_zncc              engine/write_output.py:658-690
  * compute_correlation_quality_cell   engine/write_output.py:380-447
and the part of scipy.ndimage.map_coordinates those stages use (spline
order 0 and 1, modes 'nearest' and 'constant'), which the reference calls
through engine/correlate_gpu.py:41-46.

Parity pins (tests/test_post_oracle.py): map_coordinates_01 against
scipy.ndimage.map_coordinates itself (bit-exact, every dtype / order / mode,
out-of-range and tie coordinates); interp_field_vox against affine fields
(exact) and nodal values; push_field_forward against integer translations
(closed form) and linear ramps; compute_active_mask against the solver's
sampler (dvc_oracle.sample_deformed_field, pinned by the reference's golden
solves) and against an explicit scipy evaluation for integer images;
compute_zncc against np.corrcoef and the a*f+b invariance;
compute_correlation_quality_cell against a per-element loop over np.std /
np.corrcoef.
"""
import numpy as np

from .dvc_oracle import build_N_stencil, elem_dofs, elem_voxel_origin, voxel_offsets, _chunks


def interp_field_vox(U, h, Nx, Ny, Nz, channels_first=False, chunk_size=200_000):
    """Nodal field evaluated at every voxel with the voxel-centre Hex8
    stencil, later elements overwriting shared voxels (correlate_gpu.py:160-198).
    Returns (Nvox, 3), or (3, Nvox) with channels_first."""
    Nvox = (Nx * h + 1) * (Ny * h + 1) * (Nz * h + 1)
    Nst = build_N_stencil(h)[0].astype(U.dtype)
    offs = voxel_offsets(h, Nx, Ny)
    out = np.zeros((3, Nvox) if channels_first else (Nvox, 3), dtype=U.dtype)
    for e in _chunks(Nx * Ny * Nz, chunk_size):
        u_loc = U[elem_dofs(e, Nx, Ny)].reshape(-1, 8, 3)
        vox = (elem_voxel_origin(e, h, Nx, Ny)[:, None] + offs[None, :]).reshape(-1)
        vals = np.einsum('qa,kac->kqc', Nst, u_loc).reshape(-1, 3)
        if channels_first:
            out[:, vox] = vals.T
        else:
            out[vox] = vals
    return out


def _cast_like(v, dtype):
    """scipy.ndimage's output conversion: floats round to nearest, integers
    round half away from zero, bool is v != 0."""
    dtype = np.dtype(dtype)
    if dtype == np.bool_:
        return v != 0
    if dtype.kind in 'iu':
        r = np.where(v >= 0, v + 0.5, v - 0.5)
        return np.trunc(r).astype(dtype)
    return v.astype(dtype)


def map_coordinates_01(img, coords, order=1, mode='nearest', cval=0.0):
    """scipy.ndimage.map_coordinates for spline order 0 or 1 and mode
    'nearest' or 'constant', in its arithmetic order: coordinates as
    float64; 'constant' returns cval for a point outside [0, n-1] on any
    axis; order 0 takes index floor(c + 0.5), order 1 sums the 8 corners
    (last axis fastest) of ((value * wx) * wy) * wz with weights (1 - t, t),
    t = c - floor(c), the coordinate itself unclamped and each corner index
    clamped to [0, n-1] (so outside the box 'nearest' evaluates
    v0 (1 - t) + v0 t, not v0); the float64 result is converted to img's
    dtype."""
    if order not in (0, 1) or mode not in ('nearest', 'constant'):
        raise ValueError('order 0/1 and mode nearest/constant only')
    img = np.asarray(img)
    shp = img.shape
    cs = [np.asarray(c, dtype=np.float64) for c in coords]
    inside = np.ones(cs[0].shape, dtype=bool)
    for c, n in zip(cs, shp):
        inside &= (c >= 0) & (c <= n - 1)
    I = img.astype(np.float64)
    if order == 0:
        idx = [np.clip(np.floor(c + 0.5), 0, n - 1).astype(np.int64) for c, n in zip(cs, shp)]
        v = I[idx[0], idx[1], idx[2]] * 1.0
    else:
        lo, hi, w = [], [], []
        for c, n in zip(cs, shp):
            f0 = np.floor(c)
            t = c - f0
            lo.append(np.clip(f0, 0, n - 1).astype(np.int64))
            hi.append(np.clip(f0 + 1, 0, n - 1).astype(np.int64))
            w.append((1.0 - t, t))
        v = np.zeros(cs[0].shape, dtype=np.float64)
        for ax, ix in ((0, lo[0]), (1, hi[0])):
            for ay, iy in ((0, lo[1]), (1, hi[1])):
                for az, iz in ((0, lo[2]), (1, hi[2])):
                    v = v + ((I[ix, iy, iz] * w[0][ax]) * w[1][ay]) * w[2][az]
    if mode == 'constant':
        v = np.where(inside, v, np.float64(cval))
    return _cast_like(v, img.dtype)


def _voxel_ramps(Nvx, Nvy, Nvz, dtype):
    n = np.arange(Nvx * Nvy * Nvz)
    return ((n % Nvx).astype(dtype), ((n // Nvx) % Nvy).astype(dtype),
            (n // (Nvx * Nvy)).astype(dtype))


def compute_active_mask(U, g_pix, h, Nx, Ny, Nz, g_origin=(0, 0, 0), mask0_vox=None, glt=0.0):
    """mask0 & g.InBounds(x+u) & g(x+u) > glt, the sample taken in g's own
    dtype (correlate_gpu.py:284-338).  Coordinates (u + x) - origin in U's
    dtype; bounds on the closed box [0, n-1]; clipped; order-1 'nearest'."""
    Nvx, Nvy, Nvz = Nx * h + 1, Ny * h + 1, Nz * h + 1
    dt = U.dtype
    u = interp_field_vox(U, h, Nx, Ny, Nz)
    X, Y, Z = _voxel_ramps(Nvx, Nvy, Nvz, dt)
    gnx, gny, gnz = g_pix.shape
    cx = (u[:, 0] + X) - dt.type(g_origin[0])
    cy = (u[:, 1] + Y) - dt.type(g_origin[1])
    cz = (u[:, 2] + Z) - dt.type(g_origin[2])
    inb = ((cx >= 0) & (cx <= gnx - 1) & (cy >= 0) & (cy <= gny - 1)
           & (cz >= 0) & (cz <= gnz - 1))
    cx = np.clip(cx, 0, gnx - 1)
    cy = np.clip(cy, 0, gny - 1)
    cz = np.clip(cz, 0, gnz - 1)
    gv = map_coordinates_01(g_pix, (cx, cy, cz), order=1, mode='nearest')
    m = inb & (gv > glt)
    if mask0_vox is not None:
        m &= np.asarray(mask0_vox, dtype=bool)
    return m


def push_field_forward(field, U, Nx, Ny, Nz, h, order=1, mode='nearest', cval=0.0):
    """field(x - u(x)) on the voxel grid, (Nvx, Nvy, Nvz) in field's dtype
    (write_output.py:571-608)."""
    Nvx, Nvy, Nvz = field.shape
    u = interp_field_vox(U, h, Nx, Ny, Nz)
    X, Y, Z = _voxel_ramps(Nvx, Nvy, Nvz, u.dtype)
    coords = (X - u[:, 0], Y - u[:, 1], Z - u[:, 2])
    out = map_coordinates_01(field, coords, order=order, mode=mode, cval=cval)
    return out.reshape(Nvx, Nvy, Nvz, order='F')


def compute_zncc(f_flat, g_flat, mask_vox):
    """Zero-normalised cross-correlation over mask_vox in float64
    (write_output.py:658-690); nan for an empty or constant selection."""
    f = np.asarray(f_flat, dtype=np.float64)[mask_vox]
    g = np.asarray(g_flat, dtype=np.float64)[mask_vox]
    if f.size == 0:
        return float('nan')
    f = f - f.mean()
    g = g - g.mean()
    denom = np.sqrt(np.sum(f * f) * np.sum(g * g))
    if not denom > 0:
        return float('nan')
    return float(np.sum(f * g) / denom)


def compute_correlation_quality_cell(res, ref, h, Nx, Ny, Nz, min_active=8):
    """Per-element population std of res and ZNCC of (ref, ref - res) over
    the element's (h+1)^3 voxels with res != 0; nan below min_active active
    voxels or for zero variance (write_output.py:380-447).  (Nx, Ny, Nz)."""
    res_flat = np.asarray(res).reshape(-1, order='F').astype(np.float64)
    ref_flat = np.asarray(ref).reshape(-1, order='F').astype(np.float64)
    Ne = Nx * Ny * Nz
    vox = elem_voxel_origin(np.arange(Ne), h, Nx, Ny)[:, None] + voxel_offsets(h, Nx, Ny)[None, :]
    r = res_flat[vox]
    f = ref_flat[vox]
    g = f - r
    m = (r != 0.0).astype(np.float64)
    cnt = m.sum(axis=1)
    dc = np.maximum(cnt, 1.0)
    mean_r = (r * m).sum(axis=1) / dc
    var_r = (r * r * m).sum(axis=1) / dc - mean_r ** 2
    std = np.sqrt(np.clip(var_r, 0.0, None))
    fm = (f * m).sum(axis=1) / dc
    gm = (g * m).sum(axis=1) / dc
    fc = (f - fm[:, None]) * m
    gc = (g - gm[:, None]) * m
    num = (fc * gc).sum(axis=1)
    den = np.sqrt((fc * fc).sum(axis=1) * (gc * gc).sum(axis=1))
    with np.errstate(invalid='ignore', divide='ignore'):
        zn = num / den
    std[cnt < min_active] = np.nan
    zn[(cnt < min_active) | (den <= 0.0)] = np.nan
    return std.reshape(Nx, Ny, Nz, order='F'), zn.reshape(Nx, Ny, Nz, order='F')
