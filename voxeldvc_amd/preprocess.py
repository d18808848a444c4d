"""Preprocess stages around the affine pre-alignment: drop-in equivalents of
the reference's src/voxeldvc/preprocess.py functions of the same names,
computed by libvdvc.so.

  warp_onto_ref_grid   preprocess.py:233-285
  compute_crop         preprocess.py:315-367

(The affine refinement itself, affine_refine_gn, is in
voxeldvc_amd.engine.correlate_gpu, as in the reference.)  The preprocess
command line (TIFF loading, binning, work-dir writing) is not part of this
package.
"""
import numpy as np

from . import _lib as L


def warp_onto_ref_grid(A, t, c, Nf, g_gpu, xp=np, chunk=30):
    """g sampled at x' = A (x - c) + c + t for every voxel x of the reference
    grid Nf, trilinear, 0 outside (mode 'constant'): float32 array of shape
    Nf (preprocess.py:233-285; the coordinates in float32, A (x - c) summed
    over j = 0, 1, 2 in order).  `chunk` is accepted for signature parity.
    NumPy out unless xp is torch (then a device (nx, ny, nz) view)."""
    g, gshp = L.to_dev_volume(g_gpu, 'float32')
    nx, ny, nz = (int(n) for n in Nf)
    out = L.empty(nx * ny * nz)
    A_ = np.ascontiguousarray(A, dtype=np.float64).reshape(9)
    t_ = np.ascontiguousarray(t, dtype=np.float64).reshape(3)
    c_ = np.ascontiguousarray(c, dtype=np.float64).reshape(3)
    L.call('vdvc_warp_affine', L.ctx(), nx, ny, nz, g, gshp[0], gshp[1], gshp[2],
           A_.ctypes.data_as(L._dp), t_.ctypes.data_as(L._dp), c_.ctypes.data_as(L._dp), out)
    if xp is not None and getattr(xp, '__name__', '') == 'torch':
        return out.view(nz, ny, nx).permute(2, 1, 0)
    return L.to_host(out).reshape(nx, ny, nz, order='F')


def compute_crop(f, g_warped, h, glt=0.0):
    """n*h+1 crop of the valid overlap region (f > glt and g_warped > glt):
    bounding box from the axis projections (on the GPU), each size snapped
    down to n*h+1, re-centred and shifted into f's bounds
    (preprocess.py:315-367).  Returns (lo, hi) int arrays, hi exclusive;
    RuntimeError when there is no overlap."""
    d, shp, code = L.to_dev_image(f)
    gw, gshp = L.to_dev_volume(g_warped, 'float32')
    if tuple(shp) != tuple(gshp):
        raise ValueError('f and g_warped must have the same shape')
    lo3 = (L._i * 3)()
    hi3 = (L._i * 3)()
    L.call('vdvc_overlap_bounds', L.ctx(), shp[0], shp[1], shp[2], d, code, gw, float(glt),
           lo3, hi3)
    if lo3[0] < 0:
        raise RuntimeError("No valid overlap region found")
    lo = np.array(list(lo3))
    hi = np.array(list(hi3))
    size = hi - lo + 1
    print(f"  valid overlap AABB : lo={lo}  hi={hi}  size={size}")
    snapped = ((size - 1) // h) * h + 1
    n_elems = (snapped - 1) // h
    print(f"  snapped size (h={h}): {snapped}  (n_elements = {n_elems})")
    mid = (lo + hi) // 2
    lo_out = mid - snapped // 2
    hi_out = lo_out + snapped
    Nf = np.array(shp)
    for i in range(3):
        if lo_out[i] < 0:
            s = -lo_out[i]
            lo_out[i] += s
            hi_out[i] += s
        if hi_out[i] > Nf[i]:
            s = hi_out[i] - Nf[i]
            lo_out[i] -= s
            hi_out[i] -= s
    return lo_out, hi_out
