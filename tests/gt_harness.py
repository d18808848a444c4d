"""Ground-truth accuracy harness: a NumPy restatement of the reference's
ground_truth/prepare_ground_truth.py (the deformed image from a known
displacement) and the node-level scoring of ground_truth/
analyse_ground_truth.py (truncation-recoverable node RMS).  Test
infrastructure only."""
import json
import os

import numpy as np
from scipy.interpolate import RegularGridInterpolator


def prepare(ref, u_gt_vox):
    """deformed(x) = ref(x - u_gt(x)), trilinear with linear extrapolation
    outside (RegularGridInterpolator, bounds_error False, fill_value None),
    float64 (prepare_ground_truth.py:67-89)."""
    nx, ny, nz = ref.shape
    ax = [np.arange(n, dtype=float) for n in (nx, ny, nz)]
    interp = RegularGridInterpolator(ax, ref.astype(np.float64), method='linear',
                                     bounds_error=False, fill_value=None)
    X, Y, Z = np.meshgrid(*ax, indexing='ij')
    pts = np.column_stack([(X - u_gt_vox[..., 0]).ravel(), (Y - u_gt_vox[..., 1]).ravel(),
                           (Z - u_gt_vox[..., 2]).ravel()])
    return interp(pts).reshape(ref.shape)


def voxel_in_bounds(u_gt_vox, g_shape):
    """x + u_gt(x) inside g's closed box (analyse_ground_truth.py:108-129)."""
    nx, ny, nz = u_gt_vox.shape[:3]
    X, Y, Z = np.meshgrid(*[np.arange(n, dtype=float) for n in (nx, ny, nz)], indexing='ij')
    c = [X + u_gt_vox[..., 0], Y + u_gt_vox[..., 1], Z + u_gt_vox[..., 2]]
    ok = np.ones(X.shape, dtype=bool)
    for q, n in zip(c, g_shape):
        ok &= (q >= 0) & (q <= n - 1)
    return ok


def node_oob_fraction(u_gt_vox, h, nb, g_shape, crop_lo):
    """Per node, the out-of-bounds fraction of its support box [v-h, v+h]
    clamped to the image (analyse_ground_truth.py:132-165)."""
    nx, ny, nz = u_gt_vox.shape[:3]
    inb = voxel_in_bounds(u_gt_vox, g_shape)
    out = np.zeros(nb)
    for i in range(nb[0]):
        for j in range(nb[1]):
            for k in range(nb[2]):
                v = [crop_lo[0] + i * h, crop_lo[1] + j * h, crop_lo[2] + k * h]
                sl = tuple(slice(max(v[a] - h, 0), min(v[a] + h, n - 1) + 1)
                           for a, n in enumerate((nx, ny, nz)))
                out[i, j, k] = 1.0 - inb[sl].mean()
    return out


def recoverable_node_rms(work_dir, u_gt_vox, g_shape):
    """(#recoverable nodes, RMS over their 3 components, max |error|) of the
    recovered U_recovered.npy against the ground truth at the mesh nodes
    (analyse_ground_truth.py:204-300; U_recovered is shifted by crop_lo to the
    full-image frame)."""
    with open(os.path.join(work_dir, 'affine_prealign.json')) as fh:
        meta = json.load(fh)
    h = meta['h']
    crop_lo = np.array(meta['crop_lo'])
    U = np.load(os.path.join(work_dir, 'U_recovered.npy')) - crop_lo
    nb = U.shape[:3]
    gi = [crop_lo[a] + np.arange(nb[a]) * h for a in range(3)]
    err = U - u_gt_vox[np.ix_(gi[0], gi[1], gi[2])]
    rec = node_oob_fraction(u_gt_vox, h, nb, g_shape, crop_lo) == 0
    e = err[rec]
    return int(rec.sum()), float(np.sqrt(np.mean(e ** 2))), float(np.abs(e).max())
