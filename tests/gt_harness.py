"""Ground-truth accuracy harness: a NumPy restatement of the reference's
ground_truth/prepare_ground_truth.py (the deformed image from a known
displacement) and the node-level scoring of ground_truth/
analyse_ground_truth.py (truncation-recoverable and valid node RMS).  Test
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


def _region(err, mask):
    """(#nodes, RMS over their 3 components, max node error norm)
    (analyse_ground_truth.py:280-286)."""
    if not np.any(mask):
        return 0, float('nan'), float('nan')
    e = err[mask]
    return int(mask.sum()), float(np.sqrt(np.mean(e ** 2))), float(np.linalg.norm(e, axis=-1).max())


def node_masked_out(work_dir, nb):
    """Nodes of any element flagged in overlap_lost_element_mask.npy
    (analyse_ground_truth.py:168-201); None if the file is absent."""
    path = os.path.join(work_dir, 'overlap_lost_element_mask.npy')
    if not os.path.exists(path):
        return None
    ol = np.load(path).astype(bool)
    Nx, Ny, Nz = ol.shape
    out = np.zeros(nb, dtype=bool)
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                out[di:di + Nx, dj:dj + Ny, dk:dk + Nz] |= ol
    return out


def node_error_regions(work_dir, u_gt_vox, g_shape):
    """{'recoverable': (n, rms, max), 'valid': (n, rms, max) or None} of the
    recovered U_recovered.npy against the ground truth at the mesh nodes
    (analyse_ground_truth.py:204-300; U_recovered is shifted by crop_lo to the
    full-image frame): recoverable = support box fully inside g; valid =
    recoverable and not a node of an overlap-lost element."""
    with open(os.path.join(work_dir, 'affine_prealign.json')) as fh:
        meta = json.load(fh)
    h = meta['h']
    crop_lo = np.array(meta['crop_lo'])
    U = np.load(os.path.join(work_dir, 'U_recovered.npy')) - crop_lo
    nb = U.shape[:3]
    gi = [crop_lo[a] + np.arange(nb[a]) * h for a in range(3)]
    err = U - u_gt_vox[np.ix_(gi[0], gi[1], gi[2])]
    rec = node_oob_fraction(u_gt_vox, h, nb, g_shape, crop_lo) == 0
    masked = node_masked_out(work_dir, nb)
    return {'recoverable': _region(err, rec),
            'valid': None if masked is None else _region(err, rec & ~masked)}


def recoverable_node_rms(work_dir, u_gt_vox, g_shape):
    """(#recoverable nodes, RMS, max node error norm)."""
    return node_error_regions(work_dir, u_gt_vox, g_shape)['recoverable']


def sweep(tmp_root, ref, u_gt_vox, combos):
    """The reference's h / l0_factor sweep (ground_truth/sweep_ground_truth.py:
    186-214): prepare once, then per h the preprocess stage and per
    l0_factor the run stage, each combination scored by node_error_regions.
    Returns a list of dicts (h, l0_factor, recoverable, valid)."""
    from voxeldvc_amd import preprocess
    from voxeldvc_amd.run_dvc import main as run_main
    deformed = prepare(ref, u_gt_vox)
    gin = os.path.join(tmp_root, 'gt_input')
    os.makedirs(gin, exist_ok=True)
    np.save(os.path.join(gin, 'ref.npy'), ref)
    np.save(os.path.join(gin, 'deformed.npy'), deformed)
    out = []
    done_pre = {}
    for h, l0f in combos:
        base = os.path.join(tmp_root, 'h%d' % h)
        if h not in done_pre:
            preprocess.main([os.path.join(gin, 'ref.npy'), os.path.join(gin, 'deformed.npy'),
                             '--h', str(h), '--output-dir', base])
            done_pre[h] = base
        wd = os.path.join(tmp_root, 'h%d_l0f%g' % (h, l0f))
        os.makedirs(wd, exist_ok=True)
        for name in ('affine_prealign.json', 'ref_preprocessed.npy', 'def_preprocessed.npy',
                     'def_ref_aligned.npy'):
            src = os.path.join(base, name)
            dst = os.path.join(wd, name)
            if not os.path.exists(dst):
                os.symlink(src, dst)
        run_main([wd, '--l0-factor', str(l0f)])
        reg = node_error_regions(wd, u_gt_vox, deformed.shape)
        out.append(dict(h=h, l0_factor=l0f, recoverable=reg['recoverable'], valid=reg['valid']))
    return out
