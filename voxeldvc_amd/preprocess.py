"""Preprocess stages around the affine pre-alignment: drop-in equivalents of
the reference's src/voxeldvc/preprocess.py functions of the same names,
computed by libvdvc.so.

  warp_onto_ref_grid   preprocess.py:233-285
  compute_crop         preprocess.py:315-367

  main                 preprocess.py:372-700 for .npy volumes: affine
                       pre-alignment (decimated ZNCC/Powell), full-resolution
                       affine refinement, warp, overlap crop, the deformed
                       bounding box, and the work-dir files run reads
                       (ref_preprocessed.npy, def_preprocessed.npy,
                       def_ref_aligned.npy, affine_prealign.json)

(The affine refinement itself, affine_refine_gn, is in
voxeldvc_amd.engine.correlate_gpu, as in the reference.)  TIFF inputs are not
read: volumes come as .npy files.

    python -m voxeldvc_amd preprocess REF.npy DEF.npy --output-dir DIR [--h 8]
        [--glt 0] [--bin 1] [--decimate 4] [--crop X Y Z] [--no-affine-refine]
"""
import argparse
import json
import os

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


def rotation_angles_deg(A):
    """Euler angles (deg, xyz) of the rotation of A's polar decomposition and
    the diagonal of its right stretch tensor (preprocess.py:199-219)."""
    U, _, Vt = np.linalg.svd(A)
    R = U @ Vt
    if np.linalg.det(R) < 0:
        R = R.copy()
        R[:, -1] *= -1
    rx = np.degrees(np.arctan2(R[2, 1], R[2, 2]))
    ry = np.degrees(np.arctan2(-R[2, 0], np.hypot(R[2, 1], R[2, 2])))
    rz = np.degrees(np.arctan2(R[1, 0], R[0, 0]))
    return (rx, ry, rz), np.diag(R.T @ A).copy()


def normal_strains(A):
    """diag(I - A^-1): per-axis normal strain in the reference frame
    (preprocess.py:222-230)."""
    return np.diag(np.eye(3) - np.linalg.inv(A))


def bin_volume(vol, B):
    """Average-bin by B per axis into float32 (preprocess.py:288-312)."""
    S0, S1, S2 = vol.shape
    out = np.zeros((S0 // B, S1 // B, S2 // B), dtype=np.float32)
    for i in range(B):
        for j in range(B):
            for k in range(B):
                out += vol[i::B, j::B, k::B]
    out /= B ** 3
    return out


def parse_args(argv=None):
    p = argparse.ArgumentParser(prog='voxeldvc_amd preprocess')
    p.add_argument('ref_path', type=str)
    p.add_argument('def_path', type=str)
    p.add_argument('--h', type=int, default=8, dest='h')
    p.add_argument('--glt', type=float, default=0.0, dest='glt')
    p.add_argument('--bin', type=int, default=1, dest='bin_factor')
    p.add_argument('--decimate', type=int, default=4, dest='decimate')
    p.add_argument('--no-affine-refine', action='store_true', dest='no_affine_refine')
    p.add_argument('--output-dir', type=str, required=True, dest='output_dir')
    p.add_argument('--crop', type=int, nargs=3, default=[0, 0, 0], dest='crop')
    return p.parse_args(argv)


def main(argv=None):
    """preprocess.py:372-700 for .npy volumes."""
    from .engine.correlate_gpu import affine_prealign, affine_refine_gn
    args = parse_args(argv)
    h, glt, B = args.h, args.glt, args.bin_factor
    if h < 1 or B < 1 or args.decimate < 1 or any(c < 0 for c in args.crop):
        raise ValueError('invalid --h / --bin / --decimate / --crop')
    out_dir = os.path.expanduser(args.output_dir)
    ref = np.load(args.ref_path)
    deformed = np.load(args.def_path)
    ref_dtype, def_dtype = ref.dtype, deformed.dtype
    print(f"  ref:      shape={ref.shape}  dtype={ref.dtype}")
    print(f"  deformed: shape={deformed.shape}  dtype={deformed.dtype}")
    if B > 1:
        rt = tuple((s // B) * B for s in ref.shape)
        dt = tuple((s // B) * B for s in deformed.shape)
        ref = bin_volume(ref[:rt[0], :rt[1], :rt[2]], B)
        deformed = bin_volume(deformed[:dt[0], :dt[1], :dt[2]], B)
    if any(cv > 0 for cv in args.crop):
        lo_c, hi_c = [0, 0, 0], list(ref.shape)
        for i, cv in enumerate(args.crop):
            if cv > 0:
                lo_c[i] = cv // 2
                hi_c[i] -= cv - cv // 2
                if lo_c[i] >= hi_c[i]:
                    raise ValueError('crop value %d for axis %d exceeds dimension %d'
                                     % (cv, i, ref.shape[i]))
        ref = ref[lo_c[0]:hi_c[0], lo_c[1]:hi_c[1], lo_c[2]:hi_c[2]]
        deformed = deformed[lo_c[0]:hi_c[0], lo_c[1]:hi_c[1], lo_c[2]:hi_c[2]]
    f_pix = ref.astype(np.float32)
    g_pix = deformed.astype(np.float32)
    print(f"\nRunning affine_prealign (decimate={args.decimate}) ...")
    A, t, c, info = affine_prealign(f_pix, g_pix, np, decimate=args.decimate)
    refine_info = None
    if not args.no_affine_refine:
        print("\nRefining affine by full-resolution Gauss-Newton (coarse-to-fine) ...")
        A, t, refine_info = affine_refine_gn(f_pix, g_pix, np, A, t, c)
        print(f"  refine: rms {refine_info['rms_initial']:.4f} -> "
              f"{refine_info['rms_final']:.4f}  iters {refine_info['iters']}")
    print("\nWarping deformed onto reference grid ...")
    g_warped = warp_onto_ref_grid(A, t, c, ref.shape, g_pix, np)
    print("\nFinding overlap crop ...")
    lo, hi = compute_crop(ref, g_warped, h, glt)
    del g_warped
    crop_shape = tuple(int(x) for x in (hi - lo))
    for i, s_ in enumerate(crop_shape):
        assert s_ == ((s_ - 1) // h) * h + 1, 'axis %d: %d is not n*%d+1' % (i, s_, h)
    ref_crop = ref[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
    def_ref_aligned = deformed[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
    corners_ref = np.array([[lo[0] + dx * (hi[0] - lo[0] - 1), lo[1] + dy * (hi[1] - lo[1] - 1),
                             lo[2] + dz * (hi[2] - lo[2] - 1)]
                            for dx in (0, 1) for dy in (0, 1) for dz in (0, 1)], dtype=np.float64)
    corners_def = (corners_ref - c) @ A.T + c + t
    lo_def = np.maximum(np.floor(corners_def.min(axis=0)).astype(int) - 1, 0)
    hi_def = np.minimum(np.ceil(corners_def.max(axis=0)).astype(int) + 2,
                        np.array(deformed.shape))
    def_crop = deformed[lo_def[0]:hi_def[0], lo_def[1]:hi_def[1], lo_def[2]:hi_def[2]]
    os.makedirs(out_dir, exist_ok=True)
    paths = {k: os.path.join(out_dir, k + '.npy')
             for k in ('ref_preprocessed', 'def_preprocessed', 'def_ref_aligned')}
    np.save(paths['ref_preprocessed'], ref_crop.astype(ref_dtype))
    np.save(paths['def_preprocessed'], def_crop.astype(def_dtype))
    np.save(paths['def_ref_aligned'], def_ref_aligned.astype(def_dtype))
    angles, scales = rotation_angles_deg(A)
    meta = {
        "h": h, "glt": glt, "bin_factor": B, "additional_crop_xyz": list(args.crop),
        "ref_file": args.ref_path, "def_file": args.def_path,
        "ref_crop_file": paths['ref_preprocessed'], "def_crop_file": paths['def_preprocessed'],
        "def_ref_aligned_file": paths['def_ref_aligned'],
        "crop_lo": [int(x) for x in lo], "crop_hi": [int(x) for x in hi],
        "crop_shape": list(crop_shape),
        "def_lo": [int(x) for x in lo_def], "def_hi": [int(x) for x in hi_def],
        "affine_A": A.tolist(), "affine_t": t.tolist(), "affine_c": c.tolist(),
        "affine_t_fullres": (t * B).tolist(), "affine_c_fullres": (c * B).tolist(),
        "rotation_deg_xyz": [float(a) for a in angles],
        "scale_xyz": [float(x) for x in scales],
        "normal_strain_xyz": [float(e) for e in normal_strains(A)],
        "cost_identity": float(info['cost_identity']),
        "cost_translation_only": float(info['cost_translation_only']),
        "cost_affine": float(info['cost_final']), "nfev": int(info['nfev']),
        "affine_refine": (None if refine_info is None else {
            "rms_initial": float(refine_info['rms_initial']),
            "rms_final": float(refine_info['rms_final']),
            "iters": [int(n) for n in refine_info['iters']]}),
    }
    with open(os.path.join(out_dir, 'affine_prealign.json'), 'w') as fh:
        json.dump(meta, fh, indent=2)
    print(f"\nFinal crop shape : {crop_shape}  def_lo={lo_def.tolist()}  def_hi={hi_def.tolist()}")
    return meta
