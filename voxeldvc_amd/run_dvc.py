"""`run` stage drop-in: solve a preprocessed work directory on the GPU.

Mirrors the reference's src/voxeldvc/run_dvc.py: the same work-dir inputs
(affine_prealign.json, ref_preprocessed.npy, def_preprocessed.npy,
def_ref_aligned.npy), the same affine adjustment (c_adj = c - crop_lo,
t_adj = t + crop_lo - def_lo, run_dvc.py:89-94), the same solver call
(multiscale (2,1,0), Laplacian, l0 = l0_factor*h, dynamic mask frozen after
iteration 4, ext_affine; run_dvc.py:121-133), the same post-shift of U by
def_lo and the same GN log lines mirrored into run_log.txt.

Outputs written (reference layouts, engine/write_output.py:738-752, 912-914):
U_recovered.npy (Nx+1, Ny+1, Nz+1, 3), residual.npy (Nvx, Nvy, Nvz),
U_local.npy (U minus the affine prior, flat), dU.npy (last GN update, flat),
ref.npy and deformed.npy; and the voxel-scale fields of the reference's
write_outputs (engine/write_output.py:771-789, 838-849, 883-894), computed on
the GPU: the active mask (compute_active_mask with g_origin = def_lo and
mask0 = ref > glt) and the global ZNCC over it (printed),
ref_forward_pushed.npy (ref pushed by U_local), residual_std_cell.npy and
zncc_cell.npy.  The element-scale fields (strains, unsafe / overlap-lost
masks, griddata remaps, .vti) are not produced.

    python -m voxeldvc_amd run WORK_DIR [--l0-factor F]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

DEFAULT_L0_FACTOR = 4.5  # dvc_defaults.py:8


class Tee:
    """Mirror writes to several streams (stdout + run_log.txt)."""

    def __init__(self, *streams):
        self.streams = streams

    def write(self, s):
        for st in self.streams:
            st.write(s)

    def flush(self):
        for st in self.streams:
            st.flush()


def mesh_dims(shape, h):
    """Largest N with N*h <= n-1 per axis (engine/write_output.py:228-230)."""
    return tuple((s - 1) // h for s in shape)


def parse_args(argv=None):
    p = argparse.ArgumentParser(prog='voxeldvc_amd run')
    p.add_argument('work_dir', type=str)
    p.add_argument('--l0-factor', type=float, default=DEFAULT_L0_FACTOR, dest='l0_factor',
                   help='Regularization length l0 = l0_factor * h, in voxels (default: %g).'
                        % DEFAULT_L0_FACTOR)
    p.add_argument('--storage', choices=('fp32', 'bf16'), default='fp32',
                   help='voxel storage of the device solve (default fp32)')
    return p.parse_args(argv)


def post_outputs(wd, U, U_local, res, ref, deformed, Nx, Ny, Nz, h, lo_def, glt):
    """The voxel-scale stages of the reference's write_outputs on the GPU
    (engine/write_output.py:771-789, 838-849, 883-894): active mask and
    global ZNCC, the forward-pushed reference, per-cell residual std and
    ZNCC."""
    from .engine.correlate_gpu import compute_active_mask
    from .engine import write_output as W
    ref_flat = ref.reshape(-1, order='F')
    n_active = int(np.count_nonzero(res))
    print('Nonzero (active) residual voxels: %d/%d (%.1f%%)'
          % (n_active, res.size, 100 * n_active / res.size))
    active = compute_active_mask(U, deformed, Nx, Ny, Nz, h, np,
                                 g_origin=tuple(int(x) for x in lo_def),
                                 mask0_vox=ref_flat > glt, glt=glt)
    zncc = W.zncc_of_residual(ref_flat, res, active)
    print('ZNCC (active voxels, no-ground-truth accuracy proxy): %.6f' % zncc)
    pushed = W.push_reference_forward(ref, U_local, Nx, Ny, Nz, h)
    np.save(os.path.join(wd, 'ref_forward_pushed.npy'), pushed)
    std_c, zn_c = W.compute_correlation_quality_cell(res, ref, h, Nx, Ny, Nz)
    np.save(os.path.join(wd, 'residual_std_cell.npy'), std_c)
    np.save(os.path.join(wd, 'zncc_cell.npy'), zn_c)
    if np.any(~np.isnan(std_c)):
        print('  cell residual std: mean=%.4f  min=%.4f  max=%.4f'
              % (np.nanmean(std_c), np.nanmin(std_c), np.nanmax(std_c)))
    if np.any(~np.isnan(zn_c)):
        print('  cell ZNCC: mean=%.5f  min=%.5f  max=%.5f'
              % (np.nanmean(zn_c), np.nanmin(zn_c), np.nanmax(zn_c)))
    return active, zncc


def main(argv=None):
    from .engine.correlate_gpu import multiscale_correlate_gpu
    from .engine.geometry_dvc import build_K_ref_laplacian
    t_start = time.perf_counter()
    args = parse_args(argv)
    wd = os.path.expanduser(args.work_dir)
    log_path = os.path.join(wd, 'run_log.txt')
    log_file = open(log_path, 'w')
    sys.stdout = Tee(sys.__stdout__, log_file)
    try:
        with open(os.path.join(wd, 'affine_prealign.json')) as fh:
            meta = json.load(fh)
        h = int(meta['h'])
        glt = float(meta.get('glt', 0.0))
        A = np.array(meta['affine_A'], dtype=np.float64)
        t = np.array(meta['affine_t'], dtype=np.float64)
        c = np.array(meta['affine_c'], dtype=np.float64)
        crop_lo = np.array(meta['crop_lo'])
        lo_def = np.array(meta['def_lo'])
        c_adj = c - crop_lo
        t_adj = t + crop_lo - lo_def
        ref = np.load(os.path.join(wd, 'ref_preprocessed.npy'))
        deformed = np.load(os.path.join(wd, 'def_preprocessed.npy'))
        deformed_cropped = np.load(os.path.join(wd, 'def_ref_aligned.npy'))
        Nx, Ny, Nz = mesh_dims(ref.shape, h)
        l0 = args.l0_factor * h
        K_lap = build_K_ref_laplacian()
        print('Run configuration: work_dir=%s h=%d glt=%g l0_factor=%g l0=%g mesh=%dx%dx%d '
              'ref=%s deformed=%s' % (wd, h, glt, args.l0_factor, l0, Nx, Ny, Nz,
                                      ref.shape, deformed.shape))
        print('\nGPU matrix-free DVC (multiscale, ext_affine U0, original deformed) ...')
        ref_c = ref[:Nx * h + 1, :Ny * h + 1, :Nz * h + 1]
        t_setup = time.perf_counter()
        U, res, dU = multiscale_correlate_gpu(
            ref_c, deformed, Nx, Ny, Nz, h, np, scales=(2, 1, 0), l0=l0, glt=glt,
            K_ref_laplacian=K_lap, dynamic_mask=True, fft_prealign=False,
            prealign_affine=False, ext_affine=(A, t_adj, c_adj), disp=True,
            freeze_mask_after=4, storage=args.storage)
        t_corr = time.perf_counter()
        assert np.all(np.isfinite(U)), 'Recovered displacement contains NaN/inf'
        nbx, nby, nbz = Nx + 1, Ny + 1, Nz + 1
        ii, jj, kk = np.meshgrid(np.arange(nbx), np.arange(nby), np.arange(nbz), indexing='ij')
        X = np.stack([ii, jj, kk], axis=-1).astype(np.float64) * h
        u_aff = (X - c_adj) @ A.T + c_adj + t_adj - X
        U = U.astype(np.float64)
        U_local = np.empty_like(U)
        for k in range(3):
            U_local[k::3] = (U[k::3].reshape(nbx, nby, nbz, order='F')
                             - u_aff[..., k]).flatten(order='F')
        for k in range(3):
            U[k::3] += lo_def[k]
        grid = np.stack([U[k::3].reshape(nbx, nby, nbz, order='F') for k in range(3)], axis=-1)
        shp = ref_c.shape
        np.save(os.path.join(wd, 'U_recovered.npy'), grid)
        np.save(os.path.join(wd, 'residual.npy'), res.reshape(shp, order='F'))
        np.save(os.path.join(wd, 'U_local.npy'), U_local)
        np.save(os.path.join(wd, 'dU.npy'), dU)
        np.save(os.path.join(wd, 'ref.npy'), ref_c)
        np.save(os.path.join(wd, 'deformed.npy'),
                deformed_cropped[:shp[0], :shp[1], :shp[2]])
        post_outputs(wd, U, U_local, res, ref_c, deformed, Nx, Ny, Nz, h, lo_def, glt)
        print('\nDVC COMPLETE')
        t_end = time.perf_counter()
        print('Timing: setup %.2f s | multiscale correlation %.2f s | output %.2f s | total %.2f s'
              % (t_setup - t_start, t_corr - t_setup, t_end - t_corr, t_end - t_start))
    finally:
        sys.stdout = sys.__stdout__
        log_file.close()
    print('Log written to %s' % log_path)


if __name__ == '__main__':
    main()
