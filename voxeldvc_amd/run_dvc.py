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
zncc_cell.npy; and the element-scale fields (engine/write_output.py:
790-823, 846-906): unsafe_voxel_mask.npy, overlap_lost_element_mask.npy and
its deformed-configuration remap, principal_strains_cell.npy,
equivalent_strain_cell.npy and their deformed-configuration remaps
principal_strains_cell_deformed.npy / equivalent_strain_cell_deformed.npy,
gn_convergence_cell.npy, unsafe_voxel_mask_forward_pushed.npy,
material_cell_mask.npy and material_cell_mask_eroded.npy.  Not produced:
the .vti export.

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


def post_outputs(wd, U, U_local, res, ref, deformed, Nx, Ny, Nz, h, lo_def, glt, dU=None,
                 crop_lo=(0, 0, 0), deformed_cropped=None):
    """The derived fields of the reference's write_outputs on the GPU
    (engine/write_output.py:693-929): active mask and global ZNCC, unsafe
    voxel mask and safe-cell count, overlap-lost element mask, principal and
    equivalent strains, per-cell residual std / ZNCC / GN convergence, the
    overlap-lost mask and the strains carried to the deformed configuration,
    the forward pushes of ref and of the unsafe mask, the material-cell
    masks.  Not produced: the .vti."""
    from .engine.correlate_gpu import compute_active_mask
    from .engine import kernels_dvc as K
    from .engine import write_output as W
    Nvx, Nvy, Nvz = ref.shape
    nbx, nby, nbz = Nx + 1, Ny + 1, Nz + 1
    grid = np.stack([U[c::3].reshape(nbx, nby, nbz, order='F') for c in range(3)], axis=-1)
    for c, name in enumerate(('ux', 'uy', 'uz')):
        comp = grid[..., c]
        print('  %s: mean=%.4f  std=%.4f  min=%.4f  max=%.4f'
              % (name, comp.mean(), comp.std(), comp.min(), comp.max()))
    ref_flat = ref.reshape(-1, order='F')
    n_active = int(np.count_nonzero(res))
    print('Nonzero (active) residual voxels: %d/%d (%.1f%%)'
          % (n_active, res.size, 100 * n_active / res.size))
    if n_active:
        print('residual std (active voxels): %.4f' % res[res != 0].std())
    mask0 = ref_flat > glt
    active = compute_active_mask(U, deformed, Nx, Ny, Nz, h, np,
                                 g_origin=tuple(int(x) for x in lo_def),
                                 mask0_vox=mask0, glt=glt)
    zncc = W.zncc_of_residual(ref_flat, res, active)
    print('ZNCC (active voxels, no-ground-truth accuracy proxy): %.6f' % zncc)
    unsafe = K.compute_unsafe_voxel_mask(active, h, Nx, Ny, Nz)
    unsafe_grid = unsafe.reshape(Nvx, Nvy, Nvz, order='F')
    n_unsafe = int(np.count_nonzero(unsafe))
    print('Unsafe voxels (affected by a node connected to an inactive voxel): %d/%d (%.1f%%)'
          % (n_unsafe, unsafe.size, 100 * n_unsafe / unsafe.size))
    safe = W.compute_safe_cell_mask(unsafe, h, Nx, Ny, Nz)
    print('Safe elements (no unsafe voxel): %d/%d (%.1f%%)'
          % (int(safe.sum()), safe.size, 100 * safe.sum() / safe.size))
    lost = K.compute_overlap_lost_element_mask(active, mask0, h, Nx, Ny, Nz)
    lost_grid = lost.reshape(Nx, Ny, Nz, order='F')
    print('Overlap-lost elements (lost material overlap or <50%% material): %d/%d (%.1f%%)'
          % (int(lost.sum()), lost.size, 100 * lost.sum() / lost.size))
    if deformed_cropped is not None:
        print('\ndeformed image: mean=%.2f  min=%.2f  max=%.2f'
              % (deformed_cropped.mean(), deformed_cropped.min(), deformed_cropped.max()))
    ps = W.compute_principal_strains_cell(grid, h)
    for c, name in enumerate(('e1 (max)', 'e2 (mid)', 'e3 (min)')):
        comp = ps[..., c]
        print('  cell %s: mean=%.5f  std=%.5f  min=%.5f  max=%.5f'
              % (name, comp.mean(), comp.std(), comp.min(), comp.max()))
    ps_def = W.remap_cell_field_to_deformed(ps, grid, h, crop_lo=crop_lo)
    for c, name in enumerate(('e1 (max)', 'e2 (mid)', 'e3 (min)')):
        comp = ps_def[..., c]
        print('  deformed-config cell %s: mean=%.5f  std=%.5f  min=%.5f  max=%.5f'
              % (name, comp.mean(), comp.std(), comp.min(), comp.max()))
    eq = W.compute_equivalent_strain(ps)
    eq_def = W.compute_equivalent_strain(ps_def)
    print('  cell eq. strain: mean=%.5f  std=%.5f  min=%.5f  max=%.5f'
          % (eq.mean(), eq.std(), eq.min(), eq.max()))
    std_c, zn_c = W.compute_correlation_quality_cell(res, ref, h, Nx, Ny, Nz)
    if np.any(~np.isnan(std_c)):
        print('  cell residual std: mean=%.4f  min=%.4f  max=%.4f'
              % (np.nanmean(std_c), np.nanmin(std_c), np.nanmax(std_c)))
    if np.any(~np.isnan(zn_c)):
        print('  cell ZNCC: mean=%.5f  min=%.5f  max=%.5f'
              % (np.nanmean(zn_c), np.nanmin(zn_c), np.nanmax(zn_c)))
    conv = None
    if dU is not None:
        conv = W.compute_gn_convergence_cell(dU, U, Nx, Ny, Nz)
        if np.any(~np.isnan(conv)):
            print('  cell GN convergence dU/U: mean=%.2e  min=%.2e  max=%.2e'
                  % (np.nanmean(conv), np.nanmin(conv), np.nanmax(conv)))
    lost_def = W.remap_element_mask_to_deformed(lost_grid, grid, h, crop_lo=crop_lo)
    print('Overlap-lost elements (deformed config): %d/%d (%.1f%%)'
          % (int(lost_def.sum()), lost_def.size, 100 * lost_def.sum() / lost_def.size))
    pushed = W.push_reference_forward(ref, U_local, Nx, Ny, Nz, h)
    if deformed_cropped is not None:
        diff = pushed.astype(np.float64) - deformed_cropped.astype(np.float64)
        print('\nref_forward_pushed - deformed_cropped: mean=%.4f  std=%.4f  min=%.4f  max=%.4f'
              % (diff.mean(), diff.std(), diff.min(), diff.max()))
    unsafe_pushed = W.push_field_forward(unsafe_grid, U_local, Nx, Ny, Nz, h, order=0,
                                         mode='constant', cval=1).astype(bool)
    print('Unsafe voxels (forward-pushed to deformed grid): %d/%d (%.1f%%)'
          % (int(unsafe_pushed.sum()), unsafe_pushed.size,
             100 * unsafe_pushed.sum() / unsafe_pushed.size))
    mat = W.compute_material_cell_mask(ref, h, Nx, Ny, Nz, glt=glt)
    print('Material cells (mean reference gray > glt): %d/%d (%.1f%%)'
          % (int(mat.sum()), mat.size, 100 * mat.sum() / mat.size))
    mat_ero = W.compute_material_cell_mask_eroded(mat)
    print('Material cells (eroded, no non-material neighbour): %d/%d (%.1f%%)'
          % (int(mat_ero.sum()), mat_ero.size, 100 * mat_ero.sum() / mat_ero.size))
    save = lambda name, a: np.save(os.path.join(wd, name + '.npy'), a)
    save('principal_strains_cell', ps)
    save('principal_strains_cell_deformed', ps_def)
    save('equivalent_strain_cell', eq)
    save('equivalent_strain_cell_deformed', eq_def)
    save('material_cell_mask', mat)
    save('material_cell_mask_eroded', mat_ero)
    save('residual_std_cell', std_c)
    save('zncc_cell', zn_c)
    if conv is not None:
        save('gn_convergence_cell', conv)
    save('unsafe_voxel_mask', unsafe_grid)
    save('overlap_lost_element_mask', lost_grid)
    save('overlap_lost_element_mask_deformed', lost_def)
    save('ref_forward_pushed', pushed)
    save('unsafe_voxel_mask_forward_pushed', unsafe_pushed)
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
        post_outputs(wd, U, U_local, res, ref_c, deformed, Nx, Ny, Nz, h, lo_def, glt, dU=dU,
                     crop_lo=tuple(int(x) for x in crop_lo),
                     deformed_cropped=deformed_cropped[:shp[0], :shp[1], :shp[2]])
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
