"""Write golden vectors from the reference engine (run once, in the build
container, where the reference source tree is mounted).

Only the composites that no closed form or brute-force check pins are
recorded here: the end-to-end single-level and multiscale Gauss-Newton solves
(normalisation, mask freeze, ll calibration and warm-start conventions), the
equilibrium-gap solve and its Hutchinson diagonal (RNG stream and probe
order), and the affine pre-alignment (SciPy Powell trajectory).

    PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden.py [c1]

With no argument the small-case fixtures are written; `c1` writes the 65^3
plain-method configuration (c1_*.npz: 4000-blob speckle, translation + 1 %
z-strain + N(0,1) noise, h = 8, Laplacian, 3 levels, FFT and affine
pre-alignment); `affine` writes affine_refine49.npz (the preprocess
affine_refine_gn on a 49^3 pair pulled by a known affine); `gt` copies the
reference's bundled ground-truth dataset GT_z+0.01 (assets/GT_z+0.01: the
65^3 uint16 reference and its displacement field at the 65^3 image voxels,
U[:65, :65, :65], float32) into gt_zp001.npz.  Inputs are either stored in the fixture (small) or
regenerated from the seeded generator in voxeldvc_amd/synth.py and checked by
SHA-256 (INPUT_SHA256.txt).
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, '..', '..')))
sys.path.insert(0, '/root/reference/src')
sys.dont_write_bytecode = True

from voxeldvc.engine import correlate_gpu as rc  # noqa: E402
from voxeldvc.engine import geometry_dvc as rg  # noqa: E402
from voxeldvc.engine import kernels_dvc as rk  # noqa: E402

from voxeldvc_amd import synth  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def problem17():
    """The reference test's 17^3 sinusoid texture, g = f shifted by one voxel
    in x (cubic spline) -- tests/test_dvc_gpu_multiscale.py:24-40."""
    from scipy.ndimage import shift
    n = 17
    x = np.arange(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
    f = (np.sin(2 * np.pi * X / n) + np.sin(2 * np.pi * Y / n) + np.sin(2 * np.pi * Z / n)
         + 0.5 * np.cos(2 * np.pi * (X + Y) / n))
    g = shift(f, (1.0, 0, 0), order=3, mode='nearest')
    return f, g


def speckle33(seed=7):
    """33^3 speckle (4 x 4 x 4 elements at h=8), translation + strain."""
    f = synth.speckle_numpy((33, 33, 33), 600, seed)
    g = synth.warp_backward(f, synth.translation_strain_field((1.3, -0.6, 0.4), 0.01, 16.0),
                            1.0, seed + 1000)
    return f, g


def c1_inputs():
    """The 65^3 plain-method pair as float32 (synth.c1_pair)."""
    f, g = synth.c1_pair()
    return f.astype(np.float32), g.astype(np.float32)


def main_c1():
    """Reference multiscale solves of the 65^3 configuration: FFT
    pre-alignment (run_dvc-style dynamic mask frozen after 4 iterations) and
    the affine pre-alignment.  Stores U, dU and residual statistics."""
    K = rg.build_K_ref_laplacian()
    f, g = c1_inputs()
    kw = dict(K_ref_laplacian=K, l0=36.0, glt=0.0, dynamic_mask=True, freeze_mask_after=4,
              disp=False)
    out = {}
    for name, pre in (('c1_fft', dict(fft_prealign=True)),
                      ('c1_affine', dict(fft_prealign=False, prealign_affine=True))):
        U, res, dU = rc.multiscale_correlate_gpu(f, g, 8, 8, 8, 8, np, scales=(2, 1, 0),
                                                 **pre, **kw)
        r = np.asarray(res, dtype=np.float64)
        out[name] = dict(U=U, dU=dU, res_nonzero=np.int64(np.count_nonzero(r)),
                         res_std=np.float64(r[r != 0].std()))
    for name, arrs in out.items():
        np.savez_compressed(os.path.join(HERE, name + '.npz'), **arrs)
    path = os.path.join(HERE, 'INPUT_SHA256.txt')
    lines = [ln for ln in open(path).read().splitlines() if not ln.startswith('c1_')]
    lines += ['c1_f %s' % sha(f), 'c1_g %s' % sha(g)]
    with open(path, 'w') as fh:
        fh.write('\n'.join(lines) + '\n')
    print('wrote', sorted(out))


def affine49_inputs():
    """49^3 speckle f and g(x) = f(x - M (x - c) - t_r) + N(0, 0.5): a known
    affine in the generative convention of affine_refine_gn, float32."""
    n = 49
    f = synth.speckle_numpy((n, n, n), 1500, 31)
    c = np.full(3, (n - 1) / 2.0)
    M = np.array([[0.012, -0.004, 0.003], [0.002, -0.015, 0.005], [-0.003, 0.004, 0.02]])
    tr = np.array([0.6, -0.4, 0.3])

    def u(X, Y, Z):
        d = [X - c[0], Y - c[1], Z - c[2]]
        return tuple(M[i, 0] * d[0] + M[i, 1] * d[1] + M[i, 2] * d[2] + tr[i] for i in range(3))
    g = synth.warp_backward(f, u, 0.5, 32)
    return f.astype(np.float32), g.astype(np.float32), M, tr, c


def main_affine():
    f, g, M, tr, c = affine49_inputs()
    A, t, info = rc.affine_refine_gn(f, g, np, np.eye(3), np.zeros(3), c)
    np.savez_compressed(os.path.join(HERE, 'affine_refine49.npz'), A=A, t=t, c=c, M_true=M,
                        tr_true=tr, rms_initial=np.float64(info['rms_initial']),
                        rms_final=np.float64(info['rms_final']),
                        iters=np.asarray(info['iters'], np.int64))
    path = os.path.join(HERE, 'INPUT_SHA256.txt')
    lines = [ln for ln in open(path).read().splitlines() if not ln.startswith('affine49_')]
    lines += ['affine49_f %s' % sha(f), 'affine49_g %s' % sha(g)]
    with open(path, 'w') as fh:
        fh.write('\n'.join(lines) + '\n')
    print('wrote affine_refine49', info)


def main_gt():
    d = '/root/reference/assets/GT_z+0.01'
    ref = np.load(os.path.join(d, 'ref.npy'), allow_pickle=False)
    U = np.load(os.path.join(d, 'U.npy'), allow_pickle=False)
    n = ref.shape[0]
    np.savez_compressed(os.path.join(HERE, 'gt_zp001.npz'), ref=ref,
                        u_gt_vox=np.ascontiguousarray(U[:n, :n, :n]))
    print('wrote gt_zp001', ref.shape, ref.dtype, U.shape)


def main():
    out = {}
    K = rg.build_K_ref_laplacian()
    f17, g17 = problem17()

    kw = dict(K_ref_laplacian=K, l0=36.0, maxiter=20, eps=1e-3, disp=False)
    for name, dt in (('corr17_f32', np.float32), ('corr17_f64', np.float64)):
        U, res, dU = rc.correlate_gpu(f17, g17, 2, 2, 2, 8, np, dtype=dt, **kw)
        out[name] = dict(f=f17, g=g17, U=U, res=res, dU=dU)

    U, res, dU = rc.multiscale_correlate_gpu(f17, g17, 2, 2, 2, 8, np, scales=(2, 1, 0),
                                             K_ref_laplacian=K, l0=36.0, maxiter=30,
                                             eps=1e-3, disp=False)
    out['multi17_fft'] = dict(f=f17, g=g17, U=U, res=res, dU=dU)

    f33, g33 = speckle33()
    kw33 = dict(K_ref_laplacian=K, l0=36.0, glt=0.0, dynamic_mask=True,
                freeze_mask_after=4, disp=False)
    U, res, dU = rc.multiscale_correlate_gpu(f33, g33, 4, 4, 4, 8, np, scales=(2, 1, 0),
                                             fft_prealign=True, **kw33)
    out['multi33_dyn'] = dict(f=f33.astype(np.float32), g=g33.astype(np.float32),
                              U=U, res=res, dU=dU)
    # the same solve, fed the float32-rounded inputs stored above
    U, res, dU = rc.multiscale_correlate_gpu(f33.astype(np.float32), g33.astype(np.float32),
                                             4, 4, 4, 8, np, scales=(2, 1, 0),
                                             fft_prealign=True, **kw33)
    out['multi33_dyn_f32in'] = dict(U=U, res=res, dU=dU)

    Kel = rg.build_K_ref(0.3)
    U, res, dU = rc.correlate_gpu(f33.astype(np.float32), g33.astype(np.float32), 4, 4, 4, 8,
                                  np, reg_type='equilibrium_gap', K_ref_elastic=Kel,
                                  K_ref_laplacian=K, l0=36.0, eps_reg=1e-3, maxiter=10,
                                  U0=None, disp=False)
    out['eqgap33'] = dict(U=U, res=res, dU=dU)

    # equilibrium gap through a pyramid: the coarse level (h = 4) is where the
    # Jacobi-preconditioned PCG stalls; 3 GN iterations, PCG capped at 300
    U, res, dU = rc.multiscale_correlate_gpu(f33.astype(np.float32), g33.astype(np.float32),
                                             4, 4, 4, 8, np, scales=(1, 0),
                                             reg_type='equilibrium_gap', K_ref_elastic=Kel,
                                             K_ref_laplacian=K, l0=36.0, eps_reg=1e-3,
                                             maxiter=3, pcg_max_iter=300, fft_prealign=True,
                                             dtype=np.float64, disp=False)
    out['eqgap_multi33'] = dict(U=U, res=res, dU=dU)

    interior = rg.build_interior_dof_mask(3, 3, 3).astype(np.float32)
    d = rk.build_diagonal_equilibrium_gap(Kel, interior, 8, 3, 3, 3, 3 * 64, np)
    out['eqgap_diag3'] = dict(diag=d)

    A, t, c, info = rc.affine_prealign(f33.astype(np.float32), g33.astype(np.float32), np,
                                       decimate=2)
    out['affine33'] = dict(A=A, t=t, c=c, cost_final=np.float64(info['cost_final']),
                           cost_identity=np.float64(info['cost_identity']),
                           nfev=np.int64(info['nfev']))

    d = rc.fft_rigid_shift(f33, g33, np)
    out['fft33'] = dict(d=d)

    for name, arrs in out.items():
        np.savez_compressed(os.path.join(HERE, name + '.npz'), **arrs)
    path = os.path.join(HERE, 'INPUT_SHA256.txt')
    keep = [ln for ln in open(path).read().splitlines()
            if ln.startswith('c1_') or ln.startswith('affine49_')] if os.path.exists(path) else []
    with open(path, 'w') as fh:
        fh.write('speckle33_f %s\nspeckle33_g %s\n' % (sha(f33), sha(g33)))
        fh.write('problem17_f %s\nproblem17_g %s\n' % (sha(f17), sha(g17)))
        for ln in keep:
            fh.write(ln + '\n')
    print('wrote', sorted(out))


if __name__ == '__main__':
    if sys.argv[1:] == ['c1']:
        main_c1()
    elif sys.argv[1:] == ['affine']:
        main_affine()
    elif sys.argv[1:] == ['gt']:
        main_gt()
    else:
        main()
