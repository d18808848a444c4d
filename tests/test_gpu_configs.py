"""The benchmark configurations as GPU parity cases (the 513^3, h = 8 matvec
configuration is bench.py's; these are the others):

  C1  65^3 plain method (4000-blob speckle, translation + 1 % z-strain,
      N(0,1), h = 8, Laplacian, 3 levels, FFT / affine pre-alignment, FP32):
      against the reference's own solves (golden c1_fft / c1_affine) and the
      analytic field;
  C2  257^3 tensile necking, equilibrium-gap regularizer, 4-level pyramid,
      FP32: the reference's algorithm with the equilibrium gap on every
      pyramid level stalls (its Jacobi-preconditioned PCG does not converge
      on the coarse levels, h_i <= 4, and the GN iterates run away; the
      reference's own float64 solve of a 33^3 case, golden eqgap_multi33,
      does the same and is reproduced here), so the C2 workload is solved
      with the pyramid carried by the Laplacian and the equilibrium gap on
      the finest level; FP32 against FP64 and the analytic field;
  C3  the precision form at h = 4 (257^3 8-bit speckle, 16 Fourier modes,
      peak 4 vox): FP32 and FP16/BF16 solves against the FP64 solve, and the
      FP32-against-FP64 refinement;
  C5  the combined form (necking + 8 Fourier modes up to 8 vox + N(0,2),
      h = 16, equilibrium gap on the finest level) at 257 x 257 x 193: the
      learned seed with mixed precision reaches 0.01 vox of the FP64 solve.
"""
import hashlib
import os
import time

import numpy as np
import pytest

from conftest import GOLDEN, golden
from oracle import dvc_oracle as O

pytestmark = pytest.mark.gpu


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _input_shas():
    out = {}
    for ln in open(os.path.join(GOLDEN, 'INPUT_SHA256.txt')).read().splitlines():
        k, v = ln.split()
        out[k] = v
    return out


def _rms(a, b, sel=None):
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    if sel is not None:
        d = d[sel]
    return float(np.sqrt(np.mean(d * d)))


def _interior_dofs(Nx, Ny, Nz, margin=1):
    n = np.arange((Nx + 1) * (Ny + 1) * (Nz + 1))
    i, j, k = n % (Nx + 1), (n // (Nx + 1)) % (Ny + 1), n // ((Nx + 1) * (Ny + 1))
    ok = ((i >= margin) & (i <= Nx - margin) & (j >= margin) & (j <= Ny - margin)
          & (k >= margin) & (k <= Nz - margin))
    return np.repeat(ok, 3)


@pytest.mark.parametrize('pre', ['fft', 'affine'])
def test_c1_plain_method_matches_reference(gpu, pre):
    from voxeldvc_amd import synth
    from voxeldvc_amd.engine import correlate_gpu as C
    f, g = synth.c1_pair()
    f, g = f.astype(np.float32), g.astype(np.float32)
    shas = _input_shas()
    assert _sha(f) == shas['c1_f'] and _sha(g) == shas['c1_g']
    want = golden('c1_' + pre)
    kw = dict(K_ref_laplacian=O.build_K_ref_laplacian(), l0=36.0, glt=0.0, dynamic_mask=True,
              freeze_mask_after=4, disp=False)
    pk = dict(fft_prealign=True) if pre == 'fft' else dict(fft_prealign=False,
                                                            prealign_affine=True)
    print('C1 65^3', pre, flush=True)
    U, res, _ = C.multiscale_correlate_gpu(f, g, 8, 8, 8, 8, np, scales=(2, 1, 0), **pk, **kw)
    e_ref = _rms(U, want['U'])
    d_max = float(np.abs(np.asarray(U, np.float64) - want['U']).max())
    # analytic field: translation (2.3, -1.1, 0.7) + 1 % z-strain about z = 32
    n = np.arange(9 ** 3)
    truth = np.zeros(3 * n.size)
    truth[0::3], truth[1::3] = 2.3, -1.1
    truth[2::3] = 0.7 + 0.01 * ((n // 81) * 8 - 32.0)
    inner = _interior_dofs(8, 8, 8)
    e_true = _rms(U, truth, inner)
    r = np.asarray(res, np.float64)
    print('  rms vs reference %.2e (max %.2e), vs analytic field (interior) %.4f vox, '
          'active %d vs %d' % (e_ref, d_max, e_true, np.count_nonzero(r),
                               int(want['res_nonzero'])), flush=True)
    assert e_ref < 1e-3 and d_max < 5e-3
    assert abs(np.count_nonzero(r) - int(want['res_nonzero'])) <= 1e-3 * r.size
    assert r[r != 0].std() == pytest.approx(float(want['res_std']), rel=1e-3)
    assert e_true < 0.05 and _rms(want['U'], truth, inner) < 0.05


def test_eqgap_pyramid_matches_reference_golden(gpu):
    """The equilibrium gap on a two-level pyramid (33^3, scales (1, 0), 3 GN
    iterations, PCG capped at 300), float64: the reference's own result,
    including its stalled coarse-level PCG."""
    import torch
    from voxeldvc_amd.engine import correlate_gpu as C
    d = golden('multi33_dyn')
    want = golden('eqgap_multi33')
    f, g = d['f'], d['g']
    print('equilibrium-gap pyramid 33^3 vs reference', flush=True)
    U, _, _ = C.multiscale_correlate_gpu(f, g, 4, 4, 4, 8, np, scales=(1, 0),
                                         reg_type='equilibrium_gap',
                                         K_ref_elastic=O.build_K_ref(0.3),
                                         K_ref_laplacian=O.build_K_ref_laplacian(), l0=36.0,
                                         eps_reg=1e-3, maxiter=3, pcg_max_iter=300,
                                         fft_prealign=True, dtype=np.float64, disp=False)
    e = float(np.abs(np.asarray(U) - want['U']).max())
    print('  max |U - reference| %.2e vox (|U| up to %.2f)' % (e, np.abs(want['U']).max()),
          flush=True)
    # a stalled PCG amplifies rounding: 1e-5 vox on |U| up to 2.5 vox
    assert e < 1e-5


def _lap_then_eqgap(C, torch, fv, gv, N, h, extra, scales=(3, 2, 1, 0)):
    """The C2 / C5 workload: Laplacian pyramid, then the equilibrium gap on
    the finest level from its result."""
    lap = dict(K_ref_laplacian=O.build_K_ref_laplacian(), l0=4.5 * h, glt=0.0,
               dynamic_mask=True, freeze_mask_after=4, disp=False)
    eqg = dict(lap, K_ref_elastic=O.build_K_ref(0.3), reg_type='equilibrium_gap')
    U0, _, _ = C.multiscale_correlate_gpu(fv, gv, N[0], N[1], N[2], h, torch, scales=scales,
                                          **lap, **extra)
    U, _, _ = C.correlate_gpu(fv, gv, N[0], N[1], N[2], h, torch, U0=U0, **eqg, **extra)
    return U


def test_c2_necking_equilibrium_gap(gpu):
    import torch
    from voxeldvc_amd import synth
    from voxeldvc_amd.engine import correlate_gpu as C
    n, h = 257, 8
    Ne = (n - 1) // h
    f, g, modes, extra = synth.config_pair_torch('c2', (n, n, n), seed=2)
    fv = f.view(n, n, n).permute(2, 1, 0)
    gv = g.view(n, n, n).permute(2, 1, 0)
    print('C2 257^3 necking, Laplacian pyramid + equilibrium gap', flush=True)
    N = (Ne, Ne, Ne)
    U64 = _lap_then_eqgap(C, torch, fv, gv, N, h, dict(dtype=np.float64, maxiter=60, eps=1e-6))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    U32 = _lap_then_eqgap(C, torch, fv, gv, N, h, {})
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    truth = synth.nodal_truth(modes, extra, Ne, Ne, Ne, h)
    inner = _interior_dofs(Ne, Ne, Ne, margin=2)
    e64 = _rms(U32.cpu().numpy(), U64.cpu().numpy())
    et = _rms(U32.cpu().numpy(), truth, inner)
    print('  FP32 solve %.1f ms; rms vs FP64 solve %.2e vox; vs analytic field (interior, '
          'first order) %.4f vox' % (ms, e64, et), flush=True)
    assert U32.dtype == torch.float32 and U64.dtype == torch.float64
    assert e64 < 2e-3
    assert et < 0.15


def test_c3_precision_form_h4(gpu):
    import torch
    from voxeldvc_amd import synth
    from voxeldvc_amd.engine import correlate_gpu as C
    n, h = 257, 4
    Ne = (n - 1) // h
    f, g, modes, _ = synth.config_pair_torch('c3', (n, n, n), seed=3)
    fv = f.view(n, n, n).permute(2, 1, 0)
    gv = g.view(n, n, n).permute(2, 1, 0)
    kw = dict(K_ref_laplacian=O.build_K_ref_laplacian(), l0=4.5 * h, glt=0.0,
              dynamic_mask=True, freeze_mask_after=4, disp=False)
    print('C3 257^3 h=4', flush=True)
    U64, _, _ = C.multiscale_correlate_gpu(fv, gv, Ne, Ne, Ne, h, torch, dtype=np.float64,
                                           maxiter=60, eps=1e-6, **kw)
    U64 = U64.cpu().numpy()
    out = {}
    for storage in ('fp32', 'bf16'):
        U, _, _ = C.multiscale_correlate_gpu(fv, gv, Ne, Ne, Ne, h, torch, storage=storage,
                                             **kw)
        out[storage] = _rms(U.cpu().numpy(), U64)
    log = []
    Ur, _, _ = C.multiscale_correlate_gpu(fv, gv, Ne, Ne, Ne, h, torch, storage='bf16',
                                          refine_fp64=2, log=log, **kw)
    res = [q['residual_fp64'] for lv in log for q in lv['iters'] if 'refine' in q]
    truth = synth.nodal_truth(modes, None, Ne, Ne, Ne, h)
    et = _rms(U64, truth, _interior_dofs(Ne, Ne, Ne, margin=2))
    print('  rms vs FP64 solve: fp32 %.2e, bf16 %.2e, bf16 + refinement %.2e vox; FP64 '
          'residual %s; FP64 solve vs analytic field %.4f vox'
          % (out['fp32'], out['bf16'], _rms(Ur.cpu().numpy(), U64),
             ['%.2e' % r for r in res], et), flush=True)
    assert out['fp32'] < 1e-3 and out['bf16'] < 1e-3
    # refinement continues from the dU/U < 1e-3 stop towards the FP64 fixed
    # point: never farther from the FP64 solve than the unrefined iterate
    assert Ur.dtype == torch.float64 and _rms(Ur.cpu().numpy(), U64) <= 1.05 * out['bf16']
    # the analytic field is first order in u (|u| up to 4 vox)
    assert et < 0.1


def test_c5_combined_h16_learned_mixed(gpu):
    """Reduced C5: 257 x 257 x 193 voxels (16 x 16 x 12 elements at h = 16)."""
    import torch
    from voxeldvc_amd import seednet, synth
    from voxeldvc_amd.engine import correlate_gpu as C
    shape, h = (257, 257, 193), 16
    N = tuple((s - 1) // h for s in shape)
    f, g, modes, extra = synth.config_pair_torch('c5', shape, seed=5)
    fv = f.view(shape[2], shape[1], shape[0]).permute(2, 1, 0)
    gv = g.view(shape[2], shape[1], shape[0]).permute(2, 1, 0)
    print('C5 reduced 257x257x193 h=16', flush=True)
    U64 = _lap_then_eqgap(C, torch, fv, gv, N, h, dict(dtype=np.float64, maxiter=60, eps=1e-6))
    eqg = dict(K_ref_laplacian=O.build_K_ref_laplacian(), K_ref_elastic=O.build_K_ref(0.3),
               reg_type='equilibrium_gap', l0=4.5 * h, glt=0.0, dynamic_mask=True,
               freeze_mask_after=4, disp=False)
    trace = []
    net = seednet.SeedNet.load()
    Ul, _, _ = C.multiscale_correlate_gpu(
        fv, gv, N[0], N[1], N[2], h, torch, storage='bf16', scales=(0,), learned_seed=net,
        callback=lambda s, ik, U: trace.append(float(((U.double() - U64) ** 2).mean().sqrt())),
        **eqg)
    Up = _lap_then_eqgap(C, torch, fv, gv, N, h, {})
    seed = net.predict(f, g, shape, N[0], N[1], N[2], h)
    truth = synth.nodal_truth(modes, extra, N[0], N[1], N[2], h)
    inner = _interior_dofs(N[0], N[1], N[2], margin=2)
    print('  seed rms vs FP64 %.3f vox; learned+BF16 per-iteration rms %s; plain FP32 rms '
          '%.2e; FP64 vs analytic field (interior) %.3f vox'
          % (float(((seed.double() - U64) ** 2).mean().sqrt()), ['%.2e' % r for r in trace],
             float(((Up.double() - U64) ** 2).mean().sqrt()),
             _rms(U64.cpu().numpy(), truth, inner)), flush=True)
    assert seed.numel() == 3 * (N[0] + 1) * (N[1] + 1) * (N[2] + 1)
    assert float(((Up.double() - U64) ** 2).mean().sqrt()) < 2e-3
    assert min(trace) <= 0.01
    assert float(((Ul.double() - U64) ** 2).mean().sqrt()) <= 0.01
