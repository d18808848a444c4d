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
  C3  the precision form at full size (513^3 8-bit speckle, 16 Fourier
      modes, peak 4 vox) at h = 8 and h = 4: FP32 and FP16/BF16 solves
      against the FP64 solve, the learned seed with BF16 storage, and the
      FP32-against-FP64 refinement (h = 8);
  C5  the combined form at full size (1025 x 1025 x 769, necking + 8
      Fourier modes up to 8 vox + N(0,2), h = 16, equilibrium gap on the
      finest level): the learned seed with mixed precision reaches 0.01 vox
      of the FP64 solve of a 257^3 subvolume.
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
    assert e_ref < 1e-3 and d_max < 1e-3
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


def _lap_then_eqgap(C, torch, fv, gv, N, h, extra, scales=(3, 2, 1, 0), ms_extra=None):
    """The C2 / C5 workload: Laplacian pyramid, then the equilibrium gap on
    the finest level from its result."""
    lap = dict(K_ref_laplacian=O.build_K_ref_laplacian(), l0=4.5 * h, glt=0.0,
               dynamic_mask=True, freeze_mask_after=4, disp=False)
    eqg = dict(lap, K_ref_elastic=O.build_K_ref(0.3), reg_type='equilibrium_gap')
    U0, _, _ = C.multiscale_correlate_gpu(fv, gv, N[0], N[1], N[2], h, torch, scales=scales,
                                          **lap, **extra, **(ms_extra or {}))
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


def _cube(kind, n, seed):
    from voxeldvc_amd import synth
    f, g, modes, extra = synth.config_pair_torch(kind, (n, n, n), seed=seed)
    return f, g, f.view(n, n, n).permute(2, 1, 0), g.view(n, n, n).permute(2, 1, 0), modes, extra


def _trms(a, b):
    return float(((a.double() - b.double()) ** 2).mean().sqrt())


def _c3_common(h):
    """C3 at full size: 513^3 8-bit speckle warped by 16 Fourier modes
    (wavelengths 64-256, peak 4 vox) + N(0,1), quantised; Laplacian
    l0 = 4.5 h, dynamic mask frozen after 4 iterations, 3 levels + FFT
    pre-alignment.  FP32 and BF16 solves within 1e-3 vox RMS of the FP64
    solve; the learned seed with BF16 storage reaches 0.01 vox in <= 3 GN."""
    import torch
    from voxeldvc_amd import seednet
    from voxeldvc_amd.engine import correlate_gpu as C
    n = 513
    Ne = (n - 1) // h
    f, g, fv, gv, _, _ = _cube('c3', n, 3)
    kw = dict(K_ref_laplacian=O.build_K_ref_laplacian(), l0=4.5 * h, glt=0.0,
              dynamic_mask=True, freeze_mask_after=4, disp=False)
    print('C3 513^3 h=%d' % h, flush=True)
    t0 = time.perf_counter()
    U64, _, _ = C.multiscale_correlate_gpu(fv, gv, Ne, Ne, Ne, h, torch, dtype=np.float64,
                                           maxiter=60, eps=1e-6, **kw)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    out = {}
    for storage in ('fp32', 'bf16'):
        U, _, _ = C.multiscale_correlate_gpu(fv, gv, Ne, Ne, Ne, h, torch, storage=storage, **kw)
        out[storage] = _trms(U, U64)
    net = seednet.SeedNet.load()
    trace = []
    C.multiscale_correlate_gpu(fv, gv, Ne, Ne, Ne, h, torch, storage='bf16', scales=(0,),
                               learned_seed=net, maxiter=6,
                               callback=lambda s, ik, U: trace.append(_trms(U, U64)), **kw)
    k = next((i + 1 for i, r in enumerate(trace) if r <= 0.01), None)
    print('  FP64 solve %.0f ms; rms vs FP64: fp32 %.2e, bf16 %.2e vox; learned + bf16 per GN '
          'iteration %s (0.01 vox at GN %s)'
          % (t64 * 1e3, out['fp32'], out['bf16'], ['%.2e' % r for r in trace], k), flush=True)
    assert out['fp32'] < 1e-3 and out['bf16'] < 1e-3
    assert k is not None and k <= 3
    return f, g, fv, gv, kw, U64


def test_c3_513_h8(gpu):
    _c3_common(8)


def test_c3_513_h4(gpu):
    _c3_common(4)


def test_c3_513_h8_fp64_refinement(gpu):
    """FP32-against-FP64 iterative refinement on the full C3 workload at
    h = 8.  The plain FP32 multiscale solve (dynamic mask, not frozen)
    converges to dU/U < 1e-3; the refinement then holds the b-mask at the
    converged U and iterates with FP64 residuals.  Its target is the FP64 GN
    solve of the same functional: correlate_gpu in float64 from the same U
    with the mask frozen at its first iteration.  The FP64 residual
    ||b - ll R U|| falls at every step and U ends much closer to that
    solution than the unrefined iterate."""
    import torch
    from voxeldvc_amd.engine import correlate_gpu as C
    n, h = 513, 8
    Ne = (n - 1) // h
    _, _, fv, gv, _, _ = _cube('c3', n, 3)
    kw = dict(K_ref_laplacian=O.build_K_ref_laplacian(), l0=4.5 * h, glt=0.0,
              dynamic_mask=True, freeze_mask_after=None, disp=False)
    print('C3 513^3 h=8 FP64 refinement', flush=True)
    last = {}

    def keep(s, ik, U):
        if s == 0:
            last['U'] = U.clone()
    log = []
    Ur, _, _ = C.multiscale_correlate_gpu(fv, gv, Ne, Ne, Ne, h, torch, refine_fp64=4, log=log,
                                          callback=keep, **kw)
    U32 = last['U']
    res = [q['residual_fp64'] for lv in log for q in lv['iters'] if 'refine' in q]
    U64, _, _ = C.correlate_gpu(fv, gv, Ne, Ne, Ne, h, torch, U0=U32.double(),
                                dtype=np.float64, maxiter=60, eps=1e-9,
                                **dict(kw, freeze_mask_after=0))
    e0, e1 = _trms(U32, U64), _trms(Ur, U64)
    print('  FP64 residual %s; rms vs the FP64 solve of the same functional: unrefined %.2e, '
          'refined %.2e vox' % (['%.3e' % r for r in res], e0, e1), flush=True)
    assert Ur.dtype == torch.float64 and len(res) == 5
    assert all(b < a for a, b in zip(res, res[1:]))
    assert res[-1] < 0.2 * res[0]
    assert e1 < 0.2 * e0


def test_c5_full_learned_mixed(gpu):
    """C5 at full size: 1025 x 1025 x 769 FP16 speckle (806M voxels), necking
    + 8 Fourier modes (peak 8 vox) + N(0,2), h = 16 (64 x 64 x 48
    elements), equilibrium gap on the finest level.  The learned seed with
    BF16 storage reaches 0.01 vox RMS of the FP64 solve of a 257^3
    subvolume at the volume centre (its interior nodes, 2-element margin);
    peak device memory is reported."""
    import torch
    from voxeldvc_amd import seednet, synth
    from voxeldvc_amd.engine import correlate_gpu as C
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    free0 = torch.cuda.mem_get_info()[0]
    shape, h = (1025, 1025, 769), 16
    N = tuple((s - 1) // h for s in shape)
    print('C5 1025x1025x769 h=16', flush=True)
    f, g, _, _ = synth.config_pair_torch('c5', shape, seed=5)
    f = f.half()
    g = g.half()
    fv = f.view(shape[2], shape[1], shape[0]).permute(2, 1, 0)
    gv = g.view(shape[2], shape[1], shape[0]).permute(2, 1, 0)
    kw = dict(K_ref_laplacian=O.build_K_ref_laplacian(), K_ref_elastic=O.build_K_ref(0.3),
              reg_type='equilibrium_gap', l0=4.5 * h, glt=0.0, dynamic_mask=True,
              freeze_mask_after=4, disp=False)
    o = tuple(((s - 1) // 2 - 128) // h * h for s in shape)
    sub = fv[o[0]:o[0] + 257, o[1]:o[1] + 257, o[2]:o[2] + 257]
    Ns = (16, 16, 16)
    Us = _lap_then_eqgap(C, torch, sub.float().contiguous(), gv.float(), Ns, h,
                         dict(dtype=np.float64, maxiter=60, eps=1e-6,
                              g_origin=tuple(-q for q in o)),
                         scales=(2, 1, 0), ms_extra=dict(fft_prealign=False))
    io = [q // h for q in o]
    sl = (slice(2, Ns[2] - 1), slice(2, Ns[1] - 1), slice(2, Ns[0] - 1))
    ref_nodes = Us.view(Ns[2] + 1, Ns[1] + 1, Ns[0] + 1, 3)[sl]

    def sub_rms(U):
        Uf = U.view(N[2] + 1, N[1] + 1, N[0] + 1, 3)[io[2]:io[2] + Ns[2] + 1,
                                                     io[1]:io[1] + Ns[1] + 1,
                                                     io[0]:io[0] + Ns[0] + 1][sl]
        return float(((Uf.double() - ref_nodes) ** 2).mean().sqrt())

    net = seednet.SeedNet.load()
    trace = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Ul, _, _ = C.multiscale_correlate_gpu(fv, gv, N[0], N[1], N[2], h, torch, storage='bf16',
                                          scales=(0,), learned_seed=net,
                                          callback=lambda s, ik, U: trace.append(sub_rms(U)),
                                          **kw)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3
    peak = free0 - torch.cuda.mem_get_info()[0]
    k = next((i + 1 for i, e in enumerate(trace) if e <= 0.01), None)
    print('  learned + BF16: %.0f ms incl. per-iteration checks, subvolume rms per GN %s '
          '(0.01 vox at GN %s); device memory in use %.1f GB'
          % (ms, ['%.2e' % e for e in trace], k, peak / 1e9), flush=True)
    assert Ul.numel() == 3 * (N[0] + 1) * (N[1] + 1) * (N[2] + 1)
    assert k is not None and k <= 4
    assert sub_rms(Ul) <= 0.01
