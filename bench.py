"""Benchmark of the global GN DVC hot path on one MI355X.

Step = one correlation-Hessian matvec H v over a 513^3 voxel reference volume
with a 64^3-element Hex8 mesh (h = 8), the configuration the metric
("Hessian matvec Gvox/s (% HBM peak); GN time-to-solve at 0.01-vox RMS") is
quoted on.  Synthetic speckle (Gaussian blobs, quantised to 0..255, stored
float16), gradients in bfloat16 with the mask folded in (6 B/voxel); the
float32 + uint8-mask form (13 B/voxel) is timed beside it.  The GN
time-to-solve of the plain float32 multiscale solve and of the learned-seed
mixed-precision solve on a 513^3 pair is measured after it (`time_to_solve`).
Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 513] [--h 8]
                    [--no-solve] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument('--gpus', type=int, default=1)
    p.add_argument('--steps', type=int, default=50)
    p.add_argument('--warmup', type=int, default=5)
    p.add_argument('--size', type=int, default=513)
    p.add_argument('--h', type=int, default=8)
    p.add_argument('--no-cpu', action='store_true')
    p.add_argument('--stream-probe', action='store_true',
                   help='also run one axpby over 2^28 floats (2 GiB read, 1 GiB written): a '
                        'streaming kernel of known size that scales HBM counter readings')
    p.add_argument('--only', choices=('bf16', 'f32'), default=None,
                   help='time one storage form only (for profiling runs)')
    p.add_argument('--no-solve', action='store_true',
                   help='skip the GN time-to-solve measurement')
    p.add_argument('--solve-reps', type=int, default=3)
    p.add_argument('--no-post', action='store_true',
                   help='skip the post-solve voxel stages')
    p.add_argument('--no-affine', action='store_true',
                   help='skip the preprocess affine stage')
    p.add_argument('--solve-only', choices=('plain', 'learned'), default=None,
                   help='run one GN solve of the time-to-solve pair (after a warm-up solve) '
                        'and exit: for kernel traces of a whole solve')
    p.add_argument('--summarize-trace', default=None, metavar='CSV',
                   help='print the per-kernel summary of the last --solve-only window of a '
                        'rocprofv3 kernel-trace CSV and exit (no GPU needed)')
    p.add_argument('--pcg-levels', action='store_true',
                   help='time device PCG iterations (fixed count) on the three pyramid levels '
                        'of the 513^3, h=8 solve and exit: per-iteration cost of the PCG')
    p.add_argument('--configs', action='store_true',
                   help='also measure the other benchmark configurations (C2, C3 at h=4, C5); '
                        'several minutes')
    return p.parse_args()


def _timed_trace(solve_fn, k_target, reps):
    """Median over `reps` runs of (ms to the k_target-th GN update, ms to the
    end of the solve, the solve's result): every GN update records a HIP
    event on the library's stream (the callback runs right after U += dU),
    so the time to an iterate carries no host synchronisation of its own."""
    import torch
    runs = []
    for _ in range(reps):
        evs = []
        st = torch.cuda.current_stream()

        def cb(s, ik, U):
            e = torch.cuda.Event(enable_timing=True)
            e.record(st)
            evs.append(e)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(st)
        out = solve_fn(cb)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) * 1e3
        t_k = e0.elapsed_time(evs[k_target - 1]) if k_target and k_target <= len(evs) else None
        runs.append((t_k, wall, out))
    order = np.argsort([r[1] for r in runs])
    mid = runs[int(order[len(runs) // 2])]
    tk = sorted(r[0] for r in runs if r[0] is not None)
    return (tk[len(tk) // 2] if tk else None), mid[1], mid[2]


def time_to_solve(n=513, h=8, reps=3, log=None):
    """GN time-to-solve on one 8-bit speckle pair of n^3 voxels warped by 16
    random Fourier modes (wavelengths 64-256 vox, peak 4 vox) plus N(0,1)
    noise, Laplacian regularizer (l0 = 4.5 h), dynamic mask frozen after 4
    iterations -- the production kwargs of the reference's run command.

    * oracle: the same solver in float64 (dtype=np.float64), converged to
      dU/U < 1e-6, defines the answer;
    * plain: 3-level multiscale (scales 2,1,0) with FFT pre-alignment,
      float32;
    * learned + mixed: the seed network's prediction starts the finest level
      (no coarse levels), float16 images / bfloat16 gradients, float32
      matvec arithmetic, FP64 dots and norms.
    Both solves are timed the same way: from device-resident f and g to the
    first finest-level GN iterate within 0.01 voxel RMS of the oracle (the
    iterate is found by a traced run; the timed runs record a HIP event at
    every GN update), and to the solver's own convergence (dU/U < 1e-3).
    Each timing is the median of `reps` solves."""
    import torch
    from voxeldvc_amd import seednet, synth
    from voxeldvc_amd.engine import correlate_gpu as C
    from voxeldvc_amd.engine.geometry_dvc import build_K_ref_laplacian
    Ne = (n - 1) // h
    f = synth.speckle_torch((n, n, n), synth.blobs_for((n, n, n)), seed=3).round().clamp_(0, 255)
    modes = synth.fourier_params(16, (64, 256), 4.0, seed=7, extent=(n, n, n))
    g = synth.warp_backward_torch(f, (n, n, n), modes, 1.0, seed=11, quantize=True)
    fv = f.view(n, n, n).permute(2, 1, 0)
    gv = g.view(n, n, n).permute(2, 1, 0)
    kw = dict(K_ref_laplacian=build_K_ref_laplacian(), l0=4.5 * h, dynamic_mask=True,
              freeze_mask_after=4, glt=0.0, disp=False, pcg_tol=1e-6, pcg_max_iter=2000)
    net = seednet.SeedNet.load()
    plain = dict(maxiter=30, eps=1e-3)
    learned = dict(maxiter=30, eps=1e-3, storage='bf16', scales=(0,), learned_seed=net)

    def run(extra, cb=None, lg=None):
        return C.multiscale_correlate_gpu(fv, gv, Ne, Ne, Ne, h, torch, log=lg, callback=cb,
                                          **{**kw, **extra})[0]

    t0 = time.perf_counter()
    lg_o = []
    Uo = run(dict(dtype=np.float64, maxiter=60, eps=1e-6), lg=lg_o).double()
    torch.cuda.synchronize()
    t_o = (time.perf_counter() - t0) * 1e3
    rms = lambda U: float(((U.double() - Uo) ** 2).mean().sqrt())

    def trace(extra):
        """(per-update (scale, rms) list, GN and PCG counts) of one solve."""
        tr, lg = [], []
        U = run(extra, lambda s, ik, U: tr.append((s, rms(U))), lg)
        gn = sum(len(lv['iters']) for lv in lg)
        pcg = sum(q['pcg_iters'] for lv in lg for q in lv['iters'])
        return tr, gn, pcg, rms(U)

    def first_within(tr, tol=0.01):
        return next((i + 1 for i, (s, r) in enumerate(tr) if s == 0 and r <= tol), None)

    run(plain)                      # warm-up (level buffers, kernels)
    tr_p, gn_p, pcg_p, rms_p = trace(plain)
    k_p = first_within(tr_p)
    tr_l, gn_l, pcg_l, rms_l = trace(learned)
    k_l = first_within(tr_l)
    if k_p is None or k_l is None:
        raise RuntimeError('a solve did not reach 0.01 vox RMS: plain %r, learned %r'
                           % (tr_p, tr_l))
    tp01, tp, _ = _timed_trace(lambda cb: run(plain, cb), k_p, reps)
    tl01, tl, _ = _timed_trace(lambda cb: run(learned, cb), k_l, reps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Un = net.predict(f, g, (n, n, n), Ne, Ne, Ne, h)
    torch.cuda.synchronize()
    t_net = (time.perf_counter() - t0) * 1e3
    return {
        "workload": "%d^3 8-bit speckle, 16 Fourier modes (wl 64-256, peak 4 vox) + N(0,1), "
                    "h=%d (%d^3 elements), Laplacian l0=%.0f" % (n, h, Ne, 4.5 * h),
        "criterion": "first finest-level GN iterate with nodal RMS <= 0.01 vox vs the float64 "
                     "solve (dU/U < 1e-6); both solves timed to it the same way (HIP event at "
                     "every GN update, median of %d)" % reps,
        "oracle_fp64": {"ms": round(t_o, 2), "gn_iters": sum(len(lv['iters']) for lv in lg_o)},
        "plain_fp32": {"ms_to_0.01": round(tp01, 2), "gn_to_0.01": k_p,
                       "ms_to_convergence": round(tp, 2), "gn_iters": gn_p,
                       "pcg_iters": pcg_p, "rms_vox_converged": rms_p,
                       "levels": "2,1,0 + FFT prealign",
                       "per_iter_rms_finest": [float('%.3g' % r) for s, r in tr_p if s == 0]},
        "learned_mixed": {"ms_to_0.01": round(tl01, 2), "gn_to_0.01": k_l,
                          "ms_to_convergence": round(tl, 2), "gn_iters": gn_l,
                          "pcg_iters": pcg_l, "rms_vox_converged": rms_l,
                          "network_ms": round(t_net, 2), "network_rms_vox": rms(Un),
                          "per_iter_rms": [float('%.3g' % r) for s, r in tr_l]},
        "ratio_to_0.01": round(tl01 / tp01, 4),
        "ratio_converged": round(tl / tp, 4),
        "target_ratio": 1.0 / 3.0,
    }


def solve_only(kind, n=513, h=8):
    """One plain FP32 multiscale or learned-seed BF16 solve of the
    time-to-solve pair (after one untimed warm-up solve), for kernel traces
    of a whole solve; prints its wall time."""
    import torch
    from voxeldvc_amd import seednet, synth
    from voxeldvc_amd.engine import correlate_gpu as C
    from voxeldvc_amd.engine.geometry_dvc import build_K_ref_laplacian
    Ne = (n - 1) // h
    f = synth.speckle_torch((n, n, n), synth.blobs_for((n, n, n)), seed=3).round().clamp_(0, 255)
    modes = synth.fourier_params(16, (64, 256), 4.0, seed=7, extent=(n, n, n))
    g = synth.warp_backward_torch(f, (n, n, n), modes, 1.0, seed=11, quantize=True)
    fv = f.view(n, n, n).permute(2, 1, 0)
    gv = g.view(n, n, n).permute(2, 1, 0)
    kw = dict(K_ref_laplacian=build_K_ref_laplacian(), l0=4.5 * h, dynamic_mask=True,
              freeze_mask_after=4, glt=0.0, disp=False, pcg_tol=1e-6, pcg_max_iter=2000,
              maxiter=30, eps=1e-3)
    if kind == 'learned':
        kw.update(storage='bf16', scales=(0,), learned_seed=seednet.SeedNet.load())
    for rep in range(2):
        torch.cuda.synchronize()
        # a spin kernel on each side of the solve marks its window in a
        # kernel trace (summarize_trace)
        torch.cuda._sleep(1000)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        C.multiscale_correlate_gpu(fv, gv, Ne, Ne, Ne, h, torch, **kw)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) * 1e3
        torch.cuda._sleep(1000)
        torch.cuda.synchronize()
        print('%s solve %s: %.2f ms' % (kind, 'warm-up' if rep == 0 else 'timed', dt), flush=True)


def summarize_trace(path):
    """Per-kernel summary of the last solve window of a rocprofv3 kernel
    trace (CSV) of `--solve-only`: the kernels between the last two spin
    markers, with the count of kernels that are not the library's."""
    import csv
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r['Start_Timestamp']))
    mk = [i for i, r in enumerate(rows) if 'sleep' in r['Kernel_Name'].lower()
          or 'spin' in r['Kernel_Name'].lower()]
    if len(mk) < 2:
        raise SystemExit('no marker pair in %s' % path)
    win = rows[mk[-2] + 1:mk[-1]]
    t0 = int(win[0]['Start_Timestamp'])
    t1 = int(win[-1]['End_Timestamp'])
    agg = {}
    for r in win:
        d = int(r['End_Timestamp']) - int(r['Start_Timestamp'])
        a = agg.setdefault(r['Kernel_Name'], [0, 0])
        a[0] += 1
        a[1] += d
    busy = sum(v[1] for v in agg.values())
    print('solve window: %d kernels, %.3f ms first start to last end, %.3f ms busy'
          % (len(win), (t1 - t0) / 1e6, busy / 1e6))
    print('%-8s %10s %10s  %s' % ('calls', 'total ms', 'avg us', 'kernel'))
    for name, (n, d) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        print('%-8d %10.3f %10.2f  %s' % (n, d / 1e6, d / n / 1e3, name[:110]))
    # torch's own kernels carry its namespaces; the runtime's copy kernels
    # (__amd_rocclr_*) are memory copies
    other = [nm for nm in agg if 'at::' in nm or 'c10::' in nm or 'torch' in nm]
    copies = [nm for nm in agg if nm.startswith('__amd_rocclr')]
    print('torch kernels in the window: %s; runtime copy kernels: %s'
          % (other or 'none', copies or 'none'))


def _matvec_kernel_ms(n_el, h, G, gt, mask, reps=20):
    """Mean vox_vec kernel time (HIP events on the library stream) of an H
    matvec over device gradient planes G."""
    import torch
    from voxeldvc_amd import _lib as L
    c = L.ctx()
    Ndof = 3 * (n_el[0] + 1) * (n_el[1] + 1) * (n_el[2] + 1)
    v = torch.randn(Ndof, device='cuda')
    out = L.empty(Ndof)
    args = (c, n_el[0], n_el[1], n_el[2], h, G[0], G[1], G[2], gt, mask, v, out)
    L.call('vdvc_matvec_H', *args)
    torch.cuda.synchronize()
    L.call('vdvc_timing_begin', c)
    for _ in range(reps):
        L.call('vdvc_matvec_H', *args)
    tot = (L._d * 1)()
    nl = (L._i * 1)()
    L.call('vdvc_timing_end', c, tot, nl)
    return tot[0] / max(1, nl[0])


def _level_planes(f, shape, n_el, h, gtype):
    """Pitched gradient planes (+ mask) of a dense x-fastest image, as the
    solver builds them: (planes, mask0) for F32, folded planes for BF16."""
    from voxeldvc_amd import _lib as L
    c = L.ctx()
    fp = L.pitch_dense(f, shape, 'float16')
    Np = fp.numel()
    dt = 'bfloat16' if gtype == L.BF16 else 'float32'
    G = [L.empty(Np, dt) for _ in range(3)]
    m = L.empty(Np, 'uint8')
    L.call('vdvc_image_setup', c, n_el[0], n_el[1], n_el[2], h, fp, L.F16, 0.0, G[0], G[1], G[2],
           gtype, m)
    if gtype == L.BF16:
        W = [L.empty(Np, 'bfloat16') for _ in range(3)]
        L.call('vdvc_fold_grad', c, Np, G[0], G[1], G[2], m, W[0], W[1], W[2])
        return W, None
    return G, m


def _wall_ms(fn, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps * 1e3


def config_cases():
    """Measurements the other benchmark configurations ask for (one GPU):
      C2 257^3 necking, equilibrium gap, 4 levels: H and regularizer matvec
         rates, FP32 solve time (and its distance to the FP64 solve);
      C3 513^3 at h = 4: H matvec rates, FP32 / BF16 / FP64 solve times and
         distances to the FP64 solve;
      C5 1025 x 1025 x 769 (h = 16, equilibrium gap, necking + 8 Fourier
         modes up to 8 vox + N(0,2)): BF16 H matvec % of peak, peak device
         memory, and the learned-seed mixed-precision time to 0.01 vox RMS
         against the FP64 solve of a 257^3 subvolume (interior nodes)."""
    import torch
    from voxeldvc_amd import _lib as L
    from voxeldvc_amd import seednet, synth
    from voxeldvc_amd.engine import correlate_gpu as C
    from voxeldvc_amd.engine.geometry_dvc import build_K_ref, build_K_ref_laplacian
    out = {}
    Kl, Kel = build_K_ref_laplacian(), build_K_ref(0.3)

    def rms(a, b):
        return float(((a.double() - b.double()) ** 2).mean().sqrt())

    def solve(fv, gv, N, h, **kw):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        lg = []
        U, _, _ = C.multiscale_correlate_gpu(fv, gv, N[0], N[1], N[2], h, torch, log=lg, **kw)
        torch.cuda.synchronize()
        return U, (time.perf_counter() - t0) * 1e3, sum(len(lv['iters']) for lv in lg)

    def lap_eqgap(fv, gv, N, h, g_origin=(0, 0, 0), fft_prealign=True, scales=(3, 2, 1, 0),
                  cb=None, **extra):
        """Laplacian pyramid, then the equilibrium gap on the finest level (the
        equilibrium gap on the coarse levels stalls the reference's PCG);
        cb(scale, ik, U) after every GN update of both solves."""
        lap = dict(K_ref_laplacian=Kl, l0=4.5 * h, glt=0.0, dynamic_mask=True,
                   freeze_mask_after=4, disp=False, g_origin=g_origin)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        l1, l2 = [], []
        U0, _, _ = C.multiscale_correlate_gpu(fv, gv, N[0], N[1], N[2], h, torch, log=l1,
                                              scales=scales, fft_prealign=fft_prealign,
                                              callback=cb, **lap, **extra)
        U, _, _ = C.correlate_gpu(fv, gv, N[0], N[1], N[2], h, torch, U0=U0, log=l2,
                                  K_ref_elastic=Kel, reg_type='equilibrium_gap',
                                  callback=None if cb is None else (lambda ik, Ud: cb(0, ik, Ud)),
                                  **lap, **extra)
        torch.cuda.synchronize()
        return U, (time.perf_counter() - t0) * 1e3, sum(len(lv['iters']) for lv in l1) + len(l2)

    # ---- C2 ----
    print('config C2', file=sys.stderr, flush=True)
    n, h = 257, 8
    N = ((n - 1) // h,) * 3
    Nvox = n ** 3
    Ndof = 3 * (N[0] + 1) ** 3
    f, g, _, _ = synth.config_pair_torch('c2', (n, n, n), seed=2)
    r = {}
    for name, gt, bpv in (('bf16_folded', L.BF16, 6), ('fp32_mask', L.F32, 13)):
        G, m = _level_planes(f, (n, n, n), N, h, gt)
        ms = _matvec_kernel_ms(N, h, G, gt, m)
        B = Nvox * bpv + 8 * Ndof
        r['H_matvec_' + name] = {"kernel_ms": round(ms, 4), "Gvox_s": round(Nvox / ms / 1e6, 1),
                                 "frac_8TBs": round(B / (ms * 1e-3) / 8e12, 4)}
        del G, m
    E = float(h)
    v = torch.randn(Ndof, device='cuda')
    o = torch.empty_like(v)
    kp = np.ascontiguousarray(Kel, dtype=np.float64)
    kl = np.ascontiguousarray(Kl, dtype=np.float64)
    ms = _wall_ms(lambda: L.call('vdvc_matvec_eqgap', L.ctx(), N[0], N[1], N[2], L.F32,
                                 kp.ctypes.data_as(L._dp), E, None, v, o), 50)
    nn = Ndof // 3
    r['eqgap_matvec'] = {"ms_wall": round(ms, 4), "Gnode_s": round(nn / ms / 1e6, 2),
                         "algorithmic_bytes": 48 * nn, "GB_s": round(48 * nn / ms / 1e6, 1)}
    ms = _wall_ms(lambda: L.call('vdvc_matvec_K', L.ctx(), N[0], N[1], N[2], L.F32,
                                 kl.ctypes.data_as(L._dp), E, None, v, o), 50)
    r['laplacian_matvec'] = {"ms_wall": round(ms, 4), "Gnode_s": round(nn / ms / 1e6, 2),
                             "algorithmic_bytes": 24 * nn, "GB_s": round(24 * nn / ms / 1e6, 1)}
    fv = f.view(n, n, n).permute(2, 1, 0)
    gv = g.view(n, n, n).permute(2, 1, 0)
    U64, t64, gn64 = lap_eqgap(fv, gv, N, h, dtype=np.float64, maxiter=60, eps=1e-6)
    lap_eqgap(fv, gv, N, h)
    U32, t32, gn32 = lap_eqgap(fv, gv, N, h)
    r['solve_fp32'] = {"ms": round(t32, 2), "gn_iters": gn32, "rms_vs_fp64": rms(U32, U64),
                       "levels": "Laplacian 3,2,1,0 then equilibrium gap at scale 0"}
    r['solve_fp64'] = {"ms": round(t64, 2), "gn_iters": gn64}
    out['C2_257cube_necking_eqgap_4levels'] = r
    del f, g, fv, gv, v, o

    # ---- C3 at h = 4 ----
    print('config C3 h=4', file=sys.stderr, flush=True)
    n, h = 513, 4
    N = ((n - 1) // h,) * 3
    Nvox = n ** 3
    Ndof = 3 * (N[0] + 1) ** 3
    f, g, _, _ = synth.config_pair_torch('c3', (n, n, n), seed=3)
    r = {}
    for name, gt, bpv in (('bf16_folded', L.BF16, 6), ('fp32_mask', L.F32, 13)):
        G, m = _level_planes(f, (n, n, n), N, h, gt)
        ms = _matvec_kernel_ms(N, h, G, gt, m)
        B = Nvox * bpv + 8 * Ndof
        r['H_matvec_' + name] = {"kernel_ms": round(ms, 4), "Gvox_s": round(Nvox / ms / 1e6, 1),
                                 "frac_8TBs": round(B / (ms * 1e-3) / 8e12, 4)}
        del G, m
    fv = f.view(n, n, n).permute(2, 1, 0)
    gv = g.view(n, n, n).permute(2, 1, 0)
    kw = dict(K_ref_laplacian=Kl, l0=4.5 * h, glt=0.0, dynamic_mask=True, freeze_mask_after=4,
              disp=False)
    U64, t64, gn64 = solve(fv, gv, N, h, dtype=np.float64, maxiter=60, eps=1e-6, **kw)
    solve(fv, gv, N, h, **kw)
    for st in ('fp32', 'bf16'):
        U, t, gn = solve(fv, gv, N, h, storage=st, **kw)
        r['solve_' + st] = {"ms": round(t, 2), "gn_iters": gn, "rms_vs_fp64": rms(U, U64)}
    U, t, gn = solve(fv, gv, N, h, storage='bf16', refine_fp64=2, **kw)
    r['solve_bf16_refined2'] = {"ms": round(t, 2), "rms_vs_fp64": rms(U, U64)}
    r['solve_fp64'] = {"ms": round(t64, 2), "gn_iters": gn64}
    out['C3_513cube_h4'] = r
    del f, g, fv, gv

    # ---- C5 ----
    print('config C5', file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    # device memory reserved from here on (torch's cache keeps its peak
    # reservation, the library's buffers only grow): the C5 footprint
    free0 = torch.cuda.mem_get_info()[0]
    shape, h = (1025, 1025, 769), 16
    N = tuple((s - 1) // h for s in shape)
    Nvox = shape[0] * shape[1] * shape[2]
    Ndof = 3 * (N[0] + 1) * (N[1] + 1) * (N[2] + 1)
    f, g, _, _ = synth.config_pair_torch('c5', shape, seed=5)
    f = f.half()
    g = g.half()
    r = {}
    G, m = _level_planes(f.float(), shape, N, h, L.BF16)
    ms = _matvec_kernel_ms(N, h, G, L.BF16, m)
    B = Nvox * 6 + 8 * Ndof
    r['H_matvec_bf16_folded'] = {"kernel_ms": round(ms, 4), "Gvox_s": round(Nvox / ms / 1e6, 1),
                                 "frac_8TBs": round(B / (ms * 1e-3) / 8e12, 4)}
    del G, m
    fv = f.view(shape[2], shape[1], shape[0]).permute(2, 1, 0)
    gv = g.view(shape[2], shape[1], shape[0]).permute(2, 1, 0)
    kw = dict(K_ref_laplacian=Kl, K_ref_elastic=Kel, reg_type='equilibrium_gap', l0=4.5 * h,
              glt=0.0, dynamic_mask=True, freeze_mask_after=4, disp=False)
    # FP64 answer on a 257^3 subvolume at the volume centre (g_origin = -offset)
    o = tuple(((s - 1) // 2 - 128) // h * h for s in shape)
    sub = fv[o[0]:o[0] + 257, o[1]:o[1] + 257, o[2]:o[2] + 257]
    Ns = (16, 16, 16)
    Us, _, _ = lap_eqgap(sub.float().contiguous(), gv.float(), Ns, h, dtype=np.float64,
                         maxiter=60, eps=1e-6, g_origin=tuple(-q for q in o),
                         fft_prealign=False, scales=(2, 1, 0))
    # full-mesh nodes of the subvolume's interior nodes (2-element margin)
    io = [q // h for q in o]
    sl = (slice(2, Ns[2] - 1), slice(2, Ns[1] - 1), slice(2, Ns[0] - 1))
    ref_nodes = Us.view(Ns[2] + 1, Ns[1] + 1, Ns[0] + 1, 3)[sl]

    def sub_rms(U):
        Uf = U.view(N[2] + 1, N[1] + 1, N[0] + 1, 3)[io[2]:io[2] + Ns[2] + 1,
                                                     io[1]:io[1] + Ns[1] + 1,
                                                     io[0]:io[0] + Ns[0] + 1][sl]
        return float(((Uf.double() - ref_nodes) ** 2).mean().sqrt())

    net = seednet.SeedNet.load()

    def learned_run(cb):
        return C.multiscale_correlate_gpu(fv, gv, N[0], N[1], N[2], h, torch, storage='bf16',
                                          scales=(0,), learned_seed=net, callback=cb, **kw)[0]

    def plain_run(cb):
        return lap_eqgap(fv, gv, N, h, cb=cb)[0]

    def first_within(tr, tol=0.01):
        return next((i + 1 for i, (sc, e) in enumerate(tr) if sc == 0 and e <= tol), None)

    # traced runs find the first finest-level iterate within 0.01 vox of the
    # subvolume answer; the timed runs record a HIP event at every GN update
    tr_l, tr_p = [], []
    lg_l = []
    U = C.multiscale_correlate_gpu(fv, gv, N[0], N[1], N[2], h, torch, storage='bf16',
                                   scales=(0,), learned_seed=net, log=lg_l,
                                   callback=lambda sc, ik, Ud: tr_l.append((sc, sub_rms(Ud))),
                                   **kw)[0]
    Up, _, gnp = lap_eqgap(fv, gv, N, h, cb=lambda sc, ik, Ud: tr_p.append((sc, sub_rms(Ud))))
    k_l, k_p = first_within(tr_l), first_within(tr_p)
    r['learned_bf16_per_iter_rms_subvolume'] = [float('%.3g' % e) for _, e in tr_l]
    r['plain_per_iter_rms_subvolume_finest'] = [float('%.3g' % e) for sc, e in tr_p if sc == 0]
    tl01, tl, _ = _timed_trace(learned_run, k_l, 1)
    tp01, tp, _ = _timed_trace(plain_run, k_p, 1)
    r['learned_bf16'] = {"ms_to_0.01": None if tl01 is None else round(tl01, 2), "gn_to_0.01": k_l,
                         "ms_to_convergence": round(tl, 2),
                         "gn_iters": sum(len(lv['iters']) for lv in lg_l),
                         "rms_subvolume": sub_rms(U)}
    r['plain_fp32_4levels'] = {"ms_to_0.01": None if tp01 is None else round(tp01, 2),
                               "gn_to_0.01": k_p, "ms_to_convergence": round(tp, 2),
                               "gn_iters": gnp, "rms_subvolume": sub_rms(Up),
                               "levels": "Laplacian 3,2,1,0 then equilibrium gap at scale 0"}
    if tl01 is not None and tp01 is not None:
        r['ratio_to_0.01'] = round(tl01 / tp01, 4)
    r['ratio_converged'] = round(tl / tp, 4)
    torch.cuda.synchronize()
    r['peak_device_bytes'] = int(free0 - torch.cuda.mem_get_info()[0])
    out['C5_1025x1025x769_h16'] = r
    return out


TRAFFIC_FILE = os.path.join(ROOT, 'profiles', 'hessian_matvec_traffic.json')


def post_stages(n=513, h=8, reps=10):
    """The voxel-scale post-solve stages of a correlation run (the GPU part of
    the reference's write_outputs) at n^3 voxels, h: kernel time per launch
    (HIP events on the library stream, mean of `reps` after one warm-up)
    against the stage's algorithmic HBM bytes.  Inputs are device resident:
    an 8-bit speckle reference f (uint8, as preprocess keeps the CT dtype),
    a smooth Fourier displacement U (float32, up to ~4 vox), the per-voxel
    float32 residual (zero outside the active mask) and mask0 = f > 0."""
    import torch
    from voxeldvc_amd import _lib as L
    from voxeldvc_amd import synth
    n_el = (n - 1) // h
    Nv = n_el * h + 1
    Nvox = Nv ** 3
    Nn = (n_el + 1) ** 3
    Ne = n_el ** 3
    c = L.ctx()
    f = synth.speckle_torch((Nv, Nv, Nv), synth.blobs_for((Nv, Nv, Nv)), seed=5)
    f8 = f.round().clamp_(0, 255).to(torch.uint8).view(-1)
    del f
    modes = synth.fourier_params(16, (64, 256), 4.0, 11, (Nv, Nv, Nv))
    U = torch.from_numpy(synth.fourier_u_nodes(modes, n_el, n_el, n_el, h).astype(np.float32)).cuda()
    mask0 = (f8 > 0).to(torch.uint8)
    gen = torch.Generator(device='cuda').manual_seed(0)
    res = torch.randn(Nvox, device='cuda', generator=gen) * 6
    res = res * mask0
    out = {}

    def timed(name, call, nbytes):
        call()
        torch.cuda.synchronize()
        L.call('vdvc_timing_begin', c)
        for _ in range(reps):
            call()
        tot = (L._d * 1)()
        nl = (L._i * 1)()
        L.call('vdvc_timing_end', c, tot, nl)
        ms = tot[0] / max(1, nl[0])
        gbs = nbytes / (ms * 1e-3) / 1e9
        out[name] = {"kernel_ms": round(ms, 4), "algorithmic_bytes": int(nbytes),
                     "achieved_GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}

    ui = L.empty(3 * Nvox)
    timed("interp_field_vox (float32, channels_first)",
          lambda: L.call('vdvc_interp_field', c, n_el, n_el, n_el, h, U, L.F32, ui, 1),
          12 * Nvox + 12 * Nn)
    del ui
    act = L.empty(Nvox, 'uint8')
    timed("compute_active_mask (uint8 g, mask0)",
          lambda: L.call('vdvc_active_mask', c, n_el, n_el, n_el, h, U, L.F32, f8, L.U8, Nv, Nv,
                         Nv, 0.0, 0.0, 0.0, mask0, 0.0, act),
          3 * Nvox + 12 * Nn)
    pushed = L.empty(Nvox, 'uint8')
    timed("push_reference_forward (uint8, order 1, nearest)",
          lambda: L.call('vdvc_push_forward', c, n_el, n_el, n_el, h, U, L.F32, f8, L.U8, 1, 0,
                         0.0, pushed),
          2 * Nvox + 12 * Nn)
    timed("push_field_forward (bool, order 0, constant cval=1)",
          lambda: L.call('vdvc_push_forward', c, n_el, n_el, n_el, h, U, L.F32, mask0, L.BOOL, 0,
                         1, 1.0, pushed),
          2 * Nvox + 12 * Nn)
    std = torch.empty(Ne, dtype=torch.float64, device='cuda')
    zn = torch.empty(Ne, dtype=torch.float64, device='cuda')
    timed("compute_correlation_quality_cell (float32 res, uint8 ref)",
          lambda: L.call('vdvc_quality_cell', c, n_el, n_el, n_el, h, res, L.F32, f8, L.U8, std,
                         zn),
          5 * Nvox + 16 * Ne)
    # element-scale output fields (unsafe / overlap-lost masks, safe cells,
    # strains, GN convergence, the deformed-configuration mask remap)
    ue = torch.empty(Ne, dtype=torch.uint8, device='cuda')
    un = torch.empty(Nn, dtype=torch.uint8, device='cuda')
    uv = L.empty(Nvox, 'uint8')
    ol = torch.empty(Ne, dtype=torch.uint8, device='cuda')
    timed("cell masks (unsafe element / node / voxel + overlap-lost)",
          lambda: L.call('vdvc_cell_masks', c, n_el, n_el, n_el, h, act, mask0, 0.0, 0.5, ue, un,
                         uv, ol),
          3 * Nvox + 14 * Ne + Nn)
    sc = torch.empty(Ne, dtype=torch.uint8, device='cuda')
    timed("compute_safe_cell_mask", lambda: L.call('vdvc_safe_cells', c, n_el, n_el, n_el, h,
                                                   uv, sc), Nvox + Ne)
    ps = torch.empty(3 * Ne, dtype=torch.float64, device='cuda')
    eq = torch.empty(Ne, dtype=torch.float64, device='cuda')
    timed("principal + equivalent strains (float32 U)",
          lambda: L.call('vdvc_principal_strains', c, n_el, n_el, n_el, h, U, L.F32, ps, eq),
          12 * Nn + 32 * Ne)
    conv = torch.empty(Ne, dtype=torch.float64, device='cuda')
    timed("compute_gn_convergence_cell (float32 dU, U)",
          lambda: L.call('vdvc_gn_convergence', c, n_el, n_el, n_el, U, L.F32, U, L.F32, conv),
          24 * Nn + 8 * Ne)
    crop = (L._d * 3)(0.0, 0.0, 0.0)
    old_ = torch.empty(Ne, dtype=torch.uint8, device='cuda')
    timed("remap_element_mask_to_deformed (%d elements, grid nearest, incl. one host sync)" % Ne,
          lambda: L.call('vdvc_remap_mask_deformed', c, n_el, n_el, n_el, h, U, L.F32, crop, ol,
                         old_),
          12 * Nn + 50 * Ne)
    z = (L._d * 1)()
    L.call('vdvc_zncc', c, Nvox, f8, L.U8, res, L.F32, 1, act, z)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        L.call('vdvc_zncc', c, Nvox, f8, L.U8, res, L.F32, 1, act, z)
    ms = (time.perf_counter() - t0) / reps * 1e3
    gbs = 6 * Nvox / (ms * 1e-3) / 1e9
    out["compute_zncc (uint8 f, float32 res, mask; two passes, wall incl. 2 syncs)"] = {
        "ms": round(ms, 4), "algorithmic_bytes": 6 * Nvox, "achieved_GBs": round(gbs, 1),
        "frac": round(gbs / HBM_PEAK_GBS, 4)}
    total = sum(v.get("kernel_ms", v.get("ms", 0.0)) for v in out.values())
    return {"workload": "post-solve voxel and element stages, %d^3 voxels, h=%d (%d^3 "
                        "elements), uint8 speckle, Fourier U (16 modes, 4 vox), float32 "
                        "residual" % (Nv, h, n_el),
            "stages": out, "total_ms": round(total, 3)}


def post_cpu_baseline(n=129, h=8):
    """The oracle's NumPy post-solve stages (interp_field_vox, active mask,
    push, quality cell, ZNCC) on one host thread at n^3: voxels per second
    over the five stages together."""
    from threadpoolctl import threadpool_limits
    from oracle import post_oracle as P
    from voxeldvc_amd import synth
    n_el = (n - 1) // h
    Nv = n_el * h + 1
    f = synth.speckle_numpy((Nv, Nv, Nv), synth.blobs_for((Nv, Nv, Nv)), 5)
    f = np.clip(np.round(f), 0, 255).astype(np.uint8)
    modes = synth.fourier_params(16, (64, 256), 4.0, 11, (Nv, Nv, Nv))
    U = synth.fourier_u_nodes(modes, n_el, n_el, n_el, h).astype(np.float32)
    ff = f.reshape(-1, order='F')
    m0 = ff > 0
    res = (np.random.default_rng(0).standard_normal(Nv ** 3) * 6).astype(np.float32) * m0
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        P.interp_field_vox(U, h, n_el, n_el, n_el, channels_first=True)
        act = P.compute_active_mask(U, f, h, n_el, n_el, n_el, mask0_vox=m0)
        P.push_field_forward(f, U, n_el, n_el, n_el, h)
        P.compute_correlation_quality_cell(res, f, h, n_el, n_el, n_el)
        P.compute_zncc(ff, ff.astype(np.float64) - res, act)
        t = time.perf_counter() - t0
    return {"value": round(Nv ** 3 / t / 1e6, 3), "unit": "Mvox/s (five stages)", "cores": 1,
            "kind": "port", "sample": "the five stages once at %d^3, h=%d: %.2f s; %d host cores "
                                      "available" % (Nv, h, t, os.cpu_count())}


def affine_stage(n=513, reps=5):
    """The preprocess affine stage at n^3 (device-resident float32 volumes):
    f = 8-bit-range speckle, g = f pulled by a known affine (rotation-free
    strain/shear up to 3 %, translation ~1 vox) + N(0, 0.5).  Reports the
    whole affine_refine_gn (4 scales) wall time and its recovery error, the
    normal-equation kernel's time per GN step (HIP events) against its
    algorithmic bytes (f_s, g_s, 3 gradient planes, valid: 21 B/voxel), one
    Gaussian smoothing (sigma 4, 3 passes), warp_onto_ref_grid and
    compute_crop."""
    import torch
    from voxeldvc_amd import _lib as L
    from voxeldvc_amd import preprocess as P
    from voxeldvc_amd import synth
    from voxeldvc_amd.engine import correlate_gpu as C
    c_ = L.ctx()
    N = n ** 3
    f = synth.speckle_torch((n, n, n), synth.blobs_for((n, n, n)), seed=7)
    c = np.full(3, (n - 1) / 2.0)
    M = np.array([[0.02, -0.004, 0.003], [0.002, -0.012, 0.005], [-0.003, 0.004, 0.03]])
    tr = np.array([1.2, -0.8, 0.6])
    g = synth.warp_backward_torch(f, (n, n, n), None, 0.5, seed=8, quantize=False,
                                  affine=(np.eye(3) + M, c),
                                  extra_u=lambda X, Y, Z: (tr[0] + 0 * X, tr[1] + 0 * Y,
                                                           tr[2] + 0 * Z))
    fv = f.view(n, n, n).permute(2, 1, 0)
    gv = g.view(n, n, n).permute(2, 1, 0)
    out = {}
    C.affine_refine_gn(fv, gv, torch, np.eye(3), np.zeros(3), c, sigmas=(1.0,), max_iter=2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    A_g, t_g, info = C.affine_refine_gn(fv, gv, torch, np.eye(3), np.zeros(3), c)
    torch.cuda.synchronize()
    A_true = np.linalg.inv(np.eye(3) - M)
    out['affine_refine_gn'] = {
        "ms": round((time.perf_counter() - t0) * 1e3, 2), "gn_iters": info['iters'],
        "rms": [round(info['rms_initial'], 3), round(info['rms_final'], 4)],
        "max_err_A": float('%.3g' % np.abs(A_g - A_true).max()),
        "max_err_t_vox": float('%.3g' % np.abs(t_g - A_true @ tr).max())}
    # one normal-equation step at sigma 0 (device planes)
    valid = L.empty(N, 'uint8')
    C._smooth_masked_dev(g, (n, n, n), 0.0, 0.0, valid)
    gx, gy, gz = (L.empty(N) for _ in range(3))
    L.call('vdvc_gradient3', c_, n, n, n, f, gx, gy, gz)
    H = np.zeros((12, 12))
    b = np.zeros(12)
    dp = lambda a: np.ascontiguousarray(a, np.float64).ctypes.data_as(L._dp)
    Mz, tz, cc = np.ascontiguousarray(M), np.ascontiguousarray(tr), np.ascontiguousarray(c)

    def normal():
        L.call('vdvc_affine_normal', c_, n, n, n, f, g, gx, gy, gz, valid, dp(Mz), dp(tz),
               dp(cc), 0.0, H.ctypes.data_as(L._dp), b.ctypes.data_as(L._dp))
    normal()
    L.call('vdvc_timing_begin', c_)
    for _ in range(reps):
        normal()
    tot = (L._d * 1)()
    nl = (L._i * 1)()
    L.call('vdvc_timing_end', c_, tot, nl)
    ms = tot[0] / max(1, nl[0])
    B = 21 * N
    out['normal_equations_kernel'] = {"kernel_ms": round(ms, 4), "algorithmic_bytes": B,
                                      "achieved_GBs": round(B / ms / 1e6, 1),
                                      "frac": round(B / ms / 1e6 / HBM_PEAK_GBS, 4),
                                      "Gvox_s": round(N / ms / 1e6, 2)}
    out['gaussian_filter_sigma4_ms'] = round(_wall_ms(lambda: C._gauss_dev(f, (n, n, n), 4.0),
                                                      reps), 3)
    out['warp_onto_ref_grid_ms'] = round(_wall_ms(
        lambda: P.warp_onto_ref_grid(np.eye(3) + M, tr, c, (n, n, n), gv, torch), reps), 3)
    gw = P.warp_onto_ref_grid(np.eye(3) + M, tr, c, (n, n, n), gv, torch)
    import contextlib
    import io
    with contextlib.redirect_stdout(io.StringIO()):
        out['compute_crop_ms'] = round(_wall_ms(lambda: P.compute_crop(fv, gw, 8), reps), 3)
    return {"workload": "%d^3 speckle pulled by a known affine (strain/shear <= 3 %%, t ~1 "
                        "vox) + N(0,0.5), float32, sigmas (4,2,1,0)" % n, "stages": out}


def affine_cpu_baseline(n=97):
    """The oracle's NumPy normal equations of one GN step on one host thread
    at n^3: voxels per second."""
    from threadpoolctl import threadpool_limits
    from oracle import affine_oracle as AO
    from voxeldvc_amd import synth
    f = synth.speckle_numpy((n, n, n), synth.blobs_for((n, n, n)), 7).astype(np.float32)
    c = np.full(3, (n - 1) / 2.0)
    g = f.copy()
    gg = AO.gradient(f)
    with threadpool_limits(limits=1):
        AO.normal_equations(f, g, gg, g != 0, np.zeros((3, 3)), np.zeros(3), c)
        t0 = time.perf_counter()
        AO.normal_equations(f, g, gg, g != 0, np.zeros((3, 3)), np.zeros(3), c)
        t = time.perf_counter() - t0
    return {"value": round(n ** 3 / t / 1e6, 3), "unit": "Mvox/s (one normal-equation step)",
            "cores": 1, "kind": "port",
            "sample": "oracle normal equations at %d^3: %.2f s; %d host cores available"
                      % (n, t, os.cpu_count())}


def pcg_levels(iters=60, reps=5):
    """Per-iteration time of the device PCG (vdvc_pcg with tol 0 and a fixed
    iteration count, so every iteration runs) on the three pyramid levels of
    the 513^3, h = 8 solve, all with 65^3 nodes: 129^3 voxels at h = 2 and
    257^3 at h = 4 (float32 gradients + uint8 mask, the plain path's coarse
    levels), 513^3 at h = 8 (bfloat16 folded: the learned path; float32 +
    mask: the plain path's finest level).  Gradients and mask0 of 8-bit
    speckle, Laplacian regularizer at ll = 0.5, Jacobi preconditioner.  HIP
    events around whole vdvc_pcg calls, median of `reps`."""
    import torch
    from voxeldvc_amd import _lib as L
    from voxeldvc_amd import synth
    from voxeldvc_amd.engine.geometry_dvc import build_K_ref_laplacian
    K = np.ascontiguousarray(build_K_ref_laplacian(), dtype=np.float64)
    c = L.ctx()
    out = []
    for n_el, h, store in ((64, 2, 'fp32'), (64, 4, 'fp32'), (64, 8, 'bf16'), (64, 8, 'fp32')):
        Nv = n_el * h + 1
        f = synth.speckle_torch((Nv, Nv, Nv), synth.blobs_for((Nv, Nv, Nv)), seed=5)
        f = f.round().clamp_(0, 255)
        f16 = L.pitch_dense(f, (Nv, Nv, Nv), 'float16')
        del f
        Np = f16.numel()
        gt = L.BF16 if store == 'bf16' else L.F32
        G = [L.empty(Np, 'bfloat16' if store == 'bf16' else 'float32') for _ in range(3)]
        m = L.empty(Np, 'uint8')
        L.call('vdvc_image_setup', c, n_el, n_el, n_el, h, L.ptr(f16), L.F16, 0.0,
               *[L.ptr(q) for q in G], gt, L.ptr(m))
        if store == 'bf16':
            Gf = [L.empty(Np, 'bfloat16') for _ in range(3)]
            L.call('vdvc_fold_grad', c, Np, *[L.ptr(q) for q in G], L.ptr(m),
                   *[L.ptr(q) for q in Gf])
            G, mk = Gf, None
        else:
            mk = m
        del f16
        ndof = 3 * (n_el + 1) ** 3
        diag, minv, x = L.empty(ndof), L.empty(ndof), L.empty(ndof)
        geo = (c, n_el, n_el, n_el, h, *[L.ptr(q) for q in G], gt, L.ptr(mk), L.REG_K,
               K.ctypes.data_as(L._dp), None, float(h), 0.0, 0.5)
        L.call('vdvc_jacobi', *geo, None, L.ptr(diag), L.ptr(minv))
        b = torch.randn(ndof, device='cuda', dtype=torch.float32)
        it_, rr_ = (L._i * 1)(), (L._d * 1)()

        def once():
            L.call('vdvc_pcg', *geo, L.ptr(b), L.ptr(minv), L.ptr(x), 0.0, iters, iters, it_, rr_)

        once()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            once()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1) * 1e3 / iters)
        out.append({"voxels": "%d^3" % Nv, "h": h, "storage": store, "iters": int(it_[0]),
                    "us_per_iteration": round(float(np.median(ts)), 2)})
        print('pcg %d^3 h=%d %s: %.2f us/iteration' % (Nv, h, store, out[-1]["us_per_iteration"]),
              file=sys.stderr, flush=True)
        del G, m, diag, minv, x, b
    return {"pcg_levels": out, "nodes": "65^3", "regularizer": "laplacian, ll=0.5",
            "timing": "HIP events around vdvc_pcg (tol 0, %d iterations), median of %d"
                      % (iters, reps)}


def stream_ceiling(reps=10):
    """The measured HBM streaming ceiling on this GPU: the library's read
    probe (vdvc_stream_read: 16-byte nontemporal loads, four in flight per
    lane) over 2 GiB at four grid sizes, and its byte copy (vdvc_copy) of
    2 GiB; kernel time from the library's HIP events, best of the grids."""
    from voxeldvc_amd import _lib as L
    c = L.ctx()
    nb = 1 << 31
    src = L.empty(nb // 4, 'float32')
    L.fill(src, 1.5)
    dst = L.empty(nb // 4, 'float32')
    out = L.empty(8192, 'int32')
    key = 0x12345677      # never the xor fold of equal 0x3fc00000 words

    def timed(call):
        call()
        L.call('vdvc_timing_begin', c)
        for _ in range(reps):
            call()
        tot, nl = (L._d * 1)(), (L._i * 1)()
        L.call('vdvc_timing_end', c, tot, nl)
        return tot[0] / max(1, nl[0])

    read = {}
    for grid in (1024, 2048, 4096, 8192):
        ms = timed(lambda: L.call('vdvc_stream_read', c, nb, src, grid, key, out))
        read[grid] = round(nb / (ms * 1e-3) / 1e9, 1)
    ms = timed(lambda: L.call('vdvc_copy', c, nb, src, dst))
    return {"read_GBs": max(read.values()), "read_GBs_by_grid": read,
            "copy_GBs": round(2 * nb / (ms * 1e-3) / 1e9, 1),
            "note": "2 GiB read (nontemporal 16-byte loads) / 2 GiB copied (read + write "
                    "counted), library kernels, HIP-event kernel time"}


def measured_traffic(Nv, h):
    """HBM bytes per matvec launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE measurement of the same workload (None if not measured)."""
    try:
        with open(TRAFFIC_FILE) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    for row in t.get('rows', []):
        if row.get('Nv') == Nv and row.get('h') == h and row.get('storage') == 'bf16-folded':
            return row.get('hbm_bytes')
    return None


def cpu_baseline():
    """The oracle's NumPy matvec (chunked gather, einsum, np.add.at -- the
    reference algorithm) on one host thread, on a bounded 257^3 sample."""
    from threadpoolctl import threadpool_limits
    from oracle import dvc_oracle as O
    n_el, h = 32, 8
    nv = n_el * h + 1
    rng = np.random.default_rng(0)
    g = rng.normal(size=(nv ** 3, 3)).astype(np.float32)
    m = rng.random(nv ** 3) < 0.95
    v = rng.normal(size=3 * (n_el + 1) ** 3).astype(np.float32)
    with threadpool_limits(limits=1):
        O.matvec_H(v, g, m, h, n_el, n_el, n_el)  # warm-up
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            O.matvec_H(v, g, m, h, n_el, n_el, n_el)
            ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    return {"value": round(nv ** 3 / t / 1e9, 6), "unit": "Gvox/s", "cores": 1, "kind": "port",
            "sample": "one float32 H matvec at 257^3, h=8 (32^3 elements), median of 3 after 1 "
                      "warm-up, %.2f s each; %d host cores available" % (t, os.cpu_count())}


def cpu_gn_solve():
    """The plain method's whole solve at its CPU-runnable configuration (65^3
    speckle, translation + 1 % z-strain, h = 8, Laplacian l0 = 36, 3 levels,
    FFT pre-alignment, dynamic mask frozen after 4 iterations, float32): the
    oracle's NumPy GN + PCG (the reference algorithm) on one host thread,
    beside the device solve of the same inputs."""
    import torch
    from threadpoolctl import threadpool_limits
    from oracle import dvc_oracle as O
    from voxeldvc_amd import synth
    from voxeldvc_amd.engine import correlate_gpu as C
    f, g = synth.c1_pair()
    f, g = f.astype(np.float32), g.astype(np.float32)
    kw = dict(K_ref_laplacian=O.build_K_ref_laplacian(), l0=36.0, glt=0.0, dynamic_mask=True,
              freeze_mask_after=4)
    with threadpool_limits(limits=1):
        t0 = time.perf_counter()
        U_cpu, _, _ = O.multiscale_correlate(f, g, 8, 8, 8, 8, scales=(2, 1, 0), **kw)
        t_cpu = time.perf_counter() - t0
    fd = torch.from_numpy(f).cuda()
    gd = torch.from_numpy(g).cuda()
    C.multiscale_correlate_gpu(fd, gd, 8, 8, 8, 8, torch, disp=False, **kw)
    ts = []
    for _ in range(5):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        lg = []
        U, _, _ = C.multiscale_correlate_gpu(fd, gd, 8, 8, 8, 8, torch, disp=False, log=lg, **kw)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    gn = sum(len(lv['iters']) for lv in lg)
    pcg = sum(q['pcg_iters'] for lv in lg for q in lv['iters'])
    t_gpu = float(np.median(ts))
    return {"value": round(t_cpu * 1e3, 1), "unit": "ms per GN solve", "cores": 1,
            "kind": "port", "higher_is_better": False,
            "sample": "the whole plain-method solve at 65^3 (3 levels, FFT pre-alignment), once; "
                      "%d host cores available, 1 used" % os.cpu_count(),
            "gpu_ms": round(t_gpu * 1e3, 2), "gn_iters": gn, "pcg_iters": pcg,
            "max_abs_vs_cpu_vox": float(np.abs(U.double().cpu().numpy() - U_cpu).max())}


def main():
    a = parse()
    if a.summarize_trace:
        summarize_trace(a.summarize_trace)
        return
    if a.gpus != 1:
        print('bench.py measures one GPU: this build has no multi-GPU path (--gpus %d '
              'requested)' % a.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    from voxeldvc_amd import _lib as L
    from voxeldvc_amd import synth
    assert torch.cuda.is_available(), 'bench needs a GPU'
    if a.solve_only:
        solve_only(a.solve_only, n=a.size, h=a.h)
        return
    if a.pcg_levels:
        print(json.dumps(pcg_levels()), flush=True)
        return
    n_el = (a.size - 1) // a.h
    h = a.h
    Nv = n_el * h + 1
    Nvox = Nv ** 3
    Ndof = 3 * (n_el + 1) ** 3
    print('generating %d^3 speckle' % Nv, file=sys.stderr, flush=True)
    f = synth.speckle_torch((Nv, Nv, Nv), synth.blobs_for((Nv, Nv, Nv)), seed=3)
    f = f.round().clamp_(0, 255)
    c = L.ctx()
    f16 = L.pitch_dense(f, (Nv, Nv, Nv), 'float16')   # pitched level layout
    del f
    Np = f16.numel()
    Gb = [L.empty(Np, 'bfloat16') for _ in range(3)]
    mask0 = L.empty(Np, 'uint8')
    L.call('vdvc_image_setup', c, n_el, n_el, n_el, h, L.ptr(f16), L.F16, 0.0, *[L.ptr(q) for q in Gb],
           L.BF16, L.ptr(mask0))
    Gw = [L.empty(Np, 'bfloat16') for _ in range(3)]
    L.call('vdvc_fold_grad', c, Np, *[L.ptr(q) for q in Gb], L.ptr(mask0), *[L.ptr(q) for q in Gw])
    del Gb
    v = torch.randn(Ndof, device='cuda', dtype=torch.float32)
    out = L.empty(Ndof)

    def run(G, gt, mask, steps, warm):
        args = (c, n_el, n_el, n_el, h, L.ptr(G[0]), L.ptr(G[1]), L.ptr(G[2]), gt, L.ptr(mask),
                L.ptr(v), L.ptr(out))
        for _ in range(warm):
            L.call('vdvc_matvec_H', *args)
        # the timed steps: the operator calls alone, nothing else on the stream
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            L.call('vdvc_matvec_H', *args)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        # the voxel kernel's own duration: the same calls again with HIP
        # events around each launch of it
        L.call('vdvc_timing_begin', c)
        for _ in range(steps):
            L.call('vdvc_matvec_H', *args)
        tot = (L._d * 1)()
        nl = (L._i * 1)()
        L.call('vdvc_timing_end', c, tot, nl)
        return wall / steps * 1e3, tot[0] / max(1, nl[0])

    if a.stream_probe:
        n = 1 << 28
        xs, ys = L.empty(n), L.zeros(n)
        L.call('vdvc_axpby', c, n, L.F32, 0.5, L.ptr(xs), 1.0, L.ptr(ys))
        torch.cuda.synchronize()
        del xs, ys

    print('streaming ceiling', file=sys.stderr, flush=True)
    ceiling = stream_ceiling()
    B_bf16 = Nvox * 6 + 8 * Ndof
    B_f32 = Nvox * 13 + 8 * Ndof
    res = {
        "metric": "Hessian matvec Gvox/s (% HBM peak)",
        "unit": "Gvox/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "H matvec, %d^3 voxels, h=%d (%d^3 Hex8 elements), bf16 gradients "
                               "with mask folded (6 B/voxel), fp16 speckle 0..255" % (Nv, h, n_el),
                   "Nvox": Nvox, "Ndof": Ndof, "storage": "bf16-folded"},
    }
    if a.only in (None, 'bf16'):
        print('timing bf16-folded matvec', file=sys.stderr, flush=True)
        ms_b, kms_b = run(Gw, L.BF16, None, a.steps, a.warmup)
        ach = B_bf16 / (kms_b * 1e-3) / 1e9
        res["value"] = round(Nvox / (ms_b * 1e-3) / 1e9, 3)
        res["ms_per_step"] = round(ms_b, 4)
        res["roofline"] = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                           "traffic": measured_traffic(Nv, h),
                           "kernel": "vox_vec_kernel<MATVEC,bf16,V=8>",
                           "kernel_ms": round(kms_b, 5), "algorithmic_bytes": B_bf16,
                           "frac_of_measured_read_ceiling": round(ach / ceiling["read_GBs"], 4)}
    del Gw
    if a.only in (None, 'f32'):
        G32 = [L.empty(Np) for _ in range(3)]
        L.call('vdvc_image_setup', c, n_el, n_el, n_el, h, L.ptr(f16), L.F16, 0.0,
               *[L.ptr(q) for q in G32], L.F32, L.ptr(mask0))
        print('timing fp32+mask matvec', file=sys.stderr, flush=True)
        ms_f, kms_f = run(G32, L.F32, mask0, a.steps, a.warmup)
        res["fp32_mask_form"] = {"value": round(Nvox / (ms_f * 1e-3) / 1e9, 3),
                                 "ms_per_step": round(ms_f, 4), "kernel_ms": round(kms_f, 5),
                                 "kernel": "vox_vec_kernel<MATVEC,f32,mask,V=4>",
                                 "algorithmic_bytes": B_f32,
                                 "achieved_GBs": round(B_f32 / (kms_f * 1e-3) / 1e9, 1),
                                 "frac": round(B_f32 / (kms_f * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    res["stream_ceiling"] = ceiling
    del f16, mask0
    if a.only in (None, 'f32'):
        del G32
    if not a.no_solve:
        print('time to solve', file=sys.stderr, flush=True)
        res["time_to_solve"] = time_to_solve(n=a.size, h=a.h, reps=a.solve_reps)
    if not a.no_post:
        print('post-solve stages', file=sys.stderr, flush=True)
        res["post_stages"] = post_stages(n=a.size, h=a.h)
        if not a.no_cpu:
            res["post_stages"]["cpu_baseline"] = post_cpu_baseline()
    if not a.no_affine:
        print('affine stage', file=sys.stderr, flush=True)
        res["affine_stage"] = affine_stage(n=a.size)
        if not a.no_cpu:
            res["affine_stage"]["cpu_baseline"] = affine_cpu_baseline()
    if a.configs:
        res["configs"] = config_cases()
    if not a.no_cpu:
        print('cpu baseline', file=sys.stderr, flush=True)
        res["cpu_baseline"] = cpu_baseline()
        print('cpu gn solve', file=sys.stderr, flush=True)
        res["cpu_baseline"]["gn_solve_c1"] = cpu_gn_solve()
    print(json.dumps(res), flush=True)


if __name__ == '__main__':
    main()
