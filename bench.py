"""Benchmark of the global GN DVC hot path on one MI355X.

Step = one correlation-Hessian matvec H v over a 513^3 voxel reference volume
with a 64^3-element Hex8 mesh (h = 8): the configuration the tier metric
("Hessian matvec Gvox/s (% HBM peak)") is quoted on.  Synthetic speckle
(Gaussian blobs, quantised to 0..255, stored float16), gradients in bfloat16
with the mask folded in (6 B/voxel); the float32 + uint8-mask form
(13 B/voxel) is timed beside it.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 513] [--h 8]
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
    return p.parse_args()


TRAFFIC_FILE = os.path.join(ROOT, 'profiles', 'hessian_matvec_traffic.json')


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


def main():
    a = parse()
    if a.gpus != 1:
        print('bench.py measures one GPU: this build has no multi-GPU path (--gpus %d '
              'requested)' % a.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    from voxeldvc_amd import _lib as L
    from voxeldvc_amd import synth
    assert torch.cuda.is_available(), 'bench needs a GPU'
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
        torch.cuda.synchronize()
        L.call('vdvc_timing_begin', c)
        t0 = time.perf_counter()
        for _ in range(steps):
            L.call('vdvc_matvec_H', *args)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
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
                           "kernel_ms": round(kms_b, 5), "algorithmic_bytes": B_bf16}
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
    if not a.no_cpu:
        print('cpu baseline', file=sys.stderr, flush=True)
        res["cpu_baseline"] = cpu_baseline()
    print(json.dumps(res), flush=True)


if __name__ == '__main__':
    main()
