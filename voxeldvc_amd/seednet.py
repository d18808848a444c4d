"""Learned seed for the finest-scale GN solve (an extension; the reference has
no counterpart).

A 5-layer 3-D convolutional network reads the reference and deformed images
decimated by h/2 (two standardised channels) and predicts the nodal
displacement of the h-mesh directly (its second layer has stride 2, so an
input of 2*N + 1 voxels per axis gives the N + 1 nodes).  It replaces the
coarse pyramid levels of multiscale_correlate_gpu: the finest-level GN solve
starts from its prediction.  Every forward and backward product runs in
libvdvc.so on the matrix cores (bfloat16 operands, float32 accumulation);
the parameters are float32 masters updated by Adam.

Training data: synthetic speckle pairs warped by random Fourier-mode fields
plus a random affine, labelled by the plain multiscale GN solve itself
(`make_training_set`).  `python -m voxeldvc_amd.seednet train` trains on the
GPU and writes seednet_weights.npz next to this file (its defaults, 4096
pairs and 8000 Adam steps, made the shipped weights);
`python -m voxeldvc_amd.seednet curve` trains one network per training-set
size and reports the validation RMS of each.
"""
import os
import time

import numpy as np

from . import _lib as L

# (input channels, output channels, stride) per layer; ReLU after all but
# the last
LAYERS = ((2, 16, 1), (16, 32, 2), (32, 32, 1), (32, 32, 1), (32, 3, 1))
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'seednet_weights.npz')


def out_extent(n, stride):
    return (n - 1) // stride + 1


class SeedNet:
    """Parameters, activations and optimizer state of the seed network."""

    def __init__(self, seed=0):
        t = L.require_gpu()
        rng = np.random.default_rng(seed)
        self.w, self.b = [], []
        for li, (c, o, _) in enumerate(LAYERS):
            std = np.sqrt(2.0 / (27 * c)) if li < len(LAYERS) - 1 else np.sqrt(1.0 / (27 * c))
            self.w.append(t.from_numpy(rng.normal(scale=std, size=(o, 27, c)).astype(np.float32)
                                       ).cuda())
            self.b.append(t.zeros(o, dtype=t.float32, device='cuda'))
        self.h = 8               # element size the network is trained for
        self._opt = None
        self.step_count = 0

    # ---- parameters -------------------------------------------------------
    def params(self):
        return self.w + self.b

    def state(self):
        d = {'w%d' % i: L.to_host(w) for i, w in enumerate(self.w)}
        d.update({'b%d' % i: L.to_host(b) for i, b in enumerate(self.b)})
        d['layers'] = np.array(LAYERS, dtype=np.int64)
        return d

    def save(self, path=WEIGHTS, **meta):
        np.savez(path, **self.state(), **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path=WEIGHTS):
        z = np.load(path)
        if not np.array_equal(z['layers'], np.array(LAYERS)):
            raise ValueError('%s holds a different layer layout' % path)
        net = cls()
        t = L.require_gpu()
        for i in range(len(LAYERS)):
            net.w[i] = t.from_numpy(np.ascontiguousarray(z['w%d' % i], dtype=np.float32)).cuda()
            net.b[i] = t.from_numpy(np.ascontiguousarray(z['b%d' % i], dtype=np.float32)).cuda()
        return net

    # ---- forward / backward -------------------------------------------------
    def forward(self, x, dims, keep=False):
        """x: device float32 [N][D][H][W][2]; dims = (N, D, H, W).  Returns the
        output [N][Do][Ho][Wo][3] (and the activation list if keep)."""
        t = L.require_gpu()
        c = L.ctx()
        N, D, H, W = dims
        acts = [x]
        shapes = [(N, D, H, W)]
        cur = x
        for li, (ci, co, s) in enumerate(LAYERS):
            n_, d_, h_, w_ = shapes[-1]
            do, ho, wo = out_extent(d_, s), out_extent(h_, s), out_extent(w_, s)
            y = t.empty(n_ * do * ho * wo * co, dtype=t.float32, device='cuda')
            L.call('vdvc_conv3d_forward', c, n_, d_, h_, w_, ci, co, s, L.ptr(cur),
                   L.ptr(self.w[li]), L.ptr(self.b[li]), int(li < len(LAYERS) - 1), L.ptr(y))
            cur = y
            acts.append(y)
            shapes.append((n_, do, ho, wo))
        if keep:
            return cur, acts, shapes
        return cur

    def backward(self, acts, shapes, dout):
        """Gradients of every parameter for upstream gradient dout."""
        t = L.require_gpu()
        c = L.ctx()
        gw = [None] * len(LAYERS)
        gb = [None] * len(LAYERS)
        g = dout
        for li in range(len(LAYERS) - 1, -1, -1):
            ci, co, s = LAYERS[li]
            n_, d_, h_, w_ = shapes[li]
            gw[li] = t.empty_like(self.w[li])
            gb[li] = t.empty_like(self.b[li])
            L.call('vdvc_conv3d_backward_weight', c, n_, d_, h_, w_, ci, co, s, L.ptr(acts[li]),
                   L.ptr(g), L.ptr(gw[li]), L.ptr(gb[li]))
            if li > 0:
                dx = t.empty(n_ * d_ * h_ * w_ * ci, dtype=t.float32, device='cuda')
                # acts[li] is the ReLU output that fed this layer
                L.call('vdvc_conv3d_backward_data', c, n_, d_, h_, w_, ci, co, s, L.ptr(g),
                       L.ptr(self.w[li]), L.ptr(acts[li]), L.ptr(dx))
                g = dx
        return gw + gb

    def adam(self, grads, lr, b1=0.9, b2=0.999, eps=1e-8):
        t = L.require_gpu()
        if self._opt is None:
            self._opt = [(t.zeros_like(p), t.zeros_like(p)) for p in self.params()]
        self.step_count += 1
        c = L.ctx()
        for p, g, (m, v) in zip(self.params(), grads, self._opt):
            L.call('vdvc_adam_step', c, p.numel(), L.ptr(p), L.ptr(g), L.ptr(m), L.ptr(v),
                   float(lr), float(b1), float(b2), float(eps), self.step_count)

    def train_step(self, x, dims, label, lr):
        """One forward/backward/Adam step on a batch; returns the MSE loss."""
        t = L.require_gpu()
        out, acts, shapes = self.forward(x, dims, keep=True)
        grad = t.empty_like(out)
        loss = (L._d * 1)()
        L.call('vdvc_mse', L.ctx(), out.numel(), L.ptr(out), L.ptr(label), L.ptr(grad), loss)
        self.adam(self.backward(acts, shapes, grad), lr)
        return float(loss[0])

    # ---- inference ------------------------------------------------------------
    def predict(self, f, g, shape, Nx, Ny, Nz, h, itype=None):
        """Nodal displacement (device float32, reference DOF layout) of the
        h-mesh for device volumes f, g (dense x-fastest of dims `shape`).  The
        network predicts the nodes of a mesh with its own element size
        self.h covering the same voxels; the h-mesh's nodes are picked from
        it (h a multiple of self.h) or interpolated trilinearly from it (h a
        divisor of self.h, the h-mesh then covering whole self.h elements)."""
        if h % self.h and self.h % h:
            raise ValueError('the seed network predicts nodes of an h=%d mesh; h=%d is neither '
                             'a multiple nor a divisor of it' % (self.h, h))
        ext = (Nx * h, Ny * h, Nz * h)
        if any(e % self.h for e in ext):
            raise ValueError('the mesh (%d x %d x %d elements of h=%d) does not cover whole '
                             'h=%d elements' % (Nx, Ny, Nz, h, self.h))
        mx, my, mz = (e // self.h for e in ext)
        x, dims = network_input(f, g, shape, self.h, itype)
        out = self.forward(x, dims)
        expect = 3 * (mx + 1) * (my + 1) * (mz + 1)
        if out.numel() != expect:
            raise ValueError('network output has %d values, the mesh %d' % (out.numel(), expect))
        if h == self.h:
            return out
        dst = L.empty(3 * (Nx + 1) * (Ny + 1) * (Nz + 1))
        L.call('vdvc_resample_nodes', L.ctx(), mx, my, mz, self.h, out, Nx, Ny, Nz, h, dst,
               L.F32)
        return dst


def network_input(f, g, shape, h, itype=None):
    """[1][D][H][W][2] standardised input from the stride-h/2 decimation of
    device volumes f, g (x fastest, dims `shape` = (nx, ny, nz))."""
    if h % 2:
        raise ValueError('the seed network needs an even element size (h=%d)' % h)
    t = L.require_gpu()
    s = h // 2
    nx, ny, nz = shape
    if itype is None:
        itype = L.F16 if f.dtype == t.float16 else L.F32
    m = [out_extent(n, s) for n in shape]
    x = t.empty(m[0] * m[1] * m[2] * 2, dtype=t.float32, device='cuda')
    L.call('vdvc_seed_input', L.ctx(), L.ptr(f), L.ptr(g), itype, nx, ny, nz, s, L.ptr(x))
    return x, (1, m[2], m[1], m[0])


# ---------------------------------------------------------------------------
# training data and training loop
# ---------------------------------------------------------------------------

def make_pair(n, seed, peak=6.0, rot=3.0, strain=0.02):
    """One synthetic pair of n^3 8-bit speckle volumes: g = f warped by a
    random Fourier field (16 modes, wavelength 32-128, peak `peak` vox) plus
    a random affine about the centre, N(0,1) noise, quantised."""
    from . import synth
    rng = np.random.default_rng(seed)
    f = synth.speckle_torch((n, n, n), synth.blobs_for((n, n, n)), seed=int(rng.integers(1 << 30)))
    f = f.round().clamp_(0, 255)
    modes = synth.fourier_params(16, (32, 128), peak * rng.uniform(0.3, 1.0),
                                 seed=int(rng.integers(1 << 30)), extent=(n, n, n))
    A = synth.random_affine(rng, rot, strain)
    c = ((n - 1) / 2.0,) * 3
    g = synth.warp_backward_torch(f, (n, n, n), modes, 1.0, seed=int(rng.integers(1 << 30)),
                                  quantize=True, affine=(A, c))
    return f, g


def label_pair(f, g, n, h, scales=(2, 1, 0)):
    """The plain multiscale GN solve of a pair (the training label)."""
    import torch
    from .engine import correlate_gpu as C
    from .engine.geometry_dvc import build_K_ref_laplacian
    Ne = (n - 1) // h
    fv = f.view(n, n, n).permute(2, 1, 0)
    gv = g.view(n, n, n).permute(2, 1, 0)
    U, _, _ = C.multiscale_correlate_gpu(fv, gv, Ne, Ne, Ne, h, torch, scales=scales,
                                         K_ref_laplacian=build_K_ref_laplacian(), l0=4.5 * h,
                                         dynamic_mask=True, freeze_mask_after=4, glt=0.0,
                                         disp=False)
    return U


def make_training_set(npairs, n=129, h=8, seed=0, log=print):
    """(inputs [npairs][D][H][W][2], labels [npairs][Nn][Nn][Nn][3]) on the
    device."""
    t = L.require_gpu()
    xs, ys = [], []
    t0 = time.time()
    for i in range(npairs):
        f, g = make_pair(n, seed * 1000003 + i)
        U = label_pair(f, g, n, h)
        x, _ = network_input(f, g, (n, n, n), h)
        xs.append(x)
        ys.append(U.float())
        if log and (i + 1) % 64 == 0:
            log('pairs %d/%d  %.1fs' % (i + 1, npairs, time.time() - t0))
    return t.stack(xs), t.stack(ys)


def train(net, X, Y, dims1, steps, batch=16, lr=2e-3, seed=0, log=print, every=200):
    """Adam on random mini-batches of (X, Y); returns the loss history."""
    t = L.require_gpu()
    gen = np.random.default_rng(seed)
    npairs = X.shape[0]
    hist = []
    t0 = time.time()
    for it in range(steps):
        idx = t.from_numpy(gen.choice(npairs, size=batch, replace=False)).cuda()
        lr_it = lr * 0.5 * (1 + np.cos(np.pi * it / steps))   # cosine decay
        loss = net.train_step(X[idx].contiguous().view(-1), (batch,) + tuple(dims1[1:]),
                              Y[idx].contiguous().view(-1), lr_it)
        hist.append(loss)
        if log and (it % every == 0 or it == steps - 1):
            log('step %5d  loss %.4f  lr %.2e  %.1fs' % (it, loss, lr_it, time.time() - t0))
    return hist


def evaluate(net, X, Y, dims1):
    """RMS (voxels) of the prediction against the labels, per pair."""
    out = []
    for i in range(X.shape[0]):
        p = net.forward(X[i].contiguous().view(-1), dims1)
        d = (p - Y[i].view(-1)).double()
        out.append(float((d * d).mean().sqrt()))
    return np.array(out)


def learning_curve(sizes, val=64, steps=5000, n=129, batch=16, lr=2e-3, out_dir=None,
                   log=print):
    """Validation RMS against the number of training pairs: one labelled
    set of max(sizes) pairs (the smaller sets are its prefixes) and one
    validation set of other seeds; a fresh network trained for `steps` Adam
    steps on each size.  Returns a list of dicts; with out_dir, writes
    learning_curve.json and the weights of each size there."""
    import json
    h = 8
    t0 = time.time()
    X, Y = make_training_set(max(sizes), n=n, h=h, seed=1, log=log)
    Xv, Yv = make_training_set(val, n=n, h=h, seed=2, log=log)
    log('data %.1fs' % (time.time() - t0))
    m = out_extent(n, h // 2)
    dims1 = (1, m, m, m)
    zero = float(np.sqrt(float((Yv.double() ** 2).mean())))
    rows = []
    for k in sizes:
        net = SeedNet(seed=0)
        t1 = time.time()
        hist = train(net, X[:k], Y[:k], dims1, steps, batch=batch, lr=lr, log=log,
                     every=max(steps // 4, 1))
        rv = evaluate(net, Xv, Yv, dims1)
        rt = evaluate(net, X[:min(k, val)], Y[:min(k, val)], dims1)
        row = dict(pairs=k, steps=steps, val_rms=float(rv.mean()), train_rms=float(rt.mean()),
                   zero_rms=zero, final_loss=float(np.mean(hist[-100:])),
                   train_s=round(time.time() - t1, 1))
        rows.append(row)
        log('pairs %6d: validation RMS %.4f vox, training-set RMS %.4f vox (zero %.3f)'
            % (k, row['val_rms'], row['train_rms'], zero))
        if out_dir:
            net.save(os.path.join(out_dir, 'seednet_weights_%d.npz' % k), val_rms=row['val_rms'],
                     zero_rms=zero, pairs=k, steps=steps, loss=np.array(hist, dtype=np.float32))
    if out_dir:
        with open(os.path.join(out_dir, 'learning_curve.json'), 'w') as fh:
            json.dump(dict(n=n, h=h, val_pairs=val, batch=batch, lr=lr, rows=rows), fh, indent=1)
    return rows


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(prog='python -m voxeldvc_amd.seednet')
    ap.add_argument('cmd', choices=['train', 'curve'])
    ap.add_argument('--sizes', default='256,1024,4096,8192',
                    help='curve: training-set sizes (prefixes of the largest)')
    ap.add_argument('--out-dir', default=None, help='curve: where the summary and weights go')
    ap.add_argument('--pairs', type=int, default=4096)
    ap.add_argument('--val', type=int, default=64)
    ap.add_argument('--steps', type=int, default=8000)
    ap.add_argument('--batch', type=int, default=16)
    ap.add_argument('--lr', type=float, default=2e-3)
    ap.add_argument('--n', type=int, default=129)
    ap.add_argument('--out', default=WEIGHTS)
    a = ap.parse_args(argv)
    if a.cmd == 'curve':
        sizes = [int(x) for x in a.sizes.split(',')]
        learning_curve(sizes, val=a.val, steps=a.steps, n=a.n, batch=a.batch, lr=a.lr,
                       out_dir=a.out_dir, log=lambda m: print(m, flush=True))
        return
    h = 8
    t0 = time.time()
    X, Y = make_training_set(a.pairs, n=a.n, h=h, seed=1)
    Xv, Yv = make_training_set(a.val, n=a.n, h=h, seed=2)
    m = out_extent(a.n, h // 2)
    dims1 = (1, m, m, m)
    print('data %.1fs' % (time.time() - t0), flush=True)
    net = SeedNet(seed=0)
    zero = np.sqrt(float((Yv.double() ** 2).mean()))
    hist = train(net, X, Y, dims1, a.steps, batch=a.batch, lr=a.lr)
    rms_v = evaluate(net, Xv, Yv, dims1)
    print('validation RMS %.3f vox (zero predictor %.3f vox), train time %.1fs'
          % (rms_v.mean(), zero, time.time() - t0), flush=True)
    net.save(a.out, val_rms=rms_v.mean(), zero_rms=zero, pairs=a.pairs, steps=a.steps,
             loss=np.array(hist, dtype=np.float32))
    print('wrote', a.out)


if __name__ == '__main__':
    main()
