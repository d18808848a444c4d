"""CPU oracle for the preprocess affine stage (NumPy).

TEST INFRASTRUCTURE ONLY.  The HIP kernels of the affine refinement, the
warp onto the reference grid and the overlap crop are compared against this
module; nothing in the product package imports it.

Restated from ggalu/voxelDVC (paths relative to its src/voxeldvc/):
  * gaussian_filter / smooth_plain / smooth_masked
                              engine/correlate_gpu.py:570-598 (scipy.ndimage
                              gaussian_filter, truncate 4, per-axis float32
                              passes)
  * affine_refine_gn          engine/correlate_gpu.py:600-825
  * warp_onto_ref_grid        preprocess.py:233-285
  * compute_crop              preprocess.py:315-367

Parity pins (tests/test_affine_oracle.py): gaussian_filter bit-exact against
scipy.ndimage.gaussian_filter (modes 'nearest' / 'constant', several
sigmas, lines shorter than the kernel); gradient against np.gradient;
affine_refine_gn against the reference's own result on a synthetic pair
(golden affine_refine49, tests/golden/make_golden.py) and against a known
affine (closed form); warp_onto_ref_grid against scipy.ndimage
.map_coordinates of the affine coordinates; compute_crop against brute force.
"""
import numpy as np

from .post_oracle import map_coordinates_01


def gaussian_kernel1d(sigma, truncate=4.0):
    """Normalised Gaussian weights and radius int(truncate * sigma + 0.5)."""
    r = int(truncate * float(sigma) + 0.5)
    k = np.arange(-r, r + 1)
    phi = np.exp(-0.5 / (sigma * sigma) * k ** 2)
    return phi / phi.sum(), r


def correlate1d_sym(a, w, r, axis, mode, cval=0.0):
    """scipy.ndimage.correlate1d with a symmetric kernel, in its order: the
    centre tap, then (x[i-k] + x[i+k]) * w_k for k = r .. 1, in float64; the
    line extended by edge replication ('nearest') or cval ('constant');
    result in a's dtype."""
    dt = a.dtype
    a = np.moveaxis(a, axis, 0).astype(np.float64)
    n = a.shape[0]
    if mode == 'nearest':
        lo, hi = np.repeat(a[:1], r, 0), np.repeat(a[-1:], r, 0)
    elif mode == 'constant':
        lo = hi = np.full((r,) + a.shape[1:], float(cval))
    else:
        raise ValueError(mode)
    pad = np.concatenate([lo, a, hi])
    out = pad[r:r + n] * w[r]
    for k in range(r, 0, -1):
        out = out + (pad[r - k:r - k + n] + pad[r + k:r + k + n]) * w[r + k]
    return np.moveaxis(out.astype(dt), 0, axis)


def gaussian_filter(a, sigma, mode='nearest', cval=0.0):
    """gaussian_filter(a, sigma, mode, cval) for a 3-D float array: one
    correlate1d pass per axis (0, 1, 2), each rounded to a's dtype."""
    w, r = gaussian_kernel1d(sigma)
    out = a
    for ax in range(3):
        out = correlate1d_sym(out, w, r, ax, mode, cval)
    return out


def smooth_plain(img, sigma):
    """correlate_gpu.py:579-583."""
    return img if sigma <= 0 else gaussian_filter(img, sigma, mode='nearest')


def smooth_masked(img, valid_f, sigma):
    """Normalised-convolution smoothing smooth(img*valid)/smooth(valid) where
    the denominator exceeds 1e-6, img elsewhere (correlate_gpu.py:586-598)."""
    if sigma <= 0:
        return img
    num = gaussian_filter(img * valid_f, sigma, mode='constant', cval=0.0)
    den = gaussian_filter(valid_f, sigma, mode='constant', cval=0.0)
    out = img.copy()
    good = den > 1e-6
    out[good] = num[good] / den[good]
    return out


def gradient(a):
    """np.gradient(a) (unit spacing, edge_order 1) per axis, in a's dtype."""
    out = []
    for ax in range(3):
        b = np.moveaxis(a, ax, 0)
        gd = np.empty_like(b)
        gd[1:-1] = (b[2:] - b[:-2]) / b.dtype.type(2)
        gd[0] = b[1] - b[0]
        gd[-1] = b[-1] - b[-2]
        out.append(np.moveaxis(gd, 0, ax))
    return out


def _coords(M32, t32, c32, shape):
    """Sample locations s(x) = x - M (x - c) - t_r in float32, the
    reference's element-wise order (correlate_gpu.py:704-723)."""
    Nx, Ny, Nz = shape
    X, Y, Z = np.meshgrid(np.arange(Nx, dtype=np.float32), np.arange(Ny, dtype=np.float32),
                          np.arange(Nz, dtype=np.float32), indexing='ij')
    Xc = [X - c32[0], Y - c32[1], Z - c32[2]]
    base = [X, Y, Z]
    coords = [base[i] - (M32[i, 0] * Xc[0] + M32[i, 1] * Xc[1] + M32[i, 2] * Xc[2]) - t32[i]
              for i in range(3)]
    return Xc, coords


def _inb(coords, shape):
    m = np.ones(coords[0].shape, dtype=bool)
    for q, n in zip(coords, shape):
        m &= (q >= 0) & (q <= np.float32(n - 1))
    return m


def normal_equations(f_s, g_s, gg, valid_g, M, tr, c, cval=0.0):
    """(H, b) of one GN step, float64 (correlate_gpu.py:759-793): J_k =
    dW/dp_k over M00..M22, t0..t2 in float32, e = (g_s - f_s(s)) * mask."""
    shape = f_s.shape
    M32 = np.asarray(M, np.float64).astype(np.float32)
    t32 = np.asarray(tr, np.float64).astype(np.float32)
    c32 = np.asarray(c, np.float64).astype(np.float32)
    Xc, coords = _coords(M32, t32, c32, shape)
    warped = map_coordinates_01(f_s, coords, order=1, mode='constant', cval=cval)
    mask = (_inb(coords, shape) & valid_g).astype(np.float32)
    e = (g_s - warped) * mask
    gs = [map_coordinates_01(gg[i], coords, order=1, mode='nearest') for i in range(3)]
    J = []
    for i in range(3):
        for j in range(3):
            J.append(-gs[i] * Xc[j] * mask)
    for i in range(3):
        J.append(-gs[i] * mask)
    Jf = np.stack([Jk.ravel() for Jk in J]).astype(np.float64)
    return Jf @ Jf.T, Jf @ e.ravel().astype(np.float64)


def masked_rms(f, g, valid_g, M, tr, c, cval=0.0):
    """Masked RMS of g - f(s(x)) over in-bounds voxels with valid g
    (correlate_gpu.py:725-743)."""
    shape = f.shape
    M32 = np.asarray(M, np.float64).astype(np.float32)
    t32 = np.asarray(tr, np.float64).astype(np.float32)
    c32 = np.asarray(c, np.float64).astype(np.float32)
    _, coords = _coords(M32, t32, c32, shape)
    warped = map_coordinates_01(f, coords, order=1, mode='constant', cval=cval)
    m = _inb(coords, shape) & valid_g
    e = (g - warped)[m]
    return float(np.sqrt(float((e * e).sum(dtype=np.float64)) / max(float(m.sum()), 1.0)))


def affine_refine_gn(f_pix, g_pix, A0, t0, c, sigmas=(4.0, 2.0, 1.0, 0.0), max_iter=50,
                     tol=1e-6, patience=3, cval=0.0):
    """Forward-additive Gauss-Newton refinement of x' = A (x - c) + c + t over
    a Gaussian scale space, solved in the generative convention
    g(x) ~ f(x - M (x - c) - t_r), A = (I - M)^-1, t = A t_r
    (correlate_gpu.py:600-825).  Returns (A, t, info)."""
    f = np.asarray(f_pix, dtype=np.float32)
    g = np.asarray(g_pix, dtype=np.float32)
    I3 = np.eye(3)
    A0inv = np.linalg.inv(np.asarray(A0, np.float64))
    M = I3 - A0inv
    tr = A0inv @ np.asarray(t0, np.float64)
    c = np.asarray(c, np.float64)
    valid_g = g != cval
    rms_initial = masked_rms(f, g, valid_g, M, tr, c, cval)
    iters = []
    for sigma in sigmas:
        f_s = smooth_plain(f, sigma)
        g_s = smooth_masked(g, valid_g.astype(np.float32), sigma)
        gg = gradient(f_s)
        n_it = 0
        best = np.inf
        stall = 0
        for it in range(max_iter):
            H, b = normal_equations(f_s, g_s, gg, valid_g, M, tr, c, cval)
            H = H + np.eye(12) * (1e-9 * float(np.trace(H)) / 12.0 + 1e-30)
            dp = np.linalg.solve(H, b)
            M = M + dp[:9].reshape(3, 3)
            tr = tr + dp[9:]
            step = float(np.linalg.norm(dp))
            n_it = it + 1
            if step < tol:
                break
            if step < best - 1e-12:
                best = step
                stall = 0
            else:
                stall += 1
                if stall >= patience:
                    break
        iters.append(n_it)
    rms = masked_rms(f, g, valid_g, M, tr, c, cval)
    A = np.linalg.inv(I3 - M)
    return A, A @ tr, {'rms_initial': rms_initial, 'rms_final': rms, 'iters': iters}


def warp_onto_ref_grid(A, t, c, Nf, g):
    """g sampled at x' = A (x - c) + c + t for every voxel x of an Nf grid,
    trilinear, 0 outside (preprocess.py:233-285).  The 3x3 product is
    evaluated as the float32 sum over j = 0, 1, 2 of A_ij (x_j - c_j)."""
    A32 = np.asarray(A, np.float64).astype(np.float32)
    t32 = np.asarray(t, np.float64).astype(np.float32)
    c32 = np.asarray(c, np.float64).astype(np.float32)
    X, Y, Z = np.meshgrid(*[np.arange(n, dtype=np.float32) for n in Nf], indexing='ij')
    d = [X - c32[0], Y - c32[1], Z - c32[2]]
    pts = [((A32[i, 0] * d[0] + A32[i, 1] * d[1]) + A32[i, 2] * d[2]) + c32[i] + t32[i]
           for i in range(3)]
    return map_coordinates_01(np.asarray(g, np.float32), pts, order=1, mode='constant',
                              cval=0.0)


def compute_crop(f, g_warped, h, glt=0.0):
    """n*h+1 crop of the valid overlap (f > glt and g_warped > glt): bounding
    box from axis projections, each size snapped down to n*h+1, re-centred,
    shifted into f's bounds (preprocess.py:315-367).  (lo, hi exclusive)."""
    valid = (f > glt) & (g_warped > glt)
    idx = [np.where(valid.any(axis=tuple(a for a in range(3) if a != ax)))[0] for ax in range(3)]
    if idx[0].size == 0:
        raise RuntimeError('No valid overlap region found')
    lo = np.array([i[0] for i in idx])
    hi = np.array([i[-1] for i in idx])
    size = hi - lo + 1
    snapped = ((size - 1) // h) * h + 1
    mid = (lo + hi) // 2
    lo_out = mid - snapped // 2
    hi_out = lo_out + snapped
    Nf = np.array(f.shape)
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
