"""Seeded synthetic DVC workloads: speckle reference volumes and known warps.

Speckle:  f = background + sum_i A_i exp(-|x - c_i|^2 / (2 sigma^2)),
c_i ~ U[0, n)^3, A_i ~ U(50, 255), truncated at 4 sigma.
Warp (backward, as the reference's ground-truth generator
ground_truth/prepare_ground_truth.py:63-86):  g(y) = f(y - u(y)) by
trilinear interpolation with linear extrapolation at the border, then noise.

Generation runs in NumPy for small volumes and in torch on the GPU for large
ones (input preparation only; it is not part of the timed solve).
"""
import numpy as np


def speckle_numpy(shape, n_blobs, seed, sigma=1.5, amp=(50.0, 255.0), background=20.0):
    """float64 speckle volume of `shape` (x, y, z), C-ordered [x, y, z]."""
    rng = np.random.default_rng(seed)
    shape = tuple(int(s) for s in shape)
    c = rng.uniform(0.0, 1.0, size=(n_blobs, 3)) * np.array(shape, dtype=np.float64)
    a = rng.uniform(amp[0], amp[1], size=n_blobs)
    r = int(np.ceil(4 * sigma))
    off = np.arange(-r, r + 1)
    f = np.full(shape, background, dtype=np.float64)
    chunk = 4096
    for s in range(0, n_blobs, chunk):
        cc = c[s:s + chunk]
        aa = a[s:s + chunk]
        base = np.floor(cc).astype(np.int64)
        ix = base[:, 0:1] + off[None, :]
        iy = base[:, 1:2] + off[None, :]
        iz = base[:, 2:3] + off[None, :]
        wx = np.exp(-((ix - cc[:, 0:1]) ** 2) / (2 * sigma ** 2))
        wy = np.exp(-((iy - cc[:, 1:2]) ** 2) / (2 * sigma ** 2))
        wz = np.exp(-((iz - cc[:, 2:3]) ** 2) / (2 * sigma ** 2))
        vx = (ix >= 0) & (ix < shape[0])
        vy = (iy >= 0) & (iy < shape[1])
        vz = (iz >= 0) & (iz < shape[2])
        val = (aa[:, None, None, None] * wx[:, :, None, None] * wy[:, None, :, None]
               * wz[:, None, None, :])
        ok = vx[:, :, None, None] & vy[:, None, :, None] & vz[:, None, None, :]
        IX = np.broadcast_to(ix[:, :, None, None], val.shape)[ok]
        IY = np.broadcast_to(iy[:, None, :, None], val.shape)[ok]
        IZ = np.broadcast_to(iz[:, None, None, :], val.shape)[ok]
        np.add.at(f, (IX, IY, IZ), val[ok])
    return f


def trilinear_extrap(img, px, py, pz):
    """Trilinear sample of img at float coordinates with linear extrapolation
    outside [0, n-1] (RegularGridInterpolator(fill_value=None) semantics)."""
    out = np.zeros(px.shape, dtype=np.float64)
    idx = []
    for p, n in zip((px, py, pz), img.shape):
        i0 = np.clip(np.floor(p).astype(np.int64), 0, n - 2)
        idx.append((i0, p - i0))
    (x0, fx), (y0, fy), (z0, fz) = idx
    for dx, wx in ((0, 1 - fx), (1, fx)):
        for dy, wy in ((0, 1 - fy), (1, fy)):
            for dz, wz in ((0, 1 - fz), (1, fz)):
                out += wx * wy * wz * img[x0 + dx, y0 + dy, z0 + dz]
    return out


def warp_backward(f, ufield, noise_sigma, seed):
    """g(y) = f(y - u(y)) + N(0, noise_sigma); `ufield(X, Y, Z)` returns the
    three displacement components at voxel coordinates."""
    nx, ny, nz = f.shape
    X, Y, Z = np.meshgrid(np.arange(nx, dtype=np.float64), np.arange(ny, dtype=np.float64),
                          np.arange(nz, dtype=np.float64), indexing='ij')
    ux, uy, uz = ufield(X, Y, Z)
    g = trilinear_extrap(f, X - ux, Y - uy, Z - uz)
    if noise_sigma > 0:
        g = g + np.random.default_rng(seed).normal(0.0, noise_sigma, size=g.shape)
    return g


def translation_strain_field(t=(2.3, -1.1, 0.7), ez=0.01, zc=32.0):
    """Translation plus uniform z-strain about plane z = zc."""
    def u(X, Y, Z):
        return (np.full_like(X, t[0]), np.full_like(Y, t[1]), t[2] + ez * (Z - zc))
    return u


def necking_field(center, ez=0.02, nu_r=0.3, amp=0.5, width=40.0):
    """Tensile necking about `center`: u_z = ez (z - cz) e(r),
    u_r = -nu_r ez r e(r), e(r) = 1 + amp exp(-r^2 / (2 width^2))."""
    cx, cy, cz = center

    def u(X, Y, Z):
        dx, dy = X - cx, Y - cy
        r2 = dx * dx + dy * dy
        # NumPy arrays or torch tensors
        e = 1.0 + amp * (np.exp(-r2 / (2 * width ** 2)) if isinstance(r2, np.ndarray)
                         else (-r2 / (2 * width ** 2)).exp())
        return (-nu_r * ez * dx * e, -nu_r * ez * dy * e, ez * (Z - cz) * e)
    return u


def fourier_field(n_modes, wl_range, peak, seed, extent):
    """Sum of random plane-wave modes, rescaled so that max |u| over the
    `extent` grid is `peak` voxels."""
    rng = np.random.default_rng(seed)
    k = rng.normal(size=(n_modes, 3))
    k /= np.linalg.norm(k, axis=1, keepdims=True)
    wl = rng.uniform(wl_range[0], wl_range[1], size=n_modes)
    kv = k * (2 * np.pi / wl)[:, None]
    ph = rng.uniform(0, 2 * np.pi, size=n_modes)
    av = rng.normal(size=(n_modes, 3))

    def raw(X, Y, Z):
        out = [np.zeros_like(X) for _ in range(3)]
        for m in range(n_modes):
            s = np.sin(kv[m, 0] * X + kv[m, 1] * Y + kv[m, 2] * Z + ph[m])
            for c in range(3):
                out[c] += av[m, c] * s
        return out

    g = [np.arange(0, n, max(1, n // 32), dtype=np.float64) for n in extent]
    Xs, Ys, Zs = np.meshgrid(*g, indexing='ij')
    r = raw(Xs, Ys, Zs)
    scale = peak / max(1e-12, float(np.sqrt(r[0] ** 2 + r[1] ** 2 + r[2] ** 2).max()))

    def u(X, Y, Z):
        return tuple(scale * c for c in raw(X, Y, Z))
    return u


def blobs_for(shape, density_ref=(250_000, 257 ** 3)):
    """Blob count at the density of 250k blobs per 257^3."""
    return int(round(density_ref[0] * np.prod(shape) / density_ref[1]))


def c1_pair(seed=1, n=65, n_blobs=4000):
    """65^3 speckle, translation (2.3,-1.1,0.7) + 1% z-strain, N(0,1)."""
    f = speckle_numpy((n, n, n), n_blobs, seed)
    g = warp_backward(f, translation_strain_field(zc=(n - 1) / 2.0), 1.0, seed + 1000)
    return f, g


def c2_pair(seed=2, n=257, n_blobs=None):
    """257^3 speckle with the tensile-necking field, N(0,1)."""
    n_blobs = blobs_for((n, n, n)) if n_blobs is None else n_blobs
    f = speckle_numpy((n, n, n), n_blobs, seed)
    c = (n - 1) / 2.0
    g = warp_backward(f, necking_field((c, c, c)), 1.0, seed + 1000)
    return f, g


def speckle_torch(shape, n_blobs, seed, sigma=1.5, amp=(50.0, 255.0), background=20.0,
                  device='cuda', chunk=8192):
    """Same generator as speckle_numpy, evaluated in float32 on the GPU with
    torch (for volumes too large for the host); returns a flat x-fastest
    float32 tensor (index x + y*nx + z*nx*ny)."""
    import torch
    rng = np.random.default_rng(seed)
    shape = tuple(int(s) for s in shape)
    c = rng.uniform(0.0, 1.0, size=(n_blobs, 3)) * np.array(shape, dtype=np.float64)
    a = rng.uniform(amp[0], amp[1], size=n_blobs)
    r = int(np.ceil(4 * sigma))
    nx, ny, nz = shape
    f = torch.full((nx * ny * nz,), float(background), dtype=torch.float32, device=device)
    off = torch.arange(-r, r + 1, device=device, dtype=torch.int64)
    inv = 1.0 / (2 * sigma ** 2)
    for s in range(0, n_blobs, chunk):
        cc = torch.from_numpy(c[s:s + chunk]).to(device)
        aa = torch.from_numpy(a[s:s + chunk]).to(device=device, dtype=torch.float32)
        base = torch.floor(cc).to(torch.int64)
        idx = [base[:, d:d + 1] + off[None, :] for d in range(3)]
        w = [torch.exp(-((idx[d].to(torch.float64) - cc[:, d:d + 1]) ** 2) * inv).to(torch.float32)
             for d in range(3)]
        ok = [(idx[d] >= 0) & (idx[d] < shape[d]) for d in range(3)]
        w = [w[d] * ok[d] for d in range(3)]
        idx = [idx[d].clamp(0, shape[d] - 1) for d in range(3)]
        val = (aa[:, None, None, None] * w[0][:, :, None, None] * w[1][:, None, :, None]
               * w[2][:, None, None, :])
        lin = (idx[0][:, :, None, None] + nx * (idx[1][:, None, :, None]
                                                 + ny * idx[2][:, None, None, :]))
        f.index_add_(0, lin.reshape(-1), val.reshape(-1))
    return f


def fourier_params(n_modes, wl_range, peak, seed, extent):
    """The random plane-wave modes of fourier_field and the scale that makes
    max |u| over a coarse grid of `extent` equal `peak` voxels."""
    rng = np.random.default_rng(seed)
    k = rng.normal(size=(n_modes, 3))
    k /= np.linalg.norm(k, axis=1, keepdims=True)
    wl = rng.uniform(wl_range[0], wl_range[1], size=n_modes)
    kv = k * (2 * np.pi / wl)[:, None]
    ph = rng.uniform(0, 2 * np.pi, size=n_modes)
    av = rng.normal(size=(n_modes, 3))
    g = [np.arange(0, n, max(1, n // 32), dtype=np.float64) for n in extent]
    Xs, Ys, Zs = np.meshgrid(*g, indexing='ij')
    r = [np.zeros_like(Xs) for _ in range(3)]
    for m in range(n_modes):
        s = np.sin(kv[m, 0] * Xs + kv[m, 1] * Ys + kv[m, 2] * Zs + ph[m])
        for c in range(3):
            r[c] += av[m, c] * s
    scale = peak / max(1e-12, float(np.sqrt(r[0] ** 2 + r[1] ** 2 + r[2] ** 2).max()))
    return kv, ph, av * scale


def random_affine(rng, max_rot_deg, max_strain):
    """A = R (I + E): rotation about a random axis by <= max_rot_deg degrees,
    symmetric strain E with entries <= max_strain."""
    ax = rng.normal(size=3)
    ax /= np.linalg.norm(ax)
    th = np.deg2rad(rng.uniform(-max_rot_deg, max_rot_deg))
    Kx = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    R = np.eye(3) + np.sin(th) * Kx + (1 - np.cos(th)) * Kx @ Kx
    E = rng.uniform(-max_strain, max_strain, size=(3, 3))
    E = 0.5 * (E + E.T)
    return R @ (np.eye(3) + E)


def warp_backward_torch(f, shape, modes, noise_sigma, seed, quantize=True, slab=8,
                        device='cuda', affine=None, extra_u=None):
    """GPU form of warp_backward for large volumes: g(y) = f(y - u(y)) with u
    a sum of plane-wave modes (fourier_params; None for none), trilinear
    interpolation with linear extrapolation at the border (float64
    coordinates and weights), plus N(0, noise_sigma) noise (torch generator
    `seed`), then optionally rounded and clipped to [0, 255].  f: flat
    x-fastest float32 tensor of `shape` (nx, ny, nz); returns the same
    layout.  `affine` = (A, c) adds the displacement (A - I)(y - c) to u;
    `extra_u(X, Y, Z)` (e.g. necking_field) adds its three components."""
    import torch
    nx, ny, nz = shape
    if modes is None:
        modes = (np.zeros((0, 3)), np.zeros(0), np.zeros((0, 3)))
    kv, ph, av = (torch.as_tensor(a, dtype=torch.float64, device=device) for a in modes)
    ff = f.to(device=device, dtype=torch.float64)
    out = torch.empty(nx * ny * nz, dtype=torch.float32, device=device)
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    x = torch.arange(nx, dtype=torch.float64, device=device)
    y = torch.arange(ny, dtype=torch.float64, device=device)
    for z0 in range(0, nz, slab):
        z = torch.arange(z0, min(nz, z0 + slab), dtype=torch.float64, device=device)
        Z, Y, X = torch.meshgrid(z, y, x, indexing='ij')      # x fastest
        ph_ = (X[..., None] * kv[:, 0] + Y[..., None] * kv[:, 1] + Z[..., None] * kv[:, 2]
               + ph)
        s = torch.sin(ph_)
        u = [(s * av[:, c]).sum(-1) for c in range(3)]
        if extra_u is not None:
            eu = extra_u(X, Y, Z)
            u = [u[c] + eu[c] for c in range(3)]
        if affine is not None:
            A, cen = affine
            Am = np.asarray(A, np.float64) - np.eye(3)
            d = [X - float(cen[0]), Y - float(cen[1]), Z - float(cen[2])]
            for c in range(3):
                u[c] = u[c] + Am[c, 0] * d[0] + Am[c, 1] * d[1] + Am[c, 2] * d[2]
        p = [X - u[0], Y - u[1], Z - u[2]]
        val = torch.zeros_like(X)
        idx, fr = [], []
        for q, n in zip(p, (nx, ny, nz)):
            i0 = torch.clamp(torch.floor(q), 0, n - 2).to(torch.int64)
            idx.append(i0)
            fr.append(q - i0)
        for dx in (0, 1):
            wx = fr[0] if dx else 1 - fr[0]
            for dy in (0, 1):
                wy = fr[1] if dy else 1 - fr[1]
                for dz in (0, 1):
                    wz = fr[2] if dz else 1 - fr[2]
                    lin = (idx[0] + dx) + nx * ((idx[1] + dy) + ny * (idx[2] + dz))
                    val += wx * wy * wz * ff[lin]
        if noise_sigma > 0:
            val += noise_sigma * torch.randn(val.shape, generator=gen, dtype=torch.float64,
                                             device=device)
        if quantize:
            val = val.round().clamp(0, 255)
        out[z0 * nx * ny:(z0 + len(z)) * nx * ny] = val.reshape(-1).to(torch.float32)
    return out


def fourier_u_nodes(modes, Nx, Ny, Nz, h):
    """The true displacement of a plane-wave field at the mesh nodes
    (node i at voxel i*h), reference DOF layout (3 per node, x fastest).
    The warp g(y) = f(y - u(y)) is matched by U(x) with x + U(x) = y, so the
    nodal truth solves x + U = y with u evaluated at y; this returns the
    first-order value u(x) (exact for translations)."""
    kv, ph, av = modes
    i = np.arange(Nx + 1) * h
    j = np.arange(Ny + 1) * h
    k = np.arange(Nz + 1) * h
    Z, Y, X = np.meshgrid(k, j, i, indexing='ij')
    s = np.sin(X[..., None] * kv[:, 0] + Y[..., None] * kv[:, 1] + Z[..., None] * kv[:, 2] + ph)
    U = np.zeros(3 * X.size)
    for c in range(3):
        U[c::3] = (s * av[:, c]).sum(-1).ravel()
    return U


def config_pair_torch(kind, shape, seed=3):
    """Device pairs of the benchmark configurations (flat x-fastest float32,
    8-bit speckle quantised to 0..255 at the density of 250k blobs per 257^3):
      'c2': tensile necking about the centre (necking_field) + N(0,1);
      'c3': 16 Fourier modes, wavelengths 64-256, peak 4 vox + N(0,1);
      'c5': necking + 8 Fourier modes, peak 8 vox + N(0,2).
    Returns (f, g, modes, extra_u) so callers can evaluate the truth."""
    f = speckle_torch(shape, blobs_for(shape), seed=seed).round().clamp_(0, 255)
    c = tuple((n - 1) / 2.0 for n in shape)
    if kind == 'c2':
        modes, extra, noise = None, necking_field(c), 1.0
    elif kind == 'c3':
        modes, extra, noise = fourier_params(16, (64, 256), 4.0, seed + 4, shape), None, 1.0
    elif kind == 'c5':
        modes, extra, noise = fourier_params(8, (64, 256), 8.0, seed + 4, shape), \
            necking_field(c), 2.0
    else:
        raise ValueError('unknown configuration %r' % (kind,))
    g = warp_backward_torch(f, shape, modes, noise, seed=seed + 8, quantize=True,
                            extra_u=extra)
    return f, g, modes, extra


def nodal_truth(modes, extra_u, Nx, Ny, Nz, h):
    """First-order nodal truth u(x) of a config_pair_torch warp at the mesh
    nodes (reference DOF layout, float64)."""
    U = np.zeros(3 * (Nx + 1) * (Ny + 1) * (Nz + 1))
    if modes is not None:
        U += fourier_u_nodes(modes, Nx, Ny, Nz, h)
    if extra_u is not None:
        k = np.arange(Nz + 1) * h
        j = np.arange(Ny + 1) * h
        i = np.arange(Nx + 1) * h
        Z, Y, X = np.meshgrid(k, j, i, indexing='ij')
        eu = extra_u(X.astype(np.float64), Y.astype(np.float64), Z.astype(np.float64))
        for c in range(3):
            U[c::3] += np.asarray(eu[c]).ravel()
    return U
