"""Lane-exact NumPy model of the voxel-streaming kernels' decomposition.

The Hessian kernels split the voxel grid over wavefronts (x strips x element
rows x z tiles), restrict per lane, reduce across the lanes of one element
and hand node partials to a gather kernel through per-wave slot blocks.  This
model replays that index arithmetic lane by lane (64 lanes per wave) for both
decompositions -- the vectorised one (V voxels per lane, element-aligned
strips, last voxel plane as a per-row tail; with E = 2 a lane owns two
elements, V = 2h) and the generic one (one voxel per lane, 64-voxel strips)
-- and checks on CPU that every voxel is visited exactly once, every slot the
gather reads was written, and the result equals the oracle's matvec_H /
build_diagonal_H.  It mirrors vdvc_hessian.hip,
vdvc_voxel.hip (generic kernel) and gather_partials in vdvc_node.hip.
"""
import numpy as np
import pytest

from oracle import dvc_oracle as O


def aw(t, nl):
    return np.where((t == 0) | (t == nl), 0.5, 1.0)


def plan(Nx, Ny, Nz, h, V, target, nys=1, E=1):
    if V:
        sw = 64 * V
        nsx = -(-(Nx * h) // sw)
        px = 64 * V // h + 2
    else:
        sw = 64
        nsx = -(-(Nx * h + 1) // 64)
        px = 63 // h + 3
        nys = 1
    base = nsx * Ny * nys
    nzt_t = max(1, -(-target // base))
    ez = max(1, -(-Nz // nzt_t))
    nzt = -(-Nz // ez)
    return dict(V=V, E=E, sw=sw, nsx=nsx, nys=nys, px=px, ez=ez, nzt=nzt, nwaves=base * nzt,
                stride=(ez + 1) * 2 * px * 3)


def vec_width(h, itemsize):
    if h < 1 or h > 64 or (h & (h - 1)):
        return 0
    vmax = 16 // itemsize
    return min(h, vmax)


def run_vec(mode, v, g, m, Nx, Ny, Nz, h, T, covered):
    V = T['V']
    Nvx, Nvy, Nvz = Nx * h + 1, Ny * h + 1, Nz * h + 1
    nnx, nny = Nx + 1, Ny + 1
    nid = lambda i, j, k: i + nnx * (j + nny * k)
    P = np.full((T['nwaves'], T['stride']), np.nan)
    lane = np.arange(64)
    xend = Nx * h
    E = T.get('E', 1)
    G = h // V if E == 1 else 1                       # lanes per element
    nys = T['nys']
    rpp = h // nys                                    # rows per part
    for w in range(T['nwaves']):
        sx = w % T['nsx']
        r = w // T['nsx']
        ys = r % nys
        r //= nys
        ey, zt = r % Ny, r // Ny
        x0 = sx * T['sw']
        xb = x0 + lane * V
        act = xb < xend
        xbc = np.where(act, xb, x0)
        ex = xbc // h
        tb = xbc - ex * h
        tail = act & (xb + V == xend)
        X1b = tb / h
        wx0 = np.where(xb == 0, 0.5, 1.0)
        e_lo = x0 // h
        e_hi = min(Nx, (x0 + T['sw']) // h)
        nn = e_hi - e_lo + 1
        ez0 = zt * T['ez']
        ez1 = min(Nz, ez0 + T['ez'])
        ty0 = ys * rpp                                # first local row of this part
        y0 = ey * h + ty0
        nrows = rpp + (1 if (ey == Ny - 1 and ys == nys - 1) else 0)
        C = np.zeros((E + 1, 2, 2, 3, 64))

        def emit(kzl):
            a = C[:, :, 0].copy()                     # (cx, cy, k, lane)
            grp = a.reshape(E + 1, 2, 3, 64 // G, G).sum(-1)
            base = kzl * 2 * T['px'] * 3
            for ln in range((64 // G) * E + 1):
                if ln >= nn:
                    continue
                gi, cx = divmod(ln, E)                # lane group, its node column
                val = np.zeros((2, 3))
                if gi < 64 // G:
                    val += grp[cx, :, :, gi]
                if cx == 0 and ln >= 1:
                    val += grp[E, :, :, gi - 1]
                for cy in range(2):
                    o = base + (cy * T['px'] + ln) * 3
                    P[w, o:o + 3] = val[cy]

        for ez in range(ez0, ez1):
            nzp = h + (1 if ez == Nz - 1 else 0)
            for tz in range(nzp):
                z = ez * h + tz
                Z1 = tz / h
                Z0 = 1 - Z1
                Wz = float(aw(z, Nvz - 1))
                if mode == 'mv':
                    def nv(i, j, k):
                        n = nid(i, j, k)
                        return np.stack([v[3 * n + c] for c in range(3)])
                    # u at the lane's E + 1 node columns, row 0, and y slope
                    c0 = [Z0 * nv(ex + cx, ey, ez) + Z1 * nv(ex + cx, ey, ez + 1)
                          for cx in range(E + 1)]
                    c1 = [Z0 * nv(ex + cx, ey + 1, ez) + Z1 * nv(ex + cx, ey + 1, ez + 1)
                          for cx in range(E + 1)]
                    dc = [(c1[cx] - c0[cx]) / h for cx in range(E + 1)]
                A = np.zeros((E + 1, 2, 3, 64))
                for ty in range(nrows):
                    y = y0 + ty
                    Wr = Wz * float(aw(y, Nvy - 1))
                    S = np.zeros((E, 3, 64))
                    Tt = np.zeros((E, 3, 64))
                    if mode == 'mv':
                        uc = [c0[cx] + (ty0 + ty) * dc[cx] for cx in range(E + 1)]
                        d = [uc[e + 1] - uc[e] for e in range(E)]
                    for j in range(V):
                        x = xbc + j
                        vi = x + Nvx * (y + Nvy * z)
                        covered[vi[act]] += 1
                        e = (j + tb) // h                   # the lane's element of voxel j
                        X1 = X1b + j / h - e
                        wm = (wx0 if j == 0 else 1.0) * m[vi]
                        gg = g[vi].T
                        if mode == 'mv':
                            # u of each lane's own element at its X1
                            u = np.stack([uc[q] + X1 * d[q] for q in range(E)])[e, :, lane].T
                            s = wm * (gg * u).sum(0)
                            for q in range(E):
                                sel = e == q
                                S[q][:, sel] += (s * gg)[:, sel]
                                Tt[q][:, sel] += (X1 * s * gg)[:, sel]
                        else:
                            for q in range(E):
                                sel = e == q
                                S[q][:, sel] += ((1 - X1) ** 2 * wm * gg ** 2)[:, sel]
                                Tt[q][:, sel] += (X1 ** 2 * wm * gg ** 2)[:, sel]
                    if tail.any():
                        vi = xend + Nvx * (y + Nvy * z)
                        covered[vi] += 1
                        gg = g[vi][:, None]
                        wm = 0.5 * m[vi]
                        if mode == 'mv':
                            s = wm * (gg * (uc[E - 1] + d[E - 1])).sum(0)   # t = h: X1 = 1
                            add = (s * gg)[:, tail]
                            S[E - 1][:, tail] += add
                        else:
                            add = wm * gg ** 2
                        Tt[E - 1][:, tail] += add
                    S[:, :, ~act] = 0
                    Tt[:, :, ~act] = 0
                    # node-column row values; a paired lane's middle column takes
                    # the right part of element 0 and the left part of element 1
                    R = np.zeros((E + 1, 3, 64))
                    R[0] = S[0] - Tt[0] if mode == 'mv' else S[0]
                    for c in range(1, E):
                        R[c] = Tt[c - 1] + (S[c] - Tt[c] if mode == 'mv' else S[c])
                    R[E] = Tt[E - 1]
                    Y1 = (ty0 + ty) / h
                    wy = (Wr * (1 - Y1), Wr * Y1) if mode == 'mv' else (Wr * (1 - Y1) ** 2,
                                                                        Wr * Y1 ** 2)
                    A[:, 0] += wy[0] * R
                    A[:, 1] += wy[1] * R
                zw = (Z0, Z1) if mode == 'mv' else (Z0 ** 2, Z1 ** 2)
                C[:, :, 0] += zw[0] * A
                C[:, :, 1] += zw[1] * A
            emit(ez - ez0)
            C[:, :, 0] = C[:, :, 1]
            C[:, :, 1] = 0
        emit(ez1 - ez0)
    return P


def gather(P, T, Nx, Ny, Nz, h):
    Nvx = Nx * h + 1
    nnx, nny = Nx + 1, Ny + 1
    out = np.zeros(3 * nnx * nny * (Nz + 1))
    for k in range(Nz + 1):
        for j in range(Ny + 1):
            for i in range(Nx + 1):
                s = np.zeros(3)
                za = max(0, k // T['ez'] - 1)
                zb = min(T['nzt'] - 1, k // T['ez'])
                if T['V']:
                    xa = (max(i - 1, 0) * h) // T['sw']
                    xb = (min(i, Nx - 1) * h) // T['sw']
                else:
                    xa = max(0, (i - 1) * h) // 64
                    xb = min(Nvx - 1, (i + 1) * h) // 64
                for zt in range(za, zb + 1):
                    kb = zt * T['ez']
                    if k < kb or k > min(Nz, kb + T['ez']):
                        continue
                    for cy in (1, 0):
                        ey = j - cy
                        if ey < 0 or ey >= Ny:
                            continue
                        for sx, ys in ((a, b) for a in range(xa, xb + 1)
                                       for b in range(T.get('nys', 1))):
                            x0 = sx * T['sw']
                            i_lo = x0 // h
                            if T['V']:
                                nn = min(Nx, (x0 + T['sw']) // h) - i_lo + 1
                            else:
                                nact = min(64, Nvx - x0)
                                nn = min(Nx, (x0 + nact - 1) // h + 1) - i_lo + 1
                            ln = i - i_lo
                            if ln < 0 or ln >= nn:
                                continue
                            w = sx + T['nsx'] * (ys + T.get('nys', 1) * (ey + Ny * zt))
                            o = ((k - kb) * 2 + cy) * T['px'] * 3 + ln * 3
                            blk = P[w, o:o + 3]
                            assert not np.isnan(blk).any(), ('unwritten slot', i, j, k)
                            s += blk
                n = i + nnx * (j + nny * k)
                out[3 * n:3 * n + 3] = s
    return out


CASES = [  # (Nx, Ny, Nz, h, itemsize, target)
    (2, 2, 2, 8, 2, 8192), (3, 2, 3, 8, 4, 4), (70, 1, 2, 8, 2, 3), (9, 2, 2, 16, 2, 5),
    (130, 1, 1, 4, 2, 2), (65, 2, 1, 2, 2, 3), (3, 1, 2, 1, 2, 1), (2, 1, 2, 32, 2, 1),
    (1, 1, 1, 64, 4, 1), (17, 2, 2, 4, 4, 7), (64, 1, 1, 8, 2, 1),
]


PAIRED = [  # (Nx, Ny, Nz, h, itemsize, target): Nx even, V = 2h fits 16 bytes
    (2, 2, 2, 4, 2, 8192), (4, 1, 3, 4, 2, 2), (130, 1, 1, 4, 2, 2), (6, 2, 2, 2, 4, 3),
    (66, 1, 2, 2, 2, 1), (8, 3, 2, 4, 2, 5),
]


@pytest.mark.parametrize('nys', [1, 2])
@pytest.mark.parametrize('case', PAIRED)
def test_paired_decomposition(case, nys):
    """Lanes of two elements (V = 2h): three node columns per lane, the
    middle one shared by the pair, the lanes' columns written without a
    butterfly."""
    Nx, Ny, Nz, h, isz, tg = case
    V = 2 * h
    assert V * isz <= 16 and Nx % 2 == 0
    if h % nys:
        pytest.skip('row split needs nys | h')
    T = plan(Nx, Ny, Nz, h, V, tg, nys, E=2)
    assert T['px'] == 64 * 2 + 2
    rng = np.random.default_rng(sum(case) + 1)
    Nvox = (Nx * h + 1) * (Ny * h + 1) * (Nz * h + 1)
    g = rng.normal(size=(Nvox, 3))
    m = rng.random(Nvox) < 0.8
    v = rng.normal(size=3 * (Nx + 1) * (Ny + 1) * (Nz + 1))
    for mode, ref in (('mv', O.matvec_H(v, g, m, h, Nx, Ny, Nz)),
                      ('dg', O.build_diagonal_H(g, m, h, Nx, Ny, Nz))):
        cov = np.zeros(Nvox, int)
        P = run_vec(mode, v, g, m, Nx, Ny, Nz, h, T, cov)
        assert (cov == 1).all(), np.unique(cov)
        got = gather(P, T, Nx, Ny, Nz, h)
        assert np.abs(got - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize('nys', [1, 2])
@pytest.mark.parametrize('case', CASES)
def test_vectorised_decomposition(case, nys):
    Nx, Ny, Nz, h, isz, tg = case
    V = vec_width(h, isz)
    assert V and h % V == 0
    if h % nys:
        pytest.skip('row split needs nys | h')
    T = plan(Nx, Ny, Nz, h, V, tg, nys)
    rng = np.random.default_rng(sum(case))
    Nvox = (Nx * h + 1) * (Ny * h + 1) * (Nz * h + 1)
    g = rng.normal(size=(Nvox, 3))
    m = rng.random(Nvox) < 0.8
    v = rng.normal(size=3 * (Nx + 1) * (Ny + 1) * (Nz + 1))
    for mode, ref in (('mv', O.matvec_H(v, g, m, h, Nx, Ny, Nz)),
                      ('dg', O.build_diagonal_H(g, m, h, Nx, Ny, Nz))):
        cov = np.zeros(Nvox, int)
        P = run_vec(mode, v, g, m, Nx, Ny, Nz, h, T, cov)
        assert (cov == 1).all(), np.unique(cov)
        got = gather(P, T, Nx, Ny, Nz, h)
        assert np.abs(got - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


def test_vec_width_rule():
    assert [vec_width(h, 2) for h in (1, 2, 4, 8, 16, 64, 3, 6, 128)] == [1, 2, 4, 8, 8, 8, 0, 0, 0]
    assert [vec_width(h, 4) for h in (1, 2, 4, 8, 16)] == [1, 2, 4, 4, 4]
