"""GPU parity: every libvdvc entry point against the oracle on seeded inputs
(float32 results within the tolerance written in each test; integer / mask
results exact)."""
import numpy as np
import pytest

from conftest import golden, problem17, rand_problem
from oracle import dvc_oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [(2, 2, 2, 8), (3, 2, 4, 8), (1, 1, 1, 1), (9, 2, 3, 8), (17, 2, 2, 4), (5, 3, 2, 3),
          (4, 2, 3, 5), (70, 1, 2, 1), (2, 1, 3, 16), (3, 2, 2, 32), (13, 3, 3, 6), (8, 8, 8, 4)]


def rel(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(1e-30, np.abs(b).max()))


@pytest.mark.parametrize('shape', SHAPES)
def test_matvec_H(gpu, shape):
    from voxeldvc_amd.engine import kernels_dvc as K
    Nx, Ny, Nz, h = shape
    print('matvec_H', shape, flush=True)
    g, m, v = rand_problem(1, Nx, Ny, Nz, h)
    ref = O.matvec_H(v.astype(np.float64), g.astype(np.float64), m, h, Nx, Ny, Nz)
    got = K.matvec_H(v, g, m, None, None, h, Nx, Ny, Nz)
    assert got.dtype == np.float32          # float32 inputs take the float32 path
    assert rel(got, ref) < 2e-5
    # mask folded into the gradient: same operator
    got2 = K.matvec_H(v, g * m[:, None], None, None, None, h, Nx, Ny, Nz)
    assert rel(got2, ref) < 2e-5


@pytest.mark.parametrize('nys', ['1', '2'])
@pytest.mark.parametrize('gt', ['f32', 'bf16'])
@pytest.mark.parametrize('shape', [(9, 2, 3, 8), (8, 8, 8, 4), (3, 2, 2, 32), (65, 1, 1, 8),
                                   (4, 64, 64, 2), (1, 64, 64, 8)])
def test_matvec_H_tilings(gpu, monkeypatch, shape, gt, nys):
    """The two row splits of the vectorised kernel (VDVC_NYS tuning
    override), float32 and bfloat16 gradient planes, with and without a
    separate mask.  (4, 64, 64, 2) and (1, 64, 64, 8) give 4096 wavefronts
    with one element layer each, so without the row split the plan takes
    two layers per wavefront (bfloat16, or h <= 4)."""
    import torch
    from voxeldvc_amd import _lib as L
    Nx, Ny, Nz, h = shape
    print('matvec_H tiling', shape, gt, nys, flush=True)
    monkeypatch.setenv('VDVC_NYS', nys)
    g, m, v = rand_problem(11, Nx, Ny, Nz, h)
    shp = (Nx * h + 1, Ny * h + 1, Nz * h + 1)
    tdt = torch.float32 if gt == 'f32' else torch.bfloat16
    gq = torch.from_numpy(g).to(tdt)
    g64 = gq.to(torch.float64).numpy()                 # the values the kernel sees
    ref = O.matvec_H(v.astype(np.float64), g64, m, h, Nx, Ny, Nz)
    planes = [L.to_dev_level(gq[:, c].contiguous(), shp,
                             'float32' if gt == 'f32' else 'bfloat16') for c in range(3)]
    mk = L.to_dev_level(m, shp, 'uint8')
    dv = L.to_dev_flat(v)
    out = L.empty(dv.numel())
    code = L.F32 if gt == 'f32' else L.BF16
    L.call('vdvc_matvec_H', L.ctx(), Nx, Ny, Nz, h, *[L.ptr(q) for q in planes], code, L.ptr(mk),
           L.ptr(dv), L.ptr(out))
    assert rel(L.to_host(out), ref) < 2e-5


@pytest.mark.parametrize('epl', ['2', 'auto'])
@pytest.mark.parametrize('nys', ['1', '2'])
@pytest.mark.parametrize('gt', ['f32', 'bf16'])
@pytest.mark.parametrize('shape', [(8, 8, 8, 4), (2, 1, 3, 4), (130, 2, 2, 4), (6, 3, 2, 2),
                                   (128, 2, 2, 2), (7, 2, 2, 4)])
def test_matvec_H_paired(gpu, monkeypatch, shape, gt, nys, epl):
    """Lanes of two elements (V = 2h: bfloat16 at h = 4 and 2, float32 at
    h = 2), forced (VDVC_EPL=2) and as planned (pairs where a strip fits the
    row; odd Nx keeps single elements): the Hessian product and its diagonal
    against the oracle, with and without a separate mask."""
    import torch
    from voxeldvc_amd import _lib as L
    Nx, Ny, Nz, h = shape
    print('matvec_H paired', shape, gt, nys, epl, flush=True)
    if epl == '2':
        monkeypatch.setenv('VDVC_EPL', '2')
    monkeypatch.setenv('VDVC_NYS', nys)
    g, m, v = rand_problem(14, Nx, Ny, Nz, h)
    shp = (Nx * h + 1, Ny * h + 1, Nz * h + 1)
    tdt = torch.float32 if gt == 'f32' else torch.bfloat16
    dts = 'float32' if gt == 'f32' else 'bfloat16'
    gq = torch.from_numpy(g).to(tdt)
    g64 = gq.to(torch.float64).numpy()
    ref = O.matvec_H(v.astype(np.float64), g64, m, h, Nx, Ny, Nz)
    dref = O.build_diagonal_H(g64, m, h, Nx, Ny, Nz)
    planes = [L.to_dev_level(gq[:, c].contiguous(), shp, dts) for c in range(3)]
    mk = L.to_dev_level(m, shp, 'uint8')
    folded = [L.to_dev_level((gq[:, c].float() * torch.from_numpy(m)).to(tdt).contiguous(), shp,
                             dts) for c in range(3)]
    dv = L.to_dev_flat(v)
    code = L.F32 if gt == 'f32' else L.BF16
    for G, mask in ((planes, mk), (folded, None)):
        out = L.empty(dv.numel())
        L.call('vdvc_matvec_H', L.ctx(), Nx, Ny, Nz, h, *[L.ptr(q) for q in G], code,
               None if mask is None else L.ptr(mask), L.ptr(dv), L.ptr(out))
        assert rel(L.to_host(out), ref) < 2e-5
        d = L.empty(dv.numel())
        L.call('vdvc_diag_H', L.ctx(), Nx, Ny, Nz, h, *[L.ptr(q) for q in G], code,
               None if mask is None else L.ptr(mask), L.ptr(d))
        assert rel(L.to_host(d), dref) < 2e-5


@pytest.mark.parametrize('gt', ['f32', 'bf16'])
def test_matvec_H_offset_views(gpu, gt):
    """Gradient and mask planes handed over as views whose base address is
    not 16-byte aligned take the one-voxel-per-lane decomposition and give
    the same operator and diagonal."""
    import torch
    from voxeldvc_amd import _lib as L
    Nx, Ny, Nz, h = 9, 2, 3, 8
    print('matvec_H offset views', gt, flush=True)
    g, m, v = rand_problem(12, Nx, Ny, Nz, h)
    shp = (Nx * h + 1, Ny * h + 1, Nz * h + 1)
    tdt = torch.float32 if gt == 'f32' else torch.bfloat16
    dts = 'float32' if gt == 'f32' else 'bfloat16'
    gq = torch.from_numpy(g).to(tdt)
    g64 = gq.to(torch.float64).numpy()
    ref = O.matvec_H(v.astype(np.float64), g64, m, h, Nx, Ny, Nz)
    dref = O.build_diagonal_H(g64, m, h, Nx, Ny, Nz)
    planes = [L.to_dev_level(gq[:, c].contiguous(), shp, dts) for c in range(3)]
    mk = L.to_dev_level(m, shp, 'uint8')
    views = []
    for q in planes + [mk]:
        buf = torch.zeros(q.numel() + 1, dtype=q.dtype, device=q.device)
        buf[1:].copy_(q)
        views.append(buf[1:])
    assert all(t.data_ptr() % 16 for t in views)
    dv = L.to_dev_flat(v)
    code = L.F32 if gt == 'f32' else L.BF16
    for src in (planes + [mk], views):
        out = L.empty(dv.numel())
        L.call('vdvc_matvec_H', L.ctx(), Nx, Ny, Nz, h, *[L.ptr(q) for q in src[:3]], code,
               L.ptr(src[3]), L.ptr(dv), L.ptr(out))
        assert rel(L.to_host(out), ref) < 2e-5
        d = L.empty(dv.numel())
        L.call('vdvc_diag_H', L.ctx(), Nx, Ny, Nz, h, *[L.ptr(q) for q in src[:3]], code,
               L.ptr(src[3]), L.ptr(d))
        assert rel(L.to_host(d), dref) < 2e-5


@pytest.mark.parametrize('shape', SHAPES)
def test_diag_H(gpu, shape):
    from voxeldvc_amd.engine import kernels_dvc as K
    Nx, Ny, Nz, h = shape
    print('diag_H', shape, flush=True)
    g, m, _ = rand_problem(2, Nx, Ny, Nz, h)
    ref = O.build_diagonal_H(g.astype(np.float64), m, h, Nx, Ny, Nz)
    got = K.build_diagonal_H(g, m, None, None, h, Nx, Ny, Nz, ref.size)
    assert rel(got, ref) < 2e-5


@pytest.mark.parametrize('shape', SHAPES[:8])
def test_compute_b(gpu, shape):
    from voxeldvc_amd.engine import correlate_gpu as C
    Nx, Ny, Nz, h = shape
    print('compute_b', shape, flush=True)
    g, m, _ = rand_problem(3, Nx, Ny, Nz, h)
    res = np.random.default_rng(4).normal(size=m.size).astype(np.float32)
    ref = O.compute_b(res.astype(np.float64), m, g.astype(np.float64), h, Nx, Ny, Nz)
    got = C.compute_b_gpu(res, m, g, None, None, h, Nx, Ny, Nz, ref.size)
    assert got.dtype == np.float32
    assert rel(got, ref) < 2e-5


@pytest.mark.parametrize('reg', ['laplacian', 'none'])
@pytest.mark.parametrize('form', ['f32', 'bf16', 'f64'])
@pytest.mark.parametrize('shape', [(3, 2, 2, 8), (64, 2, 1, 8), (5, 3, 4, 4), (9, 7, 5, 2),
                                   (5, 3, 4, 3)])
def test_residual_rhs_jacobi_matches_two_passes(gpu, shape, form, reg, monkeypatch):
    """vdvc_residual_rhs_jacobi = vdvc_residual_rhs + vdvc_jacobi of the
    dynamic mask's H: b and the refolded planes bit for bit, diag and M_inv
    to float32 rounding where the diagonal rides on the residual pass (the
    float32 and mixed forms at power-of-two h), bit for bit where the entry
    runs the two passes (FP64, h = 3, or VDVC_NO_RHS_DIAG=1)."""
    import torch
    from voxeldvc_amd import _lib as L
    from voxeldvc_amd.engine import kernels_dvc as Kd
    Nx, Ny, Nz, h = shape
    print('residual_rhs_jacobi', shape, form, reg, flush=True)
    rng = np.random.default_rng(14)
    shp = Kd.level_shape(h, Nx, Ny, Nz)
    nv = shp[0] * shp[1] * shp[2]
    g, m, _ = rand_problem(15, Nx, Ny, Nz, h)
    f = rng.uniform(0, 255, nv).round()
    gv = rng.uniform(0, 255, nv)
    rd, fd, gd = {'f32': ('float32', 'float32', 'float32'),
                  'bf16': ('float32', 'float16', 'bfloat16'),
                  'f64': ('float64', 'float64', 'float64')}[form]
    fold = form == 'bf16'
    gx, gy, gz = Kd.grad_planes(g, shp, gd)
    fimg = L.to_dev_level(f, shp, fd)
    gvals = L.to_dev_level(gv, shp, rd)
    bm = L.to_dev_level(m, shp, 'uint8')
    ndof = 3 * (Nx + 1) * (Ny + 1) * (Nz + 1)
    Kp = np.ascontiguousarray(O.build_K_ref_laplacian(), dtype=np.float64)
    rk = ((L.REG_K, Kp.ctypes.data_as(L._dp)) if reg == 'laplacian' else (L.REG_NONE, None))
    rk = rk + (None, float(h), 0.0, 0.5 if reg == 'laplacian' else 0.0, None)
    npl = fimg.numel()

    def run(fused):
        b = L.empty(ndof, rd)
        fo = [L.empty(npl, 'bfloat16') for _ in range(3)] if fold else [None] * 3
        d, mi = L.empty(ndof, rd), L.empty(ndof, rd)
        s4 = (L._d * 4)()
        args = (L.ctx(), Nx, Ny, Nz, h, L.ptr(gx), L.ptr(gy), L.ptr(gz), L.type_code(gd),
                L.ptr(fimg), L.type_code(fd), 97.25, L.ptr(gvals), 101.5, -1.0625, L.ptr(bm),
                L.ptr(bm), None, L.ptr(b), *[L.ptr(q) for q in fo], s4)
        if fused:
            L.call('vdvc_residual_rhs_jacobi', *args, *rk, L.ptr(d), L.ptr(mi))
        else:
            L.call('vdvc_residual_rhs', *args)
            hg = fo if fold else [gx, gy, gz]
            L.call('vdvc_jacobi', L.ctx(), Nx, Ny, Nz, h, *[L.ptr(q) for q in hg],
                   L.BF16 if fold else L.type_code(gd), None if fold else L.ptr(bm), *rk,
                   L.ptr(d), L.ptr(mi))
        return b, fo, d, mi, tuple(s4)

    b0, f0, d0, m0, s0 = run(False)
    b1, f1, d1, m1, s1 = run(True)
    assert torch.equal(b1, b0) and s1 == s0
    if fold:
        # (the row padding of the refolded planes is never written)
        assert all(torch.equal(L.unpitch(p, shp), L.unpitch(q, shp)) for p, q in zip(f1, f0))
    exact = form == 'f64' or h == 3
    if exact:
        assert torch.equal(d1, d0) and torch.equal(m1, m0)
    else:
        # sums of non-negative float32 terms in another order: a few ulps
        tol = 1e-5 * float(d0.abs().max())
        assert float((d1.double() - d0.double()).abs().max()) <= tol
        assert float(((m1.double() - m0.double()) / m0.double()).abs().max()) < 1e-4
        monkeypatch.setenv('VDVC_NO_RHS_DIAG', '1')
        b2, f2, d2, m2, s2 = run(True)
        assert torch.equal(b2, b0) and torch.equal(d2, d0) and torch.equal(m2, m0)


@pytest.mark.parametrize('rd', ['float32', 'float64'])
@pytest.mark.parametrize('n', [1, 1000, 824875])
def test_gn_update_matches_axpby_and_dots(gpu, n, rd):
    """vdvc_gn_update equals vdvc_axpby(1, dU, 1, U) followed by the two
    vdvc_dot norms, bit for bit, and the norms match NumPy's float64 sums;
    vdvc_dot2 equals its two vdvc_dot products bit for bit."""
    import torch
    from voxeldvc_amd import _lib as L
    from voxeldvc_amd.engine import kernels_dvc as K
    print('gn_update', n, rd, flush=True)
    rng = np.random.default_rng(31)
    u0, du = rng.normal(size=n), rng.normal(scale=1e-2, size=n)
    U1, dU = L.to_dev_flat(u0, rd), L.to_dev_flat(du, rd)
    U2 = L.clone(U1)
    s2 = (L._d * 2)()
    L.call('vdvc_gn_update', L.ctx(), n, L.rtype_of(U1), L.ptr(dU), L.ptr(U1), s2)
    K.axpby(1.0, dU, 1.0, U2)
    assert torch.equal(U1, U2)
    assert s2[0] == K.dot(dU, dU) and s2[1] == K.dot(U2, U2)
    d2 = (L._d * 2)()
    L.call('vdvc_dot2', L.ctx(), n, L.rtype_of(U1), L.ptr(dU), L.ptr(U1), L.ptr(U2), d2)
    assert d2[0] == K.dot(dU, U1) and d2[1] == K.dot(dU, U2)
    uh, dh = L.to_host(U2).astype(np.float64), L.to_host(dU).astype(np.float64)
    assert abs(s2[0] - dh @ dh) <= 1e-12 * (dh @ dh) and abs(s2[1] - uh @ uh) <= 1e-12 * (uh @ uh)


@pytest.mark.parametrize('types', [('float32', 'float16'), ('float32', 'float32'),
                                   ('float64', 'float64'), ('float64', 'float32')])
@pytest.mark.parametrize('shape', [(64, 2, 1, 8), (32, 1, 2, 8), (3, 2, 2, 8), (5, 3, 4, 3),
                                   (70, 1, 2, 1), (1, 1, 1, 1), (9, 2, 3, 8)])
def test_level_image_matches_copy_and_stats(gpu, shape, types):
    """vdvc_level_image writes the pitched copy of vdvc_copy_volume bit for
    bit (padding zero) and the statistics of vdvc_masked_stats over the
    stored values above glt (same count; sums to 1e-13), and the mean / std
    match the oracle's masked_quadpoint_mean_std of the stored values.
    Widths 513 and 257 (eight-chunk kernel), 25 / 71 / 2 (four-chunk)."""
    import torch
    from voxeldvc_amd import _lib as L
    from voxeldvc_amd.engine import correlate_gpu as C
    from voxeldvc_amd.engine import kernels_dvc as Kd
    Nx, Ny, Nz, h = shape
    st, dt = types
    print('level_image', shape, types, flush=True)
    rng = np.random.default_rng(21)
    shp = Kd.level_shape(h, Nx, Ny, Nz)
    nv = shp[0] * shp[1] * shp[2]
    f = rng.uniform(0, 255, nv)
    src = L.to_dev_flat(f, st)
    ref = L.pitch_dense(src, shp, dt)
    got = L.fill(L.empty(ref.numel(), dt), 7.0)
    s3 = (L._d * 3)()
    L.call('vdvc_level_image', L.ctx(), Nx, Ny, Nz, h, L.ptr(src), L.type_code(st), L.ptr(got),
           L.type_code(dt), 30.0, s3)
    assert torch.equal(got, ref)
    if dt == 'float16':
        # a destination 2 bytes off 4-byte alignment takes the one-voxel-per-
        # lane kernel instead of the paired one: same values, same statistics
        buf = L.fill(L.empty(ref.numel() + 1, dt), 7.0)
        o3 = (L._d * 3)()
        L.call('vdvc_level_image', L.ctx(), Nx, Ny, Nz, h, L.ptr(src), L.type_code(st),
               L.ptr(buf[1:]), L.type_code(dt), 30.0, o3)
        assert torch.equal(buf[1:], ref)
        assert o3[0] == s3[0] and abs(o3[1] - s3[1]) <= 1e-13 * abs(s3[1])
    r3 = (L._d * 3)()
    L.call('vdvc_masked_stats', L.ctx(), Nx, Ny, Nz, h, L.ptr(ref), L.type_code(dt), None, 1,
           30.0, r3)
    assert s3[0] == r3[0]
    assert abs(s3[1] - r3[1]) <= 1e-13 * abs(r3[1]) and abs(s3[2] - r3[2]) <= 1e-13 * r3[2]
    stored = L.to_host(L.unpitch(ref, shp)).astype(np.float64)
    rmean, rstd = O.masked_quadpoint_mean_std(stored, stored > 30.0, h, Nx, Ny, Nz)
    mean, std = C._stats_mean_std(s3)
    assert abs(mean - rmean) < 1e-9 * abs(rmean) and abs(std - rstd) < 1e-9 * rstd


@pytest.mark.parametrize('form', ['f32', 'bf16', 'f64'])
@pytest.mark.parametrize('shape', [(3, 2, 2, 8), (70, 2, 1, 8), (5, 3, 4, 3)])
def test_residual_out_matches_rhs_residual(gpu, shape, form):
    """vdvc_residual_out forms the residual of vdvc_residual_rhs bit for bit
    (dense layout) from the same inputs, and vdvc_residual_rhs gives the same
    b, bit for bit, whether or not it stores the residual: the float32
    image + gradients, the mixed form (float16 image, bfloat16 gradients:
    the vectorised kernel at h = 8, the generic one at h = 3) and FP64."""
    import torch
    from voxeldvc_amd import _lib as L
    from voxeldvc_amd.engine import kernels_dvc as Kd
    Nx, Ny, Nz, h = shape
    print('residual_out', shape, form, flush=True)
    rng = np.random.default_rng(12)
    shp = Kd.level_shape(h, Nx, Ny, Nz)
    nv = shp[0] * shp[1] * shp[2]
    g, m, _ = rand_problem(13, Nx, Ny, Nz, h)
    f = rng.uniform(0, 255, nv).round()
    gv = rng.uniform(0, 255, nv)
    rd, fd, gd = {'f32': ('float32', 'float32', 'float32'),
                  'bf16': ('float32', 'float16', 'bfloat16'),
                  'f64': ('float64', 'float64', 'float64')}[form]
    rt = L.F64 if rd == 'float64' else L.F32
    gx, gy, gz = Kd.grad_planes(g, shp, gd)
    fimg = L.to_dev_level(f, shp, fd)
    gvals = L.to_dev_level(gv, shp, rd)
    bm = L.to_dev_level(m, shp, 'uint8')
    ndof = 3 * (Nx + 1) * (Ny + 1) * (Nz + 1)
    mean0, gmean, kscale = 97.25, 101.5, -1.0625
    args = (L.ctx(), Nx, Ny, Nz, h, L.ptr(gx), L.ptr(gy), L.ptr(gz), L.type_code(gd), L.ptr(fimg),
            L.type_code(fd), mean0, L.ptr(gvals), gmean, kscale, L.ptr(bm), L.ptr(bm))
    res_p = L.empty(fimg.numel(), rd)
    b1, b2 = L.empty(ndof, rd), L.empty(ndof, rd)
    L.call('vdvc_residual_rhs', *args, L.ptr(res_p), L.ptr(b1), None, None, None, None)
    L.call('vdvc_residual_rhs', *args, None, L.ptr(b2), None, None, None, None)
    assert torch.equal(b1, b2)
    res_d = L.empty(nv, rd)
    L.call('vdvc_residual_out', L.ctx(), Nx, Ny, Nz, h, L.ptr(fimg), L.type_code(fd), mean0,
           L.ptr(gvals), rt, gmean, kscale, L.ptr(bm), L.ptr(res_d))
    assert torch.equal(res_d, L.unpitch(res_p, shp))
    # the reference's formula on the host (same single-rounded steps)
    dt = np.float64 if rd == 'float64' else np.float32
    fh = f.astype(np.float16).astype(dt) if fd == 'float16' else f.astype(dt)
    ref = ((gv.astype(dt) - dt(gmean)) * dt(kscale) + (fh - dt(mean0))) * m
    assert np.array_equal(L.to_host(res_d), ref.astype(dt))


@pytest.mark.parametrize('shape', [(2, 2, 2, 1), (3, 4, 2, 1), (5, 5, 5, 1)])
def test_matvec_K_ref(gpu, shape):
    from voxeldvc_amd.engine import kernels_dvc as K
    Nx, Ny, Nz, _ = shape
    print('matvec_K', shape, flush=True)
    Ndof = 3 * (Nx + 1) * (Ny + 1) * (Nz + 1)
    v = np.random.default_rng(5).normal(size=Ndof).astype(np.float32)
    for Kr in (O.build_K_ref_laplacian(), O.build_K_ref(0.3)):
        ref = O.matvec_K_ref(v.astype(np.float64), Kr, 8, Nx, Ny, Nz)
        assert rel(K.matvec_K_ref(v, Kr, 8, Nx, Ny, Nz), ref) < 1e-5
        dref = O.build_diagonal_K_ref(Kr, 8, Nx, Ny, Nz)
        assert rel(K.build_diagonal_K_ref(Kr, 8, Nx, Ny, Nz, Ndof), dref) < 1e-6
        E = np.random.default_rng(6).uniform(1, 3, Nx * Ny * Nz).astype(np.float32)
        ref = O.matvec_K_ref(v.astype(np.float64), Kr, E.astype(np.float64), Nx, Ny, Nz)
        assert rel(K.matvec_K_ref(v, Kr, E, Nx, Ny, Nz), ref) < 1e-5


def test_equilibrium_gap(gpu):
    from voxeldvc_amd.engine import kernels_dvc as K
    Nx, Ny, Nz = 3, 4, 5
    Kel = O.build_K_ref(0.3)
    Ndof = 3 * (Nx + 1) * (Ny + 1) * (Nz + 1)
    im = O.build_interior_dof_mask(Nx, Ny, Nz)
    v = np.random.default_rng(7).normal(size=Ndof).astype(np.float32)
    ref = O.matvec_equilibrium_gap(v.astype(np.float64), Kel, im, 8, Nx, Ny, Nz)
    assert rel(K.matvec_equilibrium_gap(v, Kel, im, 8, Nx, Ny, Nz), ref) < 1e-5
    # Hutchinson diagonal against the reference's own values (golden)
    z = golden('eqgap_diag3')
    d = K.build_diagonal_equilibrium_gap(Kel, O.build_interior_dof_mask(3, 3, 3), 8, 3, 3, 3,
                                         192)
    assert rel(d, z['diag']) < 1e-5


def test_masked_stats_and_setup(gpu):
    from voxeldvc_amd.engine import correlate_gpu as C
    Nx, Ny, Nz, h = 3, 2, 4, 4
    rng = np.random.default_rng(8)
    f = rng.integers(0, 256, size=(Nx * h + 1, Ny * h + 1, Nz * h + 1)).astype(np.float32)
    f_vox, grad = C.image_grad_and_values(f)
    rf, rg = O.image_grad_and_values(f)
    assert np.array_equal(f_vox, rf)
    assert np.array_equal(grad, rg)          # bit-exact float32 np.gradient
    m = rng.random(f_vox.size) < 0.7
    mean, std = C.masked_quadpoint_mean_std(f_vox, m, h, Nx, Ny, Nz)
    rmean, rstd = O.masked_quadpoint_mean_std(f_vox, m, h, Nx, Ny, Nz)
    assert abs(mean - rmean) < 1e-9 * abs(rmean) and abs(std - rstd) < 1e-9 * rstd
    er = C.erode_mask0_vox(m, Nx * h + 1, Ny * h + 1, Nz * h + 1)
    assert np.array_equal(er, O.erode_mask_6(m, Nx * h + 1, Ny * h + 1, Nz * h + 1))


@pytest.mark.parametrize('go', [(0, 0, 0), (2, -1, 1)])
def test_sample_deformed_field(gpu, go):
    from voxeldvc_amd.engine import correlate_gpu as C
    Nx, Ny, Nz, h = 3, 2, 2, 8
    rng = np.random.default_rng(9)
    gimg = rng.uniform(0, 255, size=(Nx * h + 3, Ny * h + 1, Nz * h + 2)).astype(np.float32)
    Ndof = 3 * (Nx + 1) * (Ny + 1) * (Nz + 1)
    U = rng.normal(scale=2.0, size=Ndof).astype(np.float32)
    m0 = rng.random((Nx * h + 1) * (Ny * h + 1) * (Nz * h + 1)) < 0.9
    rv, rm = O.sample_deformed_field(U, gimg, h, Nx, Ny, Nz, go, m0, glt=30.0)
    gv, gm = C.sample_deformed_field(U, gimg, None, h, Nx, Ny, Nz, g_origin=go, mask0_vox=m0,
                                     glt=30.0)
    # u at a voxel is formed by a different float32 summation order; allow a
    # few in-bounds / threshold flips at the boundary of the box
    assert (gm != rm).mean() < 2e-3
    ok = gm == rm
    assert np.abs(gv[ok] - rv[ok]).max() < 2e-3


def test_stride_and_phase_corr(gpu):
    from voxeldvc_amd.engine import correlate_gpu as C
    rng = np.random.default_rng(42)
    f = rng.uniform(0, 1, size=(32, 32, 32))
    g = np.roll(f, (3, -2, 1), axis=(0, 1, 2))
    assert np.allclose(C.fft_rigid_shift(f, g), (3, -2, 1))
    # non-cubic, unequal shapes (common corner crop)
    f2 = rng.uniform(0, 1, size=(20, 18, 15))
    g2 = np.roll(f2, (-4, 5, 2), axis=(0, 1, 2))
    g2 = np.pad(g2, ((0, 3), (0, 0), (0, 2)))
    assert np.allclose(C.fft_rigid_shift(f2, g2), O.fft_rigid_shift(f2, g2))
    # longer axes: the LDS-staged passes with 9, 3 and 4 outputs per thread
    # and line counts that are not multiples of the 16-line groups (129 x 37
    # x 50), and an axis over the staged pass's 227-sample limit (240, the
    # per-output pass) beside short ones
    for shp, d in (((129, 37, 50), (7, -3, 11)), ((240, 9, 7), (-17, 2, 3))):
        print('phase correlation', shp, d, flush=True)
        f3 = rng.uniform(0, 1, size=shp)
        g3 = np.roll(f3, d, axis=(0, 1, 2))
        got = C.fft_rigid_shift(f3, g3)
        assert np.allclose(got, O.fft_rigid_shift(f3, g3))
        assert np.allclose(got, d)


def test_affine_cost(gpu):
    from voxeldvc_amd.engine import correlate_gpu as C
    z = golden('multi33_dyn')
    f = z['f'][::2, ::2, ::2]
    g = z['g'][::2, ::2, ::2]
    A = O.build_affine_matrix([0.02, -0.01, 0.03, 1.01, 0.99, 1.0])
    t = np.array([0.5, -0.3, 0.2])
    c = np.array([(n - 1) / 2.0 for n in f.shape])
    ref = O.affine_cost(f, g, A, t, c)
    from voxeldvc_amd import _lib as L
    fd, fs = L.to_dev_volume(f)
    gd, gs = L.to_dev_volume(g)
    out = (L._d * 6)()
    L.call('vdvc_affine_sums', L.ctx(), L.ptr(fd), *fs, L.ptr(gd), *gs, L.dbl(*A.ravel()),
           L.dbl(*t), L.dbl(*c), out)
    assert abs(C._zncc_cost(list(out)) - ref) < 1e-6


@pytest.mark.parametrize('shape', [(4, 3, 3, 4), (9, 7, 5, 2)])
@pytest.mark.parametrize('aligned', [True, False])
def test_pcg_matches_oracle(gpu, shape, aligned):
    """Device-resident PCG on (H + ll L) against the oracle's PCG.  With the
    solution and M_inv vectors 16-byte aligned the vector updates take the
    vectorised kernels (16-byte chunks; (9, 7, 5, 2) has Ndof = 1440, a
    partial last chunk per thread and several blocks); at an offset of one
    float they take the grid-stride kernels."""
    from voxeldvc_amd import _lib as L
    Nx, Ny, Nz, h = shape
    print('pcg', shape, 'aligned' if aligned else 'offset', flush=True)
    g, m, b = rand_problem(10, Nx, Ny, Nz, h)
    Klap = O.build_K_ref_laplacian()
    ll = 0.7
    A = lambda x: (O.matvec_H(x, g.astype(np.float64), m, h, Nx, Ny, Nz)
                   + ll * O.matvec_K_ref(x, Klap, h, Nx, Ny, Nz))
    d = O.build_diagonal_H(g.astype(np.float64), m, h, Nx, Ny, Nz) + ll * O.build_diagonal_K_ref(
        Klap, h, Nx, Ny, Nz)
    Minv = 1.0 / (d + 1e-5 * d.min())
    xr, itr = O.pcg(A, b.astype(np.float64), Minv, max_iter=500, tol=1e-6)
    from voxeldvc_amd.engine import kernels_dvc as Kd
    shp = Kd.level_shape(h, Nx, Ny, Nz)
    gx, gy, gz = Kd.grad_planes(g, shp)
    md = Kd._mask_u8(m, shp)
    diag = L.empty(b.size)
    off = 0 if aligned else 1
    minv = L.empty(b.size + off)[off:]
    Kp = np.ascontiguousarray(Klap)
    L.call('vdvc_jacobi', L.ctx(), Nx, Ny, Nz, h, L.ptr(gx), L.ptr(gy), L.ptr(gz), L.F32,
           L.ptr(md), L.REG_K, Kp.ctypes.data_as(L._dp), None, float(h), 0.0, ll, None,
           L.ptr(diag), L.ptr(minv))
    assert rel(L.to_host(minv), Minv) < 1e-5
    bd = L.to_dev_flat(b)
    x = L.empty(b.size + off)[off:]
    it = (L._i * 1)()
    rr = (L._d * 1)()
    L.call('vdvc_pcg', L.ctx(), Nx, Ny, Nz, h, L.ptr(gx), L.ptr(gy), L.ptr(gz), L.F32, L.ptr(md),
           L.REG_K, Kp.ctypes.data_as(L._dp), None, float(h), 0.0, ll, L.ptr(bd), L.ptr(minv),
           L.ptr(x), 1e-6, 500, 7, it, rr)
    assert rr[0] < 1e-6
    assert abs(it[0] - itr) <= 2
    assert rel(L.to_host(x), xr) < 1e-4


@pytest.mark.parametrize('fuse', ['voxel', 'node'])
@pytest.mark.parametrize('case', [((9, 7, 5, 2), 'f32', '0'), ((64, 2, 3, 2), 'f32', '0'),
                                  ((6, 5, 4, 8), 'bf16', '0'),
                                  ((65, 2, 3, 8), 'bf16', '0'), ((3, 4, 3, 8), 'bf16', '2')])
def test_apply_A_laplacian_term_in_voxel_kernel(gpu, monkeypatch, case, fuse):
    """H + ll L with the Laplacian term added by the voxel kernel element by
    element (float32 lanes of <= 2 voxels, 2-byte gradients) against the
    oracle, and the same operator and PCG with the term in the node pass
    (VDVC_NO_REG_FUSE=1).  (64, 2, 3, 2) takes float32 lanes of 2 voxels (the
    coarse pyramid level's kernel); (65, 2, 3, 8) takes the 8-voxel lanes of two
    strips; (3, 4, 3, 8) with VDVC_NYS=2 splits each element row's rows
    between two wavefronts (the term is added by the first part only); the
    bf16 grids of 6 and 3 elements narrow the lanes (several lanes per
    element: the term is added by the first lane of each group)."""
    import torch
    from voxeldvc_amd import _lib as L
    (Nx, Ny, Nz, h), gt, nys = case
    print('apply_A laplacian term', case, fuse, flush=True)
    if nys != '0':
        monkeypatch.setenv('VDVC_NYS', nys)
    if fuse == 'node':
        monkeypatch.setenv('VDVC_NO_REG_FUSE', '1')
    g, m, v = rand_problem(21, Nx, Ny, Nz, h)
    shp = (Nx * h + 1, Ny * h + 1, Nz * h + 1)
    if gt == 'bf16':
        gq = torch.from_numpy(g * m[:, None]).to(torch.bfloat16)    # mask folded in
        g64 = gq.to(torch.float64).numpy()
        m64 = np.ones_like(m)
        planes = [L.to_dev_level(gq[:, c].contiguous(), shp, 'bfloat16') for c in range(3)]
        md, code = None, L.BF16
    else:
        g64, m64 = g.astype(np.float64), m
        planes = [L.to_dev_level(np.ascontiguousarray(g[:, c]), shp, 'float32') for c in range(3)]
        md, code = L.to_dev_level(m, shp, 'uint8'), L.F32
    Klap = O.build_K_ref_laplacian()
    ll = 0.61
    A = lambda x: (O.matvec_H(x, g64, m64, h, Nx, Ny, Nz)
                   + ll * O.matvec_K_ref(x, Klap, h, Nx, Ny, Nz))
    Kp = np.ascontiguousarray(Klap)
    kp = Kp.ctypes.data_as(L._dp)
    vd = L.to_dev_flat(v)
    out = L.empty(v.size)
    geo = (L.ctx(), Nx, Ny, Nz, h, *[L.ptr(q) for q in planes], code, L.ptr(md), L.REG_K, kp,
           None, float(h), 0.0, ll)
    L.call('vdvc_apply_A', *geo, L.ptr(vd), L.ptr(out))
    assert rel(L.to_host(out), A(v.astype(np.float64))) < 2e-5
    d = (O.build_diagonal_H(g64, m64, h, Nx, Ny, Nz)
         + ll * O.build_diagonal_K_ref(Klap, h, Nx, Ny, Nz))
    Minv = 1.0 / (d + 1e-5 * d.min())
    xr, itr = O.pcg(A, v.astype(np.float64), Minv, max_iter=500, tol=1e-6)
    diag, minv, x = L.empty(v.size), L.empty(v.size), L.empty(v.size)
    L.call('vdvc_jacobi', *geo, None, L.ptr(diag), L.ptr(minv))
    it = (L._i * 1)()
    rr = (L._d * 1)()
    L.call('vdvc_pcg', *geo, L.ptr(vd), L.ptr(minv), L.ptr(x), 1e-6, 500, 5, it, rr)
    assert rr[0] < 1e-6
    assert abs(it[0] - itr) <= 2
    assert rel(L.to_host(x), xr) < 1e-4


def _eqgap_problem():
    Nx, Ny, Nz, h = 3, 3, 2, 4
    g, m, v = rand_problem(11, Nx, Ny, Nz, h)
    return Nx, Ny, Nz, h, g, m, v


@pytest.mark.parametrize('kind', ['laplacian', 'equilibrium_gap'])
def test_apply_A_and_jacobi(gpu, kind):
    """The PCG operator H + ll R and its Jacobi preconditioner for both
    regularizers against the oracle (correlate_gpu.py:974-988, 1042-1047)."""
    from voxeldvc_amd import _lib as L
    from voxeldvc_amd.engine import kernels_dvc as Kd
    Nx, Ny, Nz, h, g, m, v = _eqgap_problem()
    Ndof = v.size
    Klap = O.build_K_ref_laplacian()
    Kel = O.build_K_ref(0.3)
    ll, eps = 0.37, 1e-3
    g64 = g.astype(np.float64)
    im = O.build_interior_dof_mask(Nx, Ny, Nz)
    if kind == 'laplacian':
        R = lambda x: O.matvec_K_ref(x, Klap, h, Nx, Ny, Nz)
        dR = O.build_diagonal_K_ref(Klap, h, Nx, Ny, Nz)
        rk, K1, K2, extra = L.REG_K, Klap, None, None
    else:
        R = lambda x: (O.matvec_equilibrium_gap(x, Kel, im, h, Nx, Ny, Nz)
                       + eps * O.matvec_K_ref(x, Klap, h, Nx, Ny, Nz))
        hutch = O.build_diagonal_equilibrium_gap(Kel, im, h, Nx, Ny, Nz)
        dR = hutch + eps * O.build_diagonal_K_ref(Klap, h, Nx, Ny, Nz)
        rk, K1, K2 = L.REG_EQGAP, Kel, Klap
        extra = L.to_dev_flat(Kd.build_diagonal_equilibrium_gap(Kel, im, h, Nx, Ny, Nz, Ndof))
        assert rel(L.to_host(extra), hutch) < 1e-5
    ref = O.matvec_H(v.astype(np.float64), g64, m, h, Nx, Ny, Nz) + ll * R(v.astype(np.float64))
    shp = Kd.level_shape(h, Nx, Ny, Nz)
    gx, gy, gz = Kd.grad_planes(g, shp)
    md = Kd._mask_u8(m, shp)
    K1p = np.ascontiguousarray(K1)
    K2p = None if K2 is None else np.ascontiguousarray(K2)
    kp = lambda a: None if a is None else a.ctypes.data_as(L._dp)
    vd = L.to_dev_flat(v)
    out = L.empty(Ndof)
    L.call('vdvc_apply_A', L.ctx(), Nx, Ny, Nz, h, L.ptr(gx), L.ptr(gy), L.ptr(gz), L.F32,
           L.ptr(md), rk, kp(K1p), kp(K2p), float(h), eps, ll, L.ptr(vd), L.ptr(out))
    assert rel(L.to_host(out), ref) < 2e-5
    d = O.build_diagonal_H(g64, m, h, Nx, Ny, Nz) + ll * dR
    Minv = 1.0 / (d + 1e-5 * d.min())
    diag = L.empty(Ndof)
    minv = L.empty(Ndof)
    L.call('vdvc_jacobi', L.ctx(), Nx, Ny, Nz, h, L.ptr(gx), L.ptr(gy), L.ptr(gz), L.F32,
           L.ptr(md), rk, kp(K1p), kp(K2p), float(h), eps, ll, L.ptr(extra), L.ptr(diag),
           L.ptr(minv))
    assert rel(L.to_host(diag), d) < 2e-5
    assert rel(L.to_host(minv), Minv) < 2e-5
    # PCG on this operator
    A = lambda x: O.matvec_H(x, g64, m, h, Nx, Ny, Nz) + ll * R(x)
    xr, itr = O.pcg(A, v.astype(np.float64), Minv, max_iter=400, tol=1e-6)
    x = L.empty(Ndof)
    it = (L._i * 1)()
    rr = (L._d * 1)()
    L.call('vdvc_pcg', L.ctx(), Nx, Ny, Nz, h, L.ptr(gx), L.ptr(gy), L.ptr(gz), L.F32, L.ptr(md),
           rk, kp(K1p), kp(K2p), float(h), eps, ll, L.ptr(vd), L.ptr(minv), L.ptr(x), 1e-6, 400, 5,
           it, rr)
    print(kind, 'pcg iters gpu', it[0], 'oracle', itr, flush=True)
    assert rr[0] < 1e-6
    assert rel(L.to_host(x), xr) < 1e-3


def test_correlate_golden_17(gpu):
    """End-to-end single-level solve on the reference test problem against the
    reference's own output (1e-3 voxel max-abs at FP32)."""
    from voxeldvc_amd.engine import correlate_gpu as C
    z = golden('corr17_f32')
    K = O.build_K_ref_laplacian()
    U, res, dU = C.correlate_gpu(z['f'], z['g'], 2, 2, 2, 8, np, K_ref_laplacian=K, l0=36.0,
                                 maxiter=20, eps=1e-3, disp=False)
    assert np.abs(U - z['U']).max() < 1e-3
    # residuals agree except where an in-bounds test flips at the box edge
    both = (res != 0) & (z['res'] != 0)
    assert ((res != 0) == (z['res'] != 0)).mean() > 0.99
    assert np.abs(res[both] - z['res'][both]).max() < 2e-3


def test_multiscale_golden_33(gpu):
    """Pyramid + FFT prealign + dynamic mask + freeze on 33^3 speckle against the
    reference's own output."""
    from voxeldvc_amd.engine import correlate_gpu as C
    z = golden('multi33_dyn')
    zr = golden('multi33_dyn_f32in')
    K = O.build_K_ref_laplacian()
    U, res, dU = C.multiscale_correlate_gpu(z['f'], z['g'], 4, 4, 4, 8, np, scales=(2, 1, 0),
                                            fft_prealign=True, K_ref_laplacian=K, l0=36.0,
                                            glt=0.0, dynamic_mask=True, freeze_mask_after=4,
                                            disp=False)
    assert np.abs(U - zr['U']).max() < 1e-3


def test_eqgap_golden_33(gpu):
    from voxeldvc_amd.engine import correlate_gpu as C
    z = golden('multi33_dyn')
    zr = golden('eqgap33')
    log, logr = [], []
    kw = dict(reg_type='equilibrium_gap', K_ref_elastic=O.build_K_ref(0.3),
              K_ref_laplacian=O.build_K_ref_laplacian(), l0=36.0, eps_reg=1e-3, maxiter=10)
    U, res, dU = C.correlate_gpu(z['f'], z['g'], 4, 4, 4, 8, np, disp=False, log=log, **kw)
    O.correlate(z['f'], z['g'], 4, 4, 4, 8, log=logr, **kw)
    for a, b in zip(log, logr):
        print('gpu', a, flush=True)
        print('ora', b, flush=True)
    assert np.abs(U - zr['U']).max() < 1e-3


@pytest.mark.parametrize('knobs', [{'VDVC_NYS': '2'}, {'VDVC_EPL': '2'},
                                   {'VDVC_EPL': '2', 'VDVC_NYS': '2'}])
def test_diag_and_pcg_with_row_split(gpu, monkeypatch, knobs):
    """The diagonal and the PCG operator read the node partials through the
    same gather: with the row split and with paired-element lanes they give
    the oracle's diag(H) and H v + ll L v."""
    from voxeldvc_amd import _lib as L
    from voxeldvc_amd.engine import kernels_dvc as Kd
    for k, val in knobs.items():
        monkeypatch.setenv(k, val)
    Nx, Ny, Nz, h = (8, 3, 4, 4) if 'VDVC_EPL' in knobs else (7, 3, 4, 8)
    print('diag / apply_A', knobs, (Nx, Ny, Nz, h), flush=True)
    g, m, v = rand_problem(13, Nx, Ny, Nz, h)
    shp = Kd.level_shape(h, Nx, Ny, Nz)
    gx, gy, gz = Kd.grad_planes(g, shp)
    md = Kd._mask_u8(m, shp)
    g64 = g.astype(np.float64)
    d = L.empty(v.size)
    L.call('vdvc_diag_H', L.ctx(), Nx, Ny, Nz, h, gx, gy, gz, L.F32, md, d)
    assert rel(L.to_host(d), O.build_diagonal_H(g64, m, h, Nx, Ny, Nz)) < 2e-5
    Kp = np.ascontiguousarray(O.build_K_ref_laplacian())
    out = L.empty(v.size)
    ll = 0.4
    L.call('vdvc_apply_A', L.ctx(), Nx, Ny, Nz, h, gx, gy, gz, L.F32, md, L.REG_K,
           Kp.ctypes.data_as(L._dp), None, float(h), 0.0, ll, L.to_dev_flat(v), out)
    ref = (O.matvec_H(v.astype(np.float64), g64, m, h, Nx, Ny, Nz)
           + ll * O.matvec_K_ref(v.astype(np.float64), Kp, h, Nx, Ny, Nz))
    assert rel(L.to_host(out), ref) < 2e-5
