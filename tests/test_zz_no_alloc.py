"""vdvc_matvec_H_noalloc (the no-allocation variant of vdvc_matvec_H) through
the C ABI: bit-identical to vdvc_matvec_H on two calls in a row, the second
on a smaller mesh, for folded bfloat16, float32 + mask and float64 storage;
meshes whose node partials exceed its static scratch are rejected before
any device work."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _level(L, Nx, Ny, Nz, h, gtype, seed):
    """Pitched gradient planes (+ mask) of a random 8-bit image on a mesh."""
    import torch
    shape = (Nx * h + 1, Ny * h + 1, Nz * h + 1)
    rng = np.random.default_rng(seed)
    f = rng.integers(0, 256, shape).astype(np.float32)
    f[rng.random(shape) < 0.05] = 0.0
    wd = 'float64' if gtype == L.F64 else 'float32'
    fp, _ = L.to_dev_volume(f, wd, pitched=True)
    gname = {L.F32: 'float32', L.BF16: 'bfloat16', L.F64: 'float64'}[gtype]
    G = [L.empty(fp.numel(), gname) for _ in range(3)]
    m = L.empty(fp.numel(), 'uint8')
    L.call('vdvc_image_setup', L.ctx(), Nx, Ny, Nz, h, fp, L.F64 if wd == 'float64' else L.F32,
           0.0, G[0], G[1], G[2], gtype, m)
    mask = None if gtype == L.BF16 else m
    v = torch.from_numpy(rng.standard_normal(3 * (Nx + 1) * (Ny + 1) * (Nz + 1))).cuda()
    v = v.to(torch.float64 if wd == 'float64' else torch.float32)
    return G, mask, v


@pytest.mark.parametrize('gname', ['bf16', 'f32', 'f64'])
def test_noalloc_matvec_matches_matvec(gpu, gname):
    import torch
    from voxeldvc_amd import _lib as L
    gtype = {'bf16': L.BF16, 'f32': L.F32, 'f64': L.F64}[gname]
    c = L.ctx()
    # two calls in a row, the second on a smaller mesh (the scratch keeps the
    # first call's contents)
    for (Nx, Ny, Nz, h) in ((24, 20, 18, 8), (5, 3, 4, 4)):
        print('noalloc', gname, (Nx, Ny, Nz, h), flush=True)
        G, mask, v = _level(L, Nx, Ny, Nz, h, gtype, Nx)
        ref = torch.empty_like(v)
        got = torch.full_like(v, float('nan'))
        L.call('vdvc_matvec_H', c, Nx, Ny, Nz, h, G[0], G[1], G[2], gtype, mask, v, ref)
        L.call('vdvc_matvec_H_noalloc', c, Nx, Ny, Nz, h, G[0], G[1], G[2], gtype, mask, v, got)
        torch.cuda.synchronize()
        assert torch.equal(got, ref)
        assert float(ref.abs().max()) > 0


def test_noalloc_rejects_meshes_beyond_its_scratch(gpu):
    from voxeldvc_amd import _lib as L
    import torch
    # 2049^3 voxels at h = 4 needs far more than 64 MiB of node partials; the
    # size check runs before anything touches the (absent) buffers
    d = torch.zeros(16, device='cuda')
    with pytest.raises(ValueError, match='no-allocation scratch'):
        L.call('vdvc_matvec_H_noalloc', L.ctx(), 300, 300, 300, 4, d, d, d, L.BF16, None, d, d)
