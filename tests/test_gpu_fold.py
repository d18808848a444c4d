"""vdvc_fold_grad (bfloat16 gradient planes times a 0/1 mask, the folded
storage of the mixed-precision matvec) against NumPy, bit for bit: the
eight-voxel vector path (length a multiple of 8, aligned pointers), a
length that is not, and pointers offset by one element (scalar path)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('n,off', [(8 * 4099, 0), (8 * 4099 + 5, 0), (8 * 512, 1), (37, 0)])
def test_fold_grad(gpu, n, off):
    from voxeldvc_amd import _lib as L
    torch = gpu
    print('fold', n, off, flush=True)
    rng = np.random.default_rng(n + off)
    raw = [rng.integers(-32768, 32768, n + 16, dtype=np.int16) for _ in range(3)]
    mask = (rng.random(n + 16) < 0.6).astype(np.uint8) * rng.integers(1, 256, n + 16).astype(np.uint8)
    gd = [torch.from_numpy(r).cuda() for r in raw]
    md = torch.from_numpy(mask).cuda()
    out = [torch.full((n + 16,), 77, dtype=torch.int16, device='cuda') for _ in range(3)]
    ptr = lambda t: L._vp(t.data_ptr() + off * t.element_size())
    L.call('vdvc_fold_grad', L.ctx(), n, ptr(gd[0]), ptr(gd[1]), ptr(gd[2]), ptr(md),
           ptr(out[0]), ptr(out[1]), ptr(out[2]))
    for c in range(3):
        got = out[c].cpu().numpy()
        want = np.where(mask[off:off + n] != 0, raw[c][off:off + n], 0).astype(np.int16)
        np.testing.assert_array_equal(got[off:off + n], want)
        assert (got[:off] == 77).all() and (got[off + n:] == 77).all()
