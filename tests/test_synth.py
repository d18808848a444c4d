"""The synthetic-workload generators: the GPU-form warp (run here on the
CPU device) equals the NumPy form on the same field."""
import numpy as np
import pytest

from voxeldvc_amd import synth


def test_fourier_warp_torch_matches_numpy():
    torch = pytest.importorskip('torch')
    shape = (13, 11, 9)
    f = synth.speckle_numpy(shape, 60, seed=4)
    modes = synth.fourier_params(5, (8, 20), 2.5, seed=5, extent=shape)
    kv, ph, av = modes

    def ufield(X, Y, Z):
        s = np.sin(X[..., None] * kv[:, 0] + Y[..., None] * kv[:, 1] + Z[..., None] * kv[:, 2]
                   + ph)
        return tuple((s * av[:, c]).sum(-1) for c in range(3))

    ref = synth.warp_backward(f, ufield, 0.0, 0)
    fx = torch.from_numpy(np.ravel(f, order='F').astype(np.float32))
    got = synth.warp_backward_torch(fx, shape, modes, 0.0, 0, quantize=False, slab=4,
                                    device='cpu').numpy()
    got = got.reshape(shape[2], shape[1], shape[0]).transpose(2, 1, 0)
    # f enters as float32, everything after in float64
    assert np.abs(got - ref).max() < 1e-4 * np.abs(ref).max()


def test_fourier_peak_amplitude():
    modes = synth.fourier_params(16, (64, 256), 4.0, seed=3, extent=(129, 129, 129))
    U = synth.fourier_u_nodes(modes, 16, 16, 16, 8)
    mag = np.sqrt(U[0::3] ** 2 + U[1::3] ** 2 + U[2::3] ** 2)
    assert 2.0 < mag.max() <= 4.0 + 1e-9
