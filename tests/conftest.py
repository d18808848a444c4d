import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), '..'))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a ROCm GPU and the built libvdvc.so')


@pytest.fixture(scope='session')
def gpu():
    """Skip unless a GPU is visible; load the library (fails loudly if it is
    missing on a GPU machine)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    from voxeldvc_amd import _lib
    _lib.load()
    return torch


def golden(name):
    return np.load(os.path.join(GOLDEN, name + '.npz'))


def problem17():
    """The reference test problem (tests/test_dvc_gpu_multiscale.py:24-40)."""
    from scipy.ndimage import shift
    n = 17
    x = np.arange(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing='ij')
    f = (np.sin(2 * np.pi * X / n) + np.sin(2 * np.pi * Y / n) + np.sin(2 * np.pi * Z / n)
         + 0.5 * np.cos(2 * np.pi * (X + Y) / n))
    g = shift(f, (1.0, 0, 0), order=3, mode='nearest')
    return f, g, 2, 2, 2, 8


def rand_problem(seed, Nx, Ny, Nz, h, mask_frac=0.8):
    rng = np.random.default_rng(seed)
    Nvox = (Nx * h + 1) * (Ny * h + 1) * (Nz * h + 1)
    Ndof = 3 * (Nx + 1) * (Ny + 1) * (Nz + 1)
    g = rng.normal(size=(Nvox, 3)).astype(np.float32)
    m = rng.random(Nvox) < mask_frac
    v = rng.normal(size=Ndof).astype(np.float32)
    return g, m, v
