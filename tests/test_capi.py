"""The C-ABI library: it loads, exports every symbol include/vdvc.h declares,
and rejects bad arguments before touching a device (CPU-only checks)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), '..'))


def declared():
    src = open(os.path.join(ROOT, 'include', 'vdvc.h')).read()
    src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
    return sorted(set(re.findall(r'\b(vdvc_[A-Za-z0-9_]+)\s*\(', src)))


@pytest.fixture(scope='module')
def lib():
    from voxeldvc_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip('libvdvc.so not built (run python -m voxeldvc_amd.build)')
    return _lib.load()


def test_exports_every_declared_symbol(lib):
    names = declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    from voxeldvc_amd import _lib
    assert set(_lib.SIGNATURES) | {'vdvc_last_error'} == set(names)


def test_abi_version(lib):
    from voxeldvc_amd import _lib
    assert lib.vdvc_abi_version() == _lib.ABI_VERSION == 12


def test_voxel_pitch_matches_host_layout(lib):
    """The library's row pitch and the host-side layout helper agree, and the
    pitch keeps every vectorised lane load naturally aligned (a multiple of 16
    voxels) with less than 16 voxels of padding."""
    from voxeldvc_amd import _lib
    for nvx in (1, 2, 7, 8, 9, 17, 33, 65, 129, 257, 513, 1025, 2049):
        p = lib.vdvc_voxel_pitch(nvx)
        assert p == _lib.pitch_of(nvx)
        assert p % 16 == 0 and nvx <= p < nvx + 16


def test_argument_errors_without_device(lib):
    from voxeldvc_amd import _lib as L
    rc = lib.vdvc_matvec_H(None, 2, 2, 2, 8, None, None, None, 0, None, None, None)
    assert rc == -1 and b'ctx' in lib.vdvc_last_error()
    with pytest.raises(ValueError):
        L.check(rc)
    assert lib.vdvc_pcg(None, 1, 1, 1, 1, None, None, None, 0, None, 0, None, None, 1.0, 0.0,
                        0.0, None, None, None, 1e-6, 10, 1, None, None) == -1
