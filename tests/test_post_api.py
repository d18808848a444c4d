"""Argument handling of the post-solve drop-ins that needs no device: the
checks run before any library call."""
import numpy as np
import pytest


def test_image_types():
    from voxeldvc_amd import _lib as L
    for dt, code in ((np.float32, L.F32), (np.float64, L.F64), (np.uint8, L.U8),
                     (np.uint16, L.U16), (np.int16, L.I16), (np.bool_, L.BOOL)):
        assert L.image_type(np.zeros(2, dt))[0] == code
    for dt in (np.float16, np.int32, np.complex64):
        with pytest.raises(ValueError):
            L.image_type(np.zeros(2, dt))


def test_push_rejects_unsupported_order_mode_and_shape():
    from voxeldvc_amd.engine import write_output as W
    f = np.zeros((9, 9, 9), np.float32)
    U = np.zeros(3 * 8, np.float32)
    with pytest.raises(ValueError):
        W.push_field_forward(f, U, 1, 1, 1, 8, order=3)
    with pytest.raises(ValueError):
        W.push_field_forward(f, U, 1, 1, 1, 8, mode='reflect')
    with pytest.raises(ValueError):
        W.push_field_forward(f, U, 1, 1, 2, 8)


def test_mesh_dims():
    from voxeldvc_amd.engine import write_output as W
    assert W.mesh_dims((513, 257, 100), 8) == (64, 32, 12)
