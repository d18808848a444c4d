"""ctypes binding of libvdvc.so (include/vdvc.h) and device plumbing.

The library is the only compute path: if it is missing or no GPU is visible,
every entry point raises.  torch is used only to own device memory and to
provide the stream handle the library enqueues on; layout changes, element
conversions, copies and fills of the solve's arrays are library kernels too.
"""
import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libvdvc.so')

F32, BF16, F16, F64 = 0, 1, 2, 3
U8, U16, I16, BOOL = 4, 5, 6, 7
ABI_VERSION = 12
REG_NONE, REG_K, REG_EQGAP = 0, 1, 2

_vp = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float
_d = ctypes.c_double
_i64 = ctypes.c_int64
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i64p = ctypes.POINTER(ctypes.c_int64)

# name -> argument types (all return int status)
SIGNATURES = {
    'vdvc_abi_version': [],
    'vdvc_voxel_pitch': [_i],
    'vdvc_ctx_create': [_i, ctypes.c_size_t, ctypes.POINTER(_vp)],
    'vdvc_ctx_destroy': [_vp],
    'vdvc_synchronize': [_vp],
    'vdvc_timing_begin': [_vp],
    'vdvc_timing_end': [_vp, _dp, _ip],
    'vdvc_stream_read': [_vp, _i64, _vp, _i, ctypes.c_uint32, _vp],
    'vdvc_image_setup': [_vp, _i, _i, _i, _i, _vp, _i, _d, _vp, _vp, _vp, _i, _vp],
    'vdvc_masked_stats': [_vp, _i, _i, _i, _i, _vp, _i, _vp, _i, _d, _dp],
    'vdvc_level_image': [_vp, _i, _i, _i, _i, _vp, _i, _vp, _i, _d, _dp],
    'vdvc_sample': [_vp, _i, _i, _i, _i, _vp, _vp, _i, _i, _i, _i, _d, _d, _d, _d, _vp, _vp,
                    _vp, _vp, _dp, _i],
    'vdvc_residual_rhs': [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _i, _vp, _i, _d, _vp, _d, _d, _vp,
                          _vp, _vp, _vp, _vp, _vp, _vp, _dp],
    'vdvc_residual_rhs_jacobi': [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _i, _vp, _i, _d, _vp, _d,
                                 _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _dp, _i, _dp, _dp, _d,
                                 _d, _d, _vp, _vp, _vp],
    'vdvc_residual_out': [_vp, _i, _i, _i, _i, _vp, _i, _d, _vp, _i, _d, _d, _vp, _vp],
    'vdvc_matvec_H': [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _i, _vp, _vp, _vp],
    'vdvc_matvec_H_noalloc': [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _i, _vp, _vp, _vp],
    'vdvc_diag_H': [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _i, _vp, _vp],
    'vdvc_matvec_K': [_vp, _i, _i, _i, _i, _dp, _d, _vp, _vp, _vp],
    'vdvc_diag_K': [_vp, _i, _i, _i, _i, _dp, _d, _vp, _vp],
    'vdvc_matvec_eqgap': [_vp, _i, _i, _i, _i, _dp, _d, _vp, _vp, _vp],
    'vdvc_hutchinson_eqgap': [_vp, _i, _i, _i, _i, _dp, _d, _vp, _i, _vp],
    'vdvc_jacobi': [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _i, _vp, _i, _dp, _dp, _d, _d, _d, _vp,
                    _vp, _vp],
    'vdvc_pcg': [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _i, _vp, _i, _dp, _dp, _d, _d, _d, _vp, _vp,
                 _vp, _d, _i, _i, _ip, _dp],
    'vdvc_apply_A': [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _i, _vp, _i, _dp, _dp, _d, _d, _d, _vp,
                     _vp],
    'vdvc_dot': [_vp, _i64, _i, _vp, _vp, _dp],
    'vdvc_gn_update': [_vp, _i64, _i, _vp, _vp, _dp],
    'vdvc_dot2': [_vp, _i64, _i, _vp, _vp, _vp, _dp],
    'vdvc_axpby': [_vp, _i64, _i, _d, _vp, _d, _vp],
    'vdvc_calibration_mode': [_vp, _i, _i, _i, _i, _i, _vp],
    'vdvc_mul': [_vp, _i64, _i, _vp, _vp, _vp],
    'vdvc_fold_grad': [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    'vdvc_erode6': [_vp, _i, _i, _i, _vp, _vp],
    'vdvc_stride': [_vp, _vp, _i, _i, _i, _i, _i, _vp],
    'vdvc_phase_corr': [_vp, _vp, _vp, _i, _i, _i, _i64p],
    'vdvc_affine_sums': [_vp, _vp, _i, _i, _i, _vp, _i, _i, _i, _dp, _dp, _dp, _dp],
    'vdvc_conv3d_forward': [_vp, _i, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _i, _vp],
    'vdvc_conv3d_backward_data': [_vp, _i, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp],
    'vdvc_conv3d_backward_weight': [_vp, _i, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp],
    'vdvc_adam_step': [_vp, _i64, _vp, _vp, _vp, _vp, _d, _d, _d, _d, _i],
    'vdvc_mse': [_vp, _i64, _vp, _vp, _vp, _dp],
    'vdvc_seed_input': [_vp, _vp, _vp, _i, _i, _i, _i, _i, _vp],
    'vdvc_interp_field': [_vp, _i, _i, _i, _i, _vp, _i, _vp, _i],
    'vdvc_active_mask': [_vp, _i, _i, _i, _i, _vp, _i, _vp, _i, _i, _i, _i, _d, _d, _d, _vp,
                         _d, _vp],
    'vdvc_push_forward': [_vp, _i, _i, _i, _i, _vp, _i, _vp, _i, _i, _i, _d, _vp],
    'vdvc_zncc': [_vp, _i64, _vp, _i, _vp, _i, _i, _vp, _dp],
    'vdvc_quality_cell': [_vp, _i, _i, _i, _i, _vp, _i, _vp, _i, _vp, _vp],
    'vdvc_gaussian_filter': [_vp, _i, _i, _i, _vp, _vp, _dp, _i, _i, _d],
    'vdvc_smooth_masked': [_vp, _i, _i, _i, _vp, _d, _dp, _i, _vp, _vp],
    'vdvc_gradient3': [_vp, _i, _i, _i, _vp, _vp, _vp, _vp],
    'vdvc_affine_normal': [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _dp, _dp, _dp, _d,
                           _dp, _dp],
    'vdvc_affine_rms': [_vp, _i, _i, _i, _vp, _vp, _vp, _dp, _dp, _dp, _d, _dp],
    'vdvc_warp_affine': [_vp, _i, _i, _i, _vp, _i, _i, _i, _dp, _dp, _dp, _vp],
    'vdvc_overlap_bounds': [_vp, _i, _i, _i, _vp, _i, _vp, _d, _ip, _ip],
    'vdvc_cell_masks': [_vp, _i, _i, _i, _i, _vp, _vp, _d, _d, _vp, _vp, _vp, _vp],
    'vdvc_safe_cells': [_vp, _i, _i, _i, _i, _vp, _vp],
    'vdvc_principal_strains': [_vp, _i, _i, _i, _i, _vp, _i, _vp, _vp],
    'vdvc_equivalent_strain': [_vp, _i64, _vp, _vp],
    'vdvc_gn_convergence': [_vp, _i, _i, _i, _vp, _i, _vp, _i, _vp],
    'vdvc_remap_mask_deformed': [_vp, _i, _i, _i, _i, _vp, _i, _dp, _vp, _vp],
    'vdvc_remap_field_deformed': [_vp, _i, _i, _i, _i, _vp, _i, _dp, _vp, _i, _vp],
    'vdvc_material_cells': [_vp, _i, _i, _i, _i, _vp, _i, _d, _vp],
    'vdvc_erode_cells': [_vp, _i, _i, _i, _vp, _vp],
    'vdvc_convert': [_vp, _i64, _vp, _i, _vp, _i],
    'vdvc_copy': [_vp, _i64, _vp, _vp],
    'vdvc_fill': [_vp, _i64, _i, _d, _vp],
    'vdvc_copy_volume': [_vp, _i, _i, _i, _vp, _i, _i64, _i64, _i64, _vp, _i, _i],
    'vdvc_axpby_x': [_vp, _i64, _d, _vp, _i, _d, _vp, _i],
    'vdvc_resample_nodes': [_vp, _i, _i, _i, _i, _vp, _i, _i, _i, _i, _vp, _i],
}

_lib = None
_lock = threading.Lock()


class LibraryMissing(RuntimeError):
    pass


def load():
    """Load libvdvc.so (raises LibraryMissing when it has not been built)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing('%s not built: run `python -m voxeldvc_amd.build`' % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        lib.vdvc_last_error.argtypes = []
        lib.vdvc_last_error.restype = ctypes.c_char_p
        if lib.vdvc_abi_version() != ABI_VERSION:
            raise LibraryMissing('%s has ABI version %d, this binding expects %d: rebuild it'
                                 % (LIB_PATH, lib.vdvc_abi_version(), ABI_VERSION))
        _lib = lib
        return lib


def check(rc):
    """Map a library status to the reference's exception types."""
    if rc == 0:
        return
    msg = _lib.vdvc_last_error().decode(errors='replace')
    if rc == -1:
        raise ValueError(msg)
    raise RuntimeError('libvdvc error %d: %s' % (rc, msg))


def call(name, *args):
    """Call a library entry point.  Device tensors may be passed directly:
    they are converted to pointers here and stay referenced until the work
    is enqueued, so a temporary cannot be freed (and its memory handed to
    another argument) before the kernel that reads it is on the stream."""
    lib = load()
    cargs = [ptr(a) if is_tensor(a) else a for a in args]
    check(getattr(lib, name)(*cargs))


# ---------------------------------------------------------------------------
# device context and arrays
# ---------------------------------------------------------------------------

_ctx = {}


def torch():
    import torch as _t
    return _t


def require_gpu():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError('voxeldvc_amd needs a ROCm GPU (torch.cuda.is_available() is False)')
    return t


def ctx():
    """The library context bound to torch's current device and stream."""
    t = require_gpu()
    load()
    dev = t.cuda.current_device()
    st = t.cuda.current_stream(dev).cuda_stream
    key = (dev, st)
    c = _ctx.get(key)
    if c is None:
        h = _vp()
        check(_lib.vdvc_ctx_create(dev, st, ctypes.byref(h)))
        c = h
        _ctx[key] = c
    return c


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    return None if t is None else _vp(t.data_ptr())


def dbl(*vals):
    arr = (ctypes.c_double * len(vals))(*vals)
    return arr


def empty(n, dtype='float32'):
    t = require_gpu()
    return t.empty(int(n), dtype=getattr(t, dtype), device='cuda')


# torch dtype name -> library element-type code
TYPE_CODES = {'float32': F32, 'bfloat16': BF16, 'float16': F16, 'float64': F64,
              'uint8': U8, 'uint16': U16, 'int16': I16, 'bool': BOOL}


def type_code(dtype):
    """Library element-type code of a torch / NumPy dtype or dtype name
    (None when the library has no such element type)."""
    name = str(dtype).replace('torch.', '')
    if name == 'bool_':
        name = 'bool'
    return TYPE_CODES.get(name)


def zeros(n, dtype='float32'):
    """Zero-filled device vector (filled by the library)."""
    d = empty(n, dtype)
    fill(d, 0.0)
    return d


def fill(d, value):
    """d[:] = value on the device (vdvc_fill)."""
    call('vdvc_fill', ctx(), d.numel(), type_code(d.dtype), float(value), d)
    return d


def copy_(dst, src):
    """dst[:] = src (same byte size) on the device (vdvc_copy)."""
    nb = src.numel() * src.element_size()
    if dst.numel() * dst.element_size() != nb:
        raise ValueError('copy_: %d bytes into %d' % (nb, dst.numel() * dst.element_size()))
    call('vdvc_copy', ctx(), nb, src, dst)
    return dst


def clone(d):
    """A device copy of a contiguous tensor."""
    return copy_(require_gpu().empty_like(d), d)


def convert(d, dtype):
    """Contiguous device tensor converted to `dtype` (the same tensor when
    it already has it)."""
    if str(d.dtype).replace('torch.', '') == dtype:
        return d
    out = empty(d.numel(), dtype).view(d.shape)
    call('vdvc_convert', ctx(), d.numel(), d, type_code(d.dtype), out, type_code(dtype))
    return out


def is_tensor(a):
    try:
        t = torch()
    except ImportError:
        return False
    return isinstance(a, t.Tensor)


def _host_upload(arr, fallback):
    """Host array -> device tensor of its own dtype (a plain copy); dtypes
    the library does not read are converted on the host to `fallback`."""
    t = require_gpu()
    arr = np.asarray(arr)
    if arr.dtype == np.bool_:
        arr = arr.view(np.uint8)
    elif type_code(arr.dtype) is None:
        arr = arr.astype(fallback)
    return t.from_numpy(np.ascontiguousarray(arr)).to(device='cuda')


def _dev_tensor(a, fallback):
    t = require_gpu()
    d = a.detach()
    if d.device.type != 'cuda':
        d = d.to(device='cuda')
    if d.dtype == t.bool:
        d = d.view(t.uint8)
    elif type_code(d.dtype) is None:
        d = d.to(getattr(t, fallback))
    return d


def _gather_dense(d, code):
    """Dense copy of a strided device tensor of up to three axes, axis 0
    fastest, in its own element type (vdvc_copy_volume)."""
    shp = list(d.shape) + [1] * (3 - d.dim())
    st = list(d.stride()) + [0] * (3 - d.dim())
    out = torch().empty(shp[0] * shp[1] * shp[2], dtype=d.dtype, device=d.device)
    call('vdvc_copy_volume', ctx(), shp[0], shp[1], shp[2], d, code, st[0], st[1], st[2], out,
         code, shp[0])
    return out


def _flat(d, code):
    """A device tensor flattened in C order (a view when it is contiguous,
    else a library copy; tensors of more than three axes are copied by
    torch)."""
    if d.is_contiguous():
        return d.view(-1)
    if d.dim() > 3:
        return d.contiguous().view(-1)
    return _gather_dense(d.permute(*reversed(range(d.dim()))), code)


def to_dev_flat(a, dtype='float32'):
    """1-D array (numpy or tensor) -> contiguous device tensor of `dtype`
    (conversion on the device)."""
    if is_tensor(a):
        d = _dev_tensor(a, dtype)
        return convert(_flat(d, type_code(d.dtype)), dtype)
    return convert(_host_upload(np.asarray(a).reshape(-1), dtype).view(-1), dtype)


def to_dev_volume(a, dtype='float32', pitched=False):
    """(nx, ny, nz) volume (NumPy C-order array or tensor of any strides) ->
    device tensor of `dtype` laid out x fastest, flat index x + p*(y + ny*z)
    with p = nx (or the level pitch when `pitched`, padding zeroed), plus its
    shape.  Host arrays are uploaded as they are and laid out on the device
    (vdvc_copy_volume: the reference's ravel(order='F') with its dtype
    conversion, engine/correlate_gpu.py:50-52, 938-939)."""
    if is_tensor(a):
        if a.dim() != 3:
            raise ValueError('volume must be 3-D')
        src = _dev_tensor(a, dtype)
    else:
        if np.ndim(a) != 3:
            raise ValueError('volume must be 3-D')
        src = _host_upload(a, dtype)
    shp = tuple(int(s) for s in src.shape)
    sx, sy, sz = src.stride()
    code = type_code(src.dtype)
    if not pitched and (sx, sy, sz) == (1, shp[0], shp[0] * shp[1]) and \
            str(src.dtype).replace('torch.', '') == dtype:
        # already x fastest and dense in the requested type
        return src.as_strided((src.numel(),), (1,)), shp
    p = pitch_of(shp[0]) if pitched else shp[0]
    out = empty(p * shp[1] * shp[2], dtype)
    call('vdvc_copy_volume', ctx(), shp[0], shp[1], shp[2], src, code, sx, sy, sz, out,
         type_code(dtype), p)
    return out, shp


# element types the post-solve stages read in the images' own dtype:
# dtype name -> (library code, device dtype name)
IMAGE_TYPES = {'float32': (F32, 'float32'), 'float64': (F64, 'float64'),
               'uint8': (U8, 'uint8'), 'uint16': (U16, 'uint16'), 'int16': (I16, 'int16'),
               'bool': (BOOL, 'uint8')}


def image_type(a):
    """(library code, device dtype name) of an image (numpy array or tensor)
    for the post-solve stages; ValueError for other dtypes."""
    name = str(a.dtype).replace('torch.', '')
    if name not in IMAGE_TYPES:
        raise ValueError('unsupported image dtype %s (float32, float64, uint8, uint16, int16 '
                         'or bool)' % name)
    return IMAGE_TYPES[name]


def _image_bits(d):
    """Device storage of an image tensor: bool as its uint8 bytes, uint16 as
    int16 bits (the library reads the raw element with the type code)."""
    t = torch()
    if d.dtype == t.bool:
        return d.view(t.uint8)
    if d.dtype == t.uint16:
        return d.view(t.int16)
    return d


def _host_bits(arr):
    if arr.dtype == np.bool_:
        return arr.view(np.uint8)
    if arr.dtype == np.uint16:
        return arr.view(np.int16)
    return arr


def to_dev_image(a):
    """(nx, ny, nz) image in its own dtype -> (device tensor laid out x
    fastest, shape, type code)."""
    code, _ = image_type(a)
    t = require_gpu()
    if is_tensor(a):
        if a.dim() != 3:
            raise ValueError('volume must be 3-D')
        d = _image_bits(a.detach().to(device='cuda'))
        # x fastest: element (x, y, z) at x sx + y sy + z sz of the source
        return _gather_dense(d, code), tuple(int(s) for s in a.shape), code
    arr = np.asarray(a)
    if arr.ndim != 3:
        raise ValueError('volume must be 3-D')
    host = _host_bits(np.ravel(arr, order='F'))
    return t.from_numpy(np.ascontiguousarray(host)).to(device='cuda'), arr.shape, code


def to_dev_image_flat(a):
    """1-D array in its own dtype -> (contiguous device tensor, type code)."""
    code, _ = image_type(a)
    t = require_gpu()
    if is_tensor(a):
        return _flat(_image_bits(a.detach().to(device='cuda')), code), code
    host = _host_bits(np.ascontiguousarray(np.asarray(a).reshape(-1)))
    return t.from_numpy(host).to(device='cuda'), code


def image_to_host(d, dtype):
    """Device image storage (see to_dev_image) -> NumPy array of `dtype`."""
    h = to_host(d)
    dtype = np.dtype(dtype)
    if dtype == np.bool_:
        return h.astype(bool)
    if dtype == np.uint16:
        return h.view(np.uint16)
    return h


def rtype_of(x):
    """Working-type code of a device vector: F64 for float64, else F32."""
    return F64 if x.dtype == torch().float64 else F32


def to_host(x):
    return x.detach().cpu().numpy()


def pitch_of(nvx):
    """Row pitch of level voxel volumes (include/vdvc.h: vdvc_voxel_pitch)."""
    return (int(nvx) + 15) & ~15


def pitch_dense(d, shape, dtype=None):
    """Dense x-fastest device volume of `shape` (nx, ny, nz) -> zero-padded
    pitched volume (index x + pitch*(y + ny*z)) of `dtype` (default: d's)."""
    nx, ny, nz = shape
    p = pitch_of(nx)
    name = str(d.dtype).replace('torch.', '') if dtype is None else dtype
    out = empty(p * ny * nz, name)
    call('vdvc_copy_volume', ctx(), nx, ny, nz, d, type_code(d.dtype), 1, nx, nx * ny, out,
         type_code(name), p)
    return out


def unpitch(d, shape):
    """Pitched device volume -> dense x-fastest (the reference's voxel order)."""
    nx, ny, nz = shape
    p = pitch_of(nx)
    out = empty(nx * ny * nz, str(d.dtype).replace('torch.', ''))
    call('vdvc_copy_volume', ctx(), nx, ny, nz, d, type_code(d.dtype), 1, p, p * ny, out,
         type_code(d.dtype), nx)
    return out


def to_dev_level(a, shape, dtype='float32'):
    """Reference-order voxel array (flat (Nvox,), or (Nvox, k) columns are not
    accepted here) or (nx, ny, nz) volume -> pitched device volume."""
    if (is_tensor(a) and a.dim() == 3) or (not is_tensor(a) and np.ndim(a) == 3):
        d, shp = to_dev_volume(a, dtype, pitched=True)
        if tuple(shp) != tuple(shape):
            raise ValueError('volume shape %r != %r' % (tuple(shp), tuple(shape)))
        return d
    d = to_dev_flat(a, dtype)
    if d.numel() != shape[0] * shape[1] * shape[2]:
        raise ValueError('voxel array has %d entries, expected %d'
                         % (d.numel(), shape[0] * shape[1] * shape[2]))
    return pitch_dense(d, shape)
