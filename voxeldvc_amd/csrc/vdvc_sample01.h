// Image element access and scipy.ndimage.map_coordinates (spline order 0 /
// 1, modes 'nearest' / 'constant') evaluated in SciPy's arithmetic order:
// float64 coordinates; order 0 at index floor(c + 0.5); order 1 summing
// ((v * wx) * wy) * wz over the 8 corners with the z index fastest, weights
// (1 - t, t) with t = c - floor(c), corner indices clamped to the image
// ('nearest'); cval for a point outside [0, n-1] on any axis ('constant');
// then the output conversion of the image dtype (floats round to nearest,
// integers round half away from zero, bool is v != 0).  All products and
// sums are explicitly rounded operations (no contraction).  Shared by the
// post-solve and the preprocess affine kernels.
#pragma once
#include <type_traits>

#include "vdvc_common.h"

namespace vdvc_img {

__device__ __forceinline__ double ld_img(const void* p, int t, long long i) {
  switch (t) {
    case GT_F32: return (double)reinterpret_cast<const float*>(p)[i];
    case GT_F64: return reinterpret_cast<const double*>(p)[i];
    case GT_U8: return (double)reinterpret_cast<const uint8_t*>(p)[i];
    case GT_U16: return (double)reinterpret_cast<const uint16_t*>(p)[i];
    case GT_I16: return (double)reinterpret_cast<const int16_t*>(p)[i];
    default: return reinterpret_cast<const uint8_t*>(p)[i] != 0 ? 1.0 : 0.0;  // GT_B8
  }
}

// integer output conversion of scipy.ndimage: round half away from zero,
// then truncate; clamped to the type's range
__device__ __forceinline__ double round_away(double v, double lo, double hi) {
  const double r = trunc(v >= 0.0 ? add_rn(v, 0.5) : sub_rn(v, 0.5));
  return fmin(fmax(r, lo), hi);
}

// value v as the image's dtype holds it, returned as double
__device__ __forceinline__ double as_img_type(double v, int t) {
  switch (t) {
    case GT_F32: return (double)__double2float_rn(v);
    case GT_F64: return v;
    case GT_U8: return round_away(v, 0.0, 255.0);
    case GT_U16: return round_away(v, 0.0, 65535.0);
    case GT_I16: return round_away(v, -32768.0, 32767.0);
    default: return v != 0.0 ? 1.0 : 0.0;  // GT_B8
  }
}

__device__ __forceinline__ void st_img(void* p, int t, long long i, double v) {
  const double w = as_img_type(v, t);
  switch (t) {
    case GT_F32: reinterpret_cast<float*>(p)[i] = (float)w; break;
    case GT_F64: reinterpret_cast<double*>(p)[i] = w; break;
    case GT_U8: reinterpret_cast<uint8_t*>(p)[i] = (uint8_t)w; break;
    case GT_U16: reinterpret_cast<uint16_t*>(p)[i] = (uint16_t)w; break;
    case GT_I16: reinterpret_cast<int16_t*>(p)[i] = (int16_t)w; break;
    default: reinterpret_cast<uint8_t*>(p)[i] = (uint8_t)w; break;
  }
}

// index clamped to [0, n-1] (image dims are below 2^31)
__device__ __forceinline__ int clamp_idx(double f, int n) {
  return (int)fmin(fmax(f, 0.0), (double)(n - 1));
}

// element i of an image of type t read through a buffer resource over it
// (32-bit byte offsets: images of fewer than 2^32 bytes)
__device__ __forceinline__ double ld_img_o32(__amdgpu_buffer_rsrc_t r, int t, unsigned i) {
  switch (t) {
    case GT_F32: return (double)__uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, i * 4u, 0, 0));
    case GT_F64: {
      const auto w = __builtin_amdgcn_raw_buffer_load_b64(r, i * 8u, 0, 0);
      return __hiloint2double((int)w[1], (int)w[0]);
    }
    case GT_U8: return (double)__builtin_amdgcn_raw_buffer_load_b8(r, i, 0, 0);
    case GT_U16: return (double)__builtin_amdgcn_raw_buffer_load_b16(r, i * 2u, 0, 0);
    case GT_I16:
      return (double)(int16_t)__builtin_amdgcn_raw_buffer_load_b16(r, i * 2u, 0, 0);
    default: return __builtin_amdgcn_raw_buffer_load_b8(r, i, 0, 0) != 0 ? 1.0 : 0.0;  // GT_B8
  }
}

// buffer resource over an image of `bytes` bytes (< 2^32); 0x00020000 is the
// gfx9 buffer format word (32-bit data format, no swizzle)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t img_rsrc(const void* img, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(img), (short)0, (int)bytes,
                                           0x00020000);
}

// map_coordinates(order 0/1, mode 'nearest' (0) / 'constant' (1)) at
// (cx, cy, cz), float64, before the output conversion; image type IT.
// O32: the image holds fewer than 2^32 bytes and is read through `rs` with
// 32-bit element indices instead of 64-bit index and address arithmetic per
// tap (rs is unused otherwise); the same elements, the same arithmetic.
template <int IT, bool O32 = false>
__device__ __forceinline__ double sample01(const void* img, int gnx, int gny, int gnz, double cx,
                                           double cy, double cz, int order, int mode,
                                           double cval, __amdgpu_buffer_rsrc_t rs) {
  constexpr int it = IT;
  if (mode == 1 && !(cx >= 0.0 && cx <= (double)(gnx - 1) && cy >= 0.0 &&
                     cy <= (double)(gny - 1) && cz >= 0.0 && cz <= (double)(gnz - 1)))
    return cval;
  using I = typename std::conditional<O32, unsigned, long long>::type;
  const I sy = (I)gnx, sz = (I)gnx * (I)gny;
  auto ld = [&](I i) -> double {
    if constexpr (O32)
      return ld_img_o32(rs, it, i);
    else
      return ld_img(img, it, i);
  };
  if (order == 0) {
    const I ix = (I)clamp_idx(floor(add_rn(cx, 0.5)), gnx);
    const I iy = (I)clamp_idx(floor(add_rn(cy, 0.5)), gny);
    const I iz = (I)clamp_idx(floor(add_rn(cz, 0.5)), gnz);
    return ld(ix + sy * iy + sz * iz);
  }
  const double fx = floor(cx), fy = floor(cy), fz = floor(cz);
  const double tx = sub_rn(cx, fx), ty = sub_rn(cy, fy), tz = sub_rn(cz, fz);
  const double wx[2] = {sub_rn(1.0, tx), tx};
  const double wy[2] = {sub_rn(1.0, ty), ty};
  const double wz[2] = {sub_rn(1.0, tz), tz};
  const I ix[2] = {(I)clamp_idx(fx, gnx), (I)clamp_idx(add_rn(fx, 1.0), gnx)};
  const I iy[2] = {(I)clamp_idx(fy, gny) * sy, (I)clamp_idx(add_rn(fy, 1.0), gny) * sy};
  const I iz[2] = {(I)clamp_idx(fz, gnz) * sz, (I)clamp_idx(add_rn(fz, 1.0), gnz) * sz};
  double v = 0.0;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const double val = ld(ix[a] + iy[b] + iz[c]);
        v = add_rn(v, mul_rn(mul_rn(mul_rn(val, wx[a]), wy[b]), wz[c]));
      }
  return v;
}

}  // namespace vdvc_img
