// Data-movement kernels of the solve's host boundary: element-type
// conversion, byte copies, fills, the strided volume copy that turns a
// caller's array (any strides, any supported element type) into the level
// layout (x fastest, padded row pitch, padding zeroed) and back, and the
// mixed-type y = a x + b y that accumulates a float32 correction into a
// float64 vector.
//
// Reference counterpart: the xp.asarray(..., dtype) conversions at the
// engine boundary (engine/correlate_gpu.py:938-939), the per-level
// f_pix[::s] copies and the float64 promotion of the solve's vectors
// (engine/correlate_gpu.py:1042-1047); the copies themselves have no
// arithmetic of their own.
#include "vdvc_kernels.h"

namespace {

// element load / store of a type code as double (exact for every type
// here: float16, bfloat16 and float32 widen exactly; the integer types are
// below 2^53)
__device__ __forceinline__ double ld_code(const void* p, int t, long long i) {
  switch (t) {
    case GT_F32: return (double)reinterpret_cast<const float*>(p)[i];
    case GT_F64: return reinterpret_cast<const double*>(p)[i];
    case GT_F16: return (double)__half2float(reinterpret_cast<const __half*>(p)[i]);
    case GT_BF16:
      return (double)__uint_as_float(((uint32_t)reinterpret_cast<const uint16_t*>(p)[i]) << 16);
    case GT_U8: case GT_B8: return (double)reinterpret_cast<const uint8_t*>(p)[i];
    case GT_U16: return (double)reinterpret_cast<const uint16_t*>(p)[i];
    case GT_I16: return (double)reinterpret_cast<const int16_t*>(p)[i];
    default: return 0.0;
  }
}

// round-to-nearest-even conversions from the double value.  float16 and
// bfloat16 are formed without double rounding: the double is first rounded
// to float with round-to-odd (toward zero, then the last bit set when
// inexact), which float's 24 bits make exact for a later rounding to 8 or 11
// significant bits.
__device__ __forceinline__ uint32_t f32_round_odd_bits(double v) {
  const float f = (float)v;
  uint32_t fb = __float_as_uint(f);
  if ((double)f != v && !isinf(f) && v == v) {
    if (fabs((double)f) > fabs(v)) fb -= 1u;  // one ulp toward zero (sign-magnitude)
    fb |= 1u;
  }
  return fb;
}

__device__ __forceinline__ void st_code(void* p, int t, long long i, double v) {
  switch (t) {
    case GT_F32: reinterpret_cast<float*>(p)[i] = (float)v; break;
    case GT_F64: reinterpret_cast<double*>(p)[i] = v; break;
    case GT_F16:
      reinterpret_cast<__half*>(p)[i] = __float2half_rn(__uint_as_float(f32_round_odd_bits(v)));
      break;
    case GT_BF16: {
      uint32_t fb = f32_round_odd_bits(v);
      if (v != v) {
        reinterpret_cast<uint16_t*>(p)[i] = (uint16_t)((fb >> 16) | 0x40u);  // quiet NaN
        break;
      }
      fb += 0x7fffu + ((fb >> 16) & 1u);
      reinterpret_cast<uint16_t*>(p)[i] = (uint16_t)(fb >> 16);
      break;
    }
    case GT_U8: case GT_B8: reinterpret_cast<uint8_t*>(p)[i] = (uint8_t)v; break;
    case GT_U16: reinterpret_cast<uint16_t*>(p)[i] = (uint16_t)v; break;
    case GT_I16: reinterpret_cast<int16_t*>(p)[i] = (int16_t)v; break;
    default: break;
  }
}

__global__ __launch_bounds__(VDVC_BLOCK) void convert_kernel(long long n, const void* src, int st,
                                                             void* dst, int dt) {
  for (long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x; q < n;
       q += (long long)gridDim.x * VDVC_BLOCK)
    st_code(dst, dt, q, ld_code(src, st, q));
}

// 16-byte copies for the aligned bulk, bytes for the rest
__global__ __launch_bounds__(VDVC_BLOCK) void copy_kernel(long long nbytes,
                                                          const uint8_t* __restrict__ src,
                                                          uint8_t* __restrict__ dst, int vec) {
  const long long tid = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  const long long stride = (long long)gridDim.x * VDVC_BLOCK;
  long long done = 0;
  if (vec) {
    const long long n16 = nbytes >> 4;
    const uint4* s = reinterpret_cast<const uint4*>(src);
    uint4* d = reinterpret_cast<uint4*>(dst);
    for (long long q = tid; q < n16; q += stride) d[q] = s[q];
    done = n16 << 4;
  }
  for (long long q = done + tid; q < nbytes; q += stride) dst[q] = src[q];
}

// Read-only streaming probe (a measurement of the HBM read ceiling the
// voxel kernels are compared against): grid-stride 16-byte nontemporal loads,
// four in flight per lane, folded with xor; a block stores its fold only if
// it equals the caller's key, so the loads cannot be dropped and, for any
// key that is not the fold, nothing is written
__global__ __launch_bounds__(VDVC_BLOCK) void stream_read_kernel(long long n16,
                                                                const uint4* __restrict__ src,
                                                                uint32_t key,
                                                                uint32_t* __restrict__ out) {
  typedef uint32_t u4 __attribute__((ext_vector_type(4)));
  const u4* s = reinterpret_cast<const u4*>(src);
  const long long stride = (long long)gridDim.x * VDVC_BLOCK;
  long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  u4 acc = {0u, 0u, 0u, 0u};
  for (; q + 3 * stride < n16; q += 4 * stride) {
    const u4 a = __builtin_nontemporal_load(s + q);
    const u4 b = __builtin_nontemporal_load(s + q + stride);
    const u4 c = __builtin_nontemporal_load(s + q + 2 * stride);
    const u4 d = __builtin_nontemporal_load(s + q + 3 * stride);
    acc ^= a ^ b ^ c ^ d;
  }
  for (; q < n16; q += stride) acc ^= __builtin_nontemporal_load(s + q);
  const uint32_t f = acc.x ^ acc.y ^ acc.z ^ acc.w;
  if (f == key) out[blockIdx.x] = f;
}

__global__ __launch_bounds__(VDVC_BLOCK) void fill_kernel(long long n, void* dst, int dt,
                                                          double v) {
  for (long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x; q < n;
       q += (long long)gridDim.x * VDVC_BLOCK)
    st_code(dst, dt, q, v);
}

// dst[x + dp (y + ny z)] = src[x sx + y sy + z sz] for x < nx, zero for
// nx <= x < dp.  One block per (row y, z), threads along x: the reads are
// contiguous when sx == 1.
__global__ __launch_bounds__(VDVC_BLOCK) void copy_rows_kernel(VolCopy C) {
  const int y = blockIdx.x, z = blockIdx.y;
  const long long sb = (long long)y * C.sy + (long long)z * C.sz;
  const long long db = (long long)C.dp * ((long long)y + (long long)C.ny * z);
  for (int x = threadIdx.x; x < C.dp; x += VDVC_BLOCK) {
    const double v = x < C.nx ? ld_code(C.src, C.st, sb + (long long)x * C.sx) : 0.0;
    st_code(C.dst, C.dt, db + x, v);
  }
}

// Element conversion of the unit-stride fast path: the same single
// rounding as st_code for the pairs it is instantiated for.
template <typename T> struct Elem;
template <> struct Elem<float> {
  static __device__ __forceinline__ float ld(const float* p, long long i) { return p[i]; }
  static __device__ __forceinline__ void st(float* p, long long i, float v) { p[i] = v; }
};
template <> struct Elem<double> {
  static __device__ __forceinline__ double ld(const double* p, long long i) { return p[i]; }
  static __device__ __forceinline__ void st(double* p, long long i, double v) { p[i] = v; }
};
template <> struct Elem<__half> {
  static __device__ __forceinline__ float ld(const __half* p, long long i) { return __half2float(p[i]); }
  static __device__ __forceinline__ void st(__half* p, long long i, float v) { p[i] = __float2half_rn(v); }
};
template <> struct Elem<uint8_t> {
  static __device__ __forceinline__ uint32_t ld(const uint8_t* p, long long i) { return p[i]; }
  static __device__ __forceinline__ void st(uint8_t* p, long long i, uint32_t v) { p[i] = (uint8_t)v; }
};

// Unit-x-stride copy (sx == 1) specialised on the element types: each
// wavefront takes whole rows (grid-stride); its lanes cover consecutive
// voxels, four 64-voxel chunks of a row in flight at once, so every load and
// store instruction moves one contiguous run; padding columns are written as
// zero.
template <typename ST, typename DT, typename W>
__global__ __launch_bounds__(VDVC_BLOCK) void copy_rows_fast_kernel(VolCopy C) {
  const ST* __restrict__ src = reinterpret_cast<const ST*>(C.src);
  DT* __restrict__ dst = reinterpret_cast<DT*>(C.dst);
  const int lane = threadIdx.x & 63;
  const long long nrows = (long long)C.ny * C.nz;
  for (long long row = (long long)blockIdx.x * VDVC_WAVES_PER_BLOCK + (threadIdx.x >> 6);
       row < nrows; row += (long long)gridDim.x * VDVC_WAVES_PER_BLOCK) {
    const long long y = row % C.ny, z = row / C.ny;
    const long long sb = y * C.sy + z * C.sz;
    const long long db = (long long)C.dp * row;
    for (int x0 = 0; x0 < C.dp; x0 += 256) {
      W v[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int x = x0 + 64 * j + lane;
        v[j] = x < C.nx ? (W)Elem<ST>::ld(src, sb + x) : (W)0;
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int x = x0 + 64 * j + lane;
        if (x < C.dp) Elem<DT>::st(dst, db + x, v[j]);
      }
    }
  }
}

// The same copy for a source whose z stride is 1 (a C-order [x][y][z]
// array): a 64 x 64 (x, z) tile of one y row is read along z, transposed in
// LDS, and written along x.
__global__ __launch_bounds__(VDVC_BLOCK) void copy_transpose_kernel(VolCopy C) {
  __shared__ double tile[64][65];
  const int y = blockIdx.y;
  const int ntz = (C.nz + 63) / 64;
  const int tx = blockIdx.x / ntz, tz = blockIdx.x % ntz;
  const int x0 = tx * 64, z0 = tz * 64;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // read: lanes along z (contiguous), waves over x rows
  for (int r = wv; r < 64; r += VDVC_WAVES_PER_BLOCK) {
    const int x = x0 + r, z = z0 + lane;
    double v = 0.0;
    if (x < C.nx && z < C.nz)
      v = ld_code(C.src, C.st, (long long)x * C.sx + (long long)y * C.sy + (long long)z * C.sz);
    tile[r][lane] = v;
  }
  __syncthreads();
  // write: lanes along x, waves over z rows; the padding columns of the
  // last x tile are written as zero
  for (int r = wv; r < 64; r += VDVC_WAVES_PER_BLOCK) {
    const int z = z0 + r, x = x0 + lane;
    if (z >= C.nz) continue;
    const long long db = (long long)C.dp * ((long long)y + (long long)C.ny * z);
    if (x < C.nx) st_code(C.dst, C.dt, db + x, tile[lane][r]);
    else if (x < C.dp && x < x0 + 64) st_code(C.dst, C.dt, db + x, 0.0);
  }
  // padding beyond the last tile's 64 columns (dp - nx can exceed the tile)
  if (tx == (C.nx + 63) / 64 - 1) {
    for (int r = wv; r < 64; r += VDVC_WAVES_PER_BLOCK) {
      const int z = z0 + r;
      if (z >= C.nz) continue;
      const long long db = (long long)C.dp * ((long long)y + (long long)C.ny * z);
      for (int x = x0 + 64 + lane; x < C.dp; x += 64) st_code(C.dst, C.dt, db + x, 0.0);
    }
  }
}

// y = a x + b y, x and y of their own types, arithmetic in double (exact
// widening of float32 into a float64 y: y + 1.0 * x is the float64 sum)
__global__ __launch_bounds__(VDVC_BLOCK) void axpby_x_kernel(long long n, double a,
                                                             const void* x, int xt, double b,
                                                             void* y, int yt) {
  for (long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x; q < n;
       q += (long long)gridDim.x * VDVC_BLOCK) {
    const double xv = ld_code(x, xt, q);
    const double yv = b == 0.0 ? 0.0 : ld_code(y, yt, q);
    st_code(y, yt, q, a * xv + b * yv);
  }
}

// dst node (i, j, k) of an (Nx, Ny, Nz)-element mesh with element size hd:
// the trilinear (Hex8) interpolation of the src nodal field of an
// (Mx, My, Mz)-element mesh with element size hs at the voxel position
// (i hd, j hd, k hd), weights in float64; positions beyond the src mesh
// are clamped to it.
__global__ __launch_bounds__(VDVC_BLOCK) void resample_nodes_kernel(NodeResample R) {
  const long long nn = (long long)(R.Nx + 1) * (R.Ny + 1) * (R.Nz + 1);
  const long long n = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  if (n >= nn) return;
  const int i = (int)(n % (R.Nx + 1));
  const int j = (int)((n / (R.Nx + 1)) % (R.Ny + 1));
  const int k = (int)(n / ((long long)(R.Nx + 1) * (R.Ny + 1)));
  const int dn[3] = {i, j, k};
  const int M[3] = {R.Mx, R.My, R.Mz};
  int a0[3];
  double t[3];
  for (int d = 0; d < 3; ++d) {
    // src coordinate dn * hd / hs, split into an element and its local t
    const long long num = (long long)dn[d] * R.hd;
    long long e = num / R.hs;
    long long rem = num - e * R.hs;
    if (e >= M[d]) { e = M[d] - 1; rem = R.hs; }
    a0[d] = (int)e;
    t[d] = (double)rem / (double)R.hs;
  }
  double acc[3] = {0.0, 0.0, 0.0};
  for (int c = 0; c < 8; ++c) {
    const int cx = c & 1, cy = (c >> 1) & 1, cz = c >> 2;
    const double w = (cx ? t[0] : 1.0 - t[0]) * (cy ? t[1] : 1.0 - t[1]) *
                     (cz ? t[2] : 1.0 - t[2]);
    if (w == 0.0) continue;
    const long long s = (long long)(a0[0] + cx) +
                        (long long)(R.Mx + 1) * ((a0[1] + cy) + (long long)(R.My + 1) * (a0[2] + cz));
    for (int q = 0; q < 3; ++q) acc[q] += w * ld_code(R.src, R.rt, 3 * s + q);
  }
  for (int q = 0; q < 3; ++q) st_code(R.dst, R.rt, 3 * n + q, acc[q]);
}

int grid_for(long long n) {
  long long nb = (n + VDVC_BLOCK - 1) / VDVC_BLOCK;
  if (nb > 8192) nb = 8192;
  if (nb < 1) nb = 1;
  return (int)nb;
}

}  // namespace

namespace vdvc {

int launch_convert(long long n, const void* src, int st, void* dst, int dt, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(convert_kernel, dim3(grid_for(n)), dim3(VDVC_BLOCK), 0, s, n, src, st, dst,
                     dt);
  return (int)hipGetLastError();
}

int launch_stream_read(long long nbytes, const void* src, int grid, uint32_t key, uint32_t* out,
                       hipStream_t s) {
  hipLaunchKernelGGL(stream_read_kernel, dim3(grid), dim3(VDVC_BLOCK), 0, s, nbytes >> 4,
                     reinterpret_cast<const uint4*>(src), key, out);
  return (int)hipGetLastError();
}

int launch_copy(long long nbytes, const void* src, void* dst, hipStream_t s) {
  if (nbytes <= 0) return 0;
  const int vec = ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15) == 0;
  hipLaunchKernelGGL(copy_kernel, dim3(grid_for(vec ? (nbytes >> 4) + 1 : nbytes)),
                     dim3(VDVC_BLOCK), 0, s, nbytes, reinterpret_cast<const uint8_t*>(src),
                     reinterpret_cast<uint8_t*>(dst), vec);
  return (int)hipGetLastError();
}

int launch_fill(long long n, void* dst, int dt, double v, hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(VDVC_BLOCK), 0, s, n, dst, dt, v);
  return (int)hipGetLastError();
}

int launch_copy_volume(const VolCopy& C, hipStream_t s) {
  if (C.sx == 1) {
    const long long nrows = (long long)C.ny * C.nz;
    long long nb = (nrows + VDVC_WAVES_PER_BLOCK - 1) / VDVC_WAVES_PER_BLOCK;
    if (nb > 8192) nb = 8192;
    const dim3 g((unsigned)nb), b(VDVC_BLOCK);
#define VC_FAST(STC, DTC, ST, DT, W)                                              \
    if (C.st == STC && C.dt == DTC) {                                             \
      hipLaunchKernelGGL((copy_rows_fast_kernel<ST, DT, W>), g, b, 0, s, C);     \
      return (int)hipGetLastError();                                            \
    }
    VC_FAST(GT_F32, GT_F32, float, float, float)
    VC_FAST(GT_F32, GT_F16, float, __half, float)
    VC_FAST(GT_F16, GT_F16, __half, __half, float)
    VC_FAST(GT_F16, GT_F32, __half, float, float)
    VC_FAST(GT_F32, GT_F64, float, double, double)
    VC_FAST(GT_F64, GT_F64, double, double, double)
    VC_FAST(GT_F64, GT_F32, double, float, double)
    VC_FAST(GT_U8, GT_U8, uint8_t, uint8_t, uint32_t)
#undef VC_FAST
  }
  if (C.sz == 1 && C.sx != 1 && C.nz > 1) {
    const int ntx = (C.nx + 63) / 64, ntz = (C.nz + 63) / 64;
    hipLaunchKernelGGL(copy_transpose_kernel, dim3(ntx * ntz, C.ny), dim3(VDVC_BLOCK), 0, s, C);
  } else {
    hipLaunchKernelGGL(copy_rows_kernel, dim3(C.ny, C.nz), dim3(VDVC_BLOCK), 0, s, C);
  }
  return (int)hipGetLastError();
}

int launch_resample_nodes(const NodeResample& R, hipStream_t s) {
  const long long nn = (long long)(R.Nx + 1) * (R.Ny + 1) * (R.Nz + 1);
  hipLaunchKernelGGL(resample_nodes_kernel, dim3((unsigned)((nn + VDVC_BLOCK - 1) / VDVC_BLOCK)),
                     dim3(VDVC_BLOCK), 0, s, R);
  return (int)hipGetLastError();
}

int launch_axpby_x(long long n, double a, const void* x, int xt, double b, void* y, int yt,
                   hipStream_t s) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(axpby_x_kernel, dim3(grid_for(n)), dim3(VDVC_BLOCK), 0, s, n, a, x, xt, b,
                     y, yt);
  return (int)hipGetLastError();
}

}  // namespace vdvc
