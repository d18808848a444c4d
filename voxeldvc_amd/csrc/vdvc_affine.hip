// Preprocess affine stage (the step before the correlation):
//   * gauss_axis_kernel: one axis pass of scipy.ndimage.gaussian_filter
//     (engine/correlate_gpu.py:570-598): symmetric correlation in float64 in
//     SciPy's order -- centre tap, then (x[i-k] + x[i+k]) * w_k for k = r..1
//     -- with 'nearest' or 'constant' line extension, stored float32;
//   * mask_prep / masked_ratio kernels: normalised-convolution smoothing
//     smooth(g * valid) / smooth(valid) where the denominator exceeds 1e-6;
//   * gradient_kernel: np.gradient (unit spacing, edge order 1), float32;
//   * affine_normal_kernel: the 12x12 Gauss-Newton normal equations of
//     affine_refine_gn (correlate_gpu.py:704-793): per voxel the sample
//     location s = x - M (x - c) - t_r in float32 (the reference's
//     element-wise order), f and the three gradient planes sampled there
//     (map_coordinates order 1, 'constant' / 'nearest'), the float32 Jacobian
//     J_k and residual e, and sum J J^T, sum J e in float64;
//   * affine_rms_kernel: the masked residual sum of squares (:725-743);
//   * warp_affine_kernel: g at A (x - c) + c + t, trilinear, 0 outside
//     (preprocess.py:233-285);
//   * overlap_proj_kernel: the axis projections of (f > glt) & (g > glt) that
//     compute_crop bounds (preprocess.py:315-367).
// Volumes are dense x fastest (index x + nx*(y + ny*z)).
#include "vdvc_common.h"
#include "vdvc_kernels.h"
#include "vdvc_sample01.h"

namespace {

using vdvc_img::as_img_type;
using vdvc_img::ld_img;
using vdvc_img::sample01;

__global__ __launch_bounds__(VDVC_BLOCK) void gauss_axis_kernel(GaussIn G,
                                                                const float* __restrict__ in,
                                                                float* __restrict__ out) {
  const int x = blockIdx.x * VDVC_BLOCK + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x >= G.nx) return;
  const long long i = (long long)x + (long long)G.nx * ((long long)y + (long long)G.ny * z);
  int len, pos;
  long long stride;
  if (G.axis == 0) {
    len = G.nx; pos = x; stride = 1;
  } else if (G.axis == 1) {
    len = G.ny; pos = y; stride = G.nx;
  } else {
    len = G.nz; pos = z; stride = (long long)G.nx * G.ny;
  }
  const long long base = i - (long long)pos * stride;
  auto at = [&](int q) -> double {
    if (q < 0 || q >= len) {
      if (G.mode == 1) return G.cval;
      q = q < 0 ? 0 : len - 1;
    }
    return (double)in[base + (long long)q * stride];
  };
  double acc = mul_rn(at(pos), G.w[0]);
  for (int k = G.r; k >= 1; --k) acc = add_rn(acc, mul_rn(add_rn(at(pos - k), at(pos + k)), G.w[k]));
  out[i] = __double2float_rn(acc);
}

// gv = g * valid, vf = valid, valid = (g != cval) in float32
__global__ __launch_bounds__(VDVC_BLOCK) void mask_prep_kernel(long long n, const float* g,
                                                               float cval, float* gv, float* vf,
                                                               uint8_t* valid) {
  const long long i = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  if (i >= n) return;
  const float v = g[i] != cval ? 1.0f : 0.0f;
  if (gv) {
    gv[i] = g[i] * v;
    vf[i] = v;
  }
  if (valid) valid[i] = v != 0.0f ? 1 : 0;
}

// out = num / den where den > 1e-6 (float32), g elsewhere
__global__ __launch_bounds__(VDVC_BLOCK) void masked_ratio_kernel(long long n, const float* num,
                                                                  const float* den,
                                                                  const float* g, float* out) {
  const long long i = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  if (i >= n) return;
  out[i] = den[i] > 1e-6f ? num[i] / den[i] : g[i];
}

// np.gradient of a float32 volume (unit spacing, edge order 1)
__global__ __launch_bounds__(VDVC_BLOCK) void gradient_kernel(int nx, int ny, int nz,
                                                              const float* __restrict__ f,
                                                              float* gx, float* gy, float* gz) {
  const int x = blockIdx.x * VDVC_BLOCK + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x >= nx) return;
  const long long sy = nx, sz = (long long)nx * ny;
  const long long i = x + sy * y + sz * z;
  auto d = [&](int p, int n, long long s) -> float {
    if (p == 0) return sub_rn(f[i + s], f[i]);
    if (p == n - 1) return sub_rn(f[i], f[i - s]);
    return sub_rn(f[i + s], f[i - s]) / 2.0f;
  };
  gx[i] = d(x, nx, 1);
  gy[i] = d(y, ny, sy);
  gz[i] = d(z, nz, sz);
}

// sample location s(x) = x - M (x - c) - t_r and the centred coordinate
// x - c, float32, the reference's element-wise order
__device__ __forceinline__ void affine_coords(const AffIn& A, int x, int y, int z, float (&Xc)[3],
                                              float (&s)[3]) {
  const float b[3] = {(float)x, (float)y, (float)z};
#pragma unroll
  for (int j = 0; j < 3; ++j) Xc[j] = sub_rn(b[j], A.c[j]);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float m = add_rn(add_rn(mul_rn(A.M[3 * i], Xc[0]), mul_rn(A.M[3 * i + 1], Xc[1])),
                           mul_rn(A.M[3 * i + 2], Xc[2]));
    s[i] = sub_rn(sub_rn(b[i], m), A.tr[i]);
  }
}

// voxel index v -> (x, y, z) of an nx * ny * nz grid: float64 quotients
// corrected to the exact integer division
__device__ __forceinline__ void split_index(long long v, int nx, int ny, int& x, int& y, int& z) {
  long long q = (long long)((double)v / (double)nx);
  long long r = v - q * nx;
  if (r < 0) { --q; r += nx; } else if (r >= nx) { ++q; r -= nx; }
  long long q2 = (long long)((double)q / (double)ny);
  long long r2 = q - q2 * ny;
  if (r2 < 0) { --q2; r2 += ny; } else if (r2 >= ny) { ++q2; r2 -= ny; }
  x = (int)r;
  y = (int)r2;
  z = (int)q2;
}

__device__ __forceinline__ bool in_box(const float (&s)[3], int nx, int ny, int nz) {
  return s[0] >= 0.0f && s[0] <= (float)(nx - 1) && s[1] >= 0.0f && s[1] <= (float)(ny - 1) &&
         s[2] >= 0.0f && s[2] <= (float)(nz - 1);
}

__device__ __forceinline__ float sample_f32(const float* img, int nx, int ny, int nz,
                                            const float (&s)[3], int mode, double cval) {
  return (float)as_img_type(
      sample01<GT_F32>(img, nx, ny, nz, (double)s[0], (double)s[1], (double)s[2], 1, mode, cval),
      GT_F32);
}

// (a, b) of normal-equation entry e: the 78 upper-triangle entries of
// H = sum J J^T row by row, then the 12 entries of b = sum J e (b = 12)
__device__ __forceinline__ void entry_ab(int e, int& a, int& b) {
  if (e >= 78) {
    a = e - 78;
    b = 12;
    return;
  }
  a = 0;
  int rem = e;
  while (rem >= 12 - a) {
    rem -= 12 - a;
    ++a;
  }
  b = a + rem;
}

// Per wavefront: tiles of 64 consecutive voxels (one per lane); each lane
// writes its voxel's 12 Jacobian channels and e to the wavefront's own LDS
// tile (row pitch 65 floats: no bank conflicts), then lane l accumulates
// entries l and l + 64 of the 90 normal-equation sums over the tile's voxels
// in order (float32 products are exact in float64).  Wavefronts of a block
// run different numbers of tiles, so only wavefront-level ordering is used
// (no workgroup barrier).  Partials per wavefront [wave][90].
constexpr int kRow = 65;
__global__ __launch_bounds__(VDVC_BLOCK) void affine_normal_kernel(AffIn A, double* part) {
  __shared__ float lds[VDVC_WAVES_PER_BLOCK][13 * kRow];
  const int lane = threadIdx.x & 63;
  const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  float* T = lds[wib];
  const long long n = (long long)A.nx * A.ny * A.nz;
  const long long ntiles = (n + 63) / 64;
  const long long wave = (long long)blockIdx.x * VDVC_WAVES_PER_BLOCK + wib;
  const long long nwaves = (long long)gridDim.x * VDVC_WAVES_PER_BLOCK;
  int a1, b1, a2 = 0, b2 = 0;
  entry_ab(lane, a1, b1);
  const bool two = lane + 64 < 90;
  if (two) entry_ab(lane + 64, a2, b2);
  double acc1 = 0.0, acc2 = 0.0;
  for (long long t = wave; t < ntiles; t += nwaves) {
    const long long v = t * 64 + lane;
    float J[13];
#pragma unroll
    for (int k = 0; k < 13; ++k) J[k] = 0.0f;
    if (v < n) {
      int x, y, z;
      split_index(v, A.nx, A.ny, x, y, z);
      float Xc[3], s[3];
      affine_coords(A, x, y, z, Xc, s);
      const float mask = (in_box(s, A.nx, A.ny, A.nz) && A.valid[v] != 0) ? 1.0f : 0.0f;
      const float warped = sample_f32(A.f, A.nx, A.ny, A.nz, s, 1, A.cval);
      const float* G[3] = {A.gx, A.gy, A.gz};
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const float ng = -sample_f32(G[i], A.nx, A.ny, A.nz, s, 0, 0.0);
#pragma unroll
        for (int j = 0; j < 3; ++j) J[3 * i + j] = mul_rn(mul_rn(ng, Xc[j]), mask);
        J[9 + i] = mul_rn(ng, mask);
      }
      J[12] = mul_rn(sub_rn(A.g[v], warped), mask);
    }
#pragma unroll
    for (int k = 0; k < 13; ++k) T[k * kRow + lane] = J[k];
    // the tile is read by the other lanes of this wavefront only
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    for (int q = 0; q < 64; ++q) {
      acc1 += (double)T[a1 * kRow + q] * (double)T[b1 * kRow + q];
      if (two) acc2 += (double)T[a2 * kRow + q] * (double)T[b2 * kRow + q];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }
  part[wave * 90 + lane] = acc1;
  if (two) part[wave * 90 + lane + 64] = acc2;
}

// masked residual sum of squares: per block (sum e^2 in float64 of float32
// e*e, count)
__global__ __launch_bounds__(VDVC_BLOCK) void affine_rms_kernel(AffIn A, double* part) {
  __shared__ double lds[4 * 2];
  const long long n = (long long)A.nx * A.ny * A.nz;
  double acc[2] = {0.0, 0.0};
  for (long long v = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x; v < n;
       v += (long long)gridDim.x * VDVC_BLOCK) {
    int x, y, z;
    split_index(v, A.nx, A.ny, x, y, z);
    float Xc[3], s[3];
    affine_coords(A, x, y, z, Xc, s);
    if (in_box(s, A.nx, A.ny, A.nz) && A.valid[v] != 0) {
      const float e = sub_rn(A.g[v], sample_f32(A.f, A.nx, A.ny, A.nz, s, 1, A.cval));
      acc[0] += (double)mul_rn(e, e);
      acc[1] += 1.0;
    }
  }
  block_sum_f64<2>(acc, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = acc[0];
    part[2 * blockIdx.x + 1] = acc[1];
  }
}

// g (float32, dims gn) at A (x - c) + c + t for every voxel x of the nx*ny*nz
// grid: per row i, ((A_i0 d0 + A_i1 d1) + A_i2 d2) + c_i + t_i in float32;
// trilinear, 0 outside
__global__ __launch_bounds__(VDVC_BLOCK) void warp_affine_kernel(WarpIn W,
                                                                 const float* __restrict__ g,
                                                                 float* __restrict__ out) {
  const int x = blockIdx.x * VDVC_BLOCK + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x >= W.nx) return;
  const float d[3] = {sub_rn((float)x, W.c[0]), sub_rn((float)y, W.c[1]), sub_rn((float)z, W.c[2])};
  double p[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float m = add_rn(add_rn(mul_rn(W.A[3 * i], d[0]), mul_rn(W.A[3 * i + 1], d[1])),
                           mul_rn(W.A[3 * i + 2], d[2]));
    p[i] = (double)add_rn(add_rn(m, W.c[i]), W.t[i]);
  }
  const long long i = (long long)x + (long long)W.nx * ((long long)y + (long long)W.ny * z);
  out[i] = (float)as_img_type(sample01<GT_F32>(g, W.gnx, W.gny, W.gnz, p[0], p[1], p[2], 1, 1, 0.0),
                              GT_F32);
}

// axis projections of (f > glt) & (g > glt): byte flags [0, nx) for x,
// then y, then z.  Every writer stores the same value 1; a block ORs its
// row before the one y / z store.
__global__ __launch_bounds__(VDVC_BLOCK) void overlap_proj_kernel(int nx, int ny, int nz,
                                                                  const void* f, int ftype,
                                                                  const float* g, double glt,
                                                                  uint8_t* flags) {
  const int x = blockIdx.x * VDVC_BLOCK + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  bool val = false;
  if (x < nx) {
    const long long i = (long long)x + (long long)nx * ((long long)y + (long long)ny * z);
    // f > glt in f's dtype for float32 f (glt a weak scalar), in double
    // otherwise; g_warped is float32
    const double fv = ld_img(f, ftype, i);
    const bool fo = ftype == GT_F32 ? ((float)fv > (float)glt) : (fv > glt);
    val = fo && g[i] > (float)glt;
    if (val) flags[x] = 1;
  }
  if (__syncthreads_or(val) && threadIdx.x == 0) {
    flags[nx + y] = 1;
    flags[nx + ny + z] = 1;
  }
}

}  // namespace

namespace vdvc {

int launch_gauss_axis(const GaussIn& G, const float* in, float* out, hipStream_t st) {
  dim3 g((unsigned)((G.nx + VDVC_BLOCK - 1) / VDVC_BLOCK), (unsigned)G.ny, (unsigned)G.nz);
  hipLaunchKernelGGL(gauss_axis_kernel, g, dim3(VDVC_BLOCK), 0, st, G, in, out);
  return (int)hipGetLastError();
}

int launch_mask_prep(long long n, const float* g, float cval, float* gv, float* vf,
                     uint8_t* valid, hipStream_t st) {
  const unsigned nb = (unsigned)((n + VDVC_BLOCK - 1) / VDVC_BLOCK);
  hipLaunchKernelGGL(mask_prep_kernel, dim3(nb), dim3(VDVC_BLOCK), 0, st, n, g, cval, gv, vf,
                     valid);
  return (int)hipGetLastError();
}

int launch_masked_ratio(long long n, const float* num, const float* den, const float* g,
                        float* out, hipStream_t st) {
  const unsigned nb = (unsigned)((n + VDVC_BLOCK - 1) / VDVC_BLOCK);
  hipLaunchKernelGGL(masked_ratio_kernel, dim3(nb), dim3(VDVC_BLOCK), 0, st, n, num, den, g, out);
  return (int)hipGetLastError();
}

int launch_gradient(int nx, int ny, int nz, const float* f, float* gx, float* gy, float* gz,
                    hipStream_t st) {
  dim3 g((unsigned)((nx + VDVC_BLOCK - 1) / VDVC_BLOCK), (unsigned)ny, (unsigned)nz);
  hipLaunchKernelGGL(gradient_kernel, g, dim3(VDVC_BLOCK), 0, st, nx, ny, nz, f, gx, gy, gz);
  return (int)hipGetLastError();
}

int launch_affine_normal(const AffIn& A, double* part, int nblocks, hipStream_t st) {
  hipLaunchKernelGGL(affine_normal_kernel, dim3(nblocks), dim3(VDVC_BLOCK), 0, st, A, part);
  return (int)hipGetLastError();
}

int launch_affine_rms(const AffIn& A, double* part, int nblocks, hipStream_t st) {
  hipLaunchKernelGGL(affine_rms_kernel, dim3(nblocks), dim3(VDVC_BLOCK), 0, st, A, part);
  return (int)hipGetLastError();
}

int launch_warp_affine(const WarpIn& W, const float* g, float* out, hipStream_t st) {
  dim3 gr((unsigned)((W.nx + VDVC_BLOCK - 1) / VDVC_BLOCK), (unsigned)W.ny, (unsigned)W.nz);
  hipLaunchKernelGGL(warp_affine_kernel, gr, dim3(VDVC_BLOCK), 0, st, W, g, out);
  return (int)hipGetLastError();
}

int launch_overlap_proj(int nx, int ny, int nz, const void* f, int ftype, const float* g,
                        double glt, uint8_t* flags, hipStream_t st) {
  dim3 gr((unsigned)((nx + VDVC_BLOCK - 1) / VDVC_BLOCK), (unsigned)ny, (unsigned)nz);
  hipLaunchKernelGGL(overlap_proj_kernel, gr, dim3(VDVC_BLOCK), 0, st, nx, ny, nz, f, ftype, g,
                     glt, flags);
  return (int)hipGetLastError();
}

}  // namespace vdvc
