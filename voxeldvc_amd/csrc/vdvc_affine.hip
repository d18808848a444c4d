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
//     (map_coordinates order 1, 'constant' / 'nearest', one set of corners
//     and weights for the four images), the float32 Jacobian J_k and
//     residual e, and sum J J^T, sum J e in float64 registers;
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

// Value at offset d (|d| <= r) along the filtered line from the voxel at
// p (line position pos of len, element stride `stride`); with EDGE the
// position is clamped ('nearest') or replaced by cval ('constant').
template <bool EDGE>
__device__ __forceinline__ double gauss_tap(const GaussIn& G, const float* p, int pos, int len,
                                            int stride, int d) {
  if (!EDGE) return (double)p[d * stride];
  const int q = pos + d;
  const int qc = min(max(q, 0), len - 1);
  const double v = (double)p[(qc - pos) * stride];
  return (G.mode == 1 && qc != q) ? G.cval : v;
}

// taps k, k-1, ..., k-C+1 of both sides loaded together, then accumulated
// in that order
template <int C, bool EDGE>
__device__ __forceinline__ void gauss_chunk(const GaussIn& G, const float* p, int pos, int len,
                                            int stride, int k, double& acc) {
  double lo[C], hi[C];
#pragma unroll
  for (int j = 0; j < C; ++j) {
    lo[j] = gauss_tap<EDGE>(G, p, pos, len, stride, -(k - j));
    hi[j] = gauss_tap<EDGE>(G, p, pos, len, stride, k - j);
  }
#pragma unroll
  for (int j = 0; j < C; ++j) acc = add_rn(acc, mul_rn(add_rn(lo[j], hi[j]), G.w[k - j]));
}

// SciPy's symmetric correlation at one voxel: the centre tap, then
// (x[i-k] + x[i+k]) * w_k for k = r..1, in float64
template <bool EDGE>
__device__ __forceinline__ double gauss_line(const GaussIn& G, const float* p, int pos, int len,
                                             int stride) {
  double acc = mul_rn((double)p[0], G.w[0]);
  int k = G.r;
  for (; k >= 8; k -= 8) gauss_chunk<8, EDGE>(G, p, pos, len, stride, k, acc);
  if (k >= 4) {
    gauss_chunk<4, EDGE>(G, p, pos, len, stride, k, acc);
    k -= 4;
  }
  for (; k >= 1; --k) gauss_chunk<1, EDGE>(G, p, pos, len, stride, k, acc);
  return acc;
}

// One output voxel per thread.  Blocks cover 256 x-consecutive voxels of one
// (y, z) row; the 1-D grid is ordered with the filtered axis fastest and
// spread so that each XCD takes a contiguous run of it (the 2r + 1 lines a
// block reads are the ones its XCD's L2 already holds for the blocks
// before).  Interior voxels take the path without the extension tests.
__global__ __launch_bounds__(VDVC_BLOCK) void gauss_axis_kernel(GaussIn G,
                                                                const float* __restrict__ in,
                                                                float* __restrict__ out) {
  const unsigned nxc = (unsigned)(G.nx + VDVC_BLOCK - 1) / VDVC_BLOCK;
  unsigned L = xcd_remap(blockIdx.x, gridDim.x);
  unsigned xc, y, z;
  if (G.axis == 2) {
    z = L % (unsigned)G.nz; L /= (unsigned)G.nz;
    y = L % (unsigned)G.ny; xc = L / (unsigned)G.ny;
  } else if (G.axis == 1) {
    y = L % (unsigned)G.ny; L /= (unsigned)G.ny;
    xc = L % nxc; z = L / nxc;
  } else {
    xc = L % nxc; L /= nxc;
    y = L % (unsigned)G.ny; z = L / (unsigned)G.ny;
  }
  const int x = (int)xc * VDVC_BLOCK + threadIdx.x;
  if (x >= G.nx) return;
  const long long i = (long long)x + (long long)G.nx * ((long long)y + (long long)G.ny * z);
  int len, pos, stride;
  if (G.axis == 0) {
    len = G.nx; pos = x; stride = 1;
  } else if (G.axis == 1) {
    len = G.ny; pos = (int)y; stride = G.nx;
  } else {
    len = G.nz; pos = (int)z; stride = G.nx * G.ny;
  }
  const float* p = in + i;
  const double acc = (pos >= G.r && pos + G.r < len) ? gauss_line<false>(G, p, pos, len, stride)
                                                     : gauss_line<true>(G, p, pos, len, stride);
  out[i] = __double2float_rn(acc);
}

// The same pass with the block's input lines staged in LDS.  A block loads
// its tile of lines once, with the r-voxel halo on both sides already
// extended as the mode says ('nearest': the end value; 'constant': cval,
// which the launcher only allows here when float32 holds it exactly), then
// forms every output of the tile from LDS in SciPy's order -- the same
// operations on the same values as gauss_axis_kernel, without its 2r + 1
// global loads per output.
//   AXIS 0: one row segment of kGaussRowTile outputs per block, thread =
//           output;
//   AXIS 1, 2: kGaussColX x-columns times kGaussColT positions along the
//           axis per block; lane = x-column, the block's four wavefronts
//           take every fourth position, so each LDS row read is 64
//           consecutive floats.
// The 1-D grid runs along the filtered axis's tiles first and is spread so
// that each XCD takes a contiguous run of it (neighbouring tiles' halos come
// from that XCD's L2).
constexpr int kGaussRowTile = 256;
constexpr int kGaussColX = 64;
constexpr int kGaussColT = 64;

// sum of the symmetric taps of the output at LDS element c, lines `ld`
// floats apart: the centre tap, then (t[c-k] + t[c+k]) * w_k for k = r..1
__device__ __forceinline__ double gauss_lds_line(const GaussIn& G, const float* t, int c,
                                                 int ld) {
  double acc = mul_rn((double)t[c * ld], G.w[0]);
  int k = G.r;
  for (; k >= 4; k -= 4) {
    const float* lo = t + (c - k) * ld;
    const float* hi = t + (c + k) * ld;
#pragma unroll
    for (int j = 0; j < 4; ++j)
      acc = add_rn(acc, mul_rn(add_rn((double)lo[j * ld], (double)hi[-j * ld]), G.w[k - j]));
  }
  for (; k >= 1; --k)
    acc = add_rn(acc, mul_rn(add_rn((double)t[(c - k) * ld], (double)t[(c + k) * ld]), G.w[k]));
  return acc;
}

template <int AXIS>
__global__ __launch_bounds__(VDVC_BLOCK) void gauss_tile_kernel(GaussIn G,
                                                                const float* __restrict__ in,
                                                                float* __restrict__ out) {
  extern __shared__ float tile[];
  const int r = G.r;
  const int tid = threadIdx.x;
  const float cv = (float)G.cval;
  unsigned L = xcd_remap(blockIdx.x, gridDim.x);
  if constexpr (AXIS == 0) {
    const unsigned nxt = (unsigned)(G.nx + kGaussRowTile - 1) / kGaussRowTile;
    const unsigned xt = L % nxt;
    L /= nxt;
    const unsigned y = L % (unsigned)G.ny, z = L / (unsigned)G.ny;
    const int x0 = (int)xt * kGaussRowTile;
    const long long rowo = ((long long)z * G.ny + y) * G.nx;
    const float* row = in + rowo;
    for (int j = tid; j < kGaussRowTile + 2 * r; j += VDVC_BLOCK) {
      const int q = x0 - r + j;
      const int qc = min(max(q, 0), G.nx - 1);
      const float v = row[qc];
      tile[j] = (G.mode == 1 && qc != q) ? cv : v;
    }
    __syncthreads();
    const int x = x0 + tid;
    if (x < G.nx) out[rowo + x] = __double2float_rn(gauss_lds_line(G, tile, tid + r, 1));
  } else {
    const int len = AXIS == 1 ? G.ny : G.nz;
    const unsigned nt = (unsigned)(len + kGaussColT - 1) / kGaussColT;
    const unsigned nxc = (unsigned)(G.nx + kGaussColX - 1) / kGaussColX;
    const unsigned t = L % nt;
    L /= nt;
    const unsigned xc = L % nxc, o = L / nxc;  // o: z (axis 1) or y (axis 2)
    const long long stride = AXIS == 1 ? (long long)G.nx : (long long)G.nx * G.ny;
    const long long obase = AXIS == 1 ? (long long)o * G.ny * G.nx : (long long)o * G.nx;
    const int p0 = (int)t * kGaussColT;
    const int lx = tid & 63, wv = tid >> 6;
    const int x = (int)xc * kGaussColX + lx;
    const bool xin = x < G.nx;
    const int xs = xin ? x : G.nx - 1;  // columns past the volume load a valid voxel, unused
    for (int j = wv; j < kGaussColT + 2 * r; j += VDVC_WAVES_PER_BLOCK) {
      const int q = p0 - r + j;
      const int qc = min(max(q, 0), len - 1);
      const float v = in[obase + (long long)qc * stride + xs];
      tile[j * kGaussColX + lx] = (G.mode == 1 && qc != q) ? cv : v;
    }
    __syncthreads();
    if (xin) {
      const int pend = min(kGaussColT, len - p0);
      for (int jj = wv; jj < pend; jj += VDVC_WAVES_PER_BLOCK)
        out[obase + (long long)(p0 + jj) * stride + x] =
            __double2float_rn(gauss_lds_line(G, tile + lx, jj + r, kGaussColX));
    }
  }
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

// Trilinear corners of an in-box sample location, shared by every image
// sampled there: base offset, the x / y / z steps to the upper corners (0 at
// the last index, where the index is clamped) and SciPy's float64 weights.
struct Corners {
  long long o;
  int ox, oy, oz;
  double wx[2], wy[2], wz[2];
};

__device__ __forceinline__ Corners in_box_corners(const float (&s)[3], int nx, int ny, int nz) {
  Corners C;
  const double fx = floor((double)s[0]), fy = floor((double)s[1]), fz = floor((double)s[2]);
  const double tx = sub_rn((double)s[0], fx), ty = sub_rn((double)s[1], fy),
               tz = sub_rn((double)s[2], fz);
  C.wx[0] = sub_rn(1.0, tx); C.wx[1] = tx;
  C.wy[0] = sub_rn(1.0, ty); C.wy[1] = ty;
  C.wz[0] = sub_rn(1.0, tz); C.wz[1] = tz;
  const int ix = (int)fx, iy = (int)fy, iz = (int)fz;
  const long long sy = nx, sz = (long long)nx * ny;
  C.o = ix + sy * iy + sz * iz;
  C.ox = ix + 1 < nx ? 1 : 0;
  C.oy = iy + 1 < ny ? nx : 0;
  C.oz = iz + 1 < nz ? (int)sz : 0;
  return C;
}

// map_coordinates order 1 of a float32 image at an in-box point: the sum of
// ((v * wx) * wy) * wz over the corners, z fastest, in float64, to float32
// (identical for 'nearest' and 'constant' inside the box)
__device__ __forceinline__ float tri_f32(const float* __restrict__ img, const Corners& C) {
  const float* p = img + C.o;
  double v = 0.0;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const double val = (double)p[a * C.ox + b * C.oy + c * C.oz];
        v = add_rn(v, mul_rn(mul_rn(mul_rn(val, C.wx[a]), C.wy[b]), C.wz[c]));
      }
  return __double2float_rn(v);
}

// Normal equations: each thread strides over voxels v, v + S, ... (S the
// grid's thread count; (x, y, z) stepped incrementally) and keeps the 90
// float64 sums in registers -- the 78 upper-triangle entries of
// H = sum J J^T row by row, then b = sum J e -- exact float32 products
// accumulated with fma.  A voxel whose sample location leaves the box or
// whose g is invalid has J = 0 and adds nothing (finite inputs), so it is
// skipped.  The wavefront's sums are combined by a fixed butterfly; lane 0
// writes the partials [wave][90].
__global__ __launch_bounds__(VDVC_BLOCK) void affine_normal_kernel(AffIn A, double* part) {
  const int lane = threadIdx.x & 63;
  const long long wave = ((long long)blockIdx.x * VDVC_BLOCK + threadIdx.x) >> 6;
  const long long n = (long long)A.nx * A.ny * A.nz;
  const long long S = (long long)gridDim.x * VDVC_BLOCK;
  long long v = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  int x, y, z, dx, dy, dz;
  split_index(v, A.nx, A.ny, x, y, z);
  split_index(S, A.nx, A.ny, dx, dy, dz);
  double acc[90];
#pragma unroll
  for (int e = 0; e < 90; ++e) acc[e] = 0.0;
  // valid and g of the next voxel are loaded one iteration ahead, so the
  // gathers of a voxel do not wait for its mask
  uint8_t vnext = v < n ? A.valid[v] : (uint8_t)0;
  float gnext = v < n ? A.g[v] : 0.0f;
  for (; v < n; v += S) {
    const uint8_t vcur = vnext;
    const float gcur = gnext;
    if (v + S < n) {
      vnext = A.valid[v + S];
      gnext = A.g[v + S];
    }
    float Xc[3], s[3];
    affine_coords(A, x, y, z, Xc, s);
    if (vcur != 0 && in_box(s, A.nx, A.ny, A.nz)) {
      const Corners C = in_box_corners(s, A.nx, A.ny, A.nz);
      const float warped = tri_f32(A.f, C);
      const float ng[3] = {-tri_f32(A.gx, C), -tri_f32(A.gy, C), -tri_f32(A.gz, C)};
      double J[13];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) J[3 * i + j] = (double)mul_rn(ng[i], Xc[j]);
        J[9 + i] = (double)ng[i];
      }
      J[12] = (double)sub_rn(gcur, warped);
      int e = 0;
#pragma unroll
      for (int a = 0; a < 12; ++a)
#pragma unroll
        for (int b = a; b < 12; ++b, ++e) acc[e] = fma(J[a], J[b], acc[e]);
#pragma unroll
      for (int a = 0; a < 12; ++a) acc[78 + a] = fma(J[a], J[12], acc[78 + a]);
    }
    x += dx; y += dy; z += dz;
    if (x >= A.nx) { x -= A.nx; ++y; }
    if (y >= A.ny) { y -= A.ny; ++z; }
  }
#pragma unroll
  for (int e = 0; e < 90; ++e) {
    double t = acc[e];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    if (lane == 0) part[wave * 90 + e] = t;
  }
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
      const float e = sub_rn(A.g[v], tri_f32(A.f, in_box_corners(s, A.nx, A.ny, A.nz)));
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
  out[i] = (float)as_img_type(sample01<GT_F32>(g, W.gnx, W.gny, W.gnz, p[0], p[1], p[2], 1, 1, 0.0,
                                               vdvc_img::img_rsrc(g, 0u)),
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
  // the LDS-tiled pass for radii up to 32 (tiles of at most 32 KB) and an
  // extension value float32 holds exactly
  if (G.r <= 32 && (G.mode == 0 || (double)(float)G.cval == G.cval)) {
    const unsigned long long nxt = (unsigned long long)(G.nx + kGaussRowTile - 1) / kGaussRowTile;
    const unsigned long long nxc = (unsigned long long)(G.nx + kGaussColX - 1) / kGaussColX;
    const int len = G.axis == 1 ? G.ny : G.nz;
    const unsigned long long nt = (unsigned long long)(len + kGaussColT - 1) / kGaussColT;
    const unsigned long long nb = G.axis == 0   ? nxt * G.ny * G.nz
                                  : G.axis == 1 ? nt * nxc * G.nz
                                                : nt * nxc * G.ny;
    if (nb > 0x7fffffffull) return (int)hipErrorInvalidValue;
    const size_t lds = G.axis == 0 ? (size_t)(kGaussRowTile + 2 * G.r) * sizeof(float)
                                   : (size_t)(kGaussColT + 2 * G.r) * kGaussColX * sizeof(float);
    if (G.axis == 0)
      hipLaunchKernelGGL(gauss_tile_kernel<0>, dim3((unsigned)nb), dim3(VDVC_BLOCK), lds, st, G,
                         in, out);
    else if (G.axis == 1)
      hipLaunchKernelGGL(gauss_tile_kernel<1>, dim3((unsigned)nb), dim3(VDVC_BLOCK), lds, st, G,
                         in, out);
    else
      hipLaunchKernelGGL(gauss_tile_kernel<2>, dim3((unsigned)nb), dim3(VDVC_BLOCK), lds, st, G,
                         in, out);
    return (int)hipGetLastError();
  }
  const unsigned long long nb =
      (unsigned long long)((G.nx + VDVC_BLOCK - 1) / VDVC_BLOCK) * G.ny * G.nz;
  // 32-bit tap offsets k * stride (k <= r) and a 32-bit 1-D grid
  if (nb > 0x7fffffffull || (long long)G.nx * G.ny * (G.r > 0 ? G.r : 1) > 0x7fffffffll)
    return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(gauss_axis_kernel, dim3((unsigned)nb), dim3(VDVC_BLOCK), 0, st, G, in, out);
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
