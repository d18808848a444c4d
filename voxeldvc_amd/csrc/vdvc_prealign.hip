// Pyramid restriction and pre-alignment kernels.
//
//   stride restriction   engine/correlate_gpu.py:1256-1258 (f[::s, ::s, ::s])
//   FFT phase shift      engine/correlate_gpu.py:381-414   (fft_rigid_shift)
//   affine warp + ZNCC   engine/correlate_gpu.py:417-433, 505-536
//
// The phase correlation runs on the coarsest pyramid level only (tens to a
// few hundred voxels per axis), so it uses separable direct DFTs in FP64
// (exact twiddles from sincospi) instead of an FFT library: per axis each
// output sample is one dot product over the line.
#include <algorithm>

#include "vdvc_kernels.h"

template <typename T>
__global__ __launch_bounds__(VDVC_BLOCK) void stride_kernel(const T* __restrict__ src, int nx,
                                                            int ny, int nz, int s,
                                                            T* __restrict__ dst, int mx, int my,
                                                            int mz) {
  const long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  const long long n = (long long)mx * my * mz;
  if (q >= n) return;
  const int x = (int)(q % mx);
  const long long r = q / mx;
  const int y = (int)(r % my);
  const int z = (int)(r / my);
  dst[q] = src[(long long)x * s + (long long)nx * ((long long)y * s + (long long)ny * ((long long)z * s))];
}

// one DFT pass along `axis` of a complex (interleaved double2) volume with
// dims (nx,ny,nz) x fastest; real input when `src_real` is set.  The block
// first fills an LDS table of the n twiddles (cos, sign * sin)(2 pi m / n)
// from sincospi; output kk then walks m = kk * t mod n incrementally.
__global__ __launch_bounds__(VDVC_BLOCK) void dft_pass_kernel(const double2* __restrict__ src,
                                                              const float* __restrict__ src_real,
                                                              double2* __restrict__ dst, int nx,
                                                              int ny, int nz, int axis,
                                                              double sign) {
  extern __shared__ double2 tw[];
  const int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
  for (int m = threadIdx.x; m < n; m += VDVC_BLOCK) {
    double sn, cs;
    sincospi(2.0 * (double)m / (double)n, &sn, &cs);
    tw[m] = make_double2(cs, sn * sign);
  }
  __syncthreads();
  const long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  const long long ntot = (long long)nx * ny * nz;
  if (q >= ntot) return;
  const int x = (int)(q % nx);
  const long long r = q / nx;
  const int y = (int)(r % ny);
  const int z = (int)(r / ny);
  int kk;
  long long stride, base;
  if (axis == 0) { kk = x; stride = 1; base = q - x; }
  else if (axis == 1) { kk = y; stride = nx; base = q - (long long)y * nx; }
  else { kk = z; stride = (long long)nx * ny; base = q - (long long)z * nx * ny; }
  double re = 0.0, im = 0.0;
  int m = 0;
  for (int t = 0; t < n; ++t) {
    const double cs = tw[m].x, sn = tw[m].y;
    if (src_real) {
      const double a = (double)src_real[base + t * stride];
      re += a * cs;
      im += a * sn;
    } else {
      const double2 a = src[base + t * stride];
      re += a.x * cs - a.y * sn;
      im += a.x * sn + a.y * cs;
    }
    m += kk;
    if (m >= n) m -= n;
  }
  dst[q] = make_double2(re, im);
}

// The same DFT pass over groups of DFT_LB lines staged in LDS.  Block b
// takes lines li = DFT_LB b + l (l < DFT_LB) of the pass's line enumeration,
// in which the fastest memory axis other than the transform axis varies
// fastest, so that a group's loads and stores are contiguous runs.  LDS
// holds the n twiddles and the lines as buf[t * DFT_LP + l] (a padded row
// per t); thread (l = tid % DFT_LB, g = tid / DFT_LB) computes outputs
// k = g + 16 j (j < KJ = ceil(n / 16); a k >= n is computed as k = 0 and
// not stored) of line l: the sum over t in order, one fused multiply-add
// per real product, the twiddle index m_j = k t mod n kept incrementally.
// The outputs go back through the LDS buffer so that the stores are as
// contiguous as the loads.  Each line is read from global memory once
// (dft_pass_kernel reads it once per output sample).  REAL: real input
// (float32 volume), imaginary part zero.
constexpr int DFT_LB = 16;
constexpr int DFT_LP = DFT_LB + 1;

template <int KJ, bool REAL>
__global__ __launch_bounds__(VDVC_BLOCK) void dft_lines_kernel(const double2* __restrict__ src,
                                                               const float* __restrict__ src_real,
                                                               double2* __restrict__ dst, int nx,
                                                               int ny, int nz, int axis,
                                                               double sign) {
  extern __shared__ double2 sm[];
  const int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
  double2* tw = sm;
  double2* buf = sm + n;
  for (int m = threadIdx.x; m < n; m += VDVC_BLOCK) {
    double sn, cs;
    sincospi(2.0 * (double)m / (double)n, &sn, &cs);
    tw[m] = make_double2(cs, sn * sign);
  }
  const long long nlines = (long long)nx * ny * nz / n;
  const long long li0 = (long long)blockIdx.x * DFT_LB;
  const long long stride = axis == 0 ? 1 : (axis == 1 ? (long long)nx : (long long)nx * ny);
  auto line_base = [&](long long li) -> long long {
    if (axis == 0) return li * nx;                       // (y, z) -> row start
    const long long x = li % nx, o = li / nx;            // o = z (axis 1) or y (axis 2)
    return axis == 1 ? x + (long long)nx * ny * o : x + (long long)nx * o;
  };
  // load: along the row for axis 0 (t fastest), across lines otherwise
  const int nq = n * DFT_LB;
  for (int q = threadIdx.x; q < nq; q += VDVC_BLOCK) {
    const int t = axis == 0 ? q % n : q / DFT_LB;
    const int l = axis == 0 ? q / n : q % DFT_LB;
    const long long li = li0 + l;
    double2 a = make_double2(0.0, 0.0);
    if (li < nlines) {
      const long long at = line_base(li) + (long long)t * stride;
      if (REAL) a.x = (double)src_real[at];
      else a = src[at];
    }
    buf[t * DFT_LP + l] = a;
  }
  __syncthreads();
  const int l = threadIdx.x % DFT_LB, g = threadIdx.x / DFT_LB;
  double re[KJ], im[KJ];
  int kk[KJ], mm[KJ];
#pragma unroll
  for (int j = 0; j < KJ; ++j) {
    re[j] = 0.0; im[j] = 0.0;
    kk[j] = g + 16 * j < n ? g + 16 * j : 0;
    mm[j] = 0;
  }
  for (int t = 0; t < n; ++t) {
    const double2 a = buf[t * DFT_LP + l];
#pragma unroll
    for (int j = 0; j < KJ; ++j) {
      const double2 w = tw[mm[j]];
      re[j] = fma(a.x, w.x, re[j]);
      im[j] = fma(a.x, w.y, im[j]);
      if (!REAL) {
        re[j] = fma(-a.y, w.y, re[j]);
        im[j] = fma(a.y, w.x, im[j]);
      }
      const int m = mm[j] + kk[j];
      mm[j] = m >= n ? m - n : m;
    }
  }
  __syncthreads();  // every thread has read the lines: the buffer takes the outputs
#pragma unroll
  for (int j = 0; j < KJ; ++j)
    if (g + 16 * j < n) buf[(g + 16 * j) * DFT_LP + l] = make_double2(re[j], im[j]);
  __syncthreads();
  for (int q = threadIdx.x; q < nq; q += VDVC_BLOCK) {
    const int k = axis == 0 ? q % n : q / DFT_LB;
    const int ll = axis == 0 ? q / n : q % DFT_LB;
    const long long li = li0 + ll;
    if (li < nlines) dst[line_base(li) + (long long)k * stride] = buf[k * DFT_LP + ll];
  }
}

// R = F conj(G) / (|F conj(G)| + 1e-12), in place into F
__global__ __launch_bounds__(VDVC_BLOCK) void cross_power_kernel(double2* __restrict__ F,
                                                                 const double2* __restrict__ G,
                                                                 long long n) {
  const long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  if (q >= n) return;
  const double2 a = F[q], b = G[q];
  const double re = a.x * b.x + a.y * b.y;
  const double im = a.y * b.x - a.x * b.y;
  const double mag = sqrt(re * re + im * im) + 1e-12;
  F[q] = make_double2(re / mag, im / mag);
}

// argmax of the real part; ties resolve to the smallest C-order index
// ((x*ny + y)*nz + z), as NumPy's argmax over a C-ordered (x,y,z) array.
// Per-block (value, index) partials, then one block over the partials; the
// lexicographic rule gives the same winner in any reduction order.
struct ArgBest {
  double v;
  long long i;
};

__device__ __forceinline__ void arg_take(ArgBest& a, double v, long long i) {
  if (v > a.v || (v == a.v && i < a.i)) { a.v = v; a.i = i; }
}

__device__ __forceinline__ ArgBest arg_block(ArgBest a) {
  __shared__ double lv[VDVC_BLOCK];
  __shared__ long long li[VDVC_BLOCK];
  lv[threadIdx.x] = a.v;
  li[threadIdx.x] = a.i;
  __syncthreads();
  for (int o = VDVC_BLOCK / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      const double v2 = lv[threadIdx.x + o];
      const long long i2 = li[threadIdx.x + o];
      if (v2 > lv[threadIdx.x] || (v2 == lv[threadIdx.x] && i2 < li[threadIdx.x])) {
        lv[threadIdx.x] = v2;
        li[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  return ArgBest{lv[0], li[0]};
}

__global__ __launch_bounds__(VDVC_BLOCK) void argmax_part_kernel(const double2* __restrict__ R,
                                                                 int nx, int ny, int nz,
                                                                 ArgBest* __restrict__ part) {
  const long long n = (long long)nx * ny * nz;
  ArgBest a{-INFINITY, 0x7fffffffffffffffLL};
  for (long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x; q < n;
       q += (long long)gridDim.x * VDVC_BLOCK) {
    const int x = (int)(q % nx);
    const long long r = q / nx;
    const int y = (int)(r % ny);
    const int z = (int)(r / ny);
    arg_take(a, R[q].x, ((long long)x * ny + y) * nz + z);
  }
  const ArgBest b = arg_block(a);
  if (threadIdx.x == 0) part[blockIdx.x] = b;
}

__global__ __launch_bounds__(VDVC_BLOCK) void argmax_final_kernel(const ArgBest* __restrict__ part,
                                                                  int nparts,
                                                                  long long* __restrict__ out) {
  ArgBest a{-INFINITY, 0x7fffffffffffffffLL};
  for (int k = threadIdx.x; k < nparts; k += VDVC_BLOCK) arg_take(a, part[k].v, part[k].i);
  const ArgBest b = arg_block(a);
  if (threadIdx.x == 0) out[0] = b.i;
}

// affine warp x' = A(x - c) + c + t of every voxel of f_d, order-1 sample of
// g_d (double weights, float result), FP64 partial sums over in-bounds
// samples: n, sum f, sum g, sum f^2, sum g^2, sum f g

__global__ __launch_bounds__(VDVC_BLOCK) void affine_cost_kernel(
    const float* __restrict__ f, int nx, int ny, int nz, const float* __restrict__ g, int gnx,
    int gny, int gnz, AffineIn P, double* __restrict__ part) {
  __shared__ double lds[4 * 6];
  const long long n = (long long)nx * ny * nz;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x; q < n;
       q += (long long)gridDim.x * VDVC_BLOCK) {
    const int x = (int)(q % nx);
    const long long r = q / nx;
    const int y = (int)(r % ny);
    const int z = (int)(r / ny);
    const double dx = x - P.c[0], dy = y - P.c[1], dz = z - P.c[2];
    const double X = P.A[0] * dx + P.A[1] * dy + P.A[2] * dz + P.c[0] + P.t[0];
    const double Y = P.A[3] * dx + P.A[4] * dy + P.A[5] * dz + P.c[1] + P.t[1];
    const double Z = P.A[6] * dx + P.A[7] * dy + P.A[8] * dz + P.c[2] + P.t[2];
    if (!(X >= 0.0 && X <= gnx - 1 && Y >= 0.0 && Y <= gny - 1 && Z >= 0.0 && Z <= gnz - 1))
      continue;
    const int ix = (int)floor(X), iy = (int)floor(Y), iz = (int)floor(Z);
    const int jx = min(ix + 1, gnx - 1), jy = min(iy + 1, gny - 1), jz = min(iz + 1, gnz - 1);
    const double fx = X - ix, fy = Y - iy, fz = Z - iz;
    const long long sy = gnx, sz = (long long)gnx * gny;
    auto G = [&](int a, int b, int cc) -> double { return (double)g[a + b * sy + cc * sz]; };
    const double gv = (1 - fz) * ((1 - fy) * ((1 - fx) * G(ix, iy, iz) + fx * G(jx, iy, iz)) +
                                  fy * ((1 - fx) * G(ix, jy, iz) + fx * G(jx, jy, iz))) +
                      fz * ((1 - fy) * ((1 - fx) * G(ix, iy, jz) + fx * G(jx, iy, jz)) +
                            fy * ((1 - fx) * G(ix, jy, jz) + fx * G(jx, jy, jz)));
    const double gs = (double)(float)gv;
    const double fv = (double)f[q];
    acc[0] += 1.0;
    acc[1] += fv;
    acc[2] += gs;
    acc[3] += fv * fv;
    acc[4] += gs * gs;
    acc[5] += fv * gs;
  }
  block_sum_f64<6>(acc, lds);
  if (threadIdx.x == 0)
    for (int k = 0; k < 6; ++k) part[6 * blockIdx.x + k] = acc[k];
}

namespace vdvc {

int launch_stride(int type, const void* src, int nx, int ny, int nz, int s, void* dst, int mx,
                  int my, int mz, hipStream_t st) {
  const long long n = (long long)mx * my * mz;
  const unsigned nb = (unsigned)((n + VDVC_BLOCK - 1) / VDVC_BLOCK);
  if (type == GT_F32)
    hipLaunchKernelGGL(stride_kernel<float>, dim3(nb), dim3(VDVC_BLOCK), 0, st,
                       (const float*)src, nx, ny, nz, s, (float*)dst, mx, my, mz);
  else if (type == GT_F64)
    hipLaunchKernelGGL(stride_kernel<double>, dim3(nb), dim3(VDVC_BLOCK), 0, st,
                       (const double*)src, nx, ny, nz, s, (double*)dst, mx, my, mz);
  else
    hipLaunchKernelGGL(stride_kernel<__half>, dim3(nb), dim3(VDVC_BLOCK), 0, st,
                       (const __half*)src, nx, ny, nz, s, (__half*)dst, mx, my, mz);
  return (int)hipGetLastError();
}

// phase correlation: f, g float32 volumes (x fastest, same dims, each axis
// <= 4096); work holds 3 complex volumes, part VDVC_ARG_PARTS (value, index)
// pairs; out_idx receives the C-order argmax index
int launch_phase_corr(const float* f, const float* g, int nx, int ny, int nz, double2* work,
                      void* part, long long* out_idx, hipStream_t st) {
  const long long n = (long long)nx * ny * nz;
  const unsigned nb = (unsigned)((n + VDVC_BLOCK - 1) / VDVC_BLOCK);
  double2* A = work;
  double2* B = work + n;
  double2* C = work + 2 * n;
  dim3 g3(nb), b3(VDVC_BLOCK);
  // one DFT pass along `axis`: lines staged in LDS in groups of DFT_LB for
  // axes whose lines and twiddles fit 64 KiB of LDS (<= 227 samples), else
  // the per-output kernel
  auto pass = [&](const double2* in, const float* in_real, double2* out, int axis,
                  double sign) {
    const int na = axis == 0 ? nx : (axis == 1 ? ny : nz);
    if ((size_t)na * (1 + DFT_LP) * sizeof(double2) <= 65536) {
      const long long nlines = n / na;
      const dim3 gl((unsigned)((nlines + DFT_LB - 1) / DFT_LB));
      const size_t lds = (size_t)na * (1 + DFT_LP) * sizeof(double2);
      const int kj = (na + 15) / 16;  // 1..15
#define VDVC_DFT_L(KJ)                                                                        \
  case KJ:                                                                                    \
    if (in_real)                                                                              \
      hipLaunchKernelGGL((dft_lines_kernel<KJ, true>), gl, b3, lds, st, in, in_real, out, nx, \
                         ny, nz, axis, sign);                                                 \
    else                                                                                      \
      hipLaunchKernelGGL((dft_lines_kernel<KJ, false>), gl, b3, lds, st, in, in_real, out, nx,\
                         ny, nz, axis, sign);                                                 \
    break;
      switch (kj) {
        VDVC_DFT_L(1) VDVC_DFT_L(2) VDVC_DFT_L(3) VDVC_DFT_L(4) VDVC_DFT_L(5)
        VDVC_DFT_L(6) VDVC_DFT_L(7) VDVC_DFT_L(8) VDVC_DFT_L(9) VDVC_DFT_L(10)
        VDVC_DFT_L(11) VDVC_DFT_L(12) VDVC_DFT_L(13) VDVC_DFT_L(14) VDVC_DFT_L(15)
        default: break;
      }
#undef VDVC_DFT_L
    } else {
      hipLaunchKernelGGL(dft_pass_kernel, g3, b3, (size_t)na * sizeof(double2), st, in, in_real,
                         out, nx, ny, nz, axis, sign);
    }
  };
  // F = DFT(f): x (real) -> B, y -> C, z -> A
  pass(nullptr, f, B, 0, -1.0);
  pass(B, nullptr, C, 1, -1.0);
  pass(C, nullptr, A, 2, -1.0);
  // G = DFT(g) into C (via B)
  pass(nullptr, g, B, 0, -1.0);
  pass(B, nullptr, C, 1, -1.0);
  pass(C, nullptr, B, 2, -1.0);
  hipLaunchKernelGGL(cross_power_kernel, g3, b3, 0, st, A, B, n);
  // inverse DFT (positive sign; the 1/N scale does not move the argmax)
  pass(A, nullptr, B, 0, 1.0);
  pass(B, nullptr, C, 1, 1.0);
  pass(C, nullptr, A, 2, 1.0);
  const int np = (int)std::min<long long>(VDVC_ARG_PARTS, (n + VDVC_BLOCK - 1) / VDVC_BLOCK);
  hipLaunchKernelGGL(argmax_part_kernel, dim3((unsigned)np), b3, 0, st, A, nx, ny, nz,
                     (ArgBest*)part);
  hipLaunchKernelGGL(argmax_final_kernel, dim3(1), b3, 0, st, (const ArgBest*)part, np, out_idx);
  return (int)hipGetLastError();
}

int launch_affine_cost(const float* f, int nx, int ny, int nz, const float* g, int gnx, int gny,
                       int gnz, const AffineIn& P, double* part, int nblocks, hipStream_t st) {
  hipLaunchKernelGGL(affine_cost_kernel, dim3(nblocks), dim3(VDVC_BLOCK), 0, st, f, nx, ny, nz,
                     g, gnx, gny, gnz, P, part);
  return (int)hipGetLastError();
}

}  // namespace vdvc
