// Voxel-streaming kernels: correlation-Hessian matvec, its diagonal, the
// fused residual + right-hand side, deformed-image sampling, reference-image
// setup (gradient + erosion) and the multiplicity-weighted masked statistics.
//
// Reference semantics followed:
//   matvec      engine/kernels_dvc.py:452-506   (matvec_H)
//   diagonal    engine/kernels_dvc.py:509-543   (build_diagonal_H)
//   rhs         engine/correlate_gpu.py:341-378, 1091-1107
//   sampling    engine/correlate_gpu.py:201-281 (sample_deformed_field)
//   setup       engine/correlate_gpu.py:49-94, 951-953
//   statistics  engine/correlate_gpu.py:113-157
//
// All restriction kernels share one decomposition (TilesDev): a wavefront
// owns a 64-voxel x strip (lanes along x, so every load instruction reads a
// contiguous 64-voxel run), one element row in y and `ez` elements in z, and
// walks rows then planes.  Per voxel it forms the scalar s (W m g.u, W m res,
// or W m g_k^2 for the diagonal), restricts s*g along y inside the lane
// (running sums), along z by carrying two node planes in registers, and
// along x by a segmented reduction across the lanes of one element at the
// end of every element layer.  Node partials go to the wave's slot block.
#include <type_traits>

#include "vdvc_kernels.h"



// Emit one node plane of a wave: c[cy*3+k] is the lane's z-complete
// y-restricted sum for node row cy; a0/a1 are the lane's x weights for the
// owning element's lower/upper node.  Writes nn node slots.
template <bool POW2, typename R>
__device__ __forceinline__ void emit_plane(const R (&c)[6], R a0, R a1,
                                           int lane, int h, int x0, int nact, int i_lo,
                                           int nn, R* __restrict__ slot, int px,
                                           R* lds /* 64*13 values */) {
  if (POW2) {
    R s0[6], s1[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) { s0[q] = a0 * c[q]; s1[q] = a1 * c[q]; }
    for (int o = 1; o < h; o <<= 1) {
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        s0[q] += __shfl_xor(s0[q], o, 64);
        s1[q] += __shfl_xor(s1[q], o, 64);
      }
    }
    R pv[6];
    const int src = lane > 0 ? lane - 1 : 0;
#pragma unroll
    for (int q = 0; q < 6; ++q) pv[q] = __shfl(s1[q], src, 64);
    if ((lane % h) == 0) {
      const int ln = lane / h;
      if (ln < nn) {
        R* d = slot + (size_t)ln * 3;
#pragma unroll
        for (int cy = 0; cy < 2; ++cy)
#pragma unroll
          for (int k = 0; k < 3; ++k)
            d[(size_t)cy * px * 3 + k] = s0[cy * 3 + k] + (ln > 0 ? pv[cy * 3 + k] : R(0));
      }
    }
    if (lane == 63) {
      const int ln = 64 / h;
      if (ln < nn) {
        R* d = slot + (size_t)ln * 3;
#pragma unroll
        for (int cy = 0; cy < 2; ++cy)
#pragma unroll
          for (int k = 0; k < 3; ++k) d[(size_t)cy * px * 3 + k] = s1[cy * 3 + k];
      }
    }
  } else {
    // generic element size: stage lane values in LDS, node lanes sum them
    // (slot 12 holds the lane's element index, an exactly representable R)
    R* L = lds;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      L[lane * 13 + q] = a0 * c[q];
      L[lane * 13 + 6 + q] = a1 * c[q];
    }
    const int x = x0 + lane;
    L[lane * 13 + 12] = (R)(x / h - i_lo);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int ln = lane; ln < nn; ln += 64) {
      const int i = i_lo + ln;
      int lo = (i - 1) * h - x0;
      int hi = (i + 1) * h - x0;
      lo = lo < 0 ? 0 : lo;
      hi = hi > nact - 1 ? nact - 1 : hi;
      R v[6] = {0, 0, 0, 0, 0, 0};
      for (int l = lo; l <= hi; ++l) {
        const int el = (int)L[l * 13 + 12];
        if (el == ln) {
#pragma unroll
          for (int q = 0; q < 6; ++q) v[q] += L[l * 13 + q];
        } else if (el == ln - 1) {
#pragma unroll
          for (int q = 0; q < 6; ++q) v[q] += L[l * 13 + 6 + q];
        }
      }
      R* d = slot + (size_t)ln * 3;
#pragma unroll
      for (int cy = 0; cy < 2; ++cy)
#pragma unroll
        for (int k = 0; k < 3; ++k) d[(size_t)cy * px * 3 + k] = v[cy * 3 + k];
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

template <int MODE, typename GT, bool HASMASK, typename FT, bool POW2, bool FOLD>
__global__ __launch_bounds__(VDVC_BLOCK) void vox_restrict_kernel(
    MeshDev M, TilesDev T, VoxIn in, const typename RealOf<GT>::type* __restrict__ v,
    typename RealOf<GT>::type* __restrict__ P) {
  using R = typename RealOf<GT>::type;
  __shared__ R lds_all[POW2 ? 1 : VDVC_WAVES_PER_BLOCK * 64 * 13];
  const int lane = threadIdx.x & 63;
  const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const unsigned blk = xcd_remap(blockIdx.x, gridDim.x);
  const long long w = (long long)blk * VDVC_WAVES_PER_BLOCK + wib;
  if (w >= T.nwaves) return;
  if (in.done && *in.done) return;
  R* lds = POW2 ? nullptr : lds_all + wib * 64 * 13;

  const int h = M.h;
  const R inv_h = R(1) / R(h);
  const int sx = (int)(w % T.nsx);
  const long long r = w / T.nsx;
  const int ey = (int)(r % M.Ny);
  const int zt = (int)(r / M.Ny);

  const int x0 = sx * 64;
  const int nact = min(64, M.Nvx - x0);
  const int x = x0 + lane;
  const bool act = lane < nact;
  const int xe = act ? x : x0 + nact - 1;
  const int ex = xe / h;               // owning element (== Nx for the last plane)
  const int tx = xe - ex * h;
  const R X1 = tx * inv_h;
  const R X0 = R(1) - X1;
  const R Wx = act ? (R)axis_weight(xe, M.Nvx - 1) : R(0);
  const int i0 = min(ex, M.Nx);
  const int i1 = min(ex + 1, M.Nx);
  const int i_lo = x0 / h;
  const int x_last = x0 + nact - 1;
  const int nn = min(M.Nx, x_last / h + 1) - i_lo + 1;
  const int mx = axis_mult(xe, h, M.Nvx - 1);

  const int ez0 = zt * T.ez;
  const int ez1 = min(M.Nz, ez0 + T.ez);
  const int y0 = ey * h;
  const int nrows = h + (ey == M.Ny - 1 ? 1 : 0);

  R a0w = X0, a1w = X1;
  if (MODE == MODE_DIAG) { a0w = X0 * X0; a1w = X1 * X1; }

  R C[2][2][3];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int k = 0; k < 3; ++k) C[a][b][k] = R(0);

  double st_c = 0.0, st_1 = 0.0, st_2 = 0.0, st_n = 0.0;

  const GT* g0 = reinterpret_cast<const GT*>(in.g[0]);
  const GT* g1 = reinterpret_cast<const GT*>(in.g[1]);
  const GT* g2 = reinterpret_cast<const GT*>(in.g[2]);
  const R* gvals = reinterpret_cast<const R*>(in.gvals);
  R* resv = reinterpret_cast<R*>(in.res);
  const R mean0 = (R)in.mean0, gmean = (R)in.gmean, kscale = (R)in.kscale;
  R* Pw = P + (size_t)w * T.wave_stride;

  for (int ez = ez0; ez < ez1; ++ez) {
    const int nzp = h + (ez == M.Nz - 1 ? 1 : 0);
    R Ux[2][2][3];
    if (MODE == MODE_MATVEC) {
#pragma unroll
      for (int cy = 0; cy < 2; ++cy)
#pragma unroll
        for (int cz = 0; cz < 2; ++cz) {
          const long long n0 = node_index(M, i0, ey + cy, ez + cz);
          const long long n1 = node_index(M, i1, ey + cy, ez + cz);
#pragma unroll
          for (int k = 0; k < 3; ++k)
            Ux[cy][cz][k] = X0 * v[3 * n0 + k] + X1 * v[3 * n1 + k];
        }
    }
    for (int tz = 0; tz < nzp; ++tz) {
      const int z = ez * h + tz;
      const R Z1 = tz * inv_h;
      const R Z0 = R(1) - Z1;
      const R Wxz = Wx * (R)axis_weight(z, M.Nvz - 1);
      const int mxz = mx * axis_mult(z, h, M.Nvz - 1);
      R u0[3] = {0, 0, 0}, du[3] = {0, 0, 0};
      if (MODE == MODE_MATVEC) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          u0[k] = Z0 * Ux[0][0][k] + Z1 * Ux[0][1][k];
          const R u1 = Z0 * Ux[1][0][k] + Z1 * Ux[1][1][k];
          du[k] = (u1 - u0[k]) * inv_h;
        }
      }
      R A[3] = {0, 0, 0}, B[3] = {0, 0, 0};
      const long long base = ((long long)z * M.Nvy + y0) * M.pitch + xe;
#pragma unroll 4
      for (int ty = 0; ty < nrows; ++ty) {
        const long long vi = base + (long long)ty * M.pitch;
        const int y = y0 + ty;
        const R W = Wxz * (R)axis_weight(y, M.Nvy - 1);
        const R gx = ld_val(g0, vi);
        const R gy = ld_val(g1, vi);
        const R gz = ld_val(g2, vi);
        R m = R(1);
        if (HASMASK) m = in.mask[vi] ? R(1) : R(0);
        const R fty = (R)ty;
        if (MODE == MODE_MATVEC) {
          const R ux = u0[0] + fty * du[0];
          const R uy = u0[1] + fty * du[1];
          const R uz = u0[2] + fty * du[2];
          const R s = W * m * (gx * ux + gy * uy + gz * uz);
          const R st = s * fty;
          A[0] += s * gx; A[1] += s * gy; A[2] += s * gz;
          B[0] += st * gx; B[1] += st * gy; B[2] += st * gz;
        } else if (MODE == MODE_RHS) {
          const uint8_t bm = in.mask[vi];
          R rv = R(0);
          if (act) {
            const R fval = ld_val(reinterpret_cast<const FT*>(in.fimg), vi);
            const R gv = gvals[vi];
            // (g - g_mean) * (-(std0/std1)) + (f - mean0), single rounded
            // steps in the reference's order, no contraction
            rv = add_rn(mul_rn(sub_rn(gv, gmean), kscale), sub_rn(fval, mean0));
            if (!bm) rv = R(0);
            if (resv) resv[vi] = rv;
            if (in.stats && in.live[vi]) {
              const double mlt = (double)(mxz * axis_mult(y, h, M.Nvy - 1));
              const double rd = (double)rv;
              st_c += mlt;
              st_1 += mlt * rd;
              st_2 += mlt * rd * rd;
              st_n += 1.0;
            }
            if constexpr (FOLD) {
              const float fm = bm ? 1.0f : 0.0f;
              in.fold_out[0][vi] = __bfloat16_as_ushort(__float2bfloat16((float)gx * fm));
              in.fold_out[1][vi] = __bfloat16_as_ushort(__float2bfloat16((float)gy * fm));
              in.fold_out[2][vi] = __bfloat16_as_ushort(__float2bfloat16((float)gz * fm));
            }
          }
          const R s = W * (bm ? R(1) : R(0)) * rv;
          const R st = s * fty;
          A[0] += s * gx; A[1] += s * gy; A[2] += s * gz;
          B[0] += st * gx; B[1] += st * gy; B[2] += st * gz;
        } else {  // MODE_DIAG
          const R Y1 = fty * inv_h;
          const R Y0 = R(1) - Y1;
          const R wm = W * m;
          const R q0 = wm * gx * gx, q1 = wm * gy * gy, q2 = wm * gz * gz;
          const R y0w = Y0 * Y0, y1w = Y1 * Y1;
          A[0] += y0w * q0; A[1] += y0w * q1; A[2] += y0w * q2;
          B[0] += y1w * q0; B[1] += y1w * q1; B[2] += y1w * q2;
        }
      }
      R Rr[2][3];
      if (MODE == MODE_DIAG) {
#pragma unroll
        for (int k = 0; k < 3; ++k) { Rr[0][k] = A[k]; Rr[1][k] = B[k]; }
      } else {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          Rr[1][k] = B[k] * inv_h;
          Rr[0][k] = A[k] - Rr[1][k];
        }
      }
      R zw0 = Z0, zw1 = Z1;
      if (MODE == MODE_DIAG) { zw0 = Z0 * Z0; zw1 = Z1 * Z1; }
#pragma unroll
      for (int cy = 0; cy < 2; ++cy)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          C[cy][0][k] += zw0 * Rr[cy][k];
          C[cy][1][k] += zw1 * Rr[cy][k];
        }
    }
    // node plane kz = ez is complete for this wave: emit and shift
    {
      R c[6];
#pragma unroll
      for (int cy = 0; cy < 2; ++cy)
#pragma unroll
        for (int k = 0; k < 3; ++k) c[cy * 3 + k] = C[cy][0][k];
      emit_plane<POW2, R>(c, a0w, a1w, lane, h, x0, nact, i_lo, nn,
                          Pw + (size_t)(ez - ez0) * 2 * T.px * 3, T.px, lds);
#pragma unroll
      for (int cy = 0; cy < 2; ++cy)
#pragma unroll
        for (int k = 0; k < 3; ++k) { C[cy][0][k] = C[cy][1][k]; C[cy][1][k] = R(0); }
    }
  }
  {
    R c[6];
#pragma unroll
    for (int cy = 0; cy < 2; ++cy)
#pragma unroll
      for (int k = 0; k < 3; ++k) c[cy * 3 + k] = C[cy][0][k];
    emit_plane<POW2, R>(c, a0w, a1w, lane, h, x0, nact, i_lo, nn,
                        Pw + (size_t)(ez1 - ez0) * 2 * T.px * 3, T.px, lds);
  }
  if (MODE == MODE_RHS && in.stats) {
    st_c = wave_sum(st_c);
    st_1 = wave_sum(st_1);
    st_2 = wave_sum(st_2);
    st_n = wave_sum(st_n);
    if (lane == 0) {
      in.stats[4 * w + 0] = st_c;
      in.stats[4 * w + 1] = st_1;
      in.stats[4 * w + 2] = st_2;
      in.stats[4 * w + 3] = st_n;
    }
  }
}

// ---------------------------------------------------------------------------
// deformed-image sampling g(x + u(x)) + active mask + statistics over b_mask
// ---------------------------------------------------------------------------

template <typename GI, typename C>
__device__ __forceinline__ double sample_tri(const GI* __restrict__ img, int nx, int ny, int nz,
                                             C cx, C cy, C cz) {
  // order-1 interpolation at coordinates inside [0, n-1], upper neighbour
  // clamped ('nearest'), weights in double
  const int ix = (int)floor(cx), iy = (int)floor(cy), iz = (int)floor(cz);
  const int jx = min(ix + 1, nx - 1), jy = min(iy + 1, ny - 1), jz = min(iz + 1, nz - 1);
  const double fx = (double)cx - ix, fy = (double)cy - iy, fz = (double)cz - iz;
  const long long sy = nx, sz = (long long)nx * ny;
  const long long b00 = iy * sy + iz * sz, b10 = jy * sy + iz * sz;
  const long long b01 = iy * sy + jz * sz, b11 = jy * sy + jz * sz;
  const double v000 = ld_val(img, b00 + ix), v100 = ld_val(img, b00 + jx);
  const double v010 = ld_val(img, b10 + ix), v110 = ld_val(img, b10 + jx);
  const double v001 = ld_val(img, b01 + ix), v101 = ld_val(img, b01 + jx);
  const double v011 = ld_val(img, b11 + ix), v111 = ld_val(img, b11 + jx);
  const double gx0 = 1.0 - fx, gy0 = 1.0 - fy, gz0 = 1.0 - fz;
  return gz0 * (gy0 * (gx0 * v000 + fx * v100) + fy * (gx0 * v010 + fx * v110)) +
         fz * (gy0 * (gx0 * v001 + fx * v101) + fy * (gx0 * v011 + fx * v111));
}

// the eight taps and weights of one order-1 sample, fetched before any use so
// that the loads of several samples overlap.  O32: buffer loads with 32-bit
// byte offsets through a resource over the image (float and half images of
// fewer than 2^32 bytes, checked by the launcher): one offset register per
// tap instead of a 64-bit address.
__device__ __forceinline__ float buf_tap(__amdgpu_buffer_rsrc_t r, unsigned i, const __half*) {
  return __half2float(__ushort_as_half(__builtin_amdgcn_raw_buffer_load_b16(r, i * 2u, 0, 0)));
}
__device__ __forceinline__ float buf_tap(__amdgpu_buffer_rsrc_t r, unsigned i, const float*) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, i * 4u, 0, 0));
}
__device__ __forceinline__ double buf_tap(__amdgpu_buffer_rsrc_t, unsigned, const double*) {
  return 0.0;  // not used: the FP64 image takes 64-bit addresses
}
// the two x-neighbours (i, j) of a float image row from one 8-byte load at
// p = clamp(i, 0, nx - 2): (p, p + 1) holds both for i <= nx - 2, and
// i = j = nx - 1 reads element p + 1 twice.  A one-voxel-wide row (nx = 1,
// p = 0, i = j = 0) takes the low word; its high word may lie past the
// image, which the buffer's range check returns as 0.  The sampler's
// gathers are bound by the texture addresser's per-lane work, so half the
// load instructions carry the same taps.  Dword alignment is what a
// multi-dword buffer load needs.
__device__ __forceinline__ void buf_tap2(__amdgpu_buffer_rsrc_t r, unsigned b, int i, int j,
                                         int nx, float& vi, float& vj) {
  const int p = max(min(i, nx - 2), 0);
  const auto w = __builtin_amdgcn_raw_buffer_load_b64(r, (b + (unsigned)p) * 4u, 0, 0);
  const float lo = __uint_as_float(w[0]), hi = __uint_as_float(w[1]);
  vi = i > p ? hi : lo;
  vj = j > p ? hi : lo;
}
template <typename GI>
struct TriTaps {
  using C = typename RealOf<GI>::type;
  C v[8];
  C fx, fy, fz;  // c - floor(c): exact in the coordinate type
};

template <bool O32, typename GI, typename C>
__device__ __forceinline__ void tri_fetch(const GI* __restrict__ img, __amdgpu_buffer_rsrc_t rs,
                                          int nx, int ny, int nz, C cx, C cy, C cz,
                                          TriTaps<GI>& t) {
  const int ix = (int)floor(cx), iy = (int)floor(cy), iz = (int)floor(cz);
  const int jx = min(ix + 1, nx - 1), jy = min(iy + 1, ny - 1), jz = min(iz + 1, nz - 1);
  // c - floor(c) of a non-negative c is exact in c's own type, so these
  // equal sample_tri's (double)c - i
  using TC = typename TriTaps<GI>::C;
  static_assert(std::is_same<C, TC>::value, "coordinates in the image's working type");
  t.fx = cx - (C)ix; t.fy = cy - (C)iy; t.fz = cz - (C)iz;
  if constexpr (O32) {
    const unsigned sy = (unsigned)nx, sz = (unsigned)nx * (unsigned)ny;
    const unsigned b00 = iy * sy + iz * sz, b10 = jy * sy + iz * sz;
    const unsigned b01 = iy * sy + jz * sz, b11 = jy * sy + jz * sz;
    if constexpr (std::is_same<GI, float>::value) {
      buf_tap2(rs, b00, ix, jx, nx, t.v[0], t.v[1]);
      buf_tap2(rs, b10, ix, jx, nx, t.v[2], t.v[3]);
      buf_tap2(rs, b01, ix, jx, nx, t.v[4], t.v[5]);
      buf_tap2(rs, b11, ix, jx, nx, t.v[6], t.v[7]);
    } else {
      t.v[0] = buf_tap(rs, b00 + ix, img); t.v[1] = buf_tap(rs, b00 + jx, img);
      t.v[2] = buf_tap(rs, b10 + ix, img); t.v[3] = buf_tap(rs, b10 + jx, img);
      t.v[4] = buf_tap(rs, b01 + ix, img); t.v[5] = buf_tap(rs, b01 + jx, img);
      t.v[6] = buf_tap(rs, b11 + ix, img); t.v[7] = buf_tap(rs, b11 + jx, img);
    }
  } else {
    const long long sy = nx, sz = (long long)nx * ny;
    const long long b00 = iy * sy + iz * sz, b10 = jy * sy + iz * sz;
    const long long b01 = iy * sy + jz * sz, b11 = jy * sy + jz * sz;
    t.v[0] = ld_val(img, b00 + ix); t.v[1] = ld_val(img, b00 + jx);
    t.v[2] = ld_val(img, b10 + ix); t.v[3] = ld_val(img, b10 + jx);
    t.v[4] = ld_val(img, b01 + ix); t.v[5] = ld_val(img, b01 + jx);
    t.v[6] = ld_val(img, b11 + ix); t.v[7] = ld_val(img, b11 + jx);
  }
}

// the same combination in float32 (the mixed-precision form's weights):
// each x, y and z interpolation one product and one fused multiply-add,
// written out so that its rounding does not depend on how the compiler
// contracts the expression where it is inlined
template <typename GI>
__device__ __forceinline__ float tri_combine_f32(const TriTaps<GI>& t) {
  const float fx = (float)t.fx, fy = (float)t.fy, fz = (float)t.fz;
  const float gx0 = 1.0f - fx, gy0 = 1.0f - fy, gz0 = 1.0f - fz;
  const float v000 = t.v[0], v100 = t.v[1], v010 = t.v[2], v110 = t.v[3];
  const float v001 = t.v[4], v101 = t.v[5], v011 = t.v[6], v111 = t.v[7];
  const float a00 = fmaf(fx, v100, gx0 * v000), a10 = fmaf(fx, v110, gx0 * v010);
  const float a01 = fmaf(fx, v101, gx0 * v001), a11 = fmaf(fx, v111, gx0 * v011);
  const float b0 = fmaf(fy, a10, gy0 * a00), b1 = fmaf(fy, a11, gy0 * a01);
  return fmaf(fz, b1, gz0 * b0);
}

// sample_tri's combination, same operations in the same order
template <typename GI>
__device__ __forceinline__ double tri_combine(const TriTaps<GI>& t) {
  const double fx = t.fx, fy = t.fy, fz = t.fz;
  const double gx0 = 1.0 - fx, gy0 = 1.0 - fy, gz0 = 1.0 - fz;
  const double v000 = t.v[0], v100 = t.v[1], v010 = t.v[2], v110 = t.v[3];
  const double v001 = t.v[4], v101 = t.v[5], v011 = t.v[6], v111 = t.v[7];
  return gz0 * (gy0 * (gx0 * v000 + fx * v100) + fy * (gx0 * v010 + fx * v110)) +
         fz * (gy0 * (gx0 * v001 + fx * v101) + fy * (gx0 * v011 + fx * v111));
}

template <typename R>
__device__ __forceinline__ R clamp_to(R c, R hi) {
  return c < R(0) ? R(0) : (c > hi ? hi : c);
}

// The working type R is double for an FP64 image (the FP64 path) and float
// otherwise; the coordinate composition and threshold test run in R.
// Rows are taken kSampleRows at a time: all their taps and mask bytes are
// loaded before any is used, and the voxels are then finished in row order,
// so every output and the statistics' summation order are those of one row
// at a time.
// (two rows: 2, 3 and 4 rows measured within 2 % of each other at 513^3;
// the kernel is bound by its per-voxel instruction count)
constexpr int kSampleRows = 2;
// The displacement at a voxel is interpolated with explicit fused
// multiply-adds, so that its rounding does not depend on how the compiler
// contracts the expression.  F32W: the eight interpolation weights and the
// combination in float32 (the mixed-precision form) instead of SciPy's
// float64 order.
// BM: a b-mask is given (statistics over it); without one no b-mask byte is
// loaded per voxel.
template <typename GI, bool O32, bool F32W, bool BM>
__global__ __launch_bounds__(VDVC_BLOCK) void sample_kernel(
    MeshDev M, TilesDev T, SampleIn S, const typename RealOf<GI>::type* __restrict__ U) {
  using R = typename RealOf<GI>::type;
  const int lane = threadIdx.x & 63;
  const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const unsigned blk = xcd_remap(blockIdx.x, gridDim.x);
  const long long w = (long long)blk * VDVC_WAVES_PER_BLOCK + wib;
  if (w >= T.nwaves) return;
  const int h = M.h;
  const R inv_h = R(1) / R(h);
  const int sx = (int)(w % T.nsx);
  const long long r = w / T.nsx;
  const int ey = (int)(r % M.Ny);
  const int zt = (int)(r / M.Ny);
  const int x0 = sx * 64;
  const int nact = min(64, M.Nvx - x0);
  const bool act = lane < nact;
  const int xe = act ? x0 + lane : x0 + nact - 1;
  const int ex = xe / h;
  const int tx = xe - ex * h;
  const R X1 = tx * inv_h, X0 = R(1) - X1;
  const int i0 = min(ex, M.Nx), i1 = min(ex + 1, M.Nx);
  const int mx = axis_mult(xe, h, M.Nvx - 1);
  const int ez0 = zt * T.ez, ez1 = min(M.Nz, ez0 + T.ez);
  const int y0 = ey * h;
  const int nrows = h + (ey == M.Ny - 1 ? 1 : 0);
  const R fgnx = (R)(S.gnx - 1), fgny = (R)(S.gny - 1), fgnz = (R)(S.gnz - 1);
  const R gox = (R)S.gox, goy = (R)S.goy, goz = (R)S.goz, glt = (R)S.glt;
  const GI* __restrict__ img = reinterpret_cast<const GI*>(S.gimg);
  R* __restrict__ gvals = reinterpret_cast<R*>(S.gvals);
  uint8_t* __restrict__ mask_out = S.mask_out;
  const uint8_t* __restrict__ mask0 = S.mask0;
  const uint8_t* __restrict__ bmask = S.bmask;
  const uint8_t* dummy = reinterpret_cast<const uint8_t*>(S.gimg);
  // 0x00020000: the gfx9 buffer format word (32-bit data format, no swizzle)
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<GI*>(img), (short)0,
      O32 ? (int)((unsigned)S.gnx * S.gny * S.gnz * sizeof(GI)) : 0, 0x00020000);
  const R vxf = (R)xe;
  double sc = 0.0, s1 = 0.0, s2 = 0.0;
  for (int ez = ez0; ez < ez1; ++ez) {
    const int nzp = h + (ez == M.Nz - 1 ? 1 : 0);
    R Ux[2][2][3];
#pragma unroll
    for (int cy = 0; cy < 2; ++cy)
#pragma unroll
      for (int cz = 0; cz < 2; ++cz) {
        const long long n0 = node_index(M, i0, ey + cy, ez + cz);
        const long long n1 = node_index(M, i1, ey + cy, ez + cz);
#pragma unroll
        for (int k = 0; k < 3; ++k) Ux[cy][cz][k] = fma(X0, U[3 * n0 + k], X1 * U[3 * n1 + k]);
      }
    for (int tz = 0; tz < nzp; ++tz) {
      const int z = ez * h + tz;
      const R Z1 = tz * inv_h, Z0 = R(1) - Z1;
      const int mxz = mx * axis_mult(z, h, M.Nvz - 1);
      R u0[3], du[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        u0[k] = fma(Z0, Ux[0][0][k], Z1 * Ux[0][1][k]);
        const R u1 = fma(Z0, Ux[1][0][k], Z1 * Ux[1][1][k]);
        du[k] = (u1 - u0[k]) * inv_h;
      }
      const long long base = ((long long)z * M.Nvy + y0) * M.pitch + xe;
      // clamped, NaN-cleared sample coordinates of row ty and the bounds test
      // of its unclamped ones ((u + x) - origin, in R, as the reference
      // composes it).  A NaN coordinate clamps to a NaN floor index in NumPy;
      // here it samples a clamped voxel.  It is out of bounds either way, so
      // the mask is false.
      auto coords = [&](int ty, R& cx, R& cy, R& cz) -> bool {
        const R fty = (R)ty;
        const R ux = fma(fty, du[0], u0[0]);
        const R uy = fma(fty, du[1], u0[1]);
        const R uz = fma(fty, du[2], u0[2]);
        cx = sub_rn(add_rn(ux, vxf), gox);
        cy = sub_rn(add_rn(uy, (R)(y0 + ty)), goy);
        cz = sub_rn(add_rn(uz, (R)z), goz);
        const bool inb = (cx >= R(0)) && (cx <= fgnx) && (cy >= R(0)) && (cy <= fgny) &&
                         (cz >= R(0)) && (cz <= fgnz);
        cx = clamp_to(cx, fgnx);
        cy = clamp_to(cy, fgny);
        cz = clamp_to(cz, fgnz);
        if (!(cx == cx)) cx = R(0);
        if (!(cy == cy)) cy = R(0);
        if (!(cz == cz)) cz = R(0);
        return inb;
      };
      auto finish = [&](int ty, const TriTaps<GI>& tp, bool inb, uint8_t m0b, uint8_t bmb) {
        const long long vi = base + (long long)ty * M.pitch;
        const int y = y0 + ty;
        const R gv = F32W ? (R)tri_combine_f32(tp) : (R)tri_combine(tp);
        const bool m0 = mask0 ? (m0b != 0) : true;
        const bool mk = inb && m0 && (gv > glt);
        if (act) {
          gvals[vi] = gv;
          mask_out[vi] = mk ? 1 : 0;
          const bool bm = BM ? (bmb != 0) : mk;
          if (bm) {
            const double mlt = (double)(mxz * axis_mult(y, h, M.Nvy - 1));
            const double gd = (double)gv;
            sc += mlt;
            s1 += mlt * gd;
            s2 += mlt * gd * gd;
          }
        }
      };
      for (int ty = 0; ty < nrows; ty += kSampleRows) {
        TriTaps<GI> tp[kSampleRows];
        bool in[kSampleRows];
        uint8_t m0[kSampleRows], bm[kSampleRows];
#pragma unroll
        for (int q = 0; q < kSampleRows; ++q) {
          // rows past the element row's last repeat its last row (loaded,
          // not finished)
          const int tq = min(ty + q, nrows - 1);
          R cx, cy, cz;
          in[q] = coords(tq, cx, cy, cz);
          tri_fetch<O32>(img, rs, S.gnx, S.gny, S.gnz, cx, cy, cz, tp[q]);
          // mask bytes: unconditional loads (an absent mask0 reads a byte of
          // the image instead, ignored) so that they are in flight with the
          // taps; the b-mask byte only when there is a b-mask
          const long long vq = base + (long long)tq * M.pitch;
          m0[q] = *(mask0 ? mask0 + vq : dummy);
          bm[q] = BM ? bmask[vq] : 0;
        }
#pragma unroll
        for (int q = 0; q < kSampleRows; ++q)
          if (ty + q < nrows) finish(ty + q, tp[q], in[q], m0[q], bm[q]);
      }
    }
  }
  if (S.tail_fold && sx == T.nsx - 1) {
    // The last voxel column x = Nvx - 1 (a strip of its own would hold one
    // active lane) is taken by the last strip's wavefront, its lanes spread
    // over the column's (row, plane) pairs of each element layer.  Per voxel
    // the same operations as the main loop with x-weights X0 = 1, X1 = 0 on
    // node column Nx.
    const int xt = M.Nvx - 1;
    const R vxt = (R)xt;
    const int mxt = axis_mult(xt, h, M.Nvx - 1);
    for (int ez = ez0; ez < ez1; ++ez) {
      const int nzp = h + (ez == M.Nz - 1 ? 1 : 0);
      R Ut[2][2][3];
#pragma unroll
      for (int cy = 0; cy < 2; ++cy)
#pragma unroll
        for (int cz = 0; cz < 2; ++cz) {
          const long long nt = node_index(M, M.Nx, ey + cy, ez + cz);
#pragma unroll
          for (int k = 0; k < 3; ++k) Ut[cy][cz][k] = fma(R(1), U[3 * nt + k], R(0) * U[3 * nt + k]);
        }
      const int npairs = nzp * nrows;
      for (int q = lane; q < npairs; q += 64) {
        const int tz = q / nrows, ty = q - tz * nrows;
        const int z = ez * h + tz, y = y0 + ty;
        const R Z1 = tz * inv_h, Z0 = R(1) - Z1;
        R ut[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const R a0 = fma(Z0, Ut[0][0][k], Z1 * Ut[0][1][k]);
          const R a1 = fma(Z0, Ut[1][0][k], Z1 * Ut[1][1][k]);
          const R d = (a1 - a0) * inv_h;
          ut[k] = fma((R)ty, d, a0);
        }
        R cx = sub_rn(add_rn(ut[0], vxt), gox);
        R cy = sub_rn(add_rn(ut[1], (R)y), goy);
        R cz = sub_rn(add_rn(ut[2], (R)z), goz);
        const bool inb = (cx >= R(0)) && (cx <= fgnx) && (cy >= R(0)) && (cy <= fgny) &&
                         (cz >= R(0)) && (cz <= fgnz);
        cx = clamp_to(cx, fgnx);
        cy = clamp_to(cy, fgny);
        cz = clamp_to(cz, fgnz);
        if (!(cx == cx)) cx = R(0);
        if (!(cy == cy)) cy = R(0);
        if (!(cz == cz)) cz = R(0);
        TriTaps<GI> tp;
        tri_fetch<O32>(img, rs, S.gnx, S.gny, S.gnz, cx, cy, cz, tp);
        const long long vi = ((long long)z * M.Nvy + y) * M.pitch + xt;
        const R gv = F32W ? (R)tri_combine_f32(tp) : (R)tri_combine(tp);
        const bool m0 = mask0 ? (mask0[vi] != 0) : true;
        const bool mk = inb && m0 && (gv > glt);
        gvals[vi] = gv;
        mask_out[vi] = mk ? 1 : 0;
        const bool bmv = BM ? (bmask[vi] != 0) : mk;
        if (bmv) {
          const double mlt =
              (double)(mxt * axis_mult(z, h, M.Nvz - 1) * axis_mult(y, h, M.Nvy - 1));
          const double gd = (double)gv;
          sc += mlt;
          s1 += mlt * gd;
          s2 += mlt * gd * gd;
        }
      }
    }
  }
  sc = wave_sum(sc);
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  if (lane == 0) {
    S.stats[3 * w + 0] = sc;
    S.stats[3 * w + 1] = s1;
    S.stats[3 * w + 2] = s2;
  }
}

// ---------------------------------------------------------------------------
// reference-image setup: np.gradient (edge_order=1) per axis, mask0 = f > glt,
// gradient zeroed where the 6-cross erosion of mask0 ('nearest') is false
// ---------------------------------------------------------------------------
template <typename GO>
__device__ __forceinline__ void st_grad(GO* p, long long i, typename RealOf<GO>::type v) {
  if constexpr (std::is_same<GO, __hip_bfloat16>::value) p[i] = __float2bfloat16(v);
  else p[i] = v;
}

// R is double for the FP64 path (double image and gradients), else float
template <typename FT, typename GO>
__global__ __launch_bounds__(VDVC_BLOCK) void setup_kernel(MeshDev M, const FT* __restrict__ f,
                                                           double glt_d, GO* __restrict__ gxo,
                                                           GO* __restrict__ gyo,
                                                           GO* __restrict__ gzo,
                                                           uint8_t* __restrict__ mask0) {
  using R = typename RealOf<GO>::type;
  const int nxb = (M.pitch + VDVC_BLOCK - 1) / VDVC_BLOCK;
  const long long row = blockIdx.x / nxb;
  const int x = (blockIdx.x % nxb) * VDVC_BLOCK + threadIdx.x;
  if (x >= M.pitch || row >= (long long)M.Nvy * M.Nvz) return;
  const long long vi = row * M.pitch + x;
  if (x >= M.Nvx) {  // row padding: zero gradients and mask
    mask0[vi] = 0;
    st_grad(gxo, vi, R(0)); st_grad(gyo, vi, R(0)); st_grad(gzo, vi, R(0));
    return;
  }
  const R glt = (R)glt_d;
  const int y = (int)(row % M.Nvy);
  const int z = (int)(row / M.Nvy);
  const long long sy = M.pitch, sz = (long long)M.pitch * M.Nvy;
  const R fc = ld_val(f, vi);
  const int xm = x > 0 ? x - 1 : 0, xp = x < M.Nvx - 1 ? x + 1 : x;
  const int ym = y > 0 ? y - 1 : 0, yp = y < M.Nvy - 1 ? y + 1 : y;
  const int zm = z > 0 ? z - 1 : 0, zp = z < M.Nvz - 1 ? z + 1 : z;
  const R fxm = ld_val(f, vi + (xm - x)), fxp = ld_val(f, vi + (xp - x));
  const R fym = ld_val(f, vi + (long long)(ym - y) * sy), fyp = ld_val(f, vi + (long long)(yp - y) * sy);
  const R fzm = ld_val(f, vi + (long long)(zm - z) * sz), fzp = ld_val(f, vi + (long long)(zp - z) * sz);
  auto grad1 = [](R fm, R fp, int t, int n) -> R {
    if (t == 0 || t == n - 1) return sub_rn(fp, fm);  // one-sided, spacing 1
    return sub_rn(fp, fm) * R(0.5);                  // central
  };
  R gx = grad1(fxm, fxp, x, M.Nvx);
  R gy = grad1(fym, fyp, y, M.Nvy);
  R gz = grad1(fzm, fzp, z, M.Nvz);
  const bool m0 = fc > glt;
  const bool er = m0 && (fxm > glt) && (fxp > glt) && (fym > glt) && (fyp > glt) && (fzm > glt) &&
                  (fzp > glt);
  if (!er) { gx = R(0); gy = R(0); gz = R(0); }
  mask0[vi] = m0 ? 1 : 0;
  st_grad(gxo, vi, gx); st_grad(gyo, vi, gy); st_grad(gzo, vi, gz);
}

// ---------------------------------------------------------------------------
// multiplicity-weighted masked statistics of a voxel field, per-block FP64
// partials (count, sum, sum of squares)
// ---------------------------------------------------------------------------
// eight consecutive values (16-byte aligned) as the working type
__device__ __forceinline__ void ld8(const __half* p, float (&v)[8]) {
  const uint4 q = *reinterpret_cast<const uint4*>(p);
  const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
  for (int j = 0; j < 8; ++j)
    v[j] = __half2float(__ushort_as_half((unsigned short)(w[j >> 1] >> (16 * (j & 1)))));
}
__device__ __forceinline__ void ld8(const float* p, float (&v)[8]) {
  const float4 a = reinterpret_cast<const float4*>(p)[0], b = reinterpret_cast<const float4*>(p)[1];
  v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
}
__device__ __forceinline__ void ld8(const double* p, double (&v)[8]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double2 a = reinterpret_cast<const double2*>(p)[k];
    v[2 * k] = a.x; v[2 * k + 1] = a.y;
  }
}

// single rounding of a float32 / float64 source value to the level's storage
// type, and the stored value in the working type
template <typename DT> struct LvlCvt;
template <> struct LvlCvt<__half> {
  static __device__ __forceinline__ __half from(float s) { return __float2half_rn(s); }
  static __device__ __forceinline__ float val(__half v) { return __half2float(v); }
};
template <> struct LvlCvt<float> {
  static __device__ __forceinline__ float from(float s) { return s; }
  static __device__ __forceinline__ float from(double s) { return (float)s; }
  static __device__ __forceinline__ float val(float v) { return v; }
};
template <> struct LvlCvt<double> {
  static __device__ __forceinline__ double from(double s) { return s; }
  static __device__ __forceinline__ double val(double v) { return v; }
};

// A dense x-fastest volume (rows of Nvx) rounded once to the level's storage
// type DT and written in the pitched level layout (padding columns zero),
// with the multiplicity-weighted (count, sum, sum of squares) of the stored
// values above glt taken in the same pass.  A wavefront takes whole rows
// (grid-stride), its lanes consecutive voxels, KCH 64-voxel chunks of a row
// loaded before any is used.  Per lane the voxels are summed in (row,
// chunk) order; lanes, wavefronts and blocks are combined in a fixed order.
template <typename ST, typename DT, int KCH>
__global__ __launch_bounds__(VDVC_BLOCK) void level_image_kernel(MeshDev M,
                                                                 const ST* __restrict__ src,
                                                                 DT* __restrict__ dst,
                                                                 double glt_d,
                                                                 double* __restrict__ part) {
  using R = typename RealOf<DT>::type;
  constexpr int kCh = KCH;
  __shared__ double lds[4 * 3];
  const R glt = (R)glt_d;
  const int lane = threadIdx.x & 63;
  const int nx = M.Nvx, dp = M.pitch, h = M.h;
  const int nch = (dp + 63) >> 6;
  const long long nrows = (long long)M.Nvy * M.Nvz;
  const int lrx = lane % h;
  double acc[3] = {0.0, 0.0, 0.0};
  for (long long row = (long long)blockIdx.x * VDVC_WAVES_PER_BLOCK +
                       __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
       row < nrows; row += (long long)gridDim.x * VDVC_WAVES_PER_BLOCK) {
    const int y = (int)(row % M.Nvy), z = (int)(row / M.Nvy);
    const int myz = axis_mult(y, h, M.Nvy - 1) * axis_mult(z, h, M.Nvz - 1);
    const ST* __restrict__ sr = src + row * nx;
    DT* __restrict__ dr = dst + row * dp;
    for (int c0 = 0; c0 < nch; c0 += kCh) {
      ST raw[kCh];
#pragma unroll
      for (int j = 0; j < kCh; ++j) {
        const int x = (c0 + j) * 64 + lane;
        raw[j] = (c0 + j < nch && x < nx) ? sr[x] : ST(0);
      }
#pragma unroll
      for (int j = 0; j < kCh; ++j) {
        if (c0 + j >= nch) break;
        const int xb = (c0 + j) * 64;
        const int x = xb + lane;
        const DT v = LvlCvt<DT>::from(raw[j]);
        if (x < dp) dr[x] = v;
        const R val = LvlCvt<DT>::val(v);
        if (x < nx && val > glt) {
          int rx = lrx + xb % h;
          rx = rx >= h ? rx - h : rx;
          const int mx = (x > 0 && x < nx - 1 && rx == 0) ? 2 : 1;
          const double mlt = (double)(mx * myz);
          const double d = (double)val;
          acc[0] += mlt;
          acc[1] += mlt * d;
          acc[2] += mlt * d * d;
        }
      }
    }
  }
  block_sum_f64<3>(acc, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 3; ++q) part[3 * (long long)blockIdx.x + q] = acc[q];
  }
}

// With `vec` (16-byte aligned field, 8-byte aligned mask; the pitch is a
// multiple of 16, so the loads stay inside the row) wavefronts stride over
// rows and lanes take eight voxels each; otherwise blocks stride over rows
// and threads take one voxel.  The row, its y / z multiplicity and base are
// uniform over the wavefront or block.
template <typename FT>
__global__ __launch_bounds__(VDVC_BLOCK) void stats_kernel(MeshDev M, const FT* __restrict__ fld,
                                                           const uint8_t* __restrict__ mask,
                                                           int use_thresh, double glt_d, int vec,
                                                           double* __restrict__ part) {
  using R = typename RealOf<FT>::type;
  __shared__ double lds[4 * 3];
  const R glt = (R)glt_d;
  const long long nrows = (long long)M.Nvy * M.Nvz;
  double acc[3] = {0.0, 0.0, 0.0};
  const int lane = threadIdx.x & 63;
  const long long wave0 =
      (long long)blockIdx.x * VDVC_WAVES_PER_BLOCK + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long long r0 = vec ? wave0 : blockIdx.x;
  const long long rstep = vec ? (long long)gridDim.x * VDVC_WAVES_PER_BLOCK : gridDim.x;
  // vectorised rows of at most two 512-voxel chunks: the loads of kRB rows
  // are issued before any is summed, then the voxels are summed in the
  // order of the one-row loop below (row, chunk, voxel), so the sums are
  // bit-identical to it and a wavefront waits on one memory round trip per
  // kRB rows instead of one per chunk
  constexpr int kRB = 4;
  if (vec && M.Nvx <= 1024) {
    for (long long rg = r0; rg < nrows; rg += kRB * rstep) {
      R v[kRB][2][8];
      uint2 mk[kRB][2];
      bool ok[kRB][2];
#pragma unroll
      for (int q = 0; q < kRB; ++q) {
        const long long row = rg + q * rstep;
        const long long rb = (row < nrows ? row : 0) * M.pitch;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const int x0 = 8 * lane + 512 * c;
          ok[q][c] = row < nrows && x0 < M.Nvx;
          mk[q][c] = make_uint2(0xffffffffu, 0xffffffffu);
          if (ok[q][c]) {
            ld8(fld + rb + x0, v[q][c]);
            if (!use_thresh && mask) mk[q][c] = *reinterpret_cast<const uint2*>(mask + rb + x0);
          }
        }
      }
#pragma unroll
      for (int q = 0; q < kRB; ++q) {
        const long long row = rg + q * rstep;
        if (row >= nrows) break;
        const int y = (int)(row % M.Nvy);
        const int z = (int)(row / M.Nvy);
        const int myz = axis_mult(y, M.h, M.Nvy - 1) * axis_mult(z, M.h, M.Nvz - 1);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          if (!ok[q][c]) continue;
          const int x0 = 8 * lane + 512 * c;
          int rx = x0 % M.h;
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const int x = x0 + j;
            const bool mb = ((j < 4 ? mk[q][c].x : mk[q][c].y) >> (8 * (j & 3))) & 0xffu;
            const bool m = x < M.Nvx && (use_thresh ? (v[q][c][j] > glt) : mb);
            if (m) {
              const int mx = (x > 0 && x < M.Nvx - 1 && rx == 0) ? 2 : 1;
              const double mlt = (double)(mx * myz);
              const double d = (double)v[q][c][j];
              acc[0] += mlt;
              acc[1] += mlt * d;
              acc[2] += mlt * d * d;
            }
            rx = rx + 1 == M.h ? 0 : rx + 1;
          }
        }
      }
    }
  }
  for (long long row = (vec && M.Nvx <= 1024) ? nrows : r0; row < nrows; row += rstep) {
    const int y = (int)(row % M.Nvy);
    const int z = (int)(row / M.Nvy);
    const int myz = axis_mult(y, M.h, M.Nvy - 1) * axis_mult(z, M.h, M.Nvz - 1);
    const long long rb = row * M.pitch;
    if (vec) {
      for (int x0 = 8 * lane; x0 < M.Nvx; x0 += 8 * 64) {
        R v[8];
        ld8(fld + rb + x0, v);
        uint2 mk = make_uint2(0xffffffffu, 0xffffffffu);
        if (!use_thresh && mask) mk = *reinterpret_cast<const uint2*>(mask + rb + x0);
        int rx = x0 % M.h;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int x = x0 + j;
          const bool mb = ((j < 4 ? mk.x : mk.y) >> (8 * (j & 3))) & 0xffu;
          const bool m = x < M.Nvx && (use_thresh ? (v[j] > glt) : mb);
          if (m) {
            const int mx = (x > 0 && x < M.Nvx - 1 && rx == 0) ? 2 : 1;
            const double mlt = (double)(mx * myz);
            const double d = (double)v[j];
            acc[0] += mlt;
            acc[1] += mlt * d;
            acc[2] += mlt * d * d;
          }
          rx = rx + 1 == M.h ? 0 : rx + 1;
        }
      }
      continue;
    }
    for (int x = threadIdx.x; x < M.Nvx; x += VDVC_BLOCK) {
      const long long vi = rb + x;
      const R val = ld_val(fld, vi);
      const bool m = use_thresh ? (val > glt) : (mask ? mask[vi] != 0 : true);
      if (m) {
        const double mlt = (double)(axis_mult(x, M.h, M.Nvx - 1) * myz);
        const double d = (double)val;
        acc[0] += mlt;
        acc[1] += mlt * d;
        acc[2] += mlt * d * d;
      }
    }
  }
  block_sum_f64<3>(acc, lds);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x + 0] = acc[0];
    part[3 * blockIdx.x + 1] = acc[1];
    part[3 * blockIdx.x + 2] = acc[2];
  }
}

// eight gradient values (16-byte aligned) stored in the plane's type
__device__ __forceinline__ void st8(__hip_bfloat16* p, const float (&v)[8]) {
  uint32_t w[4];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    w[k] = (uint32_t)__bfloat16_as_ushort(__float2bfloat16(v[2 * k])) |
           ((uint32_t)__bfloat16_as_ushort(__float2bfloat16(v[2 * k + 1])) << 16);
  *reinterpret_cast<uint4*>(p) = make_uint4(w[0], w[1], w[2], w[3]);
}
__device__ __forceinline__ void st8(float* p, const float (&v)[8]) {
  reinterpret_cast<float4*>(p)[0] = make_float4(v[0], v[1], v[2], v[3]);
  reinterpret_cast<float4*>(p)[1] = make_float4(v[4], v[5], v[6], v[7]);
}
__device__ __forceinline__ void st8(double* p, const double (&v)[8]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) reinterpret_cast<double2*>(p)[k] = make_double2(v[2 * k], v[2 * k + 1]);
}

// setup_kernel for eight consecutive voxels of a row per thread (threads
// numbered over (row, group of eight)): the row and
// its y / z neighbour rows as 16-byte vector loads, the two x neighbours
// outside the group as scalars, vector stores (16-byte aligned planes,
// 8-byte aligned mask, pitch a multiple of 8).  Per voxel the same
// operations as setup_kernel.
template <typename FT, typename GO>
__global__ __launch_bounds__(VDVC_BLOCK) void setup_vec_kernel(MeshDev M, const FT* __restrict__ f,
                                                               double glt_d, GO* __restrict__ gxo,
                                                               GO* __restrict__ gyo,
                                                               GO* __restrict__ gzo,
                                                               uint8_t* __restrict__ mask0) {
  using R = typename RealOf<GO>::type;
  const int groups = M.pitch / 8;
  const long long gid = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;  // (row, group)
  const long long row = gid / groups;
  const int gi = (int)(gid - row * groups);
  if (row >= (long long)M.Nvy * M.Nvz) return;
  const int x0 = 8 * gi;
  const long long vi = row * M.pitch + x0;
  R gx[8], gy[8], gz[8];
  uint32_t mw[2] = {0u, 0u};
  if (x0 >= M.Nvx) {  // row padding: zero gradients and mask
#pragma unroll
    for (int j = 0; j < 8; ++j) gx[j] = gy[j] = gz[j] = R(0);
  } else {
    const R glt = (R)glt_d;
    const int y = (int)(row % M.Nvy);
    const int z = (int)(row / M.Nvy);
    const long long sy = M.pitch, sz = (long long)M.pitch * M.Nvy;
    const int ym = y > 0 ? y - 1 : 0, yp = y < M.Nvy - 1 ? y + 1 : y;
    const int zm = z > 0 ? z - 1 : 0, zp = z < M.Nvz - 1 ? z + 1 : z;
    R c[8], fym[8], fyp[8], fzm[8], fzp[8];
    ld8(f + vi, c);
    ld8(f + vi + (long long)(ym - y) * sy, fym);
    ld8(f + vi + (long long)(yp - y) * sy, fyp);
    ld8(f + vi + (long long)(zm - z) * sz, fzm);
    ld8(f + vi + (long long)(zp - z) * sz, fzp);
    const R left = x0 > 0 ? (R)ld_val(f, vi - 1) : c[0];
    const R right = x0 + 8 < M.Nvx ? (R)ld_val(f, vi + 8) : c[7];
    auto grad1 = [](R fm, R fp, int t, int n) -> R {
      if (t == 0 || t == n - 1) return sub_rn(fp, fm);  // one-sided, spacing 1
      return sub_rn(fp, fm) * R(0.5);                  // central
    };
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int x = x0 + j;
      if (x >= M.Nvx) { gx[j] = gy[j] = gz[j] = R(0); continue; }
      const R fxm = x > 0 ? (j > 0 ? c[j - 1] : left) : c[j];
      const R fxp = x < M.Nvx - 1 ? (j < 7 ? c[j + 1] : right) : c[j];
      R a = grad1(fxm, fxp, x, M.Nvx);
      R b = grad1(fym[j], fyp[j], y, M.Nvy);
      R d = grad1(fzm[j], fzp[j], z, M.Nvz);
      const bool m0 = c[j] > glt;
      const bool er = m0 && (fxm > glt) && (fxp > glt) && (fym[j] > glt) && (fyp[j] > glt) &&
                      (fzm[j] > glt) && (fzp[j] > glt);
      if (!er) { a = R(0); b = R(0); d = R(0); }
      gx[j] = a; gy[j] = b; gz[j] = d;
      if (m0) mw[j >> 2] |= 1u << (8 * (j & 3));
    }
  }
  *reinterpret_cast<uint2*>(mask0 + vi) = make_uint2(mw[0], mw[1]);
  st8(gxo + vi, gx);
  st8(gyo + vi, gy);
  st8(gzo + vi, gz);
}

// fixed-order reduction of n partial records of width nv (deterministic):
// block q sums column q -- thread t over records t, t + 256, ..., then a
// tree over the block
__global__ __launch_bounds__(VDVC_BLOCK) void reduce_partials_kernel(const double* __restrict__ part,
                                                                     long long n, int nv,
                                                                     double* __restrict__ out) {
  __shared__ double lds[VDVC_BLOCK];
  const int q = blockIdx.x;
  double s = 0.0;
  for (long long i = threadIdx.x; i < n; i += VDVC_BLOCK) s += part[i * nv + q];
  lds[threadIdx.x] = s;
  __syncthreads();
  for (int o = VDVC_BLOCK / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) lds[threadIdx.x] += lds[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[q] = lds[0];
}

// bfloat16 gradient planes multiplied by a 0/1 voxel mask (the folded
// storage read by the Hessian kernels when no separate mask is given); eight
// voxels per thread (16-byte plane and 8-byte mask accesses) with `vec` (n a
// multiple of 8, aligned pointers), one otherwise
__global__ __launch_bounds__(VDVC_BLOCK) void fold_kernel(long long n, int vec,
                                                          const uint16_t* __restrict__ gx,
                                                          const uint16_t* __restrict__ gy,
                                                          const uint16_t* __restrict__ gz,
                                                          const uint8_t* __restrict__ m,
                                                          uint16_t* __restrict__ ox,
                                                          uint16_t* __restrict__ oy,
                                                          uint16_t* __restrict__ oz) {
  const long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  if (vec) {
    if (8 * q >= n) return;
    const uint2 mk = *reinterpret_cast<const uint2*>(m + 8 * q);
    // 0xffff per voxel whose mask byte is set, 0 otherwise
    uint32_t keep[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t w = j < 2 ? mk.x : mk.y;
      const int s = 16 * (j & 1);
      keep[j] = (((w >> s) & 0xffu) ? 0x0000ffffu : 0u) | (((w >> (s + 8)) & 0xffu) ? 0xffff0000u : 0u);
    }
    const uint16_t* src[3] = {gx, gy, gz};
    uint16_t* dst[3] = {ox, oy, oz};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      uint4 v = *reinterpret_cast<const uint4*>(src[c] + 8 * q);
      v.x &= keep[0]; v.y &= keep[1]; v.z &= keep[2]; v.w &= keep[3];
      *reinterpret_cast<uint4*>(dst[c] + 8 * q) = v;
    }
    return;
  }
  if (q >= n) return;
  const bool k = m[q] != 0;
  ox[q] = k ? gx[q] : (uint16_t)0;
  oy[q] = k ? gy[q] : (uint16_t)0;
  oz[q] = k ? gz[q] : (uint16_t)0;
}

// 6-cross erosion of a uint8 mask with edge replication
// (minimum_filter(mode='nearest'), engine/correlate_gpu.py:75-94)
__global__ __launch_bounds__(VDVC_BLOCK) void erode6_kernel(int nx, int ny, int nz,
                                                            const uint8_t* __restrict__ in,
                                                            uint8_t* __restrict__ out) {
  const long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  const long long n = (long long)nx * ny * nz;
  if (q >= n) return;
  const int x = (int)(q % nx);
  const long long r = q / nx;
  const int y = (int)(r % ny);
  const int z = (int)(r / ny);
  const long long sy = nx, sz = (long long)nx * ny;
  bool v = in[q] != 0;
  v = v && in[q + (x > 0 ? -1 : 0)] && in[q + (x < nx - 1 ? 1 : 0)];
  v = v && in[q + (y > 0 ? -sy : 0)] && in[q + (y < ny - 1 ? sy : 0)];
  v = v && in[q + (z > 0 ? -sz : 0)] && in[q + (z < nz - 1 ? sz : 0)];
  out[q] = v ? 1 : 0;
}

// out = a * b elementwise
template <typename R>
__global__ __launch_bounds__(VDVC_BLOCK) void mul_kernel(long long n, const R* __restrict__ a,
                                                         const R* __restrict__ b,
                                                         R* __restrict__ out) {
  const long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  if (q < n) out[q] = a[q] * b[q];
}

// ---------------------------------------------------------------------------
// host-side launchers (C++ linkage, used by the C-ABI layer)
// ---------------------------------------------------------------------------
// one instantiation of the generic voxel kernel; v and P are in its
// working type
template <int MODE, typename GT, bool HM, typename FT, bool P2, bool FO>
void launch_restrict_one(dim3 g, dim3 b, hipStream_t st, const MeshDev& M, const TilesDev& T,
                         const VoxIn& in, const void* v, void* P) {
  using R = typename RealOf<GT>::type;
  hipLaunchKernelGGL((vox_restrict_kernel<MODE, GT, HM, FT, P2, FO>), g, b, 0, st, M, T, in,
                     static_cast<const R*>(v), static_cast<R*>(P));
}

// The residual of a residual pass formed again from the same inputs with
// the same single-rounded steps, (g - g_mean) * kscale + (f - mean0) in the
// working type R and zero outside the b-mask, written in the dense layout
// x + Nvx (y + Nvy z).  The GN loop forms the residual every iteration
// (engine/correlate_gpu.py:341-378) but returns only the last one, so the
// residual passes do not store it and this pass writes the returned array
// once, from the last pass's sampled values.
template <typename R, typename FT>
__global__ __launch_bounds__(VDVC_BLOCK) void residual_out_kernel(
    MeshDev M, const FT* __restrict__ f, const R* __restrict__ gv,
    const uint8_t* __restrict__ bm, double mean0_d, double gmean_d, double kscale_d,
    R* __restrict__ out) {
  const long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  if (q >= M.Nvox) return;
  const long long r = q / M.Nvx;
  const long long vi = r * M.pitch + (q - r * M.Nvx);
  const R mean0 = (R)mean0_d, gmean = (R)gmean_d, kscale = (R)kscale_d;
  const R fval = ld_val(f, vi);
  R rv = add_rn(mul_rn(sub_rn(gv[vi], gmean), kscale), sub_rn(fval, mean0));
  if (!bm[vi]) rv = R(0);
  out[q] = rv;
}

namespace vdvc {

int launch_residual_out(int rtype, int ftype, const MeshDev& M, const void* f, const void* gv,
                        const uint8_t* bm, double mean0, double gmean, double kscale, void* out,
                        hipStream_t st) {
  const dim3 g((unsigned)((M.Nvox + VDVC_BLOCK - 1) / VDVC_BLOCK)), b(VDVC_BLOCK);
  if (rtype == GT_F64) {
    if (ftype != GT_F64) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL((residual_out_kernel<double, double>), g, b, 0, st, M, (const double*)f,
                       (const double*)gv, bm, mean0, gmean, kscale, (double*)out);
  } else if (ftype == GT_F32) {
    hipLaunchKernelGGL((residual_out_kernel<float, float>), g, b, 0, st, M, (const float*)f,
                       (const float*)gv, bm, mean0, gmean, kscale, (float*)out);
  } else if (ftype == GT_F16) {
    hipLaunchKernelGGL((residual_out_kernel<float, __half>), g, b, 0, st, M, (const __half*)f,
                       (const float*)gv, bm, mean0, gmean, kscale, (float*)out);
  } else {
    return (int)hipErrorInvalidValue;
  }
  return (int)hipGetLastError();
}

int launch_restrict(int mode, int gtype, bool hasmask, int ftype, bool fold, const MeshDev& M,
                    const TilesDev& T, const VoxIn& in, const void* vv, void* Pv,
                    hipStream_t st) {
  const unsigned nblk = (unsigned)((T.nwaves + VDVC_WAVES_PER_BLOCK - 1) / VDVC_WAVES_PER_BLOCK);
  const bool pow2 = (M.h <= 64) && ((64 % M.h) == 0);
  dim3 g(nblk), b(VDVC_BLOCK);
#define VDVC_L(MODE, GT, HM, FT, P2, FO) \
  launch_restrict_one<MODE, GT, HM, FT, P2, FO>(g, b, st, M, T, in, vv, Pv)
#define VDVC_P2(MODE, GT, HM, FT, FO) \
  do { if (pow2) VDVC_L(MODE, GT, HM, FT, true, FO); else VDVC_L(MODE, GT, HM, FT, false, FO); } while (0)
#define VDVC_HM(MODE, GT, FT, FO) \
  do { if (hasmask) VDVC_P2(MODE, GT, true, FT, FO); else VDVC_P2(MODE, GT, false, FT, FO); } while (0)
  if (mode == MODE_MATVEC || mode == MODE_DIAG) {
    if (mode == MODE_MATVEC) {
      if (gtype == GT_F32) VDVC_HM(MODE_MATVEC, float, float, false);
      else if (gtype == GT_F64) VDVC_HM(MODE_MATVEC, double, double, false);
      else VDVC_HM(MODE_MATVEC, __hip_bfloat16, float, false);
    } else {
      if (gtype == GT_F32) VDVC_HM(MODE_DIAG, float, float, false);
      else if (gtype == GT_F64) VDVC_HM(MODE_DIAG, double, double, false);
      else VDVC_HM(MODE_DIAG, __hip_bfloat16, float, false);
    }
  } else {
    // RHS always reads the b_mask explicitly
    if (gtype == GT_F64) {
      if (ftype != GT_F64 || fold) return (int)hipErrorInvalidValue;
      VDVC_P2(MODE_RHS, double, true, double, false);
    } else if (gtype == GT_F32) {
      if (ftype == GT_F32) {
        if (fold) VDVC_P2(MODE_RHS, float, true, float, true);
        else VDVC_P2(MODE_RHS, float, true, float, false);
      } else {
        if (fold) VDVC_P2(MODE_RHS, float, true, __half, true);
        else VDVC_P2(MODE_RHS, float, true, __half, false);
      }
    } else {
      if (ftype == GT_F32) {
        if (fold) VDVC_P2(MODE_RHS, __hip_bfloat16, true, float, true);
        else VDVC_P2(MODE_RHS, __hip_bfloat16, true, float, false);
      } else {
        if (fold) VDVC_P2(MODE_RHS, __hip_bfloat16, true, __half, true);
        else VDVC_P2(MODE_RHS, __hip_bfloat16, true, __half, false);
      }
    }
  }
#undef VDVC_HM
#undef VDVC_P2
#undef VDVC_L
  return (int)hipGetLastError();
}

int launch_sample(int gitype, const MeshDev& M, const TilesDev& T, const SampleIn& S,
                  const void* U, bool f32w, hipStream_t st) {
  const unsigned nblk = (unsigned)((T.nwaves + VDVC_WAVES_PER_BLOCK - 1) / VDVC_WAVES_PER_BLOCK);
  // 32-bit buffer offsets: float and half images below 2^32 bytes
  const long long nimg = (long long)S.gnx * S.gny * S.gnz;
  const bool o32 = gitype == GT_F32 ? nimg < (1LL << 30) : gitype == GT_F64 ? false : nimg < (1LL << 31);
  const bool bm = S.bmask != nullptr;
#define VDVC_SAMPLE_K(GI, W)                                                          \
  (o32 ? (bm ? sample_kernel<GI, true, W, true> : sample_kernel<GI, true, W, false>) \
       : (bm ? sample_kernel<GI, false, W, true> : sample_kernel<GI, false, W, false>))
  if (gitype == GT_F32) {
    if (f32w)
      hipLaunchKernelGGL(VDVC_SAMPLE_K(float, true), dim3(nblk), dim3(VDVC_BLOCK), 0, st, M, T, S,
                         static_cast<const float*>(U));
    else
      hipLaunchKernelGGL(VDVC_SAMPLE_K(float, false), dim3(nblk), dim3(VDVC_BLOCK), 0, st, M, T,
                         S, static_cast<const float*>(U));
  } else if (gitype == GT_F64) {
    hipLaunchKernelGGL(VDVC_SAMPLE_K(double, false), dim3(nblk), dim3(VDVC_BLOCK), 0, st, M, T, S,
                       static_cast<const double*>(U));
  } else {
    if (f32w)
      hipLaunchKernelGGL(VDVC_SAMPLE_K(__half, true), dim3(nblk), dim3(VDVC_BLOCK), 0, st, M, T,
                         S, static_cast<const float*>(U));
    else
      hipLaunchKernelGGL(VDVC_SAMPLE_K(__half, false), dim3(nblk), dim3(VDVC_BLOCK), 0, st, M, T,
                         S, static_cast<const float*>(U));
  }
#undef VDVC_SAMPLE_K
  return (int)hipGetLastError();
}

int launch_setup(int ftype, int gtype, const MeshDev& M, const void* f, double glt, void* gx,
                 void* gy, void* gz, uint8_t* mask0, hipStream_t st) {
  auto al = [](const void* p, uintptr_t a) { return (reinterpret_cast<uintptr_t>(p) & (a - 1)) == 0; };
  const bool vec = M.pitch % 8 == 0 && al(f, 16) && al(gx, 16) && al(gy, 16) && al(gz, 16) &&
                   al(mask0, 8);
  const long long nb = vec ? ((long long)(M.pitch / 8) * M.Nvy * M.Nvz + VDVC_BLOCK - 1) / VDVC_BLOCK
                           : (long long)((M.pitch + VDVC_BLOCK - 1) / VDVC_BLOCK) * M.Nvy * M.Nvz;
  if (nb > 0x7fffffffLL) return (int)hipErrorInvalidValue;
  dim3 g((unsigned)nb), b(VDVC_BLOCK);
#define VDVC_SETUP_K(FT, GO) (vec ? setup_vec_kernel<FT, GO> : setup_kernel<FT, GO>)
  if (ftype == GT_F64 || gtype == GT_F64) {
    // the FP64 path: double image, double gradients
    if (ftype != GT_F64 || gtype != GT_F64) return (int)hipErrorInvalidValue;
    hipLaunchKernelGGL(VDVC_SETUP_K(double, double), g, b, 0, st, M, (const double*)f, glt,
                       (double*)gx, (double*)gy, (double*)gz, mask0);
  } else if (ftype == GT_F32) {
    if (gtype == GT_F32)
      hipLaunchKernelGGL(VDVC_SETUP_K(float, float), g, b, 0, st, M, (const float*)f, glt,
                         (float*)gx, (float*)gy, (float*)gz, mask0);
    else
      hipLaunchKernelGGL(VDVC_SETUP_K(float, __hip_bfloat16), g, b, 0, st, M, (const float*)f,
                         glt, (__hip_bfloat16*)gx, (__hip_bfloat16*)gy, (__hip_bfloat16*)gz, mask0);
  } else {
    if (gtype == GT_F32)
      hipLaunchKernelGGL(VDVC_SETUP_K(__half, float), g, b, 0, st, M, (const __half*)f, glt,
                         (float*)gx, (float*)gy, (float*)gz, mask0);
    else
      hipLaunchKernelGGL(VDVC_SETUP_K(__half, __hip_bfloat16), g, b, 0, st, M, (const __half*)f,
                         glt, (__hip_bfloat16*)gx, (__hip_bfloat16*)gy, (__hip_bfloat16*)gz, mask0);
  }
#undef VDVC_SETUP_K
  return (int)hipGetLastError();
}

int launch_stats(int ftype, const MeshDev& M, const void* fld, const uint8_t* mask, int use_thresh,
                 double glt, double* part, int nblocks, hipStream_t st) {
  const int vec = M.pitch % 16 == 0 && (reinterpret_cast<uintptr_t>(fld) & 15) == 0 &&
                  (reinterpret_cast<uintptr_t>(mask) & 7) == 0;
  if (ftype == GT_F32)
    hipLaunchKernelGGL(stats_kernel<float>, dim3(nblocks), dim3(VDVC_BLOCK), 0, st, M,
                       (const float*)fld, mask, use_thresh, glt, vec, part);
  else if (ftype == GT_F64)
    hipLaunchKernelGGL(stats_kernel<double>, dim3(nblocks), dim3(VDVC_BLOCK), 0, st, M,
                       (const double*)fld, mask, use_thresh, glt, vec, part);
  else
    hipLaunchKernelGGL(stats_kernel<__half>, dim3(nblocks), dim3(VDVC_BLOCK), 0, st, M,
                       (const __half*)fld, mask, use_thresh, glt, vec, part);
  return (int)hipGetLastError();
}

// The same pass for a float16 level: a lane takes two neighbouring voxels of
// a 128-voxel chunk and stores them as one 4-byte pair (5 % faster than a
// lane per voxel at 513^3); per lane the voxels are summed in (row, chunk,
// pair) order.  Needs a 4-byte aligned destination.
template <typename ST>
__global__ __launch_bounds__(VDVC_BLOCK) void level_image_h2_kernel(MeshDev M,
                                                                    const ST* __restrict__ src,
                                                                    __half* __restrict__ dst,
                                                                    double glt_d,
                                                                    double* __restrict__ part) {
  constexpr int kCh = 4;
  __shared__ double lds[4 * 3];
  const float glt = (float)glt_d;
  const int lane = threadIdx.x & 63;
  const int nx = M.Nvx, dp = M.pitch, h = M.h;
  const int nch = (dp + 127) >> 7;
  const long long nrows = (long long)M.Nvy * M.Nvz;
  const int lrx = (2 * lane) % h;
  double acc[3] = {0.0, 0.0, 0.0};
  for (long long row = (long long)blockIdx.x * VDVC_WAVES_PER_BLOCK +
                       __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
       row < nrows; row += (long long)gridDim.x * VDVC_WAVES_PER_BLOCK) {
    const int y = (int)(row % M.Nvy), z = (int)(row / M.Nvy);
    const int myz = axis_mult(y, h, M.Nvy - 1) * axis_mult(z, h, M.Nvz - 1);
    const ST* __restrict__ sr = src + row * nx;
    __half2* __restrict__ dr = reinterpret_cast<__half2*>(dst + row * dp);
    for (int c0 = 0; c0 < nch; c0 += kCh) {
      ST raw[kCh][2];
      // a row that starts 8-byte aligned (every other row of an odd-width
      // float32 volume) takes each pair as one 8-byte load
      const bool al8 = ((reinterpret_cast<uintptr_t>(sr) & 7) == 0) && sizeof(ST) == 4;
#pragma unroll
      for (int j = 0; j < kCh; ++j) {
        const int x = (c0 + j) * 128 + 2 * lane;
        const bool ok = c0 + j < nch;
        if (al8 && ok && x + 1 < nx) {
          const float2 q = *reinterpret_cast<const float2*>(sr + x);
          raw[j][0] = q.x;
          raw[j][1] = q.y;
        } else {
          raw[j][0] = (ok && x < nx) ? sr[x] : ST(0);
          raw[j][1] = (ok && x + 1 < nx) ? sr[x + 1] : ST(0);
        }
      }
#pragma unroll
      for (int j = 0; j < kCh; ++j) {
        if (c0 + j >= nch) break;
        const int xb = (c0 + j) * 128;
        const int x = xb + 2 * lane;
        const __half a = LvlCvt<__half>::from(raw[j][0]), b = LvlCvt<__half>::from(raw[j][1]);
        if (x < dp) dr[x >> 1] = __halves2half2(a, b);
        const int rb = lrx + xb % h;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const float val = __half2float(e ? b : a);
          if (x + e < nx && val > glt) {
            int rx = rb + e;
            rx = rx >= h ? rx - h : rx;
            const int mx = (x + e > 0 && x + e < nx - 1 && rx == 0) ? 2 : 1;
            const double mlt = (double)(mx * myz);
            const double d = (double)val;
            acc[0] += mlt;
            acc[1] += mlt * d;
            acc[2] += mlt * d * d;
          }
        }
      }
    }
  }
  block_sum_f64<3>(acc, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 3; ++q) part[3 * (long long)blockIdx.x + q] = acc[q];
  }
}

template <typename S, typename D>
static void level_image_go(int nch, const MeshDev& M, const void* src, void* dst, double glt,
                           double* part, int nblocks, hipStream_t st) {
  // four chunks for rows of at most 256 voxels, else eight (measured at 513-
  // and 257-voxel rows: eight 2-4 % faster than four, twelve or sixteen 7-9 %
  // slower)
  auto k = nch <= 4 ? level_image_kernel<S, D, 4> : level_image_kernel<S, D, 8>;
  hipLaunchKernelGGL(k, dim3(nblocks), dim3(VDVC_BLOCK), 0, st, M, (const S*)src, (D*)dst, glt,
                     part);
}

int launch_level_image(int stype, int dtype, const MeshDev& M, const void* src, void* dst,
                       double glt, double* part, int nblocks, hipStream_t st) {
  const int nch = (M.pitch + 63) / 64;
  if (stype == GT_F32 && dtype == GT_F16 && (reinterpret_cast<uintptr_t>(dst) & 3) == 0)
    hipLaunchKernelGGL(level_image_h2_kernel<float>, dim3(nblocks), dim3(VDVC_BLOCK), 0, st, M,
                       (const float*)src, (__half*)dst, glt, part);
  else if (stype == GT_F32 && dtype == GT_F16)
    level_image_go<float, __half>(nch, M, src, dst, glt, part, nblocks, st);
  else if (stype == GT_F32 && dtype == GT_F32)
    level_image_go<float, float>(nch, M, src, dst, glt, part, nblocks, st);
  else if (stype == GT_F64 && dtype == GT_F64)
    level_image_go<double, double>(nch, M, src, dst, glt, part, nblocks, st);
  else if (stype == GT_F64 && dtype == GT_F32)
    level_image_go<double, float>(nch, M, src, dst, glt, part, nblocks, st);
  else
    return (int)hipErrorInvalidValue;
  return (int)hipGetLastError();
}

int launch_fold(long long n, const uint16_t* gx, const uint16_t* gy, const uint16_t* gz,
                const uint8_t* m, uint16_t* ox, uint16_t* oy, uint16_t* oz, hipStream_t st) {
  auto al = [](const void* p, uintptr_t a) { return (reinterpret_cast<uintptr_t>(p) & (a - 1)) == 0; };
  const int vec = n % 8 == 0 && al(gx, 16) && al(gy, 16) && al(gz, 16) && al(ox, 16) &&
                  al(oy, 16) && al(oz, 16) && al(m, 8);
  const long long work = vec ? n / 8 : n;
  const unsigned nb = (unsigned)((work + VDVC_BLOCK - 1) / VDVC_BLOCK);
  hipLaunchKernelGGL(fold_kernel, dim3(nb), dim3(VDVC_BLOCK), 0, st, n, vec, gx, gy, gz, m, ox, oy,
                     oz);
  return (int)hipGetLastError();
}

int launch_erode6(int nx, int ny, int nz, const uint8_t* in, uint8_t* out, hipStream_t st) {
  const long long n = (long long)nx * ny * nz;
  const unsigned nb = (unsigned)((n + VDVC_BLOCK - 1) / VDVC_BLOCK);
  hipLaunchKernelGGL(erode6_kernel, dim3(nb), dim3(VDVC_BLOCK), 0, st, nx, ny, nz, in, out);
  return (int)hipGetLastError();
}

template <typename R>
int launch_mul(long long n, const R* a, const R* b, R* out, hipStream_t st) {
  const unsigned nb = (unsigned)((n + VDVC_BLOCK - 1) / VDVC_BLOCK);
  hipLaunchKernelGGL(mul_kernel<R>, dim3(nb), dim3(VDVC_BLOCK), 0, st, n, a, b, out);
  return (int)hipGetLastError();
}
template int launch_mul<float>(long long, const float*, const float*, float*, hipStream_t);
template int launch_mul<double>(long long, const double*, const double*, double*, hipStream_t);

int launch_reduce(const double* part, long long n, int nv, double* out, hipStream_t st) {
  hipLaunchKernelGGL(reduce_partials_kernel, dim3((unsigned)nv), dim3(VDVC_BLOCK), 0, st, part, n, nv,
                     out);
  return (int)hipGetLastError();
}

}  // namespace vdvc
