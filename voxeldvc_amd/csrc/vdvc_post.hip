// Post-solve voxel stages of a correlation run, on dense voxel arrays (flat
// index x + Nvx*(y + Nvy*z), the reference's order, no row pitch):
//   * post_voxel_kernel<R, POST_INTERP>: nodal field at every voxel
//     (engine/correlate_gpu.py:160-198, interp_field_vox);
//   * post_voxel_kernel<R, POST_MASK>: mask0 & g.InBounds(x+u) & g(x+u) > glt
//     with the sample taken in g's own dtype (correlate_gpu.py:284-338);
//   * post_voxel_kernel<R, POST_PUSH>: field(x - u(x)) in field's dtype
//     (engine/write_output.py:571-622);
//   * quality_cell_kernel: per-element residual std and ZNCC over the active
//     voxels (write_output.py:380-447);
//   * zncc_pass_kernel: masked ZNCC over all voxels in two deterministic FP64
//     passes (write_output.py:658-690).
//
// Every stage is one sweep over the voxels.  In the voxel kernels a thread
// owns one voxel column (x, y) of one element layer: the element's 8 nodal
// values and the x/y part of the Hex8 weights are formed once, then the
// layer's z planes are walked; 64 x-consecutive columns per wavefront keep
// every plane access a coalesced dense row.
//
// Arithmetic follows the reference bit for bit where it is defined:
//   * u(x) = sum over the 8 corners a = 0..7, in that order, of
//     float(N_a) * U_a in U's dtype, N_a the voxel-centre Hex8 weight formed
//     in double as geometry_dvc.build_N_stencil forms it; the owning element
//     is the highest-numbered one containing the voxel (the reference's
//     later-element-overwrites rule);
//   * sampling is scipy.ndimage.map_coordinates for spline order 0 / 1 and
//     modes 'nearest' / 'constant': float64 coordinates, order 0 at index
//     floor(c + 0.5), order 1 summing ((v * wx) * wy) * wz over the corners
//     with the z index fastest, corner indices clamped ('nearest'), cval for a
//     point outside [0, n-1] ('constant'), then the output conversion of the
//     image dtype (round half away from zero for integers, v != 0 for bool).
// All products and sums use explicitly rounded operations (no contraction).
#include "vdvc_common.h"
#include "vdvc_kernels.h"
#include "vdvc_sample01.h"
#include "vdvc_elem.h"

namespace {

using vdvc_img::as_img_type;
using vdvc_img::ld_img;
using vdvc_img::sample01;
using vdvc_img::st_img;

// Hex8 corner offsets in the reference's local order 0=(0,0,0) 1=(1,0,0)
// 2=(1,1,0) 3=(0,1,0) 4..7 = the same with z = 1 (engine/geometry_dvc.py:14-17);
// compile-time functions so unrolled loops index registers statically
__device__ __forceinline__ constexpr int corner_x(int a) {
  return ((a & 3) == 1 || (a & 3) == 2) ? 1 : 0;
}
__device__ __forceinline__ constexpr int corner_y(int a) { return (a >> 1) & 1; }
__device__ __forceinline__ constexpr int corner_z(int a) { return a >> 2; }

// stencil coordinate 2k/h - 1 of local voxel k, formed as
// geometry_dvc.build_N_stencil forms it (float64)
__device__ __forceinline__ double stencil_xi(int k, int h) {
  return sub_rn(__ddiv_rn(2.0 * k, (double)h), 1.0);
}

// Interpolation state of one voxel column (x, y) inside element layer ez:
// the x/y part of the 8 Hex8 weights, 0.125 (1 +- xi) (1 +- eta), and the
// element's 8 nodal values.  Per voxel only the z factor remains:
// N_a = ((0.125 Fx) Fy) Fz, the reference's product order.
template <typename R>
struct Column {
  double pxy[2][2];
  R uc[8][3];
  __device__ __forceinline__ void init(const PostIn& P, int x, int y, int ez) {
    const int h = P.h;
    const int ex = min(x / h, P.Nx - 1), ey = min(y / h, P.Ny - 1);
    const double xi = stencil_xi(x - ex * h, h), et = stencil_xi(y - ey * h, h);
    const double Fx[2] = {sub_rn(1.0, xi), add_rn(1.0, xi)};
    const double Fy[2] = {sub_rn(1.0, et), add_rn(1.0, et)};
#pragma unroll
    for (int cx = 0; cx < 2; ++cx)
#pragma unroll
      for (int cy = 0; cy < 2; ++cy) pxy[cx][cy] = mul_rn(mul_rn(0.125, Fx[cx]), Fy[cy]);
    const R* U = reinterpret_cast<const R*>(P.U);
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const long long node = (long long)(ex + corner_x(a)) +
                             (long long)P.nnx * ((long long)(ey + corner_y(a)) +
                                                 (long long)P.nny * (ez + corner_z(a)));
#pragma unroll
      for (int c = 0; c < 3; ++c) uc[a][c] = U[3 * node + c];
    }
  }
  // u at local z voxel tz of the layer: sum over corners a = 0..7 in order.
  // The z factors are the same for every lane of a plane: they come from the
  // launch's table (h <= 64) instead of a float64 division per voxel.
  __device__ __forceinline__ void at(const PostIn& P, int tz, R (&u)[3]) const {
    double Fz[2];
    if (P.fz_tab) {
      Fz[0] = P.fz[tz][0];
      Fz[1] = P.fz[tz][1];
    } else {
      const double ze = stencil_xi(tz, P.h);
      Fz[0] = sub_rn(1.0, ze);
      Fz[1] = add_rn(1.0, ze);
    }
    u[0] = u[1] = u[2] = R(0);
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const R N = (R)mul_rn(pxy[corner_x(a)][corner_y(a)], Fz[corner_z(a)]);
#pragma unroll
      for (int c = 0; c < 3; ++c) u[c] = add_rn(u[c], mul_rn(N, uc[a][c]));
    }
  }
};

__device__ __forceinline__ float clip_r(float c, float hi) { return fminf(fmaxf(c, 0.0f), hi); }
__device__ __forceinline__ double clip_r(double c, double hi) { return fmin(fmax(c, 0.0), hi); }

// One thread per voxel column (x, y) of one element layer (blockIdx.z): the
// layer's h voxel planes (h + 1 for the last layer, which owns the last
// plane), 64 x-consecutive columns per wavefront.  The sampling stages work
// on two planes per step so both planes' gathers are in flight together.
// IT is the sampled image's element type (POST_PUSH, POST_MASK); O32: the
// image holds fewer than 2^32 bytes and is read with 32-bit offsets.
template <typename R, int OP, int IT, bool O32>
__global__ __launch_bounds__(VDVC_BLOCK) void post_voxel_kernel(PostIn P) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int ez = blockIdx.z;
  if (x >= P.Nvx || y >= P.Nvy) return;
  const int h = P.h;
  Column<R> col;
  col.init(P, x, y, ez);
  const int nz = h + (ez == P.Nz - 1 ? 1 : 0);
  const long long plane = (long long)P.Nvx * P.Nvy;
  const long long v0 = (long long)x + (long long)P.Nvx * y + plane * ((long long)ez * h);
  if (OP == POST_INTERP) {
    R* o = reinterpret_cast<R*>(P.out);
    long long v = v0;
    for (int tz = 0; tz < nz; ++tz, v += plane) {
      R u[3];
      col.at(P, tz, u);
      if (P.channels_first) {
        o[v] = u[0];
        o[P.Nvox + v] = u[1];
        o[2 * P.Nvox + v] = u[2];
      } else {
        o[3 * v] = u[0];
        o[3 * v + 1] = u[1];
        o[3 * v + 2] = u[2];
      }
    }
    return;
  }
  const __amdgpu_buffer_rsrc_t rs =
      vdvc_img::img_rsrc(P.img, O32 ? P.img_bytes : 0u);
  // the stage's value at local plane tz: the pushed field value (PUSH) or
  // the active-mask bit (MASK)
  auto value = [&](int tz) -> double {
    const int z = ez * h + tz;
    R u[3];
    col.at(P, tz, u);
    if (OP == POST_PUSH) {
      // coordinates x - u(x) in U's dtype
      const double cx = (double)sub_rn((R)x, u[0]);
      const double cy = (double)sub_rn((R)y, u[1]);
      const double cz = (double)sub_rn((R)z, u[2]);
      return sample01<IT, O32>(P.img, P.gnx, P.gny, P.gnz, cx, cy, cz, P.order, P.mode, P.cval,
                               rs);
    }
    // (u + x) - origin in U's dtype; bounds on the closed box; clipped
    R cx = sub_rn(add_rn(u[0], (R)x), (R)P.ox);
    R cy = sub_rn(add_rn(u[1], (R)y), (R)P.oy);
    R cz = sub_rn(add_rn(u[2], (R)z), (R)P.oz);
    const R hx = (R)(P.gnx - 1), hy = (R)(P.gny - 1), hz = (R)(P.gnz - 1);
    const bool inb = cx >= R(0) && cx <= hx && cy >= R(0) && cy <= hy && cz >= R(0) && cz <= hz;
    cx = clip_r(cx, hx);
    cy = clip_r(cy, hy);
    cz = clip_r(cz, hz);
    const double gv = as_img_type(sample01<IT, O32>(P.img, P.gnx, P.gny, P.gnz, (double)cx,
                                                    (double)cy, (double)cz, 1, 0, 0.0, rs),
                                  IT);
    // g > glt in g's dtype for a float32 image (the threshold is a weak
    // scalar there), in double otherwise
    const bool above = IT == GT_F32 ? (__double2float_rn(gv) > __double2float_rn(P.glt))
                                    : (gv > P.glt);
    return (inb && above) ? 1.0 : 0.0;
  };
  auto store = [&](long long v, double val) {
    if (OP == POST_PUSH) {
      st_img(P.out, IT, v, val);
    } else {
      const bool m = val != 0.0 && (P.mask0 == nullptr || P.mask0[v] != 0);
      reinterpret_cast<uint8_t*>(P.out)[v] = m ? 1 : 0;
    }
  };
  long long v = v0;
  for (int tz = 0; tz < nz; tz += 2, v += 2 * plane) {
    const bool two = tz + 1 < nz;
    const double va = value(tz);
    const double vb = value(two ? tz + 1 : tz);
    store(v, va);
    if (two) store(v + plane, vb);
  }
}

// One wavefront per element: two sweeps over its (h+1)^3 voxels (the second
// re-reads them from L2), FP64 sums reduced across the wave in fixed order.
// FTY: the reference image's type code, fixed at compile time so that the
// walk's loads are not separated by a type switch
template <typename RT, int FTY>
__global__ __launch_bounds__(VDVC_BLOCK) void quality_cell_kernel(QualIn Q) {
  const int lane = threadIdx.x & 63;
  const long long e = (long long)blockIdx.x * VDVC_WAVES_PER_BLOCK + (threadIdx.x >> 6);
  const long long Ne = (long long)Q.Nx * Q.Ny * Q.Nz;
  if (e >= Ne) return;
  const int i = (int)(e % Q.Nx);
  const int j = (int)((e / Q.Nx) % Q.Ny);
  const int k = (int)(e / ((long long)Q.Nx * Q.Ny));
  const long long base =
      (long long)i * Q.h + (long long)Q.Nvx * ((long long)j * Q.h + (long long)Q.Nvy * k * Q.h);
  const RT* res = reinterpret_cast<const RT*>(Q.res);
  double cnt = 0.0, sr = 0.0, srr = 0.0, sf = 0.0, sg = 0.0;
  vdvc_elem::walk(Q.h, Q.Nvx, Q.Nvy, base, lane, [&](long long vi) {
    const double r = (double)res[vi];
    const double f = ld_img(Q.ref, FTY, vi);
    const double m = r != 0.0 ? 1.0 : 0.0;
    const double g = sub_rn(f, r);
    cnt += m;
    sr = add_rn(sr, mul_rn(r, m));
    srr = add_rn(srr, mul_rn(mul_rn(r, r), m));
    sf = add_rn(sf, mul_rn(f, m));
    sg = add_rn(sg, mul_rn(g, m));
  });
  cnt = wave_sum(cnt);
  sr = wave_sum(sr);
  srr = wave_sum(srr);
  sf = wave_sum(sf);
  sg = wave_sum(sg);
  const double dc = fmax(cnt, 1.0);
  const double mean_r = sr / dc;
  const double var_r = sub_rn(srr / dc, mul_rn(mean_r, mean_r));
  const double fm = sf / dc, gm = sg / dc;
  double sfg = 0.0, sff = 0.0, sgg = 0.0;
  vdvc_elem::walk(Q.h, Q.Nvx, Q.Nvy, base, lane, [&](long long vi) {
    const double r = (double)res[vi];
    const double f = ld_img(Q.ref, FTY, vi);
    const double m = r != 0.0 ? 1.0 : 0.0;
    const double fc = mul_rn(sub_rn(f, fm), m);
    const double gc = mul_rn(sub_rn(sub_rn(f, r), gm), m);
    sfg = add_rn(sfg, mul_rn(fc, gc));
    sff = add_rn(sff, mul_rn(fc, fc));
    sgg = add_rn(sgg, mul_rn(gc, gc));
  });
  sfg = wave_sum(sfg);
  sff = wave_sum(sff);
  sgg = wave_sum(sgg);
  if (lane == 0) {
    const double nan = __builtin_nan("");
    const double den = sqrt(mul_rn(sff, sgg));
    const bool few = cnt < (double)Q.min_active;
    Q.std_out[e] = few ? nan : sqrt(fmax(var_r, 0.0));
    Q.zncc_out[e] = (few || !(den > 0.0)) ? nan : sfg / den;
  }
}

// four consecutive elements i..i+3 (i a multiple of 4, array 16-byte
// aligned) with one vector load
__device__ __forceinline__ void ld4_img(const void* p, int t, long long i, double (&v)[4]) {
  switch (t) {
    case GT_F32: {
      const float4 q = *reinterpret_cast<const float4*>(reinterpret_cast<const float*>(p) + i);
      v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
      break;
    }
    case GT_F64: {
      const double2* d = reinterpret_cast<const double2*>(reinterpret_cast<const double*>(p) + i);
      const double2 a = d[0], b = d[1];
      v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
      break;
    }
    case GT_U16:
    case GT_I16: {
      const uint2 q = *reinterpret_cast<const uint2*>(reinterpret_cast<const uint16_t*>(p) + i);
      const unsigned w[4] = {q.x & 0xffffu, q.x >> 16, q.y & 0xffffu, q.y >> 16};
#pragma unroll
      for (int k = 0; k < 4; ++k)
        v[k] = t == GT_I16 ? (double)(int16_t)(uint16_t)w[k] : (double)w[k];
      break;
    }
    default: {  // GT_U8, GT_B8
      const unsigned q = *reinterpret_cast<const unsigned*>(reinterpret_cast<const uint8_t*>(p) + i);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const unsigned b = (q >> (8 * k)) & 0xffu;
        v[k] = t == GT_B8 ? (b != 0u ? 1.0 : 0.0) : (double)b;
      }
    }
  }
}

// pass 0 accumulates (count, sum f, sum g), pass 1 the centred products, of
// one selected element
__device__ __forceinline__ void zncc_acc(int pass, double m, double f, double g, double fm,
                                         double gm, double (&acc)[3]) {
  if (pass == 0) {
    acc[0] += m;
    acc[1] = add_rn(acc[1], mul_rn(f, m));
    acc[2] = add_rn(acc[2], mul_rn(g, m));
  } else {
    const double fc = mul_rn(sub_rn(f, fm), m), gc = mul_rn(sub_rn(g, gm), m);
    acc[0] = add_rn(acc[0], mul_rn(fc, gc));
    acc[1] = add_rn(acc[1], mul_rn(fc, fc));
    acc[2] = add_rn(acc[2], mul_rn(gc, gc));
  }
}

// Block b sums voxels [b*chunk, min(n, (b+1)*chunk)); fixed-order block
// reduction; partials [nblocks][3].  Unselected voxels contribute m = 0
// (loads are unconditional, so each thread keeps several in flight).
__global__ __launch_bounds__(VDVC_BLOCK) void zncc_pass_kernel(ZnccIn Z, int pass,
                                                               const double* red, double* part) {
  __shared__ double lds[4 * 3];
  const long long b0 = (long long)blockIdx.x * Z.chunk;
  const long long b1 = max(b0, min(Z.n, b0 + Z.chunk));
  double fm = 0.0, gm = 0.0;
  if (pass == 1) {
    fm = red[1] / red[0];
    gm = red[2] / red[0];
  }
  double acc[3] = {0.0, 0.0, 0.0};
  long long i0 = b0;
  if (Z.vec) {
    const long long e4 = b0 + (b1 - b0) / 4 * 4;
    for (long long i = b0 + 4 * (long long)threadIdx.x; i < e4; i += 4 * VDVC_BLOCK) {
      double f[4], g[4], m[4] = {1.0, 1.0, 1.0, 1.0};
      ld4_img(Z.f, Z.ftype, i, f);
      ld4_img(Z.g, Z.gtype, i, g);
      if (Z.mask) ld4_img(Z.mask, GT_B8, i, m);
#pragma unroll
      for (int k = 0; k < 4; ++k)
        zncc_acc(pass, m[k], f[k], Z.g_is_residual ? sub_rn(f[k], g[k]) : g[k], fm, gm, acc);
    }
    i0 = e4;
  }
  for (long long i = i0 + threadIdx.x; i < b1; i += VDVC_BLOCK) {
    const double m = Z.mask ? (Z.mask[i] != 0 ? 1.0 : 0.0) : 1.0;
    const double f = ld_img(Z.f, Z.ftype, i);
    double g = ld_img(Z.g, Z.gtype, i);
    if (Z.g_is_residual) g = sub_rn(f, g);
    zncc_acc(pass, m, f, g, fm, gm, acc);
  }
  block_sum_f64<3>(acc, lds);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = acc[0];
    part[3 * blockIdx.x + 1] = acc[1];
    part[3 * blockIdx.x + 2] = acc[2];
  }
}

}  // namespace

namespace vdvc {

template <typename R, int OP, bool O32>
void launch_post_typed(const PostIn& P, dim3 g, dim3 b, hipStream_t st) {
  switch (P.itype) {
    case GT_F32: hipLaunchKernelGGL((post_voxel_kernel<R, OP, GT_F32, O32>), g, b, 0, st, P); break;
    case GT_F64: hipLaunchKernelGGL((post_voxel_kernel<R, OP, GT_F64, O32>), g, b, 0, st, P); break;
    case GT_U8: hipLaunchKernelGGL((post_voxel_kernel<R, OP, GT_U8, O32>), g, b, 0, st, P); break;
    case GT_U16: hipLaunchKernelGGL((post_voxel_kernel<R, OP, GT_U16, O32>), g, b, 0, st, P); break;
    case GT_I16: hipLaunchKernelGGL((post_voxel_kernel<R, OP, GT_I16, O32>), g, b, 0, st, P); break;
    default: hipLaunchKernelGGL((post_voxel_kernel<R, OP, GT_B8, O32>), g, b, 0, st, P); break;
  }
}

template <typename R, int OP>
void launch_post_typed(const PostIn& P, dim3 g, dim3 b, hipStream_t st) {
  if (P.img_bytes)
    launch_post_typed<R, OP, true>(P, g, b, st);
  else
    launch_post_typed<R, OP, false>(P, g, b, st);
}

template <typename R>
void launch_post_r(int op, const PostIn& P, dim3 g, dim3 b, hipStream_t st) {
  if (op == POST_INTERP)
    hipLaunchKernelGGL((post_voxel_kernel<R, POST_INTERP, GT_F32, false>), g, b, 0, st, P);
  else if (op == POST_PUSH)
    launch_post_typed<R, POST_PUSH>(P, g, b, st);
  else
    launch_post_typed<R, POST_MASK>(P, g, b, st);
}

int launch_post_voxel(int op, int rtype, const PostIn& P, hipStream_t st) {
  dim3 g((unsigned)((P.Nvx + 63) / 64), (unsigned)((P.Nvy + 3) / 4), (unsigned)P.Nz);
  dim3 b(VDVC_BLOCK);
  if (rtype == GT_F64)
    launch_post_r<double>(op, P, g, b, st);
  else
    launch_post_r<float>(op, P, g, b, st);
  return (int)hipGetLastError();
}

int launch_quality_cell(const QualIn& Q, hipStream_t st) {
  const long long Ne = (long long)Q.Nx * Q.Ny * Q.Nz;
  const unsigned nb = (unsigned)((Ne + VDVC_WAVES_PER_BLOCK - 1) / VDVC_WAVES_PER_BLOCK);
#define VDVC_QC(RT)                                                                        \
  switch (Q.ftype) {                                                                       \
    case GT_F32: hipLaunchKernelGGL((quality_cell_kernel<RT, GT_F32>), dim3(nb), dim3(VDVC_BLOCK), 0, st, Q); break; \
    case GT_F64: hipLaunchKernelGGL((quality_cell_kernel<RT, GT_F64>), dim3(nb), dim3(VDVC_BLOCK), 0, st, Q); break; \
    case GT_U8: hipLaunchKernelGGL((quality_cell_kernel<RT, GT_U8>), dim3(nb), dim3(VDVC_BLOCK), 0, st, Q); break;   \
    case GT_U16: hipLaunchKernelGGL((quality_cell_kernel<RT, GT_U16>), dim3(nb), dim3(VDVC_BLOCK), 0, st, Q); break; \
    case GT_I16: hipLaunchKernelGGL((quality_cell_kernel<RT, GT_I16>), dim3(nb), dim3(VDVC_BLOCK), 0, st, Q); break; \
    default: hipLaunchKernelGGL((quality_cell_kernel<RT, GT_B8>), dim3(nb), dim3(VDVC_BLOCK), 0, st, Q); break;      \
  }
  if (Q.rtype == GT_F64) {
    VDVC_QC(double)
  } else {
    VDVC_QC(float)
  }
#undef VDVC_QC
  return (int)hipGetLastError();
}

int launch_zncc_pass(int pass, const ZnccIn& Z, const double* red, double* part, int nblocks,
                     hipStream_t st) {
  hipLaunchKernelGGL(zncc_pass_kernel, dim3(nblocks), dim3(VDVC_BLOCK), 0, st, Z, pass, red,
                     part);
  return (int)hipGetLastError();
}

}  // namespace vdvc
