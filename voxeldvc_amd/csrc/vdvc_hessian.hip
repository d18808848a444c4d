// Vectorised voxel-streaming kernel for the correlation Hessian matvec, its
// diagonal and the fused residual + right-hand side (the GN hot path).
//
// Reference semantics: engine/kernels_dvc.py:452-543 (matvec_H,
// build_diagonal_H), engine/correlate_gpu.py:341-378 and 1091-1109
// (compute_b_gpu and the residual), in the unique-voxel form described in
// vdvc_common.h.
//
// Decomposition (powers of two h <= 64).  A lane owns V consecutive voxels in
// x (V | h, V * element size <= 16 bytes, so one aligned 16-byte load per
// gradient plane and row); a wavefront owns a strip of 64 * V voxels -- whole
// elements -- in x, one element row in y and ez elements in z, and walks
// rows then planes.  Per row, the lane forms s = W m (g . u) for its V voxels
// (u interpolated along x between the element's left and right node
// columns), and accumulates the x-restriction S = sum s g, T = sum (t/h) s g;
// the y-restriction is a running weighted sum per plane and the
// z-restriction carries two node planes in registers.  At the end of every
// element layer the h/V lanes of one element are reduced with a butterfly and
// the strip's node plane is written to the wave's slot block; a node gather
// kernel sums the slots of the (at most 8) waves touching a node in fixed
// order.  The last voxel plane x = Nx*h is handled by the lane that owns the
// last block of the last element, as its local coordinate t = h.
//
// Paired elements (E = 2, for h = V / 2: bfloat16 at h = 4, float32 or
// bfloat16 at h = 2).  A lane of 4-voxel rows spends as much on its per-row
// work -- the node interpolation, the y / z weights, the accumulation into
// the node planes -- as on its voxels.  With E = 2 a lane owns two adjacent
// elements (V = 2h voxels, one 16-byte load per plane): the two share their
// middle node column, so a row interpolates three node columns and
// accumulates three, for twice the voxels, and the lanes write their node
// columns without a butterfly.
#include <cstdlib>
#include <type_traits>

#include "vdvc_kernels.h"

namespace {

// ---- aligned vector stores / loads of V consecutive elements --------------
// (the index is a 32-bit element offset from a wave-uniform row pointer)
template <int V>
__device__ __forceinline__ void st_f32(float* base, unsigned i, const float (&v)[V]) {
  float* p = base + i;
  if constexpr (V == 8) {
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
    *reinterpret_cast<float4*>(p + 4) = make_float4(v[4], v[5], v[6], v[7]);
  } else if constexpr (V == 4) {
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
  } else if constexpr (V == 2) {
    *reinterpret_cast<float2*>(p) = make_float2(v[0], v[1]);
  } else {
    p[0] = v[0];
  }
}

__device__ __forceinline__ uint32_t bf16_bits(float f) {
  return (uint32_t)__bfloat16_as_ushort(__float2bfloat16(f));
}

template <int V>
__device__ __forceinline__ void st_bf16(uint16_t* base, unsigned i, const float (&v)[V]) {
  uint16_t* p = base + i;
  if constexpr (V == 8) {
    uint4 q;
    q.x = bf16_bits(v[0]) | (bf16_bits(v[1]) << 16);
    q.y = bf16_bits(v[2]) | (bf16_bits(v[3]) << 16);
    q.z = bf16_bits(v[4]) | (bf16_bits(v[5]) << 16);
    q.w = bf16_bits(v[6]) | (bf16_bits(v[7]) << 16);
    *reinterpret_cast<uint4*>(p) = q;
  } else if constexpr (V == 4) {
    uint2 q;
    q.x = bf16_bits(v[0]) | (bf16_bits(v[1]) << 16);
    q.y = bf16_bits(v[2]) | (bf16_bits(v[3]) << 16);
    *reinterpret_cast<uint2*>(p) = q;
  } else if constexpr (V == 2) {
    *reinterpret_cast<uint32_t*>(p) = bf16_bits(v[0]) | (bf16_bits(v[1]) << 16);
  } else {
    p[0] = (uint16_t)bf16_bits(v[0]);
  }
}

// raw bits of V consecutive elements of T (a 16-, 8-, 4-, 2- or 1-byte
// aligned load), unpacked separately so a row can be fetched one iteration
// ahead of its use
template <typename T, int V>
struct Raw {
  static constexpr int NB = V * (int)sizeof(T);
  static constexpr int NW = NB >= 4 ? NB / 4 : 1;
  uint32_t w[NW];
};

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// The voxel streams are read once per pass, so they are fetched with
// nontemporal loads (streaming cache policy): measured on MI355X, a 16-byte
// streaming read reaches about 6.0 TB/s this way against 5.5 TB/s with the
// default policy.
template <typename T, int V>
__device__ __forceinline__ void ld_raw(const T* base, unsigned i, Raw<T, V>& r) {
  const T* p = base + i;
  constexpr int NB = Raw<T, V>::NB;
  if constexpr (NB == 32) {
    const u32x4 q0 = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
    const u32x4 q1 = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p) + 1);
    r.w[0] = q0.x; r.w[1] = q0.y; r.w[2] = q0.z; r.w[3] = q0.w;
    r.w[4] = q1.x; r.w[5] = q1.y; r.w[6] = q1.z; r.w[7] = q1.w;
  } else if constexpr (NB == 16) {
    const u32x4 q = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p));
    r.w[0] = q.x; r.w[1] = q.y; r.w[2] = q.z; r.w[3] = q.w;
  } else if constexpr (NB == 8) {
    const u32x2 q = __builtin_nontemporal_load(reinterpret_cast<const u32x2*>(p));
    r.w[0] = q.x; r.w[1] = q.y;
  } else if constexpr (NB == 4) {
    r.w[0] = __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(p));
  } else if constexpr (NB == 2) {
    r.w[0] = __builtin_nontemporal_load(reinterpret_cast<const uint16_t*>(p));
  } else {
    r.w[0] = __builtin_nontemporal_load(reinterpret_cast<const uint8_t*>(p));
  }
}

template <typename T, int V>
__device__ __forceinline__ void unpack(const Raw<T, V>& r, float (&o)[V]) {
  if constexpr (sizeof(T) == 4) {
#pragma unroll
    for (int j = 0; j < V; ++j) o[j] = __uint_as_float(r.w[j]);
  } else if constexpr (std::is_same<T, __half>::value) {
#pragma unroll
    for (int j = 0; j < V; ++j) {
      const uint32_t wd = r.w[j / 2];
      o[j] = __half2float(__ushort_as_half((uint16_t)((j & 1) ? (wd >> 16) : (wd & 0xffffu))));
    }
  } else {  // bfloat16
#pragma unroll
    for (int j = 0; j < V; ++j) {
      const uint32_t wd = r.w[j / 2];
      o[j] = __uint_as_float((j & 1) ? (wd & 0xffff0000u) : (wd << 16));
    }
  }
}

// raw scalar of type T (the last voxel plane's values are fetched as scalars)
template <typename T>
__device__ __forceinline__ uint32_t ld_raw1(const T* base, long long i) {
  if constexpr (sizeof(T) == 4) return *reinterpret_cast<const uint32_t*>(base + i);
  else if constexpr (sizeof(T) == 2) return *reinterpret_cast<const uint16_t*>(base + i);
  else return *reinterpret_cast<const uint8_t*>(base + i);
}

template <typename T>
__device__ __forceinline__ float unpack1(uint32_t r) {
  if constexpr (sizeof(T) == 4) return __uint_as_float(r);
  else return __uint_as_float(r << 16);  // bfloat16
}

template <int V>
__device__ __forceinline__ uint32_t mask_bit(const Raw<uint8_t, V>& r, int j) {
  return (r.w[j / 4] >> (8 * (j & 3))) & 0xffu;
}

// one voxel's contribution to the lane's x-restriction sums: for the linear
// modes S += q, T += (t/h) q with q = s g; for the diagonal S += X0^2 q,
// T += X1^2 q with q = W m g^2
template <int MODE>
__device__ __forceinline__ void acc_voxel(float gx, float gy, float gz, float wm, float X1,
                                          float s_or_dot, float (&S)[3], float (&T)[3]) {
  if (MODE == MODE_DIAG) {
    const float X0 = 1.0f - X1;
    const float a0 = X0 * X0 * wm, a1 = X1 * X1 * wm;
    const float qx = gx * gx, qy = gy * gy, qz = gz * gz;
    S[0] += a0 * qx; S[1] += a0 * qy; S[2] += a0 * qz;
    T[0] += a1 * qx; T[1] += a1 * qy; T[2] += a1 * qz;
  } else {
    const float s = wm * s_or_dot;
    const float t = X1 * s;
    S[0] += s * gx; S[1] += s * gy; S[2] += s * gz;
    T[0] += t * gx; T[1] += t * gy; T[2] += t * gz;
  }
}

}  // namespace

// Blocks per CU the register budget is fitted to: 3 (occupancy 3 waves per
// SIMD, <= 168 VGPRs) for the products and the diagonal, 2-byte or 4-byte
// gradients, single or paired elements.  Fitted to 4 (<= 128 VGPRs) the
// 2-byte build spilled 28 bytes per lane; with two element layers per
// wavefront a 513^3 h = 8 grid is 2048 wavefronts, which occupancy 3 runs in
// one round without the spill (product 155.7 -> 145 us, 1025^2 x 769 h = 16
// 798 -> 787 us).  The residual pass streams three more voxel arrays (f,
// sampled g, live mask) through the row pipeline and is fitted to 2 (<= 256
// VGPRs): at 4 its registers spilled.
#ifndef VDVC_VEC_MIN_BLOCKS
#define VDVC_VEC_MIN_BLOCKS(MODE, GT, E) (MODE == MODE_RHS ? 2 : 3)
#endif

// DG (residual pass only): the Hessian diagonal of the b-mask's H is taken
// from the same voxels in the same pass -- per voxel the diagonal mode's
// operations on g with the mask weight, which equal its operations on the
// folded g * mask -- into a second slot block set (in.pdiag), so a GN
// iteration with a dynamic mask reads its gradients once for b, the refold
// and the Jacobi preconditioner.
template <int MODE, typename GT, bool HASMASK, typename FT, bool FOLD, int V, int E,
          bool DG = false>
__global__ __launch_bounds__(VDVC_BLOCK, VDVC_VEC_MIN_BLOCKS(MODE, GT, E)) void vox_vec_kernel(MeshDev M, TilesDev T, VoxIn in,
                                                            const float* __restrict__ v,
                                                            float* __restrict__ P) {
  static_assert(E == 1 || (E == 2 && MODE != MODE_RHS && V % 2 == 0),
                "paired elements: products and diagonal only");
  static_assert(!DG || (MODE == MODE_RHS && E == 1), "the diagonal rides on the residual pass");
  // element width of a paired-element lane (h = V / E); unused for E = 1
  constexpr int HE = V / E;
  __shared__ float lds_nodes[MODE == MODE_MATVEC ? VDVC_WAVES_PER_BLOCK * 64 * 12 * (E + 1) : 1];
  const int lane = threadIdx.x & 63;
  // the wave index is uniform across the wave; saying so keeps every tile,
  // row and address computation below on the scalar unit
  const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const unsigned blk = xcd_remap(blockIdx.x, gridDim.x);
  const long long w = (long long)blk * VDVC_WAVES_PER_BLOCK + wib;
  if (w >= T.nwaves) return;
  if (in.done && *in.done) return;

  const int h = M.h;
  const float inv_h = M.inv_h;
  const int sx = (int)(w % T.nsx);
  long long r = w / T.nsx;
  const int ys = (int)(r % T.nys);
  r /= T.nys;
  const int ey = (int)(r % M.Ny);
  const int zt = (int)(r / M.Ny);

  const int SW = 64 * V;
  const int x0 = sx * SW;
  const int xend = M.Nx * h;              // owned voxels [0, xend); x = xend is the tail
  const int xb = x0 + lane * V;
  const bool act = xb < xend;
  const int xbc = act ? xb : x0;          // in-range address for inactive lanes
  // the lane's offset within a row: a 32-bit offset from a wave-uniform row
  // pointer, so every row load / store is a scalar base plus this register
  const unsigned xbu = (unsigned)xbc;
  // inactive lanes stream in-tile data like the others; their node weights
  // are zero
  const float actw = act ? 1.0f : 0.0f;
  const int ex = xbc / h;                 // element of this lane's block
  const int tb = xbc - ex * h;            // local coordinate of the block's first voxel
  const bool tail = act && (xb + V == xend);
  const float X1b = tb * inv_h;
  const float wx0 = (xb == 0) ? 0.5f : 1.0f;  // x = 0 boundary weight (voxel 0 only)
  const int G = E > 1 ? 1 : h / V;       // lanes per element
  const int e_lo = x0 / h;
  const int e_hi = min(M.Nx, (x0 + SW) / h);
  const int nn = e_hi - e_lo + 1;

  const int ez0 = zt * T.ez;
  const int ez1 = min(M.Nz, ez0 + T.ez);
  // this wave's rows of the element row: [ty0, ty0 + rows per part), plus
  // the outer boundary row for the last part of the last element row
  const int rpp = h / T.nys;
  const int ty0 = ys * rpp;
  const int y0 = ey * h + ty0;
  const int nrows = rpp + ((ey == M.Ny - 1 && ys == T.nys - 1) ? 1 : 0);
  const long long pl = (long long)M.pitch * M.Nvy;  // plane stride

  const GT* g0 = reinterpret_cast<const GT*>(in.g[0]);
  const GT* g1 = reinterpret_cast<const GT*>(in.g[1]);
  const GT* g2 = reinterpret_cast<const GT*>(in.g[2]);
  const FT* fim = reinterpret_cast<const FT*>(in.fimg);
  // residual-mode inputs (this kernel computes in float)
  const float* gvals_f = reinterpret_cast<const float*>(in.gvals);
  float* res_f = reinterpret_cast<float*>(in.res);
  const float mean0_f = (float)in.mean0, gmean_f = (float)in.gmean, kscale_f = (float)in.kscale;
  float* Pw = P + (size_t)w * T.wave_stride;
  float* Pd = DG ? in.pdiag + (size_t)w * T.wave_stride : nullptr;

  // node-plane accumulators C[cx][cy][cz][k] of the lane's E + 1 node columns
  float C[E + 1][2][2][3];
#pragma unroll
  for (int a = 0; a < E + 1; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int k = 0; k < 3; ++k) C[a][b][c][k] = 0.0f;
  // DG: the diagonal's node planes
  constexpr int ND = DG ? E + 1 : 1;
  float C2[ND][2][2][3];
#pragma unroll
  for (int a = 0; a < ND; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int k = 0; k < 3; ++k) C2[a][b][c][k] = 0.0f;

  double st_c = 0.0, st_1 = 0.0, st_2 = 0.0, st_n = 0.0;
  // residual pass: display statistics requested (uniform across the grid)
  const bool stats = MODE == MODE_RHS && in.stats != nullptr;

  auto emit = [&](const auto& CC, float* PWb, int kzl) {
    // lanes of one element hold partial sums of that element's two x nodes
    // (a paired-element lane: of its three)
    float a[E + 1][2][3];
#pragma unroll
    for (int cx = 0; cx < E + 1; ++cx)
#pragma unroll
      for (int cy = 0; cy < 2; ++cy)
#pragma unroll
        for (int k = 0; k < 3; ++k) a[cx][cy][k] = CC[cx][cy][0][k];
    for (int o = 1; o < G; o <<= 1) {
#pragma unroll
      for (int cx = 0; cx < E + 1; ++cx)
#pragma unroll
        for (int cy = 0; cy < 2; ++cy)
#pragma unroll
          for (int k = 0; k < 3; ++k) a[cx][cy][k] += __shfl_xor(a[cx][cy][k], o, 64);
    }
    float prev[2][3];
    const int src = lane > 0 ? lane - 1 : 0;
#pragma unroll
    for (int cy = 0; cy < 2; ++cy)
#pragma unroll
      for (int k = 0; k < 3; ++k) prev[cy][k] = __shfl(a[E][cy][k], src, 64);
    float* slot = PWb + (size_t)kzl * 2 * T.px * 3;
    if ((lane % G) == 0) {
      // the element group's first node column ln (the previous group's last
      // column added in), then a paired lane's middle column
      const int ln = (lane / G) * E;
#pragma unroll
      for (int cx = 0; cx < E; ++cx) {
        if (ln + cx < nn) {
#pragma unroll
          for (int cy = 0; cy < 2; ++cy)
#pragma unroll
            for (int k = 0; k < 3; ++k)
              slot[((size_t)cy * T.px + ln + cx) * 3 + k] =
                  a[cx][cy][k] + (cx == 0 && ln > 0 ? prev[cy][k] : 0.0f);
        }
      }
    }
    if (lane == 63) {
      const int ln = (64 / G) * E;
      if (ln < nn) {
#pragma unroll
        for (int cy = 0; cy < 2; ++cy)
#pragma unroll
          for (int k = 0; k < 3; ++k) slot[((size_t)cy * T.px + ln) * 3 + k] = a[E][cy][k];
      }
    }
  };

  // Row pipeline.  The wave's rows are streamed in the order they are used
  // (z plane by z plane, rows within a plane); the raw gradient / mask words
  // and the last-plane scalars of the next DEPTH rows are requested before a
  // row is processed.  Every lane issues every load (inactive or non-tail
  // lanes read valid in-tile addresses), so no branch separates a load from
  // its use and the wait counts stay precise.
  // Three buffer sets, used round robin: a row is processed straight out of
  // its set, which is then re-filled with the row three ahead, so the loads
  // of the next two rows are in flight while a row is processed and no
  // register copies are needed.
  struct RowBuf {
    Raw<GT, V> g[3];
    Raw<uint8_t, V> m;
    uint32_t t[4];
    // residual pass: reference image, sampled values, live mask
    Raw<FT, V> f;
    Raw<float, V> gv;
    Raw<uint8_t, V> lv;
  };
  const int NP = (ez1 - ez0) * h + (ez1 == M.Nz ? 1 : 0);  // planes of this tile
  // pipeline depth: 2-byte gradients are unpacked into new registers before
  // the reload, so two buffer sets keep two rows in flight; 4-byte values
  // are used in place and need a third set
#ifndef VDVC_PIPE_2B
#define VDVC_PIPE_2B 2
#endif
  constexpr int PIPE = sizeof(GT) == 2 ? VDVC_PIPE_2B : 3;
  constexpr bool UNPACK_FIRST = sizeof(GT) == 2;
  RowBuf buf0, buf1, buf2;
  int cp = 0, ct = 0;  // prefetch cursor (plane, row), clamped to the tile's last row
  auto issue = [&](RowBuf& b) {
    const long long rb = (long long)(ez0 * h + cp) * pl + (long long)(y0 + ct) * M.pitch;
    ld_raw<GT, V>(g0 + rb, xbu, b.g[0]);
    ld_raw<GT, V>(g1 + rb, xbu, b.g[1]);
    ld_raw<GT, V>(g2 + rb, xbu, b.g[2]);
    if (HASMASK) ld_raw<uint8_t, V>(in.mask + rb, xbu, b.m);
    if (MODE == MODE_RHS) {
      ld_raw<FT, V>(fim + rb, xbu, b.f);
      ld_raw<float, V>(gvals_f + rb, xbu, b.gv);
      // the live mask feeds only the display statistics
      if (stats) ld_raw<uint8_t, V>(in.live + rb, xbu, b.lv);
    }
    if (MODE != MODE_RHS) {
      b.t[0] = ld_raw1<GT>(g0, rb + xend);
      b.t[1] = ld_raw1<GT>(g1, rb + xend);
      b.t[2] = ld_raw1<GT>(g2, rb + xend);
      b.t[3] = HASMASK ? ld_raw1<uint8_t>(in.mask, rb + xend) : 1u;
    }
    if (++ct == nrows) {
      ct = 0;
      if (++cp == NP) { cp = NP - 1; ct = nrows - 1; }
    }
  };
  issue(buf0);
  issue(buf1);
  if constexpr (PIPE == 3) issue(buf2);

  // plane state (updated when the flattened row walk enters a new z plane)
  int ez = ez0, tz = 0, ty = 0;
  int nzp = h + (ez == M.Nz - 1 ? 1 : 0);
  float Z0 = 1.0f, Z1 = 0.0f, Wz = 1.0f, zw0 = 1.0f, zw1 = 0.0f, zd0 = 1.0f, zd1 = 0.0f;
  int mz = 1;
  long long rowbase = 0, tailbase = 0;
  // u of the plane at the lane's node columns (row 0 of the element row)
  // and its y slope
  float u0[E + 1][3], du[E + 1][3];
#pragma unroll
  for (int cx = 0; cx < E + 1; ++cx)
#pragma unroll
    for (int k = 0; k < 3; ++k) { u0[cx][k] = 0.0f; du[cx][k] = 0.0f; }
  // the nodal values of the lane's element(s) (corner (cx,cy,cz), component
  // k; cx over the E + 1 node columns) live in LDS, slot
  // q = ((cx*2 + cy)*2 + cz)*3 + k at [q][lane], loaded once per element
  // layer; plane set-up then reads LDS only, whose wait counter is separate
  // from the row pipeline's
  float* nlds = lds_nodes + wib * 64 * 12 * (E + 1);
  // the element-stiffness term is added once per element: by the first lane
  // of its lane group, in the first row part, by an active lane
  // (single-element lanes; the host plans it for those only)
  constexpr bool KREG = MODE == MODE_MATVEC && E == 1;
  const bool kadd = KREG && in.kreg != nullptr && act && (lane % G) == 0 && ys == 0;
  auto element_setup = [&]() __attribute__((always_inline)) {
    if (MODE != MODE_MATVEC) return;
    float nvr[E + 1][2][2][3];
#pragma unroll
    for (int cx = 0; cx < E + 1; ++cx)
#pragma unroll
      for (int cy = 0; cy < 2; ++cy)
#pragma unroll
        for (int cz = 0; cz < 2; ++cz) {
          const long long n = node_index(M, ex + cx, ey + cy, ez + cz);
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            nvr[cx][cy][cz][k] = v[3 * n + k];
            nlds[((((cx * 2 + cy) * 2 + cz) * 3) + k) * 64 + lane] = nvr[cx][cy][cz][k];
          }
        }
    if constexpr (KREG) {
      if (kadd) {
        // corner a = (dx, dy, dz) of the element (reference Hex8 order)
        // receives kc sum_b kreg[a][b] v_b, its share of c E K v
#pragma unroll
        for (int a = 0; a < 8; ++a) {
          const int ax = (a == 1 || a == 2 || a == 5 || a == 6) ? 1 : 0;
          const int ay = (a == 2 || a == 3 || a == 6 || a == 7) ? 1 : 0;
          const int az = a >= 4 ? 1 : 0;
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            float acc = 0.0f;
#pragma unroll
            for (int b = 0; b < 8; ++b) {
              const int bx = (b == 1 || b == 2 || b == 5 || b == 6) ? 1 : 0;
              const int by = (b == 2 || b == 3 || b == 6 || b == 7) ? 1 : 0;
              const int bz = b >= 4 ? 1 : 0;
              acc += in.kreg[a * 8 + b] * nvr[bx][by][bz][k];
            }
            C[ax][ay][az][k] += in.kc * acc;
          }
        }
      }
    }
  };
  auto plane_setup = [&]() {
    const int z = ez * h + tz;
    Z1 = tz * inv_h;
    Z0 = 1.0f - Z1;
    Wz = axis_weight(z, M.Nvz - 1) * actw;
    mz = axis_mult(z, h, M.Nvz - 1);
    zw0 = Z0; zw1 = Z1;
    if (MODE == MODE_DIAG) { zw0 = Z0 * Z0; zw1 = Z1 * Z1; }
    if (DG) { zd0 = Z0 * Z0; zd1 = Z1 * Z1; }
    rowbase = (long long)z * pl + (long long)y0 * M.pitch;
    tailbase = (long long)z * pl + (long long)y0 * M.pitch + xend;
    if (MODE == MODE_MATVEC) {
      // u at the element's node columns (left / right; a paired lane's
      // left / middle / right), linear in y
      auto nv = [&](int cx, int cy, int cz, int k) {
        return nlds[((((cx * 2 + cy) * 2 + cz) * 3) + k) * 64 + lane];
      };
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int cx = 0; cx < E + 1; ++cx) {
          const float c0 = Z0 * nv(cx, 0, 0, k) + Z1 * nv(cx, 0, 1, k);
          const float c1 = Z0 * nv(cx, 1, 0, k) + Z1 * nv(cx, 1, 1, k);
          u0[cx][k] = c0; du[cx][k] = (c1 - c0) * inv_h;
        }
    }
  };
  // advance to the next row; closes element layers (emit + shift) and sets
  // up new planes
  auto advance = [&]() {
    if (++ty < nrows) return;
    ty = 0;
    if (++tz == nzp) {
      emit(C, Pw, ez - ez0);
#pragma unroll
      for (int cx = 0; cx < E + 1; ++cx)
#pragma unroll
        for (int cy = 0; cy < 2; ++cy)
#pragma unroll
          for (int k = 0; k < 3; ++k) { C[cx][cy][0][k] = C[cx][cy][1][k]; C[cx][cy][1][k] = 0.0f; }
      if constexpr (DG) {
        emit(C2, Pd, ez - ez0);
#pragma unroll
        for (int cx = 0; cx < E + 1; ++cx)
#pragma unroll
          for (int cy = 0; cy < 2; ++cy)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              C2[cx][cy][0][k] = C2[cx][cy][1][k];
              C2[cx][cy][1][k] = 0.0f;
            }
      }
      ++ez;
      tz = 0;
      nzp = h + (ez == M.Nz - 1 ? 1 : 0);
      if (ez >= ez1) return;
      element_setup();
    }
    plane_setup();
  };
  // a row's values unpacked from its raw words (for 2-byte types this
  // frees the raw registers, which are then re-filled at once)
  struct RowVals {
    float gx[V], gy[V], gz[V];
    uint32_t mk[V];
    uint32_t t[4];
    float fv[V], gv[V];
    uint32_t lv[V];
  };
  auto unpack_row = [&](const RowBuf& b, RowVals& o) {
    unpack<GT, V>(b.g[0], o.gx);
    unpack<GT, V>(b.g[1], o.gy);
    unpack<GT, V>(b.g[2], o.gz);
#pragma unroll
    for (int j2 = 0; j2 < V; ++j2) o.mk[j2] = HASMASK ? mask_bit<V>(b.m, j2) : 1u;
#pragma unroll
    for (int q = 0; q < 4; ++q) o.t[q] = b.t[q];
    if (MODE == MODE_RHS) {
      unpack<FT, V>(b.f, o.fv);
      unpack<float, V>(b.gv, o.gv);
#pragma unroll
      for (int j2 = 0; j2 < V; ++j2) o.lv[j2] = stats ? mask_bit<V>(b.lv, j2) : 0u;
    }
  };
  auto process = [&](const RowVals& cur) {
    const long long vi = rowbase + (long long)ty * M.pitch;
    const uint32_t* ctl = cur.t;
    const int y = y0 + ty;
    const float fty = (float)(ty0 + ty);  // row within the element
    const float Wr = Wz * axis_weight(y, M.Nvy - 1);
    const float* gx = cur.gx;
    const float* gy = cur.gy;
    const float* gz = cur.gz;
    const uint32_t* mk = cur.mk;
    // x-restriction sums of the lane's element(s)
    float S[E][3], Tt[E][3];
#pragma unroll
    for (int e = 0; e < E; ++e)
#pragma unroll
      for (int k = 0; k < 3; ++k) { S[e][k] = 0.0f; Tt[e][k] = 0.0f; }
    float S2[ND][3], T2[ND][3];
#pragma unroll
    for (int e = 0; e < ND; ++e)
#pragma unroll
      for (int k = 0; k < 3; ++k) { S2[e][k] = 0.0f; T2[e][k] = 0.0f; }
    // u of this row at the node columns, and its step across each element
    float uc[E + 1][3], dxe[E][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
#pragma unroll
      for (int cx = 0; cx < E + 1; ++cx)
        uc[cx][k] = MODE == MODE_MATVEC ? u0[cx][k] + fty * du[cx][k] : 0.0f;
#pragma unroll
      for (int e = 0; e < E; ++e) dxe[e][k] = uc[e + 1][k] - uc[e][k];
    }
    float resv[V];
    if (MODE == MODE_RHS) {
      const float* fv = cur.fv;
      const float* gv = cur.gv;
      const uint32_t* lv = cur.lv;
      const int mxy = axis_mult(y, h, M.Nvy - 1) * mz;
#pragma unroll
      for (int j = 0; j < V; ++j) {
        // (g - g_mean) * (-(std0/std1)) + (f - mean0), float32 steps in
        // the reference's order, no contraction; zero outside b_mask
        float rv = add_rn(mul_rn(sub_rn(gv[j], gmean_f), kscale_f),
                             sub_rn(fv[j], mean0_f));
        rv = mk[j] ? rv : 0.0f;
        resv[j] = rv;
        if (act && lv[j]) {
          const int mxv = (tb + j == 0 && xb + j != 0) ? 2 : 1;
          const double mlt = (double)(mxv * mxy);
          st_c += mlt;
          st_1 += mlt * (double)rv;
          st_2 += mlt * (double)rv * (double)rv;
          st_n += 1.0;
        }
      }
      if (act) {
        // the residual is stored only when the caller keeps it
        if (res_f) st_f32<V>(res_f + vi, xbu, resv);
        if (FOLD) {
          float fx[V], fy[V], fz[V];
#pragma unroll
          for (int j = 0; j < V; ++j) {
            fx[j] = mk[j] ? gx[j] : 0.0f;
            fy[j] = mk[j] ? gy[j] : 0.0f;
            fz[j] = mk[j] ? gz[j] : 0.0f;
          }
          st_bf16<V>(in.fold_out[0] + vi, xbu, fx);
          st_bf16<V>(in.fold_out[1] + vi, xbu, fy);
          st_bf16<V>(in.fold_out[2] + vi, xbu, fz);
        }
      }
    }
#pragma unroll
    for (int j = 0; j < V; ++j) {
      // when the host launched with h == V every block starts an element
      // (tb == 0) and X1 = j / V is a constant; a paired lane's voxel j lies
      // in its element j / HE at X1 = (j % HE) / HE
      const int e = E > 1 ? j / HE : 0;
      const float X1 = E > 1 ? (float)(j % HE) * (1.0f / (float)HE)
                             : (h == V) ? (float)j * (1.0f / (float)V) : X1b + (float)j * inv_h;
      float wm = (j == 0) ? wx0 : 1.0f;
      if (HASMASK) wm = mk[j] ? wm : 0.0f;
      float sd = 0.0f;
      if (MODE == MODE_MATVEC) {
        const float ux = uc[e][0] + X1 * dxe[e][0], uy = uc[e][1] + X1 * dxe[e][1],
                    uz = uc[e][2] + X1 * dxe[e][2];
        sd = gx[j] * ux + gy[j] * uy + gz[j] * uz;
      } else if (MODE == MODE_RHS) {
        sd = resv[j];
      }
      acc_voxel<MODE>(gx[j], gy[j], gz[j], wm, X1, sd, S[e], Tt[e]);
      if constexpr (DG) acc_voxel<MODE_DIAG>(gx[j], gy[j], gz[j], wm, X1, 0.0f, S2[e], T2[e]);
    }
    if (tail) {
      // last voxel plane x = Nx*h: t = h of this element, X1 = 1, W_x = 1/2
      const long long ti = tailbase + (long long)ty * M.pitch;
      float tgx, tgy, tgz;
      uint32_t tm = 1;
      if (MODE == MODE_RHS) {
        tgx = ld_val(g0, ti); tgy = ld_val(g1, ti); tgz = ld_val(g2, ti);
        if (HASMASK) tm = in.mask[ti];
      } else {
        tgx = unpack1<GT>(ctl[0]); tgy = unpack1<GT>(ctl[1]); tgz = unpack1<GT>(ctl[2]);
        tm = ctl[3];
      }
      float wm = tm ? 0.5f : 0.0f;
      float sd = 0.0f;
      if (MODE == MODE_MATVEC) {
        sd = tgx * (uc[E - 1][0] + dxe[E - 1][0]) + tgy * (uc[E - 1][1] + dxe[E - 1][1]) +
             tgz * (uc[E - 1][2] + dxe[E - 1][2]);
      } else if (MODE == MODE_RHS) {
        const float fv = ld_val(fim, ti);
        const float gv = gvals_f[ti];
        float rv = add_rn(mul_rn(sub_rn(gv, gmean_f), kscale_f),
                             sub_rn(fv, mean0_f));
        rv = tm ? rv : 0.0f;
        if (res_f) res_f[ti] = rv;
        if (stats && in.live[ti]) {
          const double mlt = (double)(axis_mult(y, h, M.Nvy - 1) * mz);
          st_c += mlt;
          st_1 += mlt * (double)rv;
          st_2 += mlt * (double)rv * (double)rv;
          st_n += 1.0;
        }
        if (FOLD) {
          in.fold_out[0][ti] = (uint16_t)bf16_bits(tm ? tgx : 0.0f);
          in.fold_out[1][ti] = (uint16_t)bf16_bits(tm ? tgy : 0.0f);
          in.fold_out[2][ti] = (uint16_t)bf16_bits(tm ? tgz : 0.0f);
        }
        sd = rv;
      }
      acc_voxel<MODE>(tgx, tgy, tgz, wm, 1.0f, sd, S[E - 1], Tt[E - 1]);
      if constexpr (DG) acc_voxel<MODE_DIAG>(tgx, tgy, tgz, wm, 1.0f, 0.0f, S2[E - 1], T2[E - 1]);
    }
    // x-restricted row values R[cx][k] (the lanes past the row end have
    // zero node weights), weighted into the node planes
    // (a paired lane's middle column takes the right part of element 0 and
    // the left part of element 1)
    float R[E + 1][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      R[0][k] = MODE == MODE_DIAG ? S[0][k] : S[0][k] - Tt[0][k];
#pragma unroll
      for (int c = 1; c < E; ++c)
        R[c][k] = Tt[c - 1][k] + (MODE == MODE_DIAG ? S[c][k] : S[c][k] - Tt[c][k]);
      R[E][k] = Tt[E - 1][k];
    }
    const float Y1 = fty * inv_h;
    float wy0 = Wr * (1.0f - Y1), wy1 = Wr * Y1;
    if (MODE == MODE_DIAG) { wy0 = Wr * (1.0f - Y1) * (1.0f - Y1); wy1 = Wr * Y1 * Y1; }
    const float w00 = wy0 * zw0, w01 = wy0 * zw1, w10 = wy1 * zw0, w11 = wy1 * zw1;
#pragma unroll
    for (int cx = 0; cx < E + 1; ++cx)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        C[cx][0][0][k] += w00 * R[cx][k];
        C[cx][0][1][k] += w01 * R[cx][k];
        C[cx][1][0][k] += w10 * R[cx][k];
        C[cx][1][1][k] += w11 * R[cx][k];
      }
    if constexpr (DG) {
      // the diagonal mode's row restriction and squared weights
      float R2[E + 1][3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        R2[0][k] = S2[0][k];
#pragma unroll
        for (int c = 1; c < E; ++c) R2[c][k] = T2[c - 1][k] + S2[c][k];
        R2[E][k] = T2[E - 1][k];
      }
      const float vy0 = Wr * (1.0f - Y1) * (1.0f - Y1), vy1 = Wr * Y1 * Y1;
      const float d00 = vy0 * zd0, d01 = vy0 * zd1, d10 = vy1 * zd0, d11 = vy1 * zd1;
#pragma unroll
      for (int cx = 0; cx < E + 1; ++cx)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          C2[cx][0][0][k] += d00 * R2[cx][k];
          C2[cx][0][1][k] += d01 * R2[cx][k];
          C2[cx][1][0][k] += d10 * R2[cx][k];
          C2[cx][1][1][k] += d11 * R2[cx][k];
        }
    }
  };

  element_setup();
  plane_setup();
  const int nrows_tile = NP * nrows;
  if constexpr (PIPE == 2) {
    // two buffer sets: a row is unpacked (its raw registers are free after
    // that), its set re-filled with the row two ahead, then processed
    for (int rr = 0; rr < nrows_tile; rr += 2) {
      RowVals cv;
      unpack_row(buf0, cv);
      issue(buf0);
      process(cv);
      advance();
      if (rr + 1 < nrows_tile) {
        unpack_row(buf1, cv);
        issue(buf1);
        process(cv);
        advance();
      }
    }
  } else {
    // three buffer sets: a row is processed straight from its set, which is
    // then re-filled with the row three ahead (2-byte values: unpacked
    // first, so the set is re-filled before the row is processed)
    for (int rr = 0; rr < nrows_tile; rr += 3) {
      RowVals cv;
      unpack_row(buf0, cv);
      if (UNPACK_FIRST) issue(buf0);
      process(cv);
      if (!UNPACK_FIRST) issue(buf0);
      advance();
      if (rr + 1 < nrows_tile) {
        unpack_row(buf1, cv);
        if (UNPACK_FIRST) issue(buf1);
        process(cv);
        if (!UNPACK_FIRST) issue(buf1);
        advance();
      }
      if (rr + 2 < nrows_tile) {
        unpack_row(buf2, cv);
        if (UNPACK_FIRST) issue(buf2);
        process(cv);
        if (!UNPACK_FIRST) issue(buf2);
        advance();
      }
    }
  }
  emit(C, Pw, ez1 - ez0);
  if constexpr (DG) emit(C2, Pd, ez1 - ez0);
  if (stats) {
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

namespace vdvc {

// voxels per lane for the vectorised kernel: the largest power of two that
// divides h and fits 16 bytes of the gradient type; 0 when h is not a power
// of two <= 64 (the generic kernel is used)
int vec_width(int h, int gtype) {
  if (h < 1 || h > 64 || (h & (h - 1)) != 0) return 0;
  if (gtype == GT_F64) return 0;  // the FP64 path uses the generic kernel
  int vmax = gtype == GT_F32 ? 4 : 8;
  // tuning override of the widest lane width (a power of two, at most the
  // 16-byte width), read at every plan: VDVC_VEC_F32 / VDVC_VEC_BF16
  const char* e = getenv(gtype == GT_F32 ? "VDVC_VEC_F32" : "VDVC_VEC_BF16");
  const int ov = e ? atoi(e) : 0;
  if ((ov == 1 || ov == 2 || ov == 4 || ov == 8) && ov <= vmax) vmax = ov;
  return h < vmax ? h : vmax;
}

// the residual + diagonal pass exists for the float32 form (float32 image
// and gradients, no refold) and the mixed form (float16 image, bfloat16
// gradients, refold), on the vectorised single-element-lane plan
bool rhs_diag_fused(int gtype, int ftype, bool fold, const TilesDev& T) {
  return T.vec != 0 && T.epl == 1 &&
         ((gtype == GT_F32 && ftype == GT_F32 && !fold) ||
          (gtype == GT_BF16 && ftype == GT_F16 && fold));
}

int launch_restrict_vec(int mode, int gtype, bool hasmask, int ftype, bool fold, const MeshDev& M,
                        const TilesDev& T, const VoxIn& in, const float* v, float* P,
                        hipStream_t st) {
  const unsigned nblk = (unsigned)((T.nwaves + VDVC_WAVES_PER_BLOCK - 1) / VDVC_WAVES_PER_BLOCK);
  dim3 g(nblk), b(VDVC_BLOCK);
#define VV_L(MODE, GT, HM, FT, FO, VV, EE) \
  hipLaunchKernelGGL((vox_vec_kernel<MODE, GT, HM, FT, FO, VV, EE>), g, b, 0, st, M, T, in, v, P)
#define VV_V(MODE, GT, HM, FT, FO)                    \
  do {                                                \
    switch (T.vec) {                                  \
      case 8: VV_L(MODE, GT, HM, FT, FO, 8, 1); break; \
      case 4: VV_L(MODE, GT, HM, FT, FO, 4, 1); break; \
      case 2: VV_L(MODE, GT, HM, FT, FO, 2, 1); break; \
      default: VV_L(MODE, GT, HM, FT, FO, 1, 1); break; \
    }                                                 \
  } while (0)
#define VV_LD(MODE, GT, HM, FT, FO, VV)                                                    \
  hipLaunchKernelGGL((vox_vec_kernel<MODE, GT, HM, FT, FO, VV, 1, true>), g, b, 0, st, M, T, in, \
                     v, P)
#define VV_VD(MODE, GT, HM, FT, FO)                    \
  do {                                                 \
    switch (T.vec) {                                   \
      case 8: VV_LD(MODE, GT, HM, FT, FO, 8); break;   \
      case 4: VV_LD(MODE, GT, HM, FT, FO, 4); break;   \
      case 2: VV_LD(MODE, GT, HM, FT, FO, 2); break;   \
      default: VV_LD(MODE, GT, HM, FT, FO, 1); break;  \
    }                                                  \
  } while (0)
  if (gtype == GT_F32 && T.vec == 8) return (int)hipErrorInvalidValue;
  if (in.pdiag) {
    // the residual pass with the Hessian diagonal (rhs_diag_fused)
    if (mode != MODE_RHS || !rhs_diag_fused(gtype, ftype, fold, T)) return (int)hipErrorInvalidValue;
    if (gtype == GT_F32) VV_VD(MODE_RHS, float, true, float, false);
    else VV_VD(MODE_RHS, __hip_bfloat16, true, __half, true);
    return (int)hipGetLastError();
  }
#undef VV_VD
#undef VV_LD
  if (T.epl == 2) {
    // paired elements: bfloat16 lanes of 8 (h = 4) or 4 (h = 2) voxels,
    // float32 lanes of 4 (h = 2); products and diagonal only
    if (mode == MODE_RHS || M.h * 2 != T.vec || (T.vec != 8 && T.vec != 4) ||
        (gtype == GT_F32 && T.vec != 4))
      return (int)hipErrorInvalidValue;
#define VV_P(MODE, GT, HM)                                          \
    do {                                                            \
      if (T.vec == 8) VV_L(MODE, GT, HM, float, false, 8, 2);       \
      else VV_L(MODE, GT, HM, float, false, 4, 2);                  \
    } while (0)
#define VV_PM(MODE, GT) \
    do { if (hasmask) VV_P(MODE, GT, true); else VV_P(MODE, GT, false); } while (0)
    if (gtype == GT_F32) {
      if (mode == MODE_MATVEC) {
        if (hasmask) VV_L(MODE_MATVEC, float, true, float, false, 4, 2);
        else VV_L(MODE_MATVEC, float, false, float, false, 4, 2);
      } else {
        if (hasmask) VV_L(MODE_DIAG, float, true, float, false, 4, 2);
        else VV_L(MODE_DIAG, float, false, float, false, 4, 2);
      }
    } else if (mode == MODE_MATVEC) {
      VV_PM(MODE_MATVEC, __hip_bfloat16);
    } else {
      VV_PM(MODE_DIAG, __hip_bfloat16);
    }
#undef VV_PM
#undef VV_P
    return (int)hipGetLastError();
  }
  if (mode == MODE_MATVEC) {
    if (gtype == GT_F32) {
      if (hasmask) VV_V(MODE_MATVEC, float, true, float, false);
      else VV_V(MODE_MATVEC, float, false, float, false);
    } else {
      if (hasmask) VV_V(MODE_MATVEC, __hip_bfloat16, true, float, false);
      else VV_V(MODE_MATVEC, __hip_bfloat16, false, float, false);
    }
  } else if (mode == MODE_DIAG) {
    if (gtype == GT_F32) {
      if (hasmask) VV_V(MODE_DIAG, float, true, float, false);
      else VV_V(MODE_DIAG, float, false, float, false);
    } else {
      if (hasmask) VV_V(MODE_DIAG, __hip_bfloat16, true, float, false);
      else VV_V(MODE_DIAG, __hip_bfloat16, false, float, false);
    }
  } else {
    if (gtype == GT_F32) {
      if (ftype == GT_F32) {
        if (fold) VV_V(MODE_RHS, float, true, float, true);
        else VV_V(MODE_RHS, float, true, float, false);
      } else {
        if (fold) VV_V(MODE_RHS, float, true, __half, true);
        else VV_V(MODE_RHS, float, true, __half, false);
      }
    } else {
      if (ftype == GT_F32) {
        if (fold) VV_V(MODE_RHS, __hip_bfloat16, true, float, true);
        else VV_V(MODE_RHS, __hip_bfloat16, true, float, false);
      } else {
        if (fold) VV_V(MODE_RHS, __hip_bfloat16, true, __half, true);
        else VV_V(MODE_RHS, __hip_bfloat16, true, __half, false);
      }
    }
  }
#undef VV_V
#undef VV_L
  return (int)hipGetLastError();
}

}  // namespace vdvc
