// Learned-seed network kernels: 3x3x3 convolutions (pad 1, stride 1 or 2)
// as implicit GEMMs on the CDNA matrix cores (v_mfma_f32_16x16x16_bf16:
// bfloat16 operands, float32 accumulation), their backward products, the
// optimizer step, the loss and the network-input builder.
//
// Tensors are float32 in memory, channels last: activations
// [n][z][y][x][c], weights [o][tap][c] with tap = (dz*3 + dy)*3 + dx.
// Operands are rounded to bfloat16 (nearest even) when they are loaded into
// the matrix core; the float32 master values are what the optimizer updates.
//
// MFMA 16x16x16 operand layout (wave64), lane l = 16*q + r (q = l / 16):
//   A (16 x 16, rows m, depth k): lane holds A[r][4q + j], j = 0..3
//   B (16 x 16, depth k, cols n): lane holds B[4q + j][r]
//   D (16 x 16):                 lane holds D[4q + j][r]
// Each GEMM below forms its A/B fragments straight from global memory
// (the whole network's tensors are a few MB and stay in L2).
#include "vdvc_kernels.h"

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

namespace {

// float -> bfloat16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ short bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

struct ConvShape {
  int N, D, H, W, C;      // input
  int Do, Ho, Wo, O;      // output
  int stride;
};

__device__ __forceinline__ f32x4_t mfma16(bf16x4_t a, bf16x4_t b, f32x4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// forward: y[p][o] = b[o] + sum_{tap,c} w[o][tap][c] x[p*s + d - 1][c]
// GEMM rows = output positions, cols = o, depth = tap*C + c.  One wave
// computes 16 positions x all output channels (NT <= 2 column tiles).
// ---------------------------------------------------------------------------
template <int NT>
__global__ __launch_bounds__(256) void conv_fwd_kernel(ConvShape S, const float* __restrict__ x,
                                                      const float* __restrict__ w,
                                                      const float* __restrict__ bias, int relu,
                                                      float* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const long long P = (long long)S.N * S.Do * S.Ho * S.Wo;
  const long long p0 = (long long)wave * 16;
  if (p0 >= P) return;
  const int r = lane & 15, q = lane >> 4;
  // A-row position of this lane
  const long long pa = p0 + r;
  const bool va = pa < P;
  int n = 0, z = 0, yy = 0, xx = 0;
  if (va) {
    long long t = pa;
    xx = (int)(t % S.Wo); t /= S.Wo;
    yy = (int)(t % S.Ho); t /= S.Ho;
    z = (int)(t % S.Do); n = (int)(t / S.Do);
  }
  const int K = 27 * S.C;
  f32x4_t acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4_t a, bv[NT];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = k0 + 4 * q + j;
      float av = 0.0f;
      if (va && k < K) {
        const int tap = k / S.C, c = k - tap * S.C;
        const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
        const int iz = z * S.stride + dz - 1, iy = yy * S.stride + dy - 1,
                  ix = xx * S.stride + dx - 1;
        if (iz >= 0 && iz < S.D && iy >= 0 && iy < S.H && ix >= 0 && ix < S.W)
          av = x[((((long long)n * S.D + iz) * S.H + iy) * S.W + ix) * S.C + c];
      }
      a[j] = bf16_rne(av);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int o = t * 16 + r;
        const float wv = (k < K && o < S.O) ? w[(long long)o * K + k] : 0.0f;
        bv[t][j] = bf16_rne(wv);
      }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = mfma16(a, bv[t], acc[t]);
  }
  // D[4q + j][r]: position p0 + 4q + j, channel t*16 + r
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int o = t * 16 + r;
    if (o >= S.O) continue;
    const float bo = bias ? bias[o] : 0.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long long p = p0 + 4 * q + j;
      if (p >= P) continue;
      float v = acc[t][j] + bo;
      if (relu) v = v > 0.0f ? v : 0.0f;
      y[p * S.O + o] = v;
    }
  }
}

// Channel-specialised forward kernels (same GEMM, depth order tap*C + c):
// with C a multiple of 16 a 16-deep chunk is 16 channels of one tap, so a
// lane's 4 operand values are one aligned 16-byte load (conv_fwd_lds_kernel);
// with C = 2 a lane's 4 values are two 8-byte loads of two taps.  Both
// replace the generic kernel's per-element index division.
__device__ __forceinline__ bf16x4_t to_bf16x4(float4 v) {
  bf16x4_t r;
  r[0] = bf16_rne(v.x); r[1] = bf16_rne(v.y); r[2] = bf16_rne(v.z); r[3] = bf16_rne(v.w);
  return r;
}

// Forward for C a multiple of 16 with the layer's weights staged once per
// block in LDS as bfloat16 B fragments: fragment (tap, cc, t) is 64 lanes x 4
// values, 512 contiguous bytes (one conflict-free 8-byte read per lane).
// kFwdWaves wavefronts per block, each computing kFwdTiles tiles of 16
// positions x all NT*16 output channels, so an A fragment feeds NT products
// and a B fragment kFwdTiles.  Per output the products accumulate over the
// taps in order, then the 16-channel chunks, as in the other forward kernels.
constexpr int kFwdWaves = 8;
constexpr int kFwdTiles = 2;

template <int NT, int CC>
__global__ __launch_bounds__(64 * kFwdWaves) void conv_fwd_lds_kernel(
    ConvShape S, const float* __restrict__ x, const float* __restrict__ w,
    const float* __restrict__ bias, int relu, float* __restrict__ y) {
  constexpr int C = 16 * CC;
  constexpr int NF = 27 * CC * NT;
  __shared__ bf16x4_t wl[NF * 64];
  const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int e = threadIdx.x; e < NF * 64; e += 64 * kFwdWaves) {
    const int f = e >> 6, l = e & 63;
    const int t = f % NT, cc = (f / NT) % CC, tap = f / (NT * CC);
    const int o = t * 16 + (l & 15), c0 = cc * 16 + 4 * (l >> 4);
    float4 v = zero4;
    if (o < S.O) v = *reinterpret_cast<const float4*>(w + ((long long)o * 27 + tap) * C + c0);
    wl[e] = to_bf16x4(v);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r = lane & 15, q = lane >> 4;
  const long long P = (long long)S.N * S.Do * S.Ho * S.Wo;
  const long long pbase = ((long long)blockIdx.x * kFwdWaves + wv) * (16 * kFwdTiles);
  if (pbase >= P) return;
  int bz[kFwdTiles], by[kFwdTiles], bx[kFwdTiles], nn[kFwdTiles];
  bool va[kFwdTiles];
#pragma unroll
  for (int u = 0; u < kFwdTiles; ++u) {
    const long long pa = pbase + 16 * u + r;
    va[u] = pa < P;
    long long t = va[u] ? pa : 0;
    const int xx = (int)(t % S.Wo); t /= S.Wo;
    const int yy = (int)(t % S.Ho); t /= S.Ho;
    const int zz = (int)(t % S.Do);
    nn[u] = (int)(t / S.Do);
    bz[u] = zz * S.stride - 1; by[u] = yy * S.stride - 1; bx[u] = xx * S.stride - 1;
  }
  f32x4_t acc[kFwdTiles][NT];
#pragma unroll
  for (int u = 0; u < kFwdTiles; ++u)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[u][t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll 3
  for (int tap = 0; tap < 27; ++tap) {
    const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
    bf16x4_t a[kFwdTiles][CC];
#pragma unroll
    for (int u = 0; u < kFwdTiles; ++u) {
      const int iz = bz[u] + dz, iy = by[u] + dy, ix = bx[u] + dx;
      const bool v = va[u] && iz >= 0 && iz < S.D && iy >= 0 && iy < S.H && ix >= 0 && ix < S.W;
      const float* xp = x +
          ((((long long)nn[u] * S.D + (v ? iz : 0)) * S.H + (v ? iy : 0)) * S.W + (v ? ix : 0)) * C +
          4 * q;
#pragma unroll
      for (int cc = 0; cc < CC; ++cc) {
        float4 av = *reinterpret_cast<const float4*>(xp + cc * 16);
        if (!v) av = zero4;
        a[u][cc] = to_bf16x4(av);
      }
    }
#pragma unroll
    for (int cc = 0; cc < CC; ++cc)
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const bf16x4_t b = wl[((tap * CC + cc) * NT + t) * 64 + lane];
#pragma unroll
        for (int u = 0; u < kFwdTiles; ++u) acc[u][t] = mfma16(a[u][cc], b, acc[u][t]);
      }
  }
#pragma unroll
  for (int u = 0; u < kFwdTiles; ++u)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int o = t * 16 + r;
      if (o >= S.O) continue;
      const float bo = bias ? bias[o] : 0.0f;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const long long p = pbase + 16 * u + 4 * q + j;
        if (p >= P) continue;
        float v = acc[u][t][j] + bo;
        if (relu) v = v > 0.0f ? v : 0.0f;
        y[p * S.O + o] = v;
      }
    }
}

// Forward with the input neighbourhood of each output box staged in LDS as
// bfloat16.  An output box is kTileZ x kTileY x 16 (z, y, x) positions --
// kFwdWaves waves x kFwdTiles rows of 16 x-consecutive positions -- by all
// output channels.  Its input box, ((kTileZ-1)s+3) x ((kTileY-1)s+3) x
// (15s+3) positions x C channels, is read from global memory once (16-byte
// loads along the rows), rounded to bfloat16 and stored with a position
// stride of C + 4 values, so the 16 rows of an A fragment spread over the
// banks; each of the 27 taps then reads its A fragments from LDS instead of
// re-gathering and re-rounding them from L2.  Blocks are persistent (one per
// CU): a block stages the layer's weights once, as conv_fwd_lds_kernel does,
// then walks the output boxes with a grid stride, with two input-box buffers:
// the next box's loads are issued into registers before the current box is
// computed and written to the other buffer after it.  Per output the
// products accumulate over the taps in order, then the 16-channel chunks, as
// in the other forward kernels.
constexpr int kTileZ = 4, kTileY = 4;
static_assert(kTileZ * kTileY == kFwdWaves * kFwdTiles, "a box's rows are its waves' tiles");

template <int NT, int CC, int STR>
__global__ __launch_bounds__(64 * kFwdWaves) void conv_fwd_tile_kernel(
    ConvShape S, const float* __restrict__ x, const float* __restrict__ w,
    const float* __restrict__ bias, int relu, float* __restrict__ y) {
  constexpr int C = 16 * CC;
  constexpr int NF = 27 * CC * NT;
  constexpr int HZ = (kTileZ - 1) * STR + 3, HY = (kTileY - 1) * STR + 3, HX = 15 * STR + 3;
  constexpr int CS4 = (C + 4) / 4;  // position stride in 4-value groups
  constexpr int NPOS = HZ * HY * HX;
  constexpr int NQ = NPOS * (C / 4);  // 16-byte loads per input box
  constexpr int NT_ = 64 * kFwdWaves;
  constexpr int NLD = (NQ + NT_ - 1) / NT_;  // per thread
  __shared__ bf16x4_t wl[NF * 64];
  __shared__ bf16x4_t xl[2][NPOS * CS4];
  const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int e = threadIdx.x; e < NF * 64; e += NT_) {
    const int f = e >> 6, l = e & 63;
    const int t = f % NT, cc = (f / NT) % CC, tap = f / (NT * CC);
    const int o = t * 16 + (l & 15), c0 = cc * 16 + 4 * (l >> 4);
    float4 v = zero4;
    if (o < S.O) v = *reinterpret_cast<const float4*>(w + ((long long)o * 27 + tap) * C + c0);
    wl[e] = to_bf16x4(v);
  }
  // output boxes, x fastest, then y, z, n
  const int nbx = (S.Wo + 15) / 16, nby = (S.Ho + kTileY - 1) / kTileY,
            nbz = (S.Do + kTileZ - 1) / kTileZ;
  const int nboxes = S.N * nbz * nby * nbx;
  float4 pre[NLD];
  auto load_box = [&](int box) {
    const int bx = box % nbx, by = (box / nbx) % nby, bz = (box / (nbx * nby)) % nbz;
    const int n = box / (nbx * nby * nbz);
    const int iz0 = bz * kTileZ * STR - 1, iy0 = by * kTileY * STR - 1, ix0 = bx * 16 * STR - 1;
#pragma unroll
    for (int i = 0; i < NLD; ++i) {
      const int e = threadIdx.x + i * NT_;
      const int pos = e / (C / 4), c4 = e - pos * (C / 4);
      const int hx = pos % HX, hy = (pos / HX) % HY, hz = pos / (HX * HY);
      const int iz = iz0 + hz, iy = iy0 + hy, ix = ix0 + hx;
      pre[i] = zero4;
      if (e < NQ && iz >= 0 && iz < S.D && iy >= 0 && iy < S.H && ix >= 0 && ix < S.W)
        pre[i] = *reinterpret_cast<const float4*>(
            x + ((((long long)n * S.D + iz) * S.H + iy) * S.W + ix) * C + 4 * c4);
    }
  };
  auto store_box = [&](bf16x4_t* buf) {
#pragma unroll
    for (int i = 0; i < NLD; ++i) {
      const int e = threadIdx.x + i * NT_;
      const int pos = e / (C / 4), c4 = e - pos * (C / 4);
      if (e < NQ) buf[pos * CS4 + c4] = to_bf16x4(pre[i]);
    }
  };
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r = lane & 15, q = lane >> 4;
  int tz[kFwdTiles], ty[kFwdTiles], base[kFwdTiles];
#pragma unroll
  for (int u = 0; u < kFwdTiles; ++u) {
    const int rr = wv * kFwdTiles + u;
    tz[u] = rr / kTileY;
    ty[u] = rr % kTileY;
    // input-box position of this lane's row at tap (0, 0, 0)
    base[u] = ((tz[u] * STR) * HY + ty[u] * STR) * HX + r * STR;
  }
  int box = blockIdx.x;
  if (box < nboxes) {
    load_box(box);
    store_box(xl[0]);
  }
  __syncthreads();
  for (int k = 0; box < nboxes; box += gridDim.x, ++k) {
    const int nxt = box + gridDim.x;
    if (nxt < nboxes) load_box(nxt);  // in flight while this box is computed
    const bf16x4_t* cur = xl[k & 1];
    f32x4_t acc[kFwdTiles][NT];
#pragma unroll
    for (int u = 0; u < kFwdTiles; ++u)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[u][t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll 3
    for (int tap = 0; tap < 27; ++tap) {
      const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
      const int toff = (dz * HY + dy) * HX + dx;
      bf16x4_t a[kFwdTiles][CC];
#pragma unroll
      for (int u = 0; u < kFwdTiles; ++u)
#pragma unroll
        for (int cc = 0; cc < CC; ++cc) a[u][cc] = cur[(base[u] + toff) * CS4 + cc * 4 + q];
#pragma unroll
      for (int cc = 0; cc < CC; ++cc)
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const bf16x4_t bw = wl[((tap * CC + cc) * NT + t) * 64 + lane];
#pragma unroll
          for (int u = 0; u < kFwdTiles; ++u) acc[u][t] = mfma16(a[u][cc], bw, acc[u][t]);
        }
    }
    // D[4q + j][r]: output x = bx*16 + 4q + j of row u, channel t*16 + r
    const int bx = box % nbx, by = (box / nbx) % nby, bz = (box / (nbx * nby)) % nbz;
    const int n = box / (nbx * nby * nbz);
#pragma unroll
    for (int u = 0; u < kFwdTiles; ++u) {
      const int oz = bz * kTileZ + tz[u], oy = by * kTileY + ty[u];
      if (oz >= S.Do || oy >= S.Ho) continue;
      const long long prow = (((long long)n * S.Do + oz) * S.Ho + oy) * S.Wo;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int o = t * 16 + r;
        if (o >= S.O) continue;
        const float bo = bias ? bias[o] : 0.0f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int ox = bx * 16 + 4 * q + j;
          if (ox >= S.Wo) continue;
          float v = acc[u][t][j] + bo;
          if (relu) v = v > 0.0f ? v : 0.0f;
          y[(prow + ox) * S.O + o] = v;
        }
      }
    }
    // the other buffer was last read before the previous barrier
    if (nxt < nboxes) store_box(xl[(k + 1) & 1]);
    __syncthreads();
  }
}

template <int NT>
__global__ __launch_bounds__(256) void conv_fwd_c2_kernel(ConvShape S, const float* __restrict__ x,
                                                         const float* __restrict__ w,
                                                         const float* __restrict__ bias, int relu,
                                                         float* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const long long P = (long long)S.N * S.Do * S.Ho * S.Wo;
  const long long p0 = (long long)wave * 16;
  if (p0 >= P) return;
  const int r = lane & 15, q = lane >> 4;
  const long long pa = p0 + r;
  const bool va = pa < P;
  int n = 0, z = 0, yy = 0, xx = 0;
  if (va) {
    long long t = pa;
    xx = (int)(t % S.Wo); t /= S.Wo;
    yy = (int)(t % S.Ho); t /= S.Ho;
    z = (int)(t % S.Do); n = (int)(t / S.Do);
  }
  const int bz = z * S.stride - 1, by = yy * S.stride - 1, bx = xx * S.stride - 1;
  f32x4_t acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  const float2 zero2 = make_float2(0.f, 0.f);
  // depth k = 2*tap + c: chunk k0 holds taps k0/2 .. k0/2 + 7; lane q holds
  // taps k0/2 + 2q and k0/2 + 2q + 1 (both channels each)
#pragma unroll
  for (int k0 = 0; k0 < 54; k0 += 16) {
    float2 av[2], wv[NT][2];
#pragma unroll
    for (int h2 = 0; h2 < 2; ++h2) {
      const int tap = k0 / 2 + 2 * q + h2;
      av[h2] = zero2;
      if (tap < 27) {
        const int iz = bz + tap / 9, iy = by + (tap / 3) % 3, ix = bx + tap % 3;
        if (va && iz >= 0 && iz < S.D && iy >= 0 && iy < S.H && ix >= 0 && ix < S.W)
          av[h2] = *reinterpret_cast<const float2*>(
              x + ((((long long)n * S.D + iz) * S.H + iy) * S.W + ix) * 2);
      }
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int o = t * 16 + r;
        wv[t][h2] = (tap < 27 && o < S.O)
                        ? *reinterpret_cast<const float2*>(w + ((long long)o * 27 + tap) * 2)
                        : zero2;
      }
    }
    bf16x4_t a;
    a[0] = bf16_rne(av[0].x); a[1] = bf16_rne(av[0].y);
    a[2] = bf16_rne(av[1].x); a[3] = bf16_rne(av[1].y);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      bf16x4_t b;
      b[0] = bf16_rne(wv[t][0].x); b[1] = bf16_rne(wv[t][0].y);
      b[2] = bf16_rne(wv[t][1].x); b[3] = bf16_rne(wv[t][1].y);
      acc[t] = mfma16(a, b, acc[t]);
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int o = t * 16 + r;
    if (o >= S.O) continue;
    const float bo = bias ? bias[o] : 0.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long long p = p0 + 4 * q + j;
      if (p >= P) continue;
      float v = acc[t][j] + bo;
      if (relu) v = v > 0.0f ? v : 0.0f;
      y[p * S.O + o] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// backward data: dx[qpos][c] = sum_{tap,o} w[o][tap][c] dy[p][o] over the
// outputs p with p*s + d - 1 = qpos; optionally times (act[qpos][c] > 0),
// the derivative of the ReLU that produced this layer's input.
// GEMM rows = input positions, cols = c, depth = tap*O + o.
// ---------------------------------------------------------------------------
template <int NT>
__global__ __launch_bounds__(256) void conv_bwd_data_kernel(ConvShape S,
                                                           const float* __restrict__ dy,
                                                           const float* __restrict__ w,
                                                           const float* __restrict__ act,
                                                           float* __restrict__ dx) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const long long P = (long long)S.N * S.D * S.H * S.W;
  const long long p0 = (long long)wave * 16;
  if (p0 >= P) return;
  const int r = lane & 15, q = lane >> 4;
  const long long pa = p0 + r;
  const bool va = pa < P;
  int n = 0, z = 0, yy = 0, xx = 0;
  if (va) {
    long long t = pa;
    xx = (int)(t % S.W); t /= S.W;
    yy = (int)(t % S.H); t /= S.H;
    z = (int)(t % S.D); n = (int)(t / S.D);
  }
  const int K = 27 * S.O;
  const int KW = 27 * S.C;  // weight row length
  f32x4_t acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4_t a, bv[NT];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int k = k0 + 4 * q + j;
      float av = 0.0f;
      int tap = 0, o = 0;
      if (k < K) { tap = k / S.O; o = k - tap * S.O; }
      if (va && k < K) {
        const int dz = tap / 9, dyy = (tap / 3) % 3, dx_ = tap % 3;
        // output position whose tap lands on (z, yy, xx)
        const int tz = z + 1 - dz, ty = yy + 1 - dyy, tx = xx + 1 - dx_;
        if (tz >= 0 && ty >= 0 && tx >= 0 && tz % S.stride == 0 && ty % S.stride == 0 &&
            tx % S.stride == 0) {
          const int oz = tz / S.stride, oy = ty / S.stride, ox = tx / S.stride;
          if (oz < S.Do && oy < S.Ho && ox < S.Wo)
            av = dy[((((long long)n * S.Do + oz) * S.Ho + oy) * S.Wo + ox) * S.O + o];
        }
      }
      a[j] = bf16_rne(av);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int c = t * 16 + r;
        const float wv = (k < K && c < S.C) ? w[(long long)o * KW + tap * S.C + c] : 0.0f;
        bv[t][j] = bf16_rne(wv);
      }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = mfma16(a, bv[t], acc[t]);
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int c = t * 16 + r;
    if (c >= S.C) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long long p = p0 + 4 * q + j;
      if (p >= P) continue;
      float v = acc[t][j];
      if (act && !(act[p * S.C + c] > 0.0f)) v = 0.0f;
      dx[p * S.C + c] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// backward weights: dw[o][tap*C + c] = sum_p dy[p][o] x[p*s + d - 1][c].
// GEMM rows = o, cols = tap*C + c, depth = output positions.  Wave
// (blockIdx.x, blockIdx.y*4 + w) sums positions [pb*chunk, (pb+1)*chunk) for
// 4 column tiles and writes partials part[pb][o][col]; a fixed-order
// reduction finishes (deterministic, no atomics).
// ---------------------------------------------------------------------------
template <int MT>
__global__ __launch_bounds__(256) void conv_bwd_weight_kernel(ConvShape S,
                                                             const float* __restrict__ x,
                                                             const float* __restrict__ dy,
                                                             long long chunk,
                                                             float* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int r = lane & 15, q = lane >> 4;
  const int KW = 27 * S.C;
  const int ntiles = (KW + 15) / 16;
  const int tile0 = (blockIdx.y * 4 + (threadIdx.x >> 6)) * 4;  // 4 column tiles per wave
  if (tile0 >= ntiles) return;
  const long long P = (long long)S.N * S.Do * S.Ho * S.Wo;
  const long long pb = blockIdx.x;
  const long long pbeg = pb * chunk;
  const long long pend = min(P, pbeg + chunk);
  f32x4_t acc[MT][4];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[m][t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  // this lane's B columns (tap, c) per tile
  int ctap[4], cc[4];
  bool cv[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int col = (tile0 + t) * 16 + r;
    cv[t] = col < KW && (tile0 + t) < ntiles;
    const int tap = cv[t] ? col / S.C : 0;
    ctap[t] = tap;
    cc[t] = cv[t] ? col - tap * S.C : 0;
  }
  for (long long k0 = pbeg; k0 < pend; k0 += 16) {
    bf16x4_t a[MT], bv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long long p = k0 + 4 * q + j;
      const bool vp = p < pend;
      int n = 0, z = 0, yy = 0, xx = 0;
      if (vp) {
        long long t = p;
        xx = (int)(t % S.Wo); t /= S.Wo;
        yy = (int)(t % S.Ho); t /= S.Ho;
        z = (int)(t % S.Do); n = (int)(t / S.Do);
      }
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        const int o = m * 16 + r;
        a[m][j] = bf16_rne((vp && o < S.O) ? dy[p * S.O + o] : 0.0f);
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        float bvv = 0.0f;
        if (vp && cv[t]) {
          const int tap = ctap[t];
          const int dz = tap / 9, dyy = (tap / 3) % 3, dx_ = tap % 3;
          const int iz = z * S.stride + dz - 1, iy = yy * S.stride + dyy - 1,
                    ix = xx * S.stride + dx_ - 1;
          if (iz >= 0 && iz < S.D && iy >= 0 && iy < S.H && ix >= 0 && ix < S.W)
            bvv = x[((((long long)n * S.D + iz) * S.H + iy) * S.W + ix) * S.C + cc[t]];
        }
        bv[t][j] = bf16_rne(bvv);
      }
    }
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[m][t] = mfma16(a[m], bv[t], acc[m][t]);
  }
  // D[4q + j][r]: row o = m*16 + 4q + j, column (tile0 + t)*16 + r
  float* dst = part + pb * (long long)S.O * KW;
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (!cv[t]) continue;
      const int col = (tile0 + t) * 16 + r;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int o = m * 16 + 4 * q + j;
        if (o < S.O) dst[(long long)o * KW + col] = acc[m][t][j];
      }
    }
}

// dw[i] = sum_b part[b][i] (fixed order); db[o] = sum_p dy[p][o]
__global__ __launch_bounds__(256) void reduce_dw_kernel(const float* __restrict__ part, int nb,
                                                       long long n, float* __restrict__ dw) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float s = 0.0f;
  for (int b = 0; b < nb; ++b) s += part[(long long)b * n + i];
  dw[i] = s;
}

__global__ __launch_bounds__(256) void bias_grad_kernel(const float* __restrict__ dy,
                                                       long long P, int O,
                                                       float* __restrict__ db) {
  // one block per output channel, fixed-order tree
  __shared__ float lds[256];
  const int o = blockIdx.x;
  float s = 0.0f;
  for (long long p = threadIdx.x; p < P; p += 256) s += dy[p * O + o];
  lds[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) lds[threadIdx.x] += lds[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[o] = lds[0];
}

// Adam (bias-corrected), float32 master parameters
__global__ __launch_bounds__(256) void adam_kernel(long long n, float* __restrict__ prm,
                                                  const float* __restrict__ g,
                                                  float* __restrict__ m, float* __restrict__ v,
                                                  float lr, float b1, float b2, float eps,
                                                  float c1, float c2) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  prm[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// mean squared error: per-block FP64 partial of sum (pred - label)^2 and
// grad = 2 (pred - label) / n
__global__ __launch_bounds__(256) void mse_kernel(long long n, const float* __restrict__ pred,
                                                 const float* __restrict__ label,
                                                 float* __restrict__ grad,
                                                 double* __restrict__ part) {
  __shared__ double lds[4];
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  double v[1] = {0.0};
  if (i < n) {
    const float d = pred[i] - label[i];
    grad[i] = 2.0f * d / (float)n;
    v[0] = (double)d * d;
  }
  block_sum_f64<1>(v, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

// per-block FP64 (count, sum, sum sq) of the stride-s decimated volume
template <typename T>
__global__ __launch_bounds__(256) void dec_stats_kernel(const T* __restrict__ img, int nx, int ny,
                                                       int nz, int s, int mx, int my, int mz,
                                                       double* __restrict__ part) {
  __shared__ double lds[12];
  double acc[3] = {0.0, 0.0, 0.0};
  const long long n = (long long)mx * my * mz;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (long long)gridDim.x * 256) {
    const int a = (int)(i % mx);
    const long long t = i / mx;
    const int b = (int)(t % my), c = (int)(t / my);
    const double v =
        (double)ld_val(img, (long long)a * s + (long long)nx * ((long long)b * s + (long long)ny * c * s));
    acc[0] += 1.0;
    acc[1] += v;
    acc[2] += v * v;
  }
  block_sum_f64<3>(acc, lds);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = acc[0];
    part[3 * blockIdx.x + 1] = acc[1];
    part[3 * blockIdx.x + 2] = acc[2];
  }
}

// network input [z][y][x][2]: channel 0 = (f - mean_f) / std_f, channel 1 =
// (g - mean_g) / std_g on the stride-s decimated grids; stats[0..5] are the
// two images' reduced (count, sum, sum sq)
template <typename T>
__global__ __launch_bounds__(256) void seed_input_kernel(const T* __restrict__ f,
                                                        const T* __restrict__ g, int nx, int ny,
                                                        int nz, int s, int mx, int my, int mz,
                                                        const double* __restrict__ st,
                                                        float* __restrict__ out) {
  const long long n = (long long)mx * my * mz;
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int a = (int)(i % mx);
  const long long t = i / mx;
  const int b = (int)(t % my), c = (int)(t / my);
  const long long src = (long long)a * s + (long long)nx * ((long long)b * s + (long long)ny * c * s);
  double mf = st[1] / st[0], mg = st[4] / st[3];
  double vf = st[2] / st[0] - mf * mf, vg = st[5] / st[3] - mg * mg;
  const double sf = vf > 1e-20 ? sqrt(vf) : 1.0, sg = vg > 1e-20 ? sqrt(vg) : 1.0;
  out[2 * i] = (float)(((double)ld_val(f, src) - mf) / sf);
  out[2 * i + 1] = (float)(((double)ld_val(g, src) - mg) / sg);
}

}  // namespace

namespace vdvc {

static ConvShape conv_shape(int N, int D, int H, int W, int C, int O, int stride) {
  ConvShape S;
  S.N = N; S.D = D; S.H = H; S.W = W; S.C = C; S.O = O; S.stride = stride;
  S.Do = (D - 1) / stride + 1; S.Ho = (H - 1) / stride + 1; S.Wo = (W - 1) / stride + 1;
  return S;
}

int launch_conv_fwd(int N, int D, int H, int W, int C, int O, int stride, const float* x,
                    const float* w, const float* b, int relu, float* y, hipStream_t st) {
  const ConvShape S = conv_shape(N, D, H, W, C, O, stride);
  const long long P = (long long)N * S.Do * S.Ho * S.Wo;
  const long long waves = (P + 15) / 16;
  const unsigned nb = (unsigned)((waves + 3) / 4);
#define VDVC_FWD(K) hipLaunchKernelGGL(K, dim3(nb), dim3(256), 0, st, S, x, w, b, relu, y)
  const bool two = O > 16;
  const unsigned nbl = (unsigned)((P + 16 * kFwdWaves * kFwdTiles - 1) / (16 * kFwdWaves * kFwdTiles));
#define VDVC_FWDL(K) hipLaunchKernelGGL(K, dim3(nbl), dim3(64 * kFwdWaves), 0, st, S, x, w, b, relu, y)
  // the LDS-staged kernel for stride-1 inputs of 16 or 32 channels (a
  // stride-2 input box is eight times the output box: the per-tap gathers
  // of conv_fwd_lds_kernel measured faster there)
  const long long nboxes = (long long)N * ((S.Do + kTileZ - 1) / kTileZ) *
                           ((S.Ho + kTileY - 1) / kTileY) * ((S.Wo + 15) / 16);
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      ncu < 1)
    ncu = 256;
  const unsigned nbt = (unsigned)(nboxes < ncu ? nboxes : ncu);
#define VDVC_FWDT(K) hipLaunchKernelGGL(K, dim3(nbt), dim3(64 * kFwdWaves), 0, st, S, x, w, b, relu, y)
  if (C == 16 && stride == 1) {
    if (two) VDVC_FWDT((conv_fwd_tile_kernel<2, 1, 1>)); else VDVC_FWDT((conv_fwd_tile_kernel<1, 1, 1>));
  } else if (C == 16) {
    if (two) VDVC_FWDL((conv_fwd_lds_kernel<2, 1>)); else VDVC_FWDL((conv_fwd_lds_kernel<1, 1>));
  } else if (C == 32 && stride == 1) {
    if (two) VDVC_FWDT((conv_fwd_tile_kernel<2, 2, 1>)); else VDVC_FWDT((conv_fwd_tile_kernel<1, 2, 1>));
  } else if (C == 32) {
    if (two) VDVC_FWDL((conv_fwd_lds_kernel<2, 2>)); else VDVC_FWDL((conv_fwd_lds_kernel<1, 2>));
  } else if (C == 2) {
    if (two) VDVC_FWD(conv_fwd_c2_kernel<2>); else VDVC_FWD(conv_fwd_c2_kernel<1>);
  } else {
    if (two) VDVC_FWD(conv_fwd_kernel<2>); else VDVC_FWD(conv_fwd_kernel<1>);
  }
#undef VDVC_FWD
#undef VDVC_FWDL
#undef VDVC_FWDT
  return (int)hipGetLastError();
}

int launch_conv_bwd_data(int N, int D, int H, int W, int C, int O, int stride, const float* dy,
                         const float* w, const float* act, float* dx, hipStream_t st) {
  const ConvShape S = conv_shape(N, D, H, W, C, O, stride);
  const long long P = (long long)N * D * H * W;
  const long long waves = (P + 15) / 16;
  const unsigned nb = (unsigned)((waves + 3) / 4);
  if (C <= 16)
    hipLaunchKernelGGL(conv_bwd_data_kernel<1>, dim3(nb), dim3(256), 0, st, S, dy, w, act, dx);
  else
    hipLaunchKernelGGL(conv_bwd_data_kernel<2>, dim3(nb), dim3(256), 0, st, S, dy, w, act, dx);
  return (int)hipGetLastError();
}

// positions summed per partial in the weight gradient
long long conv_wgrad_chunk(long long P) {
  long long c = (P + 255) / 256;
  c = (c + 15) / 16 * 16;
  return c < 256 ? 256 : c;
}

int conv_wgrad_parts(int N, int D, int H, int W, int C, int O, int stride) {
  const ConvShape S = conv_shape(N, D, H, W, C, O, stride);
  const long long P = (long long)N * S.Do * S.Ho * S.Wo;
  const long long ch = conv_wgrad_chunk(P);
  return (int)((P + ch - 1) / ch);
}

int launch_conv_bwd_weight(int N, int D, int H, int W, int C, int O, int stride, const float* x,
                           const float* dy, float* part, float* dw, float* db, hipStream_t st) {
  const ConvShape S = conv_shape(N, D, H, W, C, O, stride);
  const long long P = (long long)N * S.Do * S.Ho * S.Wo;
  const long long ch = conv_wgrad_chunk(P);
  const int nparts = (int)((P + ch - 1) / ch);
  const int KW = 27 * C;
  const int ntiles = (KW + 15) / 16;
  const int ngroups = (ntiles + 15) / 16;  // 4 waves x 4 tiles per block
  dim3 g((unsigned)nparts, (unsigned)ngroups);
  if (O <= 16)
    hipLaunchKernelGGL(conv_bwd_weight_kernel<1>, g, dim3(256), 0, st, S, x, dy, ch, part);
  else
    hipLaunchKernelGGL(conv_bwd_weight_kernel<2>, g, dim3(256), 0, st, S, x, dy, ch, part);
  const long long n = (long long)O * KW;
  hipLaunchKernelGGL(reduce_dw_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, part,
                     nparts, n, dw);
  hipLaunchKernelGGL(bias_grad_kernel, dim3(O), dim3(256), 0, st, dy, P, O, db);
  return (int)hipGetLastError();
}

int launch_adam(long long n, float* p, const float* g, float* m, float* v, float lr, float b1,
                float b2, float eps, float c1, float c2, hipStream_t st) {
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, p, g,
                     m, v, lr, b1, b2, eps, c1, c2);
  return (int)hipGetLastError();
}

int launch_mse(long long n, const float* pred, const float* label, float* grad, double* part,
               hipStream_t st) {
  hipLaunchKernelGGL(mse_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, pred,
                     label, grad, part);
  return (int)hipGetLastError();
}

int launch_seed_input(int itype, const void* f, const void* g, int nx, int ny, int nz, int s,
                      double* part, int nparts, double* stats6, float* out, hipStream_t st) {
  const int mx = (nx - 1) / s + 1, my = (ny - 1) / s + 1, mz = (nz - 1) / s + 1;
  const long long n = (long long)mx * my * mz;
  if (itype == GT_F16) {
    hipLaunchKernelGGL(dec_stats_kernel<__half>, dim3(nparts), dim3(256), 0, st,
                       (const __half*)f, nx, ny, nz, s, mx, my, mz, part);
    hipLaunchKernelGGL(dec_stats_kernel<__half>, dim3(nparts), dim3(256), 0, st,
                       (const __half*)g, nx, ny, nz, s, mx, my, mz, part + 3 * nparts);
  } else {
    hipLaunchKernelGGL(dec_stats_kernel<float>, dim3(nparts), dim3(256), 0, st, (const float*)f,
                       nx, ny, nz, s, mx, my, mz, part);
    hipLaunchKernelGGL(dec_stats_kernel<float>, dim3(nparts), dim3(256), 0, st, (const float*)g,
                       nx, ny, nz, s, mx, my, mz, part + 3 * nparts);
  }
  int e = launch_reduce(part, nparts, 3, stats6, st);
  if (e) return e;
  e = launch_reduce(part + 3 * nparts, nparts, 3, stats6 + 3, st);
  if (e) return e;
  const unsigned nb = (unsigned)((n + 255) / 256);
  if (itype == GT_F16)
    hipLaunchKernelGGL(seed_input_kernel<__half>, dim3(nb), dim3(256), 0, st, (const __half*)f,
                       (const __half*)g, nx, ny, nz, s, mx, my, mz, stats6, out);
  else
    hipLaunchKernelGGL(seed_input_kernel<float>, dim3(nb), dim3(256), 0, st, (const float*)f,
                       (const float*)g, nx, ny, nz, s, mx, my, mz, stats6, out);
  return (int)hipGetLastError();
}

}  // namespace vdvc
