// Node-level kernels: gather of the voxel kernels' node partials, the
// element-stiffness regularizer applied as a node stencil, Jacobi
// preconditioner, the device-resident PCG vector updates with FP64 dots, and
// small vector utilities.
//
// Reference semantics followed:
//   regularizer   engine/kernels_dvc.py:349-437 (matvec_K_ref, diagonals,
//                 equilibrium gap)
//   Jacobi        engine/correlate_gpu.py:1042-1047, 1113-1120
//   PCG           engine/kernels_dvc.py:550-583
#include <algorithm>
#include <cstdlib>

#include "vdvc_kernels.h"

// corner index of local node offset (dx,dy,dz) in the reference Hex8 order
__device__ __forceinline__ int corner_of(int dx, int dy, int dz) {
  return dz * 4 + (dy ? (dx ? 2 : 3) : (dx ? 1 : 0));
}


// Hex8 corner offsets of the reference order as compile-time functions
__device__ __forceinline__ constexpr int cdx(int a) { return (a == 1 || a == 2 || a == 5 || a == 6) ? 1 : 0; }
__device__ __forceinline__ constexpr int cdy(int a) { return (a == 2 || a == 3 || a == 6 || a == 7) ? 1 : 0; }
__device__ __forceinline__ constexpr int cdz(int a) { return a >= 4 ? 1 : 0; }

// the three components of one node: one 12-byte (float) / 24-byte (double)
// load, so a wavefront's load of 64 consecutive nodes is one instruction
// instead of three strided ones
template <typename R>
struct Node3 {
  R c[3];
};
template <typename R>
__device__ __forceinline__ void ld3(const R* __restrict__ p, R& a, R& b, R& c) {
  const Node3<R> t = *reinterpret_cast<const Node3<R>*>(p);
  a = t.c[0]; b = t.c[1]; c = t.c[2];
}
template <typename R>
__device__ __forceinline__ void st3(R* __restrict__ p, R a, R b, R c) {
  Node3<R> t;
  t.c[0] = a; t.c[1] = b; t.c[2] = c;
  *reinterpret_cast<Node3<R>*>(p) = t;
}

// Element-stiffness operator at node (i,j,k): sum over the (up to 8)
// adjacent elements e, in order, of E_e * sum over the element's nodes lb,
// in order, of K[la][lb] v_lb (la the node's corner in e).  The 27
// neighbouring node values are loaded once (zero outside the mesh; they are
// only used by elements inside it) and both loops are unrolled, so every
// neighbour index and stiffness offset is a compile-time constant.  Each
// neighbour is one 3-component load at a 32-bit DOF offset (Ndof < 2^31 is
// checked when a mesh is set up); a neighbour outside the mesh loads the
// node itself and is replaced by zero, so no load is predicated.  The
// element matrix has one of two compile-time forms KM: 1 = block-diagonal
// with equal 3x3 blocks, read from the launch arguments (Ks = KOp::kc,
// Ks[la * 8 + lb] = K[3 la][3 lb]); 2 = the full 24 x 24 matrix, read from
// the block's LDS copy (stage_K).
template <typename R, int KM>
__device__ __forceinline__ void kop_apply(const MeshDev& M, const R* Ks, R E,
                                          const R* __restrict__ Earr,
                                          const R* __restrict__ v, int i, int j, int k,
                                          R (&out)[3]) {
  const unsigned sy = 3u * (unsigned)M.nnx, sz = sy * (unsigned)M.nny;
  const unsigned o = 3u * (unsigned)i + sy * (unsigned)j + sz * (unsigned)k;
  const bool xl = i > 0, xh = i + 1 < M.nnx, yl = j > 0, yh = j + 1 < M.nny, zl = k > 0,
             zh = k + 1 < M.nnz;
  R V[27][3];
#pragma unroll
  for (int m = 0; m < 27; ++m) {
    const int dx = (m % 3) - 1, dy = (m / 3) % 3 - 1, dz = m / 9 - 1;
    const bool in = (dx < 0 ? xl : dx > 0 ? xh : true) && (dy < 0 ? yl : dy > 0 ? yh : true) &&
                    (dz < 0 ? zl : dz > 0 ? zh : true);
    unsigned off = o;
    if (dx) off = dx > 0 ? off + 3u : off - 3u;
    if (dy) off = dy > 0 ? off + sy : off - sy;
    if (dz) off = dz > 0 ? off + sz : off - sz;
    R a, b, c;
    ld3(v + (in ? off : o), a, b, c);
    V[m][0] = in ? a : R(0);
    V[m][1] = in ? b : R(0);
    V[m][2] = in ? c : R(0);
  }
  out[0] = out[1] = out[2] = R(0);
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int ox = (e & 1) - 1, oy = ((e >> 1) & 1) - 1, oz = (e >> 2) - 1;  // element offset
    const int ei = i + ox, ej = j + oy, ek = k + oz;
    if (ei < 0 || ej < 0 || ek < 0 || ei >= M.Nx || ej >= M.Ny || ek >= M.Nz) continue;
    const int la = corner_of(-ox, -oy, -oz);
    const R Ee = Earr ? Earr[(long long)ei + (long long)M.Nx * (ej + (long long)M.Ny * ek)] : E;
    R acc[3] = {R(0), R(0), R(0)};
#pragma unroll
    for (int lb = 0; lb < 8; ++lb) {
      const int m = (ox + cdx(lb) + 1) + 3 * (oy + cdy(lb) + 1) + 9 * (oz + cdz(lb) + 1);
      const R vb0 = V[m][0], vb1 = V[m][1], vb2 = V[m][2];
      if (KM == 1) {
        const R l = Ks[la * 8 + lb];
        acc[0] += l * vb0; acc[1] += l * vb1; acc[2] += l * vb2;
      } else {
        const R* Kr = Ks + (3 * la) * 24 + 3 * lb;
#pragma unroll
        for (int c = 0; c < 3; ++c)
          acc[c] += Kr[c * 24 + 0] * vb0 + Kr[c * 24 + 1] * vb1 + Kr[c * 24 + 2] * vb2;
      }
    }
    out[0] += Ee * acc[0]; out[1] += Ee * acc[1]; out[2] += Ee * acc[2];
  }
}

// size of the LDS copy of an element matrix in form KM (only the full form
// KM = 2 is staged; 1 word otherwise)
template <int KM>
struct KsLen {
  static constexpr int n = KM == 2 ? 576 : 1;
};

// the block's LDS copy of a full element matrix (form 2); the caller
// synchronises the block before the first use
template <typename R>
__device__ __forceinline__ void stage_K(const R* __restrict__ K, R* Ks) {
  for (int t = threadIdx.x; t < 576; t += VDVC_BLOCK) Ks[t] = K[t];
}

template <typename R>
__device__ __forceinline__ void kop_diag(const MeshDev& M, const KOp<R>& op, int i, int j, int k,
                                         R (&out)[3]) {
  out[0] = out[1] = out[2] = R(0);
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int ei = i - 1 + (e & 1), ej = j - 1 + ((e >> 1) & 1), ek = k - 1 + (e >> 2);
    if (ei < 0 || ej < 0 || ek < 0 || ei >= M.Nx || ej >= M.Ny || ek >= M.Nz) continue;
    const int la = corner_of(i - ei, j - ej, k - ek);
    const R Ee = op.Earr ? op.Earr[(long long)ei + (long long)M.Nx * (ej + (long long)M.Ny * ek)]
                             : op.E;
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c] += Ee * op.K[(3 * la + c) * 24 + 3 * la + c];
  }
}

// sum of the node partials written by the voxel kernels for node (i,j,k),
// in fixed order (z tiles, then y rows, then x strips).  SINGLE: the level
// has one x strip and one row part per element row (T.nsx == T.nys == 1), so
// only strip 0, row part 0 of each (z tile, element row) can hold a partial
// and the other slots' loads are not issued (the same sum: those slots are
// absent and add nothing).
template <typename R, bool SINGLE = false>
__device__ __forceinline__ void gather_partials(const MeshDev& M, const TilesDev& T,
                                                const R* __restrict__ P, int i, int j, int k,
                                                R (&s)[3]) {
  const int h = M.h;
  const int kz = k / T.ez;
  const int za = max(0, kz - 1), zb = min(T.nzt - 1, kz);
  int xa, xb;
  if (T.vec) {
    // strips of sw voxels over [0, Nx*h) holding element i-1 or element i
    xa = (max(i - 1, 0) * h) / T.sw;
    xb = (min(i, M.Nx - 1) * h) / T.sw;
  } else {
    xa = max(0, (i - 1) * h) / 64;
    xb = min(M.Nvx - 1, (i + 1) * h) / 64;
  }
  // the node's slot in each of the (at most two) strips sx = xa, xa + 1
  int ln[2];
  bool okx[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int sx = xa + q;
    const int x0 = sx * T.sw;
    const int i_lo = x0 / h;
    int nn;
    if (T.vec) {
      nn = min(M.Nx, (x0 + T.sw) / h) - i_lo + 1;
    } else {
      const int nact = min(64, M.Nvx - x0);
      nn = min(M.Nx, (x0 + nact - 1) / h + 1) - i_lo + 1;
    }
    ln[q] = i - i_lo;
    okx[q] = sx <= xb && ln[q] >= 0 && ln[q] < nn;
  }
  // At most 2 z tiles x 2 element rows x 2 strips x 2 row parts hold
  // partials of a node.  All slot addresses are formed first and every load
  // is issued before the first sum, so their latencies overlap; absent slots
  // read P[0..2] and are dropped.  Summation order is fixed: z tiles, then
  // element rows (upper first), then strips and row parts.
  constexpr int NC = SINGLE ? 1 : 4;  // (strip, row part) slots per (z tile, element row)
  const R* ptr[4 * NC];
  bool ok[4 * NC];
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int zt = za + a;
    const int kb = zt * T.ez;
    const bool okz = zt <= zb && k >= kb && k <= min(M.Nz, kb + T.ez);
    const int kzl = k - kb;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int cy = 1 - b;
      const int ey = j - cy;
      const bool oky = okz && ey >= 0 && ey < M.Ny;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int q = c >> 1, ys = c & 1;  // strip xa + q, row part ys
        const bool okc = oky && okx[q] && ys < T.nys;
        const long long w = (long long)(xa + q) +
                            (long long)T.nsx * (ys + (long long)T.nys * (ey + (long long)M.Ny * zt));
        const int e = (a * 2 + b) * NC + c;
        ok[e] = okc;
        ptr[e] = okc ? P + (size_t)w * T.wave_stride + ((size_t)(kzl * 2 + cy) * T.px + ln[q]) * 3
                     : P;
      }
    }
  }
  R val[4 * NC][3];
#pragma unroll
  for (int e = 0; e < 4 * NC; ++e) ld3(ptr[e], val[e][0], val[e][1], val[e][2]);
  s[0] = s[1] = s[2] = R(0);
#pragma unroll
  for (int e = 0; e < 4 * NC; ++e)
    if (ok[e]) { s[0] += val[e][0]; s[1] += val[e][1]; s[2] += val[e][2]; }
}

// gather_partials with its loads issued in two batches (one z tile each, 8
// slots): the same slots summed in the same order, half the values live
template <typename R, bool SINGLE = false>
__device__ __forceinline__ void gather_partials_lean(const MeshDev& M, const TilesDev& T,
                                                     const R* __restrict__ P, int i, int j,
                                                     int k, R (&s)[3]) {
  const int h = M.h;
  const int kz = k / T.ez;
  const int za = max(0, kz - 1), zb = min(T.nzt - 1, kz);
  int xa, xb;
  if (T.vec) {
    xa = (max(i - 1, 0) * h) / T.sw;
    xb = (min(i, M.Nx - 1) * h) / T.sw;
  } else {
    xa = max(0, (i - 1) * h) / 64;
    xb = min(M.Nvx - 1, (i + 1) * h) / 64;
  }
  int ln[2];
  bool okx[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int sx = xa + q;
    const int x0 = sx * T.sw;
    const int i_lo = x0 / h;
    int nn;
    if (T.vec) {
      nn = min(M.Nx, (x0 + T.sw) / h) - i_lo + 1;
    } else {
      const int nact = min(64, M.Nvx - x0);
      nn = min(M.Nx, (x0 + nact - 1) / h + 1) - i_lo + 1;
    }
    ln[q] = i - i_lo;
    okx[q] = sx <= xb && ln[q] >= 0 && ln[q] < nn;
  }
  s[0] = s[1] = s[2] = R(0);
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int zt = za + a;
    const int kb = zt * T.ez;
    const bool okz = zt <= zb && k >= kb && k <= min(M.Nz, kb + T.ez);
    const int kzl = k - kb;
    constexpr int NC = SINGLE ? 1 : 4;
    const R* ptr[2 * NC];
    bool ok[2 * NC];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int cy = 1 - b;
      const int ey = j - cy;
      const bool oky = okz && ey >= 0 && ey < M.Ny;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int q = c >> 1, ys = c & 1;
        const bool okc = oky && okx[q] && ys < T.nys;
        const long long w = (long long)(xa + q) +
                            (long long)T.nsx * (ys + (long long)T.nys * (ey + (long long)M.Ny * zt));
        const int e = b * NC + c;
        ok[e] = okc;
        ptr[e] = okc ? P + (size_t)w * T.wave_stride + ((size_t)(kzl * 2 + cy) * T.px + ln[q]) * 3
                     : P;
      }
    }
    R val[2 * NC][3];
#pragma unroll
    for (int e = 0; e < 2 * NC; ++e) ld3(ptr[e], val[e][0], val[e][1], val[e][2]);
#pragma unroll
    for (int e = 0; e < 2 * NC; ++e)
      if (ok[e]) { s[0] += val[e][0]; s[1] += val[e][1]; s[2] += val[e][2]; }
  }
}

// node (i, j, k) of flat node index n (< 2^31: checked when a level is set up)
__device__ __forceinline__ void node_ijk(const MeshDev& M, unsigned n, int& i, int& j, int& k) {
  const unsigned r = n / (unsigned)M.nnx;
  i = (int)(n - r * (unsigned)M.nnx);
  k = (int)(r / (unsigned)M.nny);
  j = (int)(r - (unsigned)k * (unsigned)M.nny);
}

// KM1 / KM2: compile-time forms of the two element matrices (0: the term
// is off, 1: block-diagonal, 2: full; see kop_apply)
template <typename R, int KM1, int KM2>
__global__ __launch_bounds__(VDVC_BLOCK) void node_op_kernel(MeshDev M, NodeOp<R> op) {
  __shared__ double lds[4];
  __shared__ R ks1[KsLen<KM1>::n];
  __shared__ R ks2[KsLen<KM2>::n];
  if (op.done && *op.done) return;
  if (KM1 == 2) stage_K(op.k1.K, ks1);
  if (KM2 == 2) stage_K(op.k2.K, ks2);
  if (KM1 == 2 || KM2 == 2) __syncthreads();
  const long long n = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  double d = 0.0;
  if (n < M.Nnodes) {
    int i, j, k;
    node_ijk(M, (unsigned)n, i, j, k);
    // the dot operand (never the output vector) is loaded first, so its
    // latency overlaps the gather and the stencil
    R d0 = R(0), d1 = R(0), d2 = R(0);
    if (op.dot_part) ld3(op.dotv + 3 * n, d0, d1, d2);
    R acc[3] = {R(0), R(0), R(0)};
    if (op.P) {
      R s[3];
      if (op.T.nsx == 1 && op.T.nys == 1)
        gather_partials<R, true>(M, op.T, op.P, i, j, k, s);
      else
        gather_partials<R, false>(M, op.T, op.P, i, j, k, s);
      acc[0] = op.cH * s[0]; acc[1] = op.cH * s[1]; acc[2] = op.cH * s[2];
    }
    if (op.addin) {
      R a0, a1, a2;
      ld3(op.addin + 3 * n, a0, a1, a2);
      acc[0] += op.cadd * a0;
      acc[1] += op.cadd * a1;
      acc[2] += op.cadd * a2;
    }
    if (KM1) {
      R t[3];
      kop_apply<R, KM1>(M, KM1 == 1 ? op.k1.kc : ks1, op.k1.E, op.k1.Earr, op.v1, i, j, k, t);
      acc[0] += op.c1 * t[0]; acc[1] += op.c1 * t[1]; acc[2] += op.c1 * t[2];
    }
    if (KM2) {
      R t[3];
      kop_apply<R, KM2>(M, KM2 == 1 ? op.k2.kc : ks2, op.k2.E, op.k2.Earr, op.v2, i, j, k, t);
      acc[0] += op.c2 * t[0]; acc[1] += op.c2 * t[1]; acc[2] += op.c2 * t[2];
    }
    if (op.interior) {
      // interior nodes: strictly inside the mesh bounding box
      // (engine/geometry_dvc.py:201-217)
      const bool in = i > 0 && i < M.Nx && j > 0 && j < M.Ny && k > 0 && k < M.Nz;
      if (!in) { acc[0] = R(0); acc[1] = R(0); acc[2] = R(0); }
    }
    st3(op.out + 3 * n, acc[0], acc[1], acc[2]);
    if (op.dot_part) d = (double)d0 * acc[0] + (double)d1 * acc[1] + (double)d2 * acc[2];
  }
  if (op.dot_part) {
    double v[1] = {d};
    block_sum_f64<1>(v, lds);
    if (threadIdx.x == 0) op.dot_part[blockIdx.x] = v[0];
  }
}


// The node gather of the Hessian matvec alone (no regularizer): the
// partials' loads in two batches (gather_partials_lean), so the kernel
// needs few registers; the same sums in the same order as node_op_kernel.
template <typename R>
__global__ __launch_bounds__(VDVC_BLOCK) void node_gather_kernel(MeshDev M, NodeOp<R> op) {
  __shared__ double lds[4];
  if (op.done && *op.done) return;
  const long long n = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  double d = 0.0;
  if (n < M.Nnodes) {
    int i, j, k;
    node_ijk(M, (unsigned)n, i, j, k);
    R sp[3];
    if (op.T.nsx == 1 && op.T.nys == 1)
      gather_partials_lean<R, true>(M, op.T, op.P, i, j, k, sp);
    else
      gather_partials_lean<R, false>(M, op.T, op.P, i, j, k, sp);
    const R acc[3] = {op.cH * sp[0], op.cH * sp[1], op.cH * sp[2]};
    st3(op.out + 3 * n, acc[0], acc[1], acc[2]);
    if (op.dot_part) {
      R d0, d1, d2;
      ld3(op.dotv + 3 * n, d0, d1, d2);
      d = (double)d0 * acc[0] + (double)d1 * acc[1] + (double)d2 * acc[2];
    }
  }
  if (op.dot_part) {
    double vv[1] = {d};
    block_sum_f64<1>(vv, lds);
    if (threadIdx.x == 0) op.dot_part[blockIdx.x] = vv[0];
  }
}

template <typename R>
__global__ __launch_bounds__(VDVC_BLOCK) void diag_op_kernel(MeshDev M, DiagOp<R> op) {
  __shared__ R lds[VDVC_BLOCK];
  const long long n = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  R mn = R(INFINITY);
  if (n < M.Nnodes) {
    int i, j, k;
    node_ijk(M, (unsigned)n, i, j, k);
    R acc[3] = {R(0), R(0), R(0)};
    if (op.P) gather_partials(M, op.T, op.P, i, j, k, acc);
    if (op.k1_on) {
      R t[3];
      kop_diag(M, op.k1, i, j, k, t);
      acc[0] += op.c1 * t[0]; acc[1] += op.c1 * t[1]; acc[2] += op.c1 * t[2];
    }
    if (op.extra) {
      R e0, e1, e2;
      ld3(op.extra + 3 * n, e0, e1, e2);
      acc[0] += op.cextra * e0;
      acc[1] += op.cextra * e1;
      acc[2] += op.cextra * e2;
    }
    st3(op.out + 3 * n, acc[0], acc[1], acc[2]);
    mn = rmin(acc[0], rmin(acc[1], acc[2]));
  }
  if (op.min_part) {
    lds[threadIdx.x] = mn;
    __syncthreads();
    for (int o = VDVC_BLOCK / 2; o > 0; o >>= 1) {
      if (threadIdx.x < o) lds[threadIdx.x] = rmin(lds[threadIdx.x], lds[threadIdx.x + o]);
      __syncthreads();
    }
    if (threadIdx.x == 0) op.min_part[blockIdx.x] = lds[0];
  }
}

// M_inv = 1 / (diag + 1e-5 * min(diag)); every block reduces the partial
// minima itself (fixed order), so no host round trip is needed
template <typename R>
__global__ __launch_bounds__(VDVC_BLOCK) void jacobi_kernel(const R* __restrict__ diag,
                                                            const R* __restrict__ min_part,
                                                            int nparts, long long ndof,
                                                            R* __restrict__ minv) {
  __shared__ R lds[VDVC_BLOCK];
  R mn = R(INFINITY);
  for (int q = threadIdx.x; q < nparts; q += VDVC_BLOCK) mn = rmin(mn, min_part[q]);
  lds[threadIdx.x] = mn;
  __syncthreads();
  for (int o = VDVC_BLOCK / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) lds[threadIdx.x] = rmin(lds[threadIdx.x], lds[threadIdx.x + o]);
    __syncthreads();
  }
  const R eps0 = (R)(1e-5 * (double)lds[0]);
  for (long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x; q < ndof;
       q += (long long)gridDim.x * VDVC_BLOCK)
    minv[q] = R(1) / (diag[q] + eps0);
}

// ---------------------------------------------------------------------------
// PCG (x0 = 0) with FP64 dot products; state lives in device memory
// ---------------------------------------------------------------------------

// Fixed-order sums of NV interleaved partials, out[t] = sum over q < n of
// part[q * NV + t]: a strided sum per thread, a butterfly across the
// wavefront (every lane ends with the same value), then the block's
// wavefronts in order.  Every thread of every block computes the same sums
// in the same order, so all blocks take the same step length.
template <int NV>
__device__ __forceinline__ void sum_parts(const double* __restrict__ part, int n,
                                          double (&out)[NV], double* lds /* >= 4 * NV */) {
  double s[NV];
#pragma unroll
  for (int t = 0; t < NV; ++t) s[t] = 0.0;
  // batches of U strided partials: every load of a batch is issued before
  // the first add (absent ones read as 0.0, which leaves the sum unchanged),
  // so a block of up to U * VDVC_BLOCK partials costs one memory round trip
  constexpr int U = 8;
  for (int q0 = threadIdx.x; q0 < n; q0 += U * VDVC_BLOCK) {
    double tmp[U][NV];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int q = q0 + u * VDVC_BLOCK;
#pragma unroll
      for (int t = 0; t < NV; ++t) tmp[u][t] = q < n ? part[(size_t)q * NV + t] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int t = 0; t < NV; ++t) s[t] += tmp[u][t];
  }
#pragma unroll
  for (int t = 0; t < NV; ++t) s[t] = wave_sum(s[t]);
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int t = 0; t < NV; ++t) lds[(threadIdx.x >> 6) * NV + t] = s[t];
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < NV; ++t) {
    double r = 0.0;
#pragma unroll
    for (int w = 0; w < VDVC_WAVES_PER_BLOCK; ++w) r += lds[w * NV + t];
    out[t] = r;
  }
  __syncthreads();  // lds is free again
}

template <typename R>
__global__ __launch_bounds__(VDVC_BLOCK) void pcg_init_kernel(long long ndof,
                                                              const R* __restrict__ b,
                                                              const R* __restrict__ minv,
                                                              R* __restrict__ x,
                                                              R* __restrict__ r,
                                                              R* __restrict__ z,
                                                              R* __restrict__ p,
                                                              double* __restrict__ part) {
  __shared__ double lds[8];
  const long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (q < ndof) {
    const R bq = b[q];
    const R zq = minv[q] * bq;
    x[q] = R(0); r[q] = bq; z[q] = zq; p[q] = zq;
    v[0] = (double)bq * zq;
    v[1] = (double)bq * bq;
  }
  block_sum_f64<2>(v, lds);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = v[0]; part[2 * blockIdx.x + 1] = v[1]; }
}

__global__ __launch_bounds__(VDVC_BLOCK) void pcg_init_state_kernel(const double* __restrict__ part,
                                                                    int nparts,
                                                                    PcgState* __restrict__ S) {
  __shared__ double lds[VDVC_BLOCK];
  double rzbb[2];
  sum_parts<2>(part, nparts, rzbb, lds);
  const double rz = rzbb[0], bb = rzbb[1];
  if (threadIdx.x == 0) {
    S->rz[0] = rz; S->rz[1] = 0.0;
    S->r0 = bb + 1e-30;
    S->relres = 1.0; S->done = 0; S->iters = 0;
  }
}

// alpha = rz / (p.Ap + 1e-30); x += alpha p; r -= alpha Ap; z = M r;
// partial dots (r.r, r.z)
template <typename R>
__global__ __launch_bounds__(VDVC_BLOCK) void pcg_update_kernel(
    long long ndof, int it, const double* __restrict__ pap_part, int npap,
    PcgState* __restrict__ S, const R* __restrict__ minv, const R* __restrict__ p,
    const R* __restrict__ Ap, R* __restrict__ x, R* __restrict__ r,
    R* __restrict__ z, double* __restrict__ part) {
  __shared__ double lds[VDVC_BLOCK];
  if (S->done) return;
  double pAp1[1];
  sum_parts<1>(pap_part, npap, pAp1, lds);
  const double pAp = pAp1[0];
  const double alpha = S->rz[it & 1] / (pAp + 1e-30);
  const R af = (R)alpha;
  double v[2] = {0.0, 0.0};
  for (long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x; q < ndof;
       q += (long long)gridDim.x * VDVC_BLOCK) {
    const R pq = p[q];
    x[q] = x[q] + af * pq;
    const R rq = r[q] - af * Ap[q];
    r[q] = rq;
    const R zq = minv[q] * rq;
    z[q] = zq;
    v[0] += (double)rq * rq;
    v[1] += (double)rq * zq;
  }
  block_sum_f64<2>(v, lds);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = v[0]; part[2 * blockIdx.x + 1] = v[1]; }
}

// convergence test sqrt(r.r / r0) < tol (then stop: x is final), else
// beta = rz_new / (rz + 1e-30); p = z + beta p
template <typename R>
__global__ __launch_bounds__(VDVC_BLOCK) void pcg_pupdate_kernel(
    long long ndof, int it, double tol, const double* __restrict__ part, int nparts,
    PcgState* __restrict__ S, const R* __restrict__ z, R* __restrict__ p) {
  __shared__ double lds[VDVC_BLOCK];
  if (S->done) return;
  double rrz[2];
  sum_parts<2>(part, nparts, rrz, lds);
  const double rr = rrz[0], rzn = rrz[1];
  const double res = sqrt(rr / S->r0);
  const bool conv = res < tol;
  if (conv) {
    // every block sees the same sums; block 0 records the stop
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      S->relres = res;
      S->iters = it + 1;
      __threadfence();
      S->done = 1;
    }
    return;
  }
  const double beta = rzn / (S->rz[it & 1] + 1e-30);
  const R bf = (R)beta;
  for (long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x; q < ndof;
       q += (long long)gridDim.x * VDVC_BLOCK)
    p[q] = fma(bf, p[q], z[q]);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S->rz[(it + 1) & 1] = rzn;
    S->relres = res;
    S->iters = it + 1;
  }
}

// Vectorised forms of the two PCG updates: a thread owns PCG_EPT entries as
// 16-byte chunks (every pointer 16-byte aligned), chunk c of thread t at
// entry b0 + (c * VDVC_BLOCK + t) * W (W entries per chunk, b0 the block's
// first entry), so each load instruction of a wavefront reads 1 KiB
// contiguously; it issues all of its loads before the block reduces the
// previous kernel's partial dots, and the grid covers the vector once, so a
// kernel costs one memory round trip instead of one per grid-stride step.
// The same arithmetic per entry as the scalar kernels; the partial dots are
// summed per thread in chunk and entry order, then per wavefront and block
// in fixed order.
constexpr int PCG_EPT = 8;

template <typename R>
struct PackW {
  static constexpr int W = 16 / sizeof(R);     // entries per 16-byte chunk
  static constexpr int NC = PCG_EPT / W;       // chunks per thread
};

template <typename R>
struct alignas(16) Chunk {
  R v[PackW<R>::W];
};

// first entry of chunk c of this thread
template <typename R>
__device__ __forceinline__ long long chunk_base(int c) {
  return (long long)blockIdx.x * VDVC_BLOCK * PCG_EPT +
         ((long long)c * VDVC_BLOCK + threadIdx.x) * PackW<R>::W;
}

template <typename R>
__device__ __forceinline__ void ld_pack(const R* __restrict__ a, long long n, R (&o)[PCG_EPT]) {
  constexpr int W = PackW<R>::W;
#pragma unroll
  for (int c = 0; c < PackW<R>::NC; ++c) {
    const long long q0 = chunk_base<R>(c);
    if (q0 + W <= n) {
      const Chunk<R> t = *reinterpret_cast<const Chunk<R>*>(a + q0);
#pragma unroll
      for (int e = 0; e < W; ++e) o[c * W + e] = t.v[e];
    } else {
#pragma unroll
      for (int e = 0; e < W; ++e) o[c * W + e] = q0 + e < n ? a[q0 + e] : R(0);
    }
  }
}

template <typename R>
__device__ __forceinline__ void st_pack(R* __restrict__ a, long long n, const R (&o)[PCG_EPT]) {
  constexpr int W = PackW<R>::W;
#pragma unroll
  for (int c = 0; c < PackW<R>::NC; ++c) {
    const long long q0 = chunk_base<R>(c);
    if (q0 + W <= n) {
      Chunk<R> t;
#pragma unroll
      for (int e = 0; e < W; ++e) t.v[e] = o[c * W + e];
      *reinterpret_cast<Chunk<R>*>(a + q0) = t;
    } else {
#pragma unroll
      for (int e = 0; e < W; ++e)
        if (q0 + e < n) a[q0 + e] = o[c * W + e];
    }
  }
}

// is entry (chunk c, e) of this thread inside the vector
template <typename R>
__device__ __forceinline__ bool pack_in(long long n, int idx) {
  return chunk_base<R>(idx / PackW<R>::W) + idx % PackW<R>::W < n;
}

template <typename R>
__global__ __launch_bounds__(VDVC_BLOCK) void pcg_update_vec_kernel(
    long long ndof, int it, const double* __restrict__ pap_part, int npap,
    PcgState* __restrict__ S, const R* __restrict__ minv, const R* __restrict__ p,
    const R* __restrict__ Ap, R* __restrict__ x, R* __restrict__ r,
    R* __restrict__ z, double* __restrict__ part) {
  __shared__ double lds[VDVC_BLOCK];
  R pv[PCG_EPT], av[PCG_EPT], xv[PCG_EPT], rv[PCG_EPT], mv[PCG_EPT], zv[PCG_EPT];
  ld_pack(p, ndof, pv);
  ld_pack(Ap, ndof, av);
  ld_pack(x, ndof, xv);
  ld_pack(r, ndof, rv);
  ld_pack(minv, ndof, mv);
  if (S->done) return;
  const double rz = S->rz[it & 1];
  double pAp1[1];
  sum_parts<1>(pap_part, npap, pAp1, lds);
  const double alpha = rz / (pAp1[0] + 1e-30);
  const R af = (R)alpha;
  double v[2] = {0.0, 0.0};
#pragma unroll
  for (int e = 0; e < PCG_EPT; ++e) {
    xv[e] = xv[e] + af * pv[e];
    rv[e] = rv[e] - af * av[e];
    zv[e] = mv[e] * rv[e];
    if (pack_in<R>(ndof, e)) {
      v[0] += (double)rv[e] * rv[e];
      v[1] += (double)rv[e] * zv[e];
    }
  }
  st_pack(x, ndof, xv);
  st_pack(r, ndof, rv);
  st_pack(z, ndof, zv);
  block_sum_f64<2>(v, lds);
  if (threadIdx.x == 0) { part[2 * blockIdx.x] = v[0]; part[2 * blockIdx.x + 1] = v[1]; }
}

template <typename R>
__global__ __launch_bounds__(VDVC_BLOCK) void pcg_pupdate_vec_kernel(
    long long ndof, int it, double tol, const double* __restrict__ part, int nparts,
    PcgState* __restrict__ S, const R* __restrict__ z, R* __restrict__ p) {
  __shared__ double lds[VDVC_BLOCK];
  R zv[PCG_EPT], pv[PCG_EPT];
  ld_pack(z, ndof, zv);
  ld_pack(p, ndof, pv);
  if (S->done) return;
  const double r0 = S->r0, rzo = S->rz[it & 1];
  double rrz[2];
  sum_parts<2>(part, nparts, rrz, lds);
  const double rr = rrz[0], rzn = rrz[1];
  const double res = sqrt(rr / r0);
  if (res < tol) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      S->relres = res;
      S->iters = it + 1;
      __threadfence();
      S->done = 1;
    }
    return;
  }
  const double beta = rzn / (rzo + 1e-30);
  const R bf = (R)beta;
#pragma unroll
  for (int e = 0; e < PCG_EPT; ++e) pv[e] = fma(bf, pv[e], zv[e]);
  st_pack(p, ndof, pv);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    S->rz[(it + 1) & 1] = rzn;
    S->relres = res;
    S->iters = it + 1;
  }
}

// ---------------------------------------------------------------------------
// small vector utilities
// ---------------------------------------------------------------------------
// y = a*x + b*y
template <typename R>
__global__ __launch_bounds__(VDVC_BLOCK) void axpby_kernel(long long n, R a,
                                                           const R* __restrict__ x, R b,
                                                           R* __restrict__ y) {
  const long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  if (q < n) y[q] = a * x[q] + b * y[q];
}

// partial FP64 dot products of two vectors
template <typename R>
__global__ __launch_bounds__(VDVC_BLOCK) void dot_kernel(long long n, const R* __restrict__ a,
                                                         const R* __restrict__ b,
                                                         double* __restrict__ part) {
  __shared__ double lds[4];
  double v[1] = {0.0};
  for (long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x; q < n;
       q += (long long)gridDim.x * VDVC_BLOCK)
    v[0] += (double)a[q] * b[q];
  block_sum_f64<1>(v, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

// The GN update U <- U + dU with the partial FP64 dot products dU.dU and
// U.U of the updated U: the same element operations, grid and per-thread
// summation order as axpby_kernel (a = b = 1) followed by dot_kernel on each
// pair, so the two sums equal those of separate dot products bit for bit.
template <typename R>
__global__ __launch_bounds__(VDVC_BLOCK) void gn_update_kernel(long long n,
                                                               const R* __restrict__ dU,
                                                               R* __restrict__ U,
                                                               double* __restrict__ part) {
  __shared__ double lds[4 * 2];
  double v[2] = {0.0, 0.0};
  for (long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x; q < n;
       q += (long long)gridDim.x * VDVC_BLOCK) {
    const R d = dU[q];
    const R u = d + U[q];
    U[q] = u;
    v[0] += (double)d * d;
    v[1] += (double)u * u;
  }
  block_sum_f64<2>(v, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = v[0];
    part[2 * blockIdx.x + 1] = v[1];
  }
}

// two partial FP64 dot products a.b and a.c in one pass: per value the grid,
// per-thread order and block reduction of dot_kernel, so each equals the
// corresponding dot product bit for bit
template <typename R>
__global__ __launch_bounds__(VDVC_BLOCK) void dot2_kernel(long long n, const R* __restrict__ a,
                                                          const R* __restrict__ b,
                                                          const R* __restrict__ c,
                                                          double* __restrict__ part) {
  __shared__ double lds[4 * 2];
  double v[2] = {0.0, 0.0};
  for (long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x; q < n;
       q += (long long)gridDim.x * VDVC_BLOCK) {
    const R x = a[q];
    v[0] += (double)x * b[q];
    v[1] += (double)x * c[q];
  }
  block_sum_f64<2>(v, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = v[0];
    part[2 * blockIdx.x + 1] = v[1];
  }
}

// Hutchinson accumulation acc += z * Rz (FP64), z a Rademacher probe
template <typename R>
__global__ __launch_bounds__(VDVC_BLOCK) void hutch_acc_kernel(long long n,
                                                               const int8_t* __restrict__ z,
                                                               const R* __restrict__ Rz,
                                                               double* __restrict__ acc) {
  const long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  if (q < n) acc[q] += (double)z[q] * (double)Rz[q];
}

// an int8 probe vector converted to the working type
template <typename R>
__global__ __launch_bounds__(VDVC_BLOCK) void expand_i8_kernel(long long n,
                                                               const int8_t* __restrict__ z,
                                                               R* __restrict__ out) {
  const long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  if (q < n) out[q] = (R)z[q];
}

// max |acc| / nprobes per block
__global__ __launch_bounds__(VDVC_BLOCK) void hutch_max_kernel(long long n,
                                                               const double* __restrict__ acc,
                                                               double* __restrict__ part) {
  __shared__ double lds[VDVC_BLOCK];
  double m = 0.0;
  for (long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x; q < n;
       q += (long long)gridDim.x * VDVC_BLOCK)
    m = fmax(m, fabs(acc[q]));
  lds[threadIdx.x] = m;
  __syncthreads();
  for (int o = VDVC_BLOCK / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) lds[threadIdx.x] = fmax(lds[threadIdx.x], lds[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = lds[0];
}

// d = acc / nprobes clipped below at 1e-12 * max|d| (every block reduces
// the block maxima itself)
template <typename R>
__global__ __launch_bounds__(VDVC_BLOCK) void hutch_final_kernel(long long n,
                                                                 const double* __restrict__ acc,
                                                                 int nprobes,
                                                                 const double* __restrict__ part,
                                                                 int nparts,
                                                                 R* __restrict__ out) {
  __shared__ double lds[VDVC_BLOCK];
  double m = 0.0;
  for (int q = threadIdx.x; q < nparts; q += VDVC_BLOCK) m = fmax(m, part[q]);
  lds[threadIdx.x] = m;
  __syncthreads();
  for (int o = VDVC_BLOCK / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) lds[threadIdx.x] = fmax(lds[threadIdx.x], lds[threadIdx.x + o]);
    __syncthreads();
  }
  const double lo = 1e-12 * (lds[0] / nprobes);
  for (long long q = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x; q < n;
       q += (long long)gridDim.x * VDVC_BLOCK) {
    const double d = acc[q] / nprobes;
    out[q] = (R)(d < lo ? lo : d);
  }
}

namespace vdvc {

int vec_blocks(long long n) { return (int)((n + VDVC_BLOCK - 1) / VDVC_BLOCK); }

// grid of the PCG vector updates: grid-stride over at most 512 blocks, so
// the next kernel's per-block sum over the partial dots stays short
int pcg_blocks(long long n) { return std::min(vec_blocks(n), 512); }

int node_blocks(const MeshDev& M) { return (int)((M.Nnodes + VDVC_BLOCK - 1) / VDVC_BLOCK); }

template <typename R>
int launch_hutch_acc(long long ndof, const int8_t* z, const R* Rz, double* acc, hipStream_t st) {
  hipLaunchKernelGGL(hutch_acc_kernel<R>, dim3(vec_blocks(ndof)), dim3(VDVC_BLOCK), 0, st, ndof,
                     z, Rz, acc);
  return (int)hipGetLastError();
}

template <typename R>
int launch_expand_i8(long long n, const int8_t* z, R* out, hipStream_t st) {
  hipLaunchKernelGGL(expand_i8_kernel<R>, dim3(vec_blocks(n)), dim3(VDVC_BLOCK), 0, st, n, z,
                     out);
  return (int)hipGetLastError();
}

template <typename R>
int launch_hutch_final(long long ndof, const double* acc, int nprobes, R* out, double* scratch,
                       hipStream_t st) {
  int nb = vec_blocks(ndof);
  if (nb > 1024) nb = 1024;
  hipLaunchKernelGGL(hutch_max_kernel, dim3(nb), dim3(VDVC_BLOCK), 0, st, ndof, acc, scratch);
  hipLaunchKernelGGL(hutch_final_kernel<R>, dim3(nb), dim3(VDVC_BLOCK), 0, st, ndof, acc,
                     nprobes, scratch, nb, out);
  return (int)hipGetLastError();
}

template <typename R>
int launch_node_op(const MeshDev& M, const NodeOp<R>& op, hipStream_t st) {
  // the gather alone takes the low-register kernel
  const bool gather_only = op.P && !op.addin && !op.k1_on && !op.k2_on && !op.interior;
  if (gather_only)
    hipLaunchKernelGGL(node_gather_kernel<R>, dim3(node_blocks(M)), dim3(VDVC_BLOCK), 0, st, M,
                       op);
  else {
    const int km1 = op.k1_on ? (op.k1.blockdiag ? 1 : 2) : 0;
    const int km2 = op.k2_on ? (op.k2.blockdiag ? 1 : 2) : 0;
    const dim3 g(node_blocks(M)), b(VDVC_BLOCK);
#define VDVC_NODE_OP(a, c)                                                     \
  if (km1 == a && km2 == c) {                                                  \
    hipLaunchKernelGGL((node_op_kernel<R, a, c>), g, b, 0, st, M, op);         \
    return (int)hipGetLastError();                                             \
  }
    VDVC_NODE_OP(0, 0)
    VDVC_NODE_OP(1, 0)
    VDVC_NODE_OP(2, 0)
    VDVC_NODE_OP(1, 1)
    VDVC_NODE_OP(1, 2)
    VDVC_NODE_OP(2, 1)
    VDVC_NODE_OP(2, 2)
#undef VDVC_NODE_OP
    // a second matrix without a first is not formed by any entry point
    return (int)hipErrorInvalidValue;
  }
  return (int)hipGetLastError();
}

template <typename R>
int launch_diag_op(const MeshDev& M, const DiagOp<R>& op, hipStream_t st) {
  hipLaunchKernelGGL(diag_op_kernel<R>, dim3(node_blocks(M)), dim3(VDVC_BLOCK), 0, st, M, op);
  return (int)hipGetLastError();
}

template <typename R>
int launch_jacobi(const R* diag, const R* min_part, int nparts, long long ndof, R* minv,
                  hipStream_t st) {
  long long nb = (ndof + VDVC_BLOCK - 1) / VDVC_BLOCK;
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(jacobi_kernel<R>, dim3((unsigned)nb), dim3(VDVC_BLOCK), 0, st, diag,
                     min_part, nparts, ndof, minv);
  return (int)hipGetLastError();
}

template <typename R>
int launch_pcg_init(long long ndof, const R* b, const R* minv, R* x, R* r, R* z, R* p,
                    double* part, PcgState* S, hipStream_t st) {
  const int nb = vec_blocks(ndof);
  hipLaunchKernelGGL(pcg_init_kernel<R>, dim3(nb), dim3(VDVC_BLOCK), 0, st, ndof, b, minv, x, r,
                     z, p, part);
  hipLaunchKernelGGL(pcg_init_state_kernel, dim3(1), dim3(VDVC_BLOCK), 0, st, part, nb, S);
  return (int)hipGetLastError();
}

// the vectorised PCG updates cover the vector with one grid; beyond 4096
// blocks (ndof > 8.4M) the grid-stride kernels keep the partial count short
bool pcg_vec_ok(long long ndof) {
  return (ndof + (long long)VDVC_BLOCK * PCG_EPT - 1) / ((long long)VDVC_BLOCK * PCG_EPT) <= 4096;
}

int pcg_upd_blocks(long long ndof, bool vec) {
  return vec ? (int)((ndof + (long long)VDVC_BLOCK * PCG_EPT - 1) /
                     ((long long)VDVC_BLOCK * PCG_EPT))
             : pcg_blocks(ndof);
}

template <typename R>
int launch_pcg_update(long long ndof, int it, const double* pap_part, int npap, PcgState* S,
                      const R* minv, const R* p, const R* Ap, R* x, R* r, R* z, double* part,
                      bool vec, hipStream_t st) {
  if (vec)
    hipLaunchKernelGGL(pcg_update_vec_kernel<R>, dim3(pcg_upd_blocks(ndof, true)),
                       dim3(VDVC_BLOCK), 0, st, ndof, it, pap_part, npap, S, minv, p, Ap, x, r,
                       z, part);
  else
    hipLaunchKernelGGL(pcg_update_kernel<R>, dim3(pcg_blocks(ndof)), dim3(VDVC_BLOCK), 0, st,
                       ndof, it, pap_part, npap, S, minv, p, Ap, x, r, z, part);
  return (int)hipGetLastError();
}

template <typename R>
int launch_pcg_pupdate(long long ndof, int it, double tol, const double* part, PcgState* S,
                       const R* z, R* p, bool vec, hipStream_t st) {
  const int nb = pcg_upd_blocks(ndof, vec);
  if (vec)
    hipLaunchKernelGGL(pcg_pupdate_vec_kernel<R>, dim3(nb), dim3(VDVC_BLOCK), 0, st, ndof, it,
                       tol, part, nb, S, z, p);
  else
    hipLaunchKernelGGL(pcg_pupdate_kernel<R>, dim3(nb), dim3(VDVC_BLOCK), 0, st, ndof, it, tol,
                       part, nb, S, z, p);
  return (int)hipGetLastError();
}

template <typename R>
int launch_axpby(long long n, double a, const R* x, double b, R* y, hipStream_t st) {
  hipLaunchKernelGGL(axpby_kernel<R>, dim3(vec_blocks(n)), dim3(VDVC_BLOCK), 0, st, n, (R)a, x,
                     (R)b, y);
  return (int)hipGetLastError();
}

// the regularization weight's plane-wave calibration mode: x component
// cos(j h / T * 2 pi) at node (i, j, k), T = 10 h, y and z components zero;
// one correctly rounded operation at a time, left to right, in R
template <typename R>
__global__ __launch_bounds__(VDVC_BLOCK) void calib_mode_kernel(long long nn, int nnx, int nny,
                                                                R hR, R TR, R* __restrict__ out) {
  const long long n = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  if (n >= nn) return;
  const long long j = (n / nnx) % nny;
  R v = (R)j * hR;
  v = v / TR;
  v = v * R(2);
  v = v * (R)3.141592653589793;
  out[3 * n] = cos(v);
  out[3 * n + 1] = R(0);
  out[3 * n + 2] = R(0);
}

template <typename R>
int launch_calib_mode(int Nx, int Ny, int Nz, int h, R* out, hipStream_t st) {
  const long long nn = (long long)(Nx + 1) * (Ny + 1) * (Nz + 1);
  const unsigned nb = (unsigned)((nn + VDVC_BLOCK - 1) / VDVC_BLOCK);
  hipLaunchKernelGGL(calib_mode_kernel<R>, dim3(nb), dim3(VDVC_BLOCK), 0, st, nn, Nx + 1, Ny + 1,
                     (R)h, (R)(10 * h), out);
  return (int)hipGetLastError();
}

template <typename R>
int launch_dot(long long n, const R* a, const R* b, double* part, int nparts, hipStream_t st) {
  hipLaunchKernelGGL(dot_kernel<R>, dim3(nparts), dim3(VDVC_BLOCK), 0, st, n, a, b, part);
  return (int)hipGetLastError();
}

template <typename R>
int launch_dot2(long long n, const R* a, const R* b, const R* c, double* part, int nparts,
                hipStream_t st) {
  hipLaunchKernelGGL(dot2_kernel<R>, dim3(nparts), dim3(VDVC_BLOCK), 0, st, n, a, b, c, part);
  return (int)hipGetLastError();
}

template <typename R>
int launch_gn_update(long long n, const R* dU, R* U, double* part, int nparts, hipStream_t st) {
  hipLaunchKernelGGL(gn_update_kernel<R>, dim3(nparts), dim3(VDVC_BLOCK), 0, st, n, dU, U, part);
  return (int)hipGetLastError();
}

#define VDVC_INST(R)                                                                         \
  template int launch_hutch_acc<R>(long long, const int8_t*, const R*, double*, hipStream_t); \
  template int launch_expand_i8<R>(long long, const int8_t*, R*, hipStream_t);               \
  template int launch_hutch_final<R>(long long, const double*, int, R*, double*, hipStream_t); \
  template int launch_node_op<R>(const MeshDev&, const NodeOp<R>&, hipStream_t);             \
  template int launch_diag_op<R>(const MeshDev&, const DiagOp<R>&, hipStream_t);             \
  template int launch_jacobi<R>(const R*, const R*, int, long long, R*, hipStream_t);        \
  template int launch_pcg_init<R>(long long, const R*, const R*, R*, R*, R*, R*, double*,    \
                                  PcgState*, hipStream_t);                                   \
  template int launch_pcg_update<R>(long long, int, const double*, int, PcgState*, const R*, \
                                    const R*, const R*, R*, R*, R*, double*, bool,           \
                                    hipStream_t);                                            \
  template int launch_pcg_pupdate<R>(long long, int, double, const double*, PcgState*,       \
                                     const R*, R*, bool, hipStream_t);                       \
  template int launch_axpby<R>(long long, double, const R*, double, R*, hipStream_t);        \
  template int launch_dot<R>(long long, const R*, const R*, double*, int, hipStream_t);      \
  template int launch_gn_update<R>(long long, const R*, R*, double*, int, hipStream_t);      \
  template int launch_dot2<R>(long long, const R*, const R*, const R*, double*, int,          \
                              hipStream_t);                                                  \
  template int launch_calib_mode<R>(int, int, int, int, R*, hipStream_t);
VDVC_INST(float)
VDVC_INST(double)
#undef VDVC_INST

}  // namespace vdvc
