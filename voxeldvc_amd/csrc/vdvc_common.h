// Shared definitions for the gfx950 DVC kernels.
//
// Grid conventions (reference engine/kernels_dvc.py:9-21):
//   voxel (x,y,z), flat index x + y*Nvx + z*Nvx*Nvy, Nv* = N*·h + 1;
//   node (i,j,k), flat i + j*(Nx+1) + k*(Nx+1)*(Ny+1), DOF 3*node + c;
//   element (i,j,k) spans voxels [i*h, i*h+h] per axis.
//
// Unique-voxel form.  The reference sums over (element, stencil point) pairs
// with trapezoid weights w_q (engine/geometry_dvc.py:26-51).  A voxel on a
// shared element face appears once per element with weight 1/2 per shared
// axis, and the Hex8 shape functions agree there, so every element sum equals
// a sum over unique voxels with weight W(v) = prod_axis (1/2 if the voxel lies
// on the outer boundary plane of that axis, else 1) and global trilinear hat
// functions phi_node.  Each voxel is owned by element floor(coord / h) per
// axis (the reference's "highest element id wins" rule of
// engine/correlate_gpu.py:276-279); the last voxel plane of an axis is owned
// by the last element as local coordinate t = h.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <stdint.h>

#define VDVC_WAVE 64
#define VDVC_BLOCK 256
#define VDVC_WAVES_PER_BLOCK (VDVC_BLOCK / VDVC_WAVE)

// Voxel volumes of a mesh level are stored x fastest with a padded row
// pitch: index x + pitch * (y + Nvy * z).  The pitch is a multiple of 16
// voxels, so every row starts 16-byte aligned for every element type and
// each vectorised lane load (4-8 consecutive voxels) is naturally aligned.
// (A multiple of 8 would also keep the loads aligned and pad less, but
// measured on MI355X at 513^3 it streams 0.6-2% slower.)  Padding voxels
// are never read as data.
__host__ __device__ inline int vdvc_pitch_of(int nvx) { return (nvx + 15) & ~15; }

struct MeshDev {
  int Nx, Ny, Nz, h;
  int Nvx, Nvy, Nvz;
  int pitch;              // row pitch of voxel volumes (elements)
  int nnx, nny, nnz;      // nodes per axis
  long long Nvox;         // unique voxels
  long long Nvox_p;       // stored voxels (pitch * Nvy * Nvz)
  long long Nnodes;
  float inv_h;
};

// Voxel-tile decomposition shared by every streaming kernel.  One wavefront
// owns an x-strip of 64 voxel columns, one element row in y and ez elements in
// z; it walks y rows and z planes.  Node partial sums of a wave are written to
// a per-wave slot block P[w][kz_local][cy][ln][3] (ln < px) and summed by the
// node gather kernel in fixed order, so results are deterministic and no
// global atomics are used.
struct TilesDev {
  int vec;      // voxels per lane of the vectorised kernel (0: generic kernel)
  int sw;       // strip width in voxels (64 * vec, or 64 for the generic kernel)
  int nsx;      // x strips
  int nys;      // parts an element row's h rows are split into (vectorised kernel)
  int ez;       // elements per z tile
  int nzt;      // z tiles
  int px;       // node slots per strip (>= nodes touched by 64 voxels)
  int epl;      // elements per lane of the vectorised kernel: 1, or 2 (paired, V = 2h)
  long long nwaves;
  long long wave_stride;  // floats per wave in the partial buffer
};

enum { GT_F32 = 0, GT_BF16 = 1, GT_F16 = 2, GT_F64 = 3 };
// image element types of the post-solve stages (the images' own dtypes)
enum { GT_U8 = 4, GT_U16 = 5, GT_I16 = 6, GT_B8 = 7 };

// Arithmetic type of a storage type: double storage selects the FP64 path
// (every voxel and node computation in double, the reference's
// dtype=float64 solve); every other storage type computes in float.
template <typename T> struct RealOf { using type = float; };
template <> struct RealOf<double> { using type = double; };

// Correctly rounded single operations that are never fused with a
// neighbouring one (the reference's element-wise NumPy steps).  HIP's
// __fadd_rn / __fmul_rn compile to plain operations, which FP contraction
// fuses into FMAs; the pragma takes the contract flag off these.
#define VDVC_RN_OP(T, name, op)                                    \
  __device__ __forceinline__ T name(T a, T b) {                     \
    _Pragma("clang fp contract(off)") return a op b;                \
  }
VDVC_RN_OP(float, add_rn, +)
VDVC_RN_OP(float, sub_rn, -)
VDVC_RN_OP(float, mul_rn, *)
VDVC_RN_OP(double, add_rn, +)
VDVC_RN_OP(double, sub_rn, -)
VDVC_RN_OP(double, mul_rn, *)
#undef VDVC_RN_OP

__device__ __forceinline__ float rmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ double rmin(double a, double b) { return fmin(a, b); }

__device__ __forceinline__ double ld_val(const double* p, long long i) { return p[i]; }
__device__ __forceinline__ float ld_val(const float* p, long long i) { return p[i]; }
__device__ __forceinline__ float ld_val(const __hip_bfloat16* p, long long i) {
  // bf16 -> f32 is a 16-bit left shift of the raw bits
  const unsigned short* r = reinterpret_cast<const unsigned short*>(p);
  return __uint_as_float(((unsigned int)r[i]) << 16);
}
__device__ __forceinline__ float ld_val(const __half* p, long long i) { return __half2float(p[i]); }

__device__ __forceinline__ float axis_weight(int t, int n_last) {
  return (t == 0 || t == n_last) ? 0.5f : 1.0f;
}

// multiplicity of a voxel coordinate in the (element, stencil point)
// multiset: 2 on interior element faces, 1 elsewhere
// (engine/correlate_gpu.py:97-157).
__device__ __forceinline__ int axis_mult(int t, int h, int n_last) {
  return (t > 0 && t < n_last && (t % h) == 0) ? 2 : 1;
}

__device__ __forceinline__ long long node_index(const MeshDev& M, int i, int j, int k) {
  return (long long)i + (long long)M.nnx * ((long long)j + (long long)M.nny * k);
}

// map a linear block index so that consecutive work lands on one XCD: the
// dispatcher sends block b to XCD b % 8; this gives each XCD a contiguous run
// of tiles (neighbouring tiles share nodal data in that XCD's L2).
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned nblocks) {
  const unsigned nx = 8;
  unsigned per = nblocks / nx;
  unsigned full = per * nx;
  if (b >= full) return b;
  return (b % nx) * per + b / nx;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// block-wide FP64 sum of up to 4 values; returns valid results in thread 0
template <int NV>
__device__ __forceinline__ void block_sum_f64(double (&v)[NV], double* lds /* >= 4*NV */) {
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = wave_sum(v[q]);
  const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
  __syncthreads();
  if (ln == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) lds[wv * NV + q] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double s = 0.0;
      for (int k = 0; k < nw; ++k) s += lds[k * NV + q];
      v[q] = s;
    }
  }
}
