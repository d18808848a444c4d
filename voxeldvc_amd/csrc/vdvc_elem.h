// Element voxel blocks of the Hex8 mesh: element (i, j, k) covers voxels
// [i*h, i*h + h] x [j*h, j*h + h] x [k*h, k*h + h] of the (Nx*h+1) x
// (Ny*h+1) x (Nz*h+1) voxel grid (x fastest), faces shared with its
// neighbours.  Shared by the post-solve and element-scale output kernels.
#pragma once
#include "vdvc_common.h"

namespace vdvc_elem {

// flat index of element e's first voxel (elements x fastest)
__device__ __forceinline__ long long elem_base(long long e, int Nx, int Ny, int h, int Nvx,
                                               int Nvy) {
  const long long ex = e % Nx, ey = (e / Nx) % Ny, ez = e / ((long long)Nx * Ny);
  return ex * h + (long long)Nvx * (ey * h + (long long)Nvy * (ez * h));
}

// Visit the (h+1)^3 voxels of an element with one wavefront: lane -> (row,
// qx) once, rows (qy, qz) of h+1 voxels stepping by 64 / (h+1) rows; for
// h + 1 > 64 each lane walks voxels directly.
template <typename F>
__device__ __forceinline__ void walk(int h, int Nvx, int Nvy, long long base, int lane, F f) {
  const int h1 = h + 1;
  const int nrows = h1 * h1;
  if (h1 <= 64) {
    const int rpi = 64 / h1;
    const int lr = lane / h1, qx = lane - lr * h1;
    if (lr >= rpi) return;
    for (int r = lr; r < nrows; r += rpi) {
      const int qz = r / h1, qy = r - qz * h1;
      f(base + qx + (long long)Nvx * (qy + (long long)Nvy * qz));
    }
  } else {
    const int nq = nrows * h1;
    for (int q = lane; q < nq; q += 64) {
      const int qx = q % h1, qy = (q / h1) % h1, qz = q / nrows;
      f(base + qx + (long long)Nvx * (qy + (long long)Nvy * qz));
    }
  }
}

}  // namespace vdvc_elem
