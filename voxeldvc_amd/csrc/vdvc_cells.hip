// Element-scale write-output fields of the `run` stage:
//   * elem_count_kernel: per element, counts over its (h+1)^3 voxels (one
//     wavefront per element) -- inactive voxels, material voxels, material
//     voxels that lost overlap (engine/kernels_dvc.py:118-163, 236-340), or
//     unsafe voxels (engine/write_output.py:352-377);
//   * elem_flags_kernel: unsafe / overlap-lost elements and unsafe nodes.
//     The reference dilates through the 8 corner nodes (an element is
//     flagged when it shares a node with a defective one); two elements share
//     a node exactly when their indices differ by at most 1 on every axis, so
//     the dilation is a 27-neighbourhood test on the counts;
//   * unsafe_voxel_kernel: a voxel is unsafe when an element holding it is
//     (kernels_dvc.py:166-195); voxel v of an axis lies in elements v / h and,
//     on an element face (v % h == 0), v / h - 1;
//   * strain_kernel: grad u at the element centre from the 8 corner nodes
//     (Hex8 gradients +-0.25 / h), its symmetric part, eigenvalues by cyclic
//     Jacobi rotations in float64, descending, and the von Mises equivalent
//     strain (write_output.py:256-306);
//   * gn_convergence_kernel: ||dU|| / ||U|| over the element's 24 corner
//     DOFs in NumPy's summation order (write_output.py:450-491);
//   * elem_centres_kernel + the nn_* kernels: the nearest-neighbour remap of
//     an element mask to the deformed configuration (write_output.py:
//     494-530): element centres moved by the mean corner displacement, and
//     for every regular centre (+ crop offset) the value of the nearest
//     moved centre, found exactly through a uniform grid of cells (ties:
//     the lowest element index).
// Elements x fastest (e = i + Nx*(j + Ny*k)), nodes n = i + (Nx+1)*(j +
// (Ny+1)*k), voxels x fastest on the (Nx*h+1) x (Ny*h+1) x (Nz*h+1) grid.
#include "vdvc_common.h"
#include "vdvc_kernels.h"
#include "vdvc_elem.h"
#include "vdvc_sample01.h"

namespace {

using vdvc_img::ld_img;

// HASB: a second mask is given (fixed at compile time, so that no branch
// separates the walk's loads)
template <bool HASB>
__global__ __launch_bounds__(VDVC_BLOCK) void elem_count_kernel(CellIn C) {
  const int lane = threadIdx.x & 63;
  const long long e = (long long)blockIdx.x * VDVC_WAVES_PER_BLOCK + (threadIdx.x >> 6);
  const long long Ne = (long long)C.Nx * C.Ny * C.Nz;
  if (e >= Ne) return;  // whole wavefronts
  const long long base = vdvc_elem::elem_base(e, C.Nx, C.Ny, C.h, C.Nvx, C.Nvy);
  int n0 = 0, n1 = 0, n2 = 0;
  vdvc_elem::walk(C.h, C.Nvx, C.Nvy, base, lane, [&](long long vi) {
    const bool a0 = C.a[vi] == 0;
    const bool b1 = HASB ? C.b[vi] != 0 : true;
    n0 += a0 ? 1 : 0;
    n1 += b1 ? 1 : 0;
    n2 += (a0 && b1) ? 1 : 0;
  });
  n0 = wave_sum(n0);
  n1 = wave_sum(n1);
  n2 = wave_sum(n2);
  if (lane == 0) {
    if (C.c0) C.c0[e] = n0;
    if (C.c1) C.c1[e] = n1;
    if (C.c2) C.c2[e] = n2;
    if (C.all_a0) C.all_a0[e] = n0 == (C.h + 1) * (C.h + 1) * (C.h + 1) ? 1 : 0;
  }
}

// The same counts with one block per element row (ey, ez): the block's four
// wavefronts take every fourth voxel row (qy, qz) of the row's (h+1)^2, a
// lane takes element columns lane, lane + 64, ... and reads its element's
// h+1 voxels of the row, so each voxel row is read by one block with
// neighbouring lanes on neighbouring bytes (the per-element kernel above
// fetches a cache line once per element it spans).  Integer totals are
// combined with LDS atomics (exact, so the result is order-independent).
// Dynamic LDS: 3 * Nx ints.
template <bool HASB>
__global__ __launch_bounds__(VDVC_BLOCK) void elem_row_count_kernel(CellIn C) {
  extern __shared__ int rc_lds[];
  const int Nx = C.Nx, h = C.h, h1 = C.h + 1;
  const int ey = (int)(blockIdx.x % C.Ny), ez = (int)(blockIdx.x / C.Ny);
  const int lane = threadIdx.x & 63, grp = threadIdx.x >> 6;
  for (int q = threadIdx.x; q < 3 * Nx; q += VDVC_BLOCK) rc_lds[q] = 0;
  __syncthreads();
  const int nrows = h1 * h1;
  for (int ex = lane; ex < Nx; ex += 64) {
    int n0 = 0, n1 = 0, n2 = 0;
    for (int r = grp; r < nrows; r += VDVC_WAVES_PER_BLOCK) {
      const int qz = r / h1, qy = r - qz * h1;
      const long long vb = (long long)ex * h +
                           (long long)C.Nvx * ((long long)(ey * h + qy) + (long long)C.Nvy * (ez * h + qz));
      for (int x = 0; x < h1; ++x) {
        const bool a0 = C.a[vb + x] == 0;
        const bool b1 = HASB ? C.b[vb + x] != 0 : true;
        n0 += a0 ? 1 : 0;
        n1 += b1 ? 1 : 0;
        n2 += (a0 && b1) ? 1 : 0;
      }
    }
    atomicAdd(&rc_lds[ex], n0);
    atomicAdd(&rc_lds[Nx + ex], n1);
    atomicAdd(&rc_lds[2 * Nx + ex], n2);
  }
  __syncthreads();
  const long long e0 = (long long)Nx * ((long long)ey + (long long)C.Ny * ez);
  for (int ex = threadIdx.x; ex < Nx; ex += VDVC_BLOCK) {
    const int n0 = rc_lds[ex], n1 = rc_lds[Nx + ex], n2 = rc_lds[2 * Nx + ex];
    if (C.c0) C.c0[e0 + ex] = n0;
    if (C.c1) C.c1[e0 + ex] = n1;
    if (C.c2) C.c2[e0 + ex] = n2;
    if (C.all_a0) C.all_a0[e0 + ex] = n0 == h1 * h1 * h1 ? 1 : 0;
  }
}

__device__ __forceinline__ bool on_shell(int i, int j, int k, int Nx, int Ny, int Nz) {
  return i == 0 || i == Nx - 1 || j == 0 || j == Ny - 1 || k == 0 || k == Nz - 1;
}

// material lost overlap: c1 > 0 and c2 / c1 > frac_threshold (float64)
__device__ __forceinline__ bool elem_lost(const CellFlagsIn& F, long long e) {
  const int m = F.c1[e];
  return m > 0 && (double)F.c2[e] / (double)m > F.frac_threshold;
}

__global__ __launch_bounds__(VDVC_BLOCK) void elem_flags_kernel(CellFlagsIn F) {
  const long long e = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  const long long Ne = (long long)F.Nx * F.Ny * F.Nz;
  if (e >= Ne) return;
  const int i = (int)(e % F.Nx), j = (int)((e / F.Nx) % F.Ny), k = (int)(e / ((long long)F.Nx * F.Ny));
  const bool shell = on_shell(i, j, k, F.Nx, F.Ny, F.Nz);
  bool inact = false, lost = false;
  for (int dk = -1; dk <= 1; ++dk) {
    const int kk = k + dk;
    if (kk < 0 || kk >= F.Nz) continue;
    for (int dj = -1; dj <= 1; ++dj) {
      const int jj = j + dj;
      if (jj < 0 || jj >= F.Ny) continue;
      for (int di = -1; di <= 1; ++di) {
        const int ii = i + di;
        if (ii < 0 || ii >= F.Nx) continue;
        const long long n = ii + (long long)F.Nx * (jj + (long long)F.Ny * kk);
        if (F.unsafe_elem) inact = inact || F.c0[n] > 0;
        if (F.lost_elem) lost = lost || elem_lost(F, n);
      }
    }
  }
  if (F.unsafe_elem) F.unsafe_elem[e] = (shell || inact) ? 1 : 0;
  if (F.lost_elem) {
    const double nper = (double)((F.h + 1) * (F.h + 1) * (F.h + 1));
    const bool sparse = (double)F.c1[e] / nper < F.min_material_frac;
    F.lost_elem[e] = (lost || sparse || shell) ? 1 : 0;
  }
}

// unsafe nodes: an adjacent element holds an inactive voxel or is on the shell
__global__ __launch_bounds__(VDVC_BLOCK) void node_flags_kernel(CellFlagsIn F) {
  const long long n = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  const int nnx = F.Nx + 1, nny = F.Ny + 1;
  const long long Nn = (long long)nnx * nny * (F.Nz + 1);
  if (n >= Nn) return;
  const int i = (int)(n % nnx), j = (int)((n / nnx) % nny), k = (int)(n / ((long long)nnx * nny));
  bool u = false;
  for (int ek = k - 1; ek <= k; ++ek) {
    if (ek < 0 || ek >= F.Nz) continue;
    for (int ej = j - 1; ej <= j; ++ej) {
      if (ej < 0 || ej >= F.Ny) continue;
      for (int ei = i - 1; ei <= i; ++ei) {
        if (ei < 0 || ei >= F.Nx) continue;
        const long long e = ei + (long long)F.Nx * (ej + (long long)F.Ny * ek);
        u = u || F.c0[e] > 0 || on_shell(ei, ej, ek, F.Nx, F.Ny, F.Nz);
      }
    }
  }
  F.unsafe_node[n] = u ? 1 : 0;
}

// the one or two elements of an axis holding voxel v: [lo, hi]
__device__ __forceinline__ void axis_elems(int v, int h, int N, int& lo, int& hi) {
  const int q = v / h;
  hi = q < N - 1 ? q : N - 1;
  lo = (v - q * h == 0 && q > 0) ? q - 1 : hi;
}

// One wavefront per voxel row (y, z): the flags of the (at most 2 x 2)
// element rows holding the row are OR-ed per element column into LDS, then
// each voxel ORs the columns of its (at most two) elements.  Dynamic LDS:
// Nx bytes per wavefront.
__global__ __launch_bounds__(VDVC_BLOCK) void unsafe_voxel_kernel(int Nx, int Ny, int Nz, int h,
                                                                  const uint8_t* __restrict__ ue,
                                                                  uint8_t* __restrict__ out) {
  extern __shared__ uint8_t colf_all[];
  const int Nvx = Nx * h + 1, Nvy = Ny * h + 1, Nvz = Nz * h + 1;
  const int lane = threadIdx.x & 63, wib = threadIdx.x >> 6;
  uint8_t* colf = colf_all + (size_t)wib * Nx;
  const long long row = (long long)blockIdx.x * VDVC_WAVES_PER_BLOCK + wib;
  const bool live = row < (long long)Nvy * Nvz;  // uniform per wavefront
  int y = 0, z = 0, y0 = 0, y1 = 0, z0 = 0, z1 = 0;
  if (live) {
    y = (int)(row % Nvy);
    z = (int)(row / Nvy);
    axis_elems(y, h, Ny, y0, y1);
    axis_elems(z, h, Nz, z0, z1);
    for (int i = lane; i < Nx; i += 64) {
      const uint8_t a = ue[i + (long long)Nx * (y0 + (long long)Ny * z0)];
      const uint8_t b = ue[i + (long long)Nx * (y1 + (long long)Ny * z0)];
      const uint8_t c = ue[i + (long long)Nx * (y0 + (long long)Ny * z1)];
      const uint8_t d = ue[i + (long long)Nx * (y1 + (long long)Ny * z1)];
      colf[i] = (a | b | c | d) != 0 ? 1 : 0;
    }
  }
  __syncthreads();
  if (!live) return;
  uint8_t* orow = out + (long long)Nvx * ((long long)y + (long long)Nvy * z);
  for (int x = lane; x < Nvx; x += 64) {
    int x0, x1;
    axis_elems(x, h, Nx, x0, x1);
    orow[x] = colf[x0] | colf[x1];
  }
}

// one Jacobi rotation zeroing A[p][q] of a symmetric 3x3 matrix (r the
// third index)
template <int p, int q, int r>
__device__ __forceinline__ void jacobi_rot(double (&A)[3][3]) {
  const double apq = A[p][q];
  if (apq == 0.0) return;
  const double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
  double t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
  if (theta < 0.0) t = -t;
  const double c = 1.0 / sqrt(t * t + 1.0), s = t * c, tau = s / (1.0 + c);
  A[p][p] -= t * apq;
  A[q][q] += t * apq;
  A[p][q] = A[q][p] = 0.0;
  const double arp = A[r][p], arq = A[r][q];
  A[r][p] = A[p][r] = arp - s * (arq + arp * tau);
  A[r][q] = A[q][r] = arq + s * (arp - arq * tau);
}

// eigenvalues of a symmetric 3x3 matrix, descending
__device__ __forceinline__ void sym3_eigvals(double (&A)[3][3], double (&e)[3]) {
  for (int sweep = 0; sweep < 32; ++sweep) {
    const double off = fabs(A[0][1]) + fabs(A[0][2]) + fabs(A[1][2]);
    const double dia = fabs(A[0][0]) + fabs(A[1][1]) + fabs(A[2][2]);
    if (off == 0.0 || off <= 1e-19 * dia) break;
    jacobi_rot<0, 1, 2>(A);
    jacobi_rot<0, 2, 1>(A);
    jacobi_rot<1, 2, 0>(A);
  }
  double a = A[0][0], b = A[1][1], c = A[2][2], t;
  if (a < b) { t = a; a = b; b = t; }
  if (b < c) { t = b; b = c; c = t; }
  if (a < b) { t = a; a = b; b = t; }
  e[0] = a; e[1] = b; e[2] = c;
}

// sqrt(2)/3 * sqrt(((e1-e2)^2 + (e2-e3)^2) + (e3-e1)^2), NumPy's order
__device__ __forceinline__ double equiv_of(const double (&e)[3], double k_eq) {
  const double d12 = sub_rn(e[0], e[1]), d23 = sub_rn(e[1], e[2]), d31 = sub_rn(e[2], e[0]);
  const double s = add_rn(add_rn(mul_rn(d12, d12), mul_rn(d23, d23)), mul_rn(d31, d31));
  return mul_rn(k_eq, sqrt(s));
}

// Hex8 corner a = 0..7: (x, y, z) offsets of the reference's corner order
__device__ __forceinline__ constexpr int hx(int a) { return (a == 1 || a == 2 || a == 5 || a == 6) ? 1 : 0; }
__device__ __forceinline__ constexpr int hy(int a) { return (a == 2 || a == 3 || a == 6 || a == 7) ? 1 : 0; }
__device__ __forceinline__ constexpr int hz(int a) { return a >= 4 ? 1 : 0; }

__global__ __launch_bounds__(VDVC_BLOCK) void strain_kernel(StrainIn S) {
  const long long e = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  const long long Ne = (long long)S.Nx * S.Ny * S.Nz;
  if (e >= Ne) return;
  const int i = (int)(e % S.Nx), j = (int)((e / S.Nx) % S.Ny), k = (int)(e / ((long long)S.Nx * S.Ny));
  const int nnx = S.Nx + 1, nny = S.Ny + 1;
  // grad[c][d] = sum_a U_a,c dN_a/dx_d with dN_a/dx_d = +-g (+ on the
  // corner's upper face of axis d)
  double G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const long long n = (i + hx(a)) + (long long)nnx * ((j + hy(a)) + (long long)nny * (k + hz(a)));
    const double sd[3] = {hx(a) ? S.g : -S.g, hy(a) ? S.g : -S.g, hz(a) ? S.g : -S.g};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double u = ld_img(S.U, S.ucode, 3 * n + c);
#pragma unroll
      for (int d = 0; d < 3; ++d) G[c][d] = add_rn(G[c][d], mul_rn(u, sd[d]));
    }
  }
  double A[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int d = 0; d < 3; ++d) A[c][d] = mul_rn(0.5, add_rn(G[c][d], G[d][c]));
  double ev[3];
  sym3_eigvals(A, ev);
#pragma unroll
  for (int c = 0; c < 3; ++c) S.principal[3 * e + c] = ev[c];
  if (S.equiv) S.equiv[e] = equiv_of(ev, S.k_eq);
}

__global__ __launch_bounds__(VDVC_BLOCK) void equiv_kernel(long long n,
                                                           const double* __restrict__ p,
                                                           double k_eq, double* __restrict__ out) {
  const long long i = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  if (i >= n) return;
  const double ev[3] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
  out[i] = equiv_of(ev, k_eq);
}

// sum over the 8 corners (Hex8 order) of ((u0^2 + u1^2) + u2^2), float64
__device__ __forceinline__ double corner_sq(const void* U, int code, int i, int j, int k, int nnx,
                                            int nny) {
  double acc = 0.0;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const long long n = (i + hx(a)) + (long long)nnx * ((j + hy(a)) + (long long)nny * (k + hz(a)));
    const double u0 = ld_img(U, code, 3 * n), u1 = ld_img(U, code, 3 * n + 1),
                 u2 = ld_img(U, code, 3 * n + 2);
    acc = add_rn(acc, add_rn(add_rn(mul_rn(u0, u0), mul_rn(u1, u1)), mul_rn(u2, u2)));
  }
  return acc;
}

__global__ __launch_bounds__(VDVC_BLOCK) void gn_convergence_kernel(int Nx, int Ny, int Nz,
                                                                    const void* dU, int dcode,
                                                                    const void* U, int ucode,
                                                                    double* out) {
  const long long e = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  const long long Ne = (long long)Nx * Ny * Nz;
  if (e >= Ne) return;
  const int i = (int)(e % Nx), j = (int)((e / Nx) % Ny), k = (int)(e / ((long long)Nx * Ny));
  const double num = sqrt(corner_sq(dU, dcode, i, j, k, Nx + 1, Ny + 1));
  const double den = sqrt(corner_sq(U, ucode, i, j, k, Nx + 1, Ny + 1));
  out[e] = den <= 0.0 ? __longlong_as_double(0x7ff8000000000000LL) : num / den;
}

// moved centres pts = (i + 0.5) h + mean corner displacement (the corners
// summed with the z offset fastest, then / 8) and regular centres
// qry = (i + 0.5) h + crop
__global__ __launch_bounds__(VDVC_BLOCK) void elem_centres_kernel(int Nx, int Ny, int Nz, int h,
                                                                  const void* U, int ucode,
                                                                  double c0, double c1, double c2,
                                                                  double* pts, double* qry) {
  const long long e = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  const long long Ne = (long long)Nx * Ny * Nz;
  if (e >= Ne) return;
  const int ijk[3] = {(int)(e % Nx), (int)((e / Nx) % Ny), (int)(e / ((long long)Nx * Ny))};
  const int nnx = Nx + 1, nny = Ny + 1;
  double u[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int di = 0; di < 2; ++di)
#pragma unroll
    for (int dj = 0; dj < 2; ++dj)
#pragma unroll
      for (int dk = 0; dk < 2; ++dk) {
        const long long n = (ijk[0] + di) + (long long)nnx * ((ijk[1] + dj) + (long long)nny * (ijk[2] + dk));
#pragma unroll
        for (int c = 0; c < 3; ++c) u[c] = add_rn(u[c], ld_img(U, ucode, 3 * n + c));
      }
  const double crop[3] = {c0, c1, c2};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double xr = mul_rn(add_rn((double)ijk[c], 0.5), (double)h);
    pts[3 * e + c] = add_rn(xr, u[c] / 8.0);
    qry[3 * e + c] = add_rn(xr, crop[c]);
  }
}

// Nearest moved centre per query on a uniform grid of cells (edge cs) over
// the points' bounding box: counting sort of the point indices by cell
// (integer atomics; the order inside a cell does not matter), then each
// query walks Chebyshev rings of cells around its own cell.  After rings
// 0..r every unvisited point lies at least r * cs away, so the walk stops
// once the best squared distance is below (r * cs)^2 (with a 1e-9 relative
// margin for the cell assignment's rounding).  Points and queries outside
// the box are clamped into the border cells, which keeps that bound.  The
// squared distance is ((d0^2 + d1^2) + d2^2) and ties go to the lowest
// index, so the result equals an exhaustive search.
__device__ __forceinline__ int nn_cell1(double p, double o, double inv_cs, int G) {
  const double f = floor((p - o) * inv_cs);
  return !(f >= 0.0) ? 0 : (f > (double)(G - 1) ? G - 1 : (int)f);
}

// min / max that propagate NaN (a non-finite point makes the box non-finite,
// which the host rejects)
__device__ __forceinline__ double nmin(double a, double b) { return (a != a || b != b) ? a + b : fmin(a, b); }
__device__ __forceinline__ double nmax(double a, double b) { return (a != a || b != b) ? a + b : fmax(a, b); }

__global__ __launch_bounds__(VDVC_BLOCK) void bbox_part_kernel(long long n,
                                                               const double* __restrict__ pts,
                                                               double* __restrict__ part) {
  __shared__ double lds[6][VDVC_BLOCK];
  double v[6] = {__longlong_as_double(0x7ff0000000000000LL), __longlong_as_double(0x7ff0000000000000LL),
                 __longlong_as_double(0x7ff0000000000000LL), -__longlong_as_double(0x7ff0000000000000LL),
                 -__longlong_as_double(0x7ff0000000000000LL), -__longlong_as_double(0x7ff0000000000000LL)};
  for (long long i = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x; i < n;
       i += (long long)gridDim.x * VDVC_BLOCK) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double x = pts[3 * i + c];
      v[c] = nmin(v[c], x);
      v[3 + c] = nmax(v[3 + c], x);
    }
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) lds[c][threadIdx.x] = v[c];
  __syncthreads();
  for (int o = VDVC_BLOCK / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        lds[c][threadIdx.x] = nmin(lds[c][threadIdx.x], lds[c][threadIdx.x + o]);
        lds[3 + c][threadIdx.x] = nmax(lds[3 + c][threadIdx.x], lds[3 + c][threadIdx.x + o]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < 6; ++c) part[6 * blockIdx.x + c] = lds[c][0];
  }
}

__global__ __launch_bounds__(VDVC_BLOCK) void bbox_final_kernel(int nparts,
                                                                const double* __restrict__ part,
                                                                double* __restrict__ out) {
  __shared__ double lds[6][VDVC_BLOCK];
  const double inf = __longlong_as_double(0x7ff0000000000000LL);
#pragma unroll
  for (int c = 0; c < 6; ++c)
    lds[c][threadIdx.x] = threadIdx.x < nparts ? part[6 * threadIdx.x + c] : (c < 3 ? inf : -inf);
  __syncthreads();
  for (int o = VDVC_BLOCK / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        lds[c][threadIdx.x] = nmin(lds[c][threadIdx.x], lds[c][threadIdx.x + o]);
        lds[3 + c][threadIdx.x] = nmax(lds[3 + c][threadIdx.x], lds[3 + c][threadIdx.x + o]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) out[threadIdx.x] = lds[threadIdx.x][0];
}

__global__ __launch_bounds__(VDVC_BLOCK) void nn_count_kernel(long long n,
                                                              const double* __restrict__ pts,
                                                              NNGrid g, int* __restrict__ cnt) {
  const long long i = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  if (i >= n) return;
  const int cx = nn_cell1(pts[3 * i], g.o[0], g.inv_cs, g.G[0]);
  const int cy = nn_cell1(pts[3 * i + 1], g.o[1], g.inv_cs, g.G[1]);
  const int cz = nn_cell1(pts[3 * i + 2], g.o[2], g.inv_cs, g.G[2]);
  atomicAdd(&cnt[cx + (long long)g.G[0] * (cy + (long long)g.G[1] * cz)], 1);
}

// Exclusive prefix sum of the cell counts in three passes: per-chunk sums
// (chunks of kScanChunk cells, 8 consecutive cells per thread), a scan of the
// chunk sums in one block (also writing off[ncell] = n), and per chunk the
// local scan plus the chunk's offset.  off[c] = first slot of cell c; the
// counts are replaced by the same offsets (the fill cursors).
constexpr int kScanItems = 8;
constexpr int kScanChunk = VDVC_BLOCK * kScanItems;

__device__ __forceinline__ int block_excl_scan(int v, int& total) {
  __shared__ int wsum[VDVC_WAVES_PER_BLOCK];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  int base = 0;
  total = 0;
#pragma unroll
  for (int k = 0; k < VDVC_WAVES_PER_BLOCK; ++k) {
    if (k < w) base += wsum[k];
    total += wsum[k];
  }
  __syncthreads();
  return base + x - v;
}

__global__ __launch_bounds__(VDVC_BLOCK) void nn_chunk_sum_kernel(long long ncell,
                                                                  const int* __restrict__ cnt,
                                                                  int* __restrict__ csum) {
  const long long b0 = (long long)blockIdx.x * kScanChunk + (long long)threadIdx.x * kScanItems;
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (b0 + k < ncell) s += cnt[b0 + k];
  int total;
  block_excl_scan(s, total);
  if (threadIdx.x == 0) csum[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void nn_chunk_scan_kernel(int nchunk, int* __restrict__ csum,
                                                             long long ncell,
                                                             int* __restrict__ off) {
  __shared__ int lds[1024];
  const int t = threadIdx.x;
  const int per = (nchunk + 1023) / 1024;
  const int b = t * per, e = b + per < nchunk ? b + per : nchunk;
  int s = 0;
  for (int i = b; i < e; ++i) s += csum[i];
  lds[t] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int y = t >= o ? lds[t - o] : 0;
    __syncthreads();
    lds[t] += y;
    __syncthreads();
  }
  int acc = lds[t] - s;
  for (int i = b; i < e; ++i) {
    const int c = csum[i];
    csum[i] = acc;
    acc += c;
  }
  if (t == 1023) off[ncell] = lds[1023];
}

__global__ __launch_bounds__(VDVC_BLOCK) void nn_chunk_apply_kernel(long long ncell,
                                                                    int* __restrict__ cnt,
                                                                    const int* __restrict__ csum,
                                                                    int* __restrict__ off) {
  const long long b0 = (long long)blockIdx.x * kScanChunk + (long long)threadIdx.x * kScanItems;
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = b0 + k < ncell ? cnt[b0 + k] : 0;
    s += v[k];
  }
  int total;
  int acc = csum[blockIdx.x] + block_excl_scan(s, total);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (b0 + k < ncell) {
      off[b0 + k] = acc;
      cnt[b0 + k] = acc;
    }
    acc += v[k];
  }
}

__global__ __launch_bounds__(VDVC_BLOCK) void nn_fill_kernel(long long n,
                                                             const double* __restrict__ pts,
                                                             NNGrid g, int* __restrict__ cur,
                                                             int* __restrict__ list) {
  const long long i = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  if (i >= n) return;
  const int cx = nn_cell1(pts[3 * i], g.o[0], g.inv_cs, g.G[0]);
  const int cy = nn_cell1(pts[3 * i + 1], g.o[1], g.inv_cs, g.G[1]);
  const int cz = nn_cell1(pts[3 * i + 2], g.o[2], g.inv_cs, g.G[2]);
  const int slot = atomicAdd(&cur[cx + (long long)g.G[0] * (cy + (long long)g.G[1] * cz)], 1);
  list[slot] = (int)i;
}

__global__ __launch_bounds__(VDVC_BLOCK) void nn_query_kernel(long long n,
                                                              const double* __restrict__ pts,
                                                              const double* __restrict__ qry,
                                                              NNGrid g, const int* __restrict__ off,
                                                              const int* __restrict__ list,
                                                              const uint8_t* __restrict__ vals,
                                                              uint8_t* __restrict__ out,
                                                              int* __restrict__ idx_out) {
  const long long qi = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  if (qi >= n) return;
  const double q0 = qry[3 * qi], q1 = qry[3 * qi + 1], q2 = qry[3 * qi + 2];
  const int c0 = nn_cell1(q0, g.o[0], g.inv_cs, g.G[0]);
  const int c1 = nn_cell1(q1, g.o[1], g.inv_cs, g.G[1]);
  const int c2 = nn_cell1(q2, g.o[2], g.inv_cs, g.G[2]);
  int rmax = c0 > g.G[0] - 1 - c0 ? c0 : g.G[0] - 1 - c0;
  rmax = max(rmax, max(c1, g.G[1] - 1 - c1));
  rmax = max(rmax, max(c2, g.G[2] - 1 - c2));
  double best = __longlong_as_double(0x7ff0000000000000LL);
  int bi = -1;
  for (int r = 0; r <= rmax; ++r) {
    for (int dz = -r; dz <= r; ++dz) {
      const int z = c2 + dz;
      if (z < 0 || z >= g.G[2]) continue;
      for (int dy = -r; dy <= r; ++dy) {
        const int y = c1 + dy;
        if (y < 0 || y >= g.G[1]) continue;
        const bool face = dz == -r || dz == r || dy == -r || dy == r;
        const int step = (face || r == 0) ? 1 : 2 * r;
        for (int dx = -r; dx <= r; dx += step) {
          const int x = c0 + dx;
          if (x < 0 || x >= g.G[0]) continue;
          const long long c = x + (long long)g.G[0] * (y + (long long)g.G[1] * z);
          for (int k = off[c]; k < off[c + 1]; ++k) {
            const int i = list[k];
            const double d0 = sub_rn(q0, pts[3 * (long long)i]);
            const double d1 = sub_rn(q1, pts[3 * (long long)i + 1]);
            const double d2 = sub_rn(q2, pts[3 * (long long)i + 2]);
            const double dd = add_rn(add_rn(mul_rn(d0, d0), mul_rn(d1, d1)), mul_rn(d2, d2));
            if (dd < best || (dd == best && i < bi)) {
              best = dd;
              bi = i;
            }
          }
        }
      }
    }
    const double lim = (double)r * g.cs * (1.0 - 1e-9);
    if (bi >= 0 && best < lim * lim) break;
  }
  if (out) out[qi] = bi >= 0 ? vals[bi] : (uint8_t)0;
  if (idx_out) idx_out[qi] = bi;
}


// Material cells (engine/write_output.py:309-338): the mean reference gray
// level over an element's (h+1)^3 voxels (shared faces included) exceeds
// glt.  One wavefront per element, lanes over the block's (h+1)^2 voxel
// rows, float64 sums (exact for the integer image types, whose block sums
// stay below 2^53); mean = sum / (h+1)^3 as NumPy's mean forms it.
__global__ __launch_bounds__(VDVC_BLOCK) void material_cells_kernel(int Nx, int Ny, int Nz, int h,
                                                                    const void* img, int itype,
                                                                    double glt,
                                                                    uint8_t* __restrict__ out) {
  const long long e = ((long long)blockIdx.x * VDVC_BLOCK + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const long long Ne = (long long)Nx * Ny * Nz;
  if (e >= Ne) return;  // uniform across the wavefront
  const int i = (int)(e % Nx), j = (int)((e / Nx) % Ny), k = (int)(e / ((long long)Nx * Ny));
  const long long Nvx = (long long)Nx * h + 1, Nvy = (long long)Ny * h + 1;
  const int m = h + 1;
  double s = 0.0;
  for (int r = lane; r < m * m; r += 64) {
    const int dy = r % m, dz = r / m;
    const long long base = (long long)i * h + Nvx * ((long long)j * h + dy + Nvy * ((long long)k * h + dz));
    for (int dx = 0; dx < m; ++dx) s += ld_img(img, itype, base + dx);
  }
  s = wave_sum(s);
  if (lane == 0) out[e] = (s / (double)((long long)m * m * m) > glt) ? 1 : 0;
}

// One-element erosion of an element mask over the 26-neighbourhood with the
// mesh border counted as set (scipy binary_erosion with a 3x3x3 structure
// and border_value=1, engine/write_output.py:341-349).
__global__ __launch_bounds__(VDVC_BLOCK) void erode_cells_kernel(int Nx, int Ny, int Nz,
                                                                 const uint8_t* __restrict__ in,
                                                                 uint8_t* __restrict__ out) {
  const long long e = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  const long long Ne = (long long)Nx * Ny * Nz;
  if (e >= Ne) return;
  const int i = (int)(e % Nx), j = (int)((e / Nx) % Ny), k = (int)(e / ((long long)Nx * Ny));
  uint8_t v = in[e] ? 1 : 0;
  for (int dz = -1; dz <= 1 && v; ++dz)
    for (int dy = -1; dy <= 1 && v; ++dy)
      for (int dx = -1; dx <= 1 && v; ++dx) {
        const int a = i + dx, b = j + dy, c = k + dz;
        if (a < 0 || b < 0 || c < 0 || a >= Nx || b >= Ny || c >= Nz) continue;
        if (!in[a + (long long)Nx * (b + (long long)Ny * c)]) v = 0;
      }
  out[e] = v;
}

// Barycentric coordinates of q in the tetrahedron (v0, v1, v2, v3) by
// Cramer's rule; false for a degenerate tetrahedron.
__device__ __forceinline__ bool bary4(const double (&q)[3], const double (&v0)[3],
                                      const double (&v1)[3], const double (&v2)[3],
                                      const double (&v3)[3], double (&l)[4]) {
  double a[3], b[3], c[3], d[3];
  for (int t = 0; t < 3; ++t) {
    a[t] = v1[t] - v0[t]; b[t] = v2[t] - v0[t]; c[t] = v3[t] - v0[t]; d[t] = q[t] - v0[t];
  }
  const double det = a[0] * (b[1] * c[2] - b[2] * c[1]) - b[0] * (a[1] * c[2] - a[2] * c[1]) +
                     c[0] * (a[1] * b[2] - a[2] * b[1]);
  if (!(fabs(det) > 1e-300)) return false;
  const double l1 = (d[0] * (b[1] * c[2] - b[2] * c[1]) - b[0] * (d[1] * c[2] - d[2] * c[1]) +
                     c[0] * (d[1] * b[2] - d[2] * b[1])) / det;
  const double l2 = (a[0] * (d[1] * c[2] - d[2] * c[1]) - d[0] * (a[1] * c[2] - a[2] * c[1]) +
                     c[0] * (a[1] * d[2] - a[2] * d[1])) / det;
  const double l3 = (a[0] * (b[1] * d[2] - b[2] * d[1]) - b[0] * (a[1] * d[2] - a[2] * d[1]) +
                     d[0] * (a[1] * b[2] - a[2] * b[1])) / det;
  l[0] = 1.0 - l1 - l2 - l3; l[1] = l1; l[2] = l2; l[3] = l3;
  return true;
}

// Linear remap of a per-element field to the deformed configuration
// (engine/write_output.py:533-568): the moved element centres pts form a
// deformed lattice; each lattice cell (8 neighbouring centres) is split
// into the 6 Kuhn tetrahedra along its (0,0,0)-(1,1,1) diagonal (a
// conforming split of the whole lattice), and the query takes the
// barycentric interpolation of the field over the tetrahedron containing
// it.  The cell is found by a walk from the reference cell of the query,
// moved by the displacement of its nearest reference centre: each step
// maps the query into the cell's mean affine frame and moves by the whole
// cells it lies beyond.  A query in no tetrahedron within 64 steps (outside
// the moved lattice) takes the value of its nearest moved centre (nn_idx),
// the reference's nearest fill.
__global__ __launch_bounds__(VDVC_BLOCK) void remap_linear_kernel(int Nx, int Ny, int Nz, double h,
                                                                  const double* __restrict__ pts,
                                                                  const double* __restrict__ qry,
                                                                  const double* __restrict__ field,
                                                                  int C,
                                                                  const int* __restrict__ nn_idx,
                                                                  double* __restrict__ out) {
  const long long qi = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  const long long Ne = (long long)Nx * Ny * Nz;
  if (qi >= Ne) return;
  const int N[3] = {Nx, Ny, Nz};
  const double q[3] = {qry[3 * qi], qry[3 * qi + 1], qry[3 * qi + 2]};
  auto pidx = [&](int a, int b, int c) { return (long long)a + (long long)Nx * (b + (long long)Ny * c); };
  bool found = false;
  long long vid[4] = {0, 0, 0, 0};
  double lam[4] = {0.0, 0.0, 0.0, 0.0};
  if (Nx >= 2 && Ny >= 2 && Nz >= 2) {
    int nr[3], cell[3];
    for (int d = 0; d < 3; ++d) {
      const double f = rint(q[d] / h - 0.5);
      nr[d] = (int)fmin(fmax(f, 0.0), (double)(N[d] - 1));
    }
    const long long n0 = pidx(nr[0], nr[1], nr[2]);
    for (int d = 0; d < 3; ++d) {
      const double u0 = pts[3 * n0 + d] - ((double)nr[d] + 0.5) * h;
      const double sd = floor((q[d] - u0) / h - 0.5);
      cell[d] = (int)fmin(fmax(sd, 0.0), (double)(N[d] - 2));
    }
    for (int it = 0; it < 64 && !found; ++it) {
      double V[8][3];
      for (int cc = 0; cc < 8; ++cc) {
        const long long n = pidx(cell[0] + (cc & 1), cell[1] + ((cc >> 1) & 1), cell[2] + (cc >> 2));
        for (int d = 0; d < 3; ++d) V[cc][d] = pts[3 * n + d];
      }
      // the 6 Kuhn tetrahedra: corner 000, + e_p0, + e_p1, then 111
      const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
      for (int t = 0; t < 6 && !found; ++t) {
        const int c1 = 1 << perm[t][0];
        const int c2 = c1 | (1 << perm[t][1]);
        double l[4];
        if (!bary4(q, V[0], V[c1], V[c2], V[7], l)) continue;
        if (fmin(fmin(l[0], l[1]), fmin(l[2], l[3])) >= -1e-10) {
          found = true;
          const int cs[4] = {0, c1, c2, 7};
          for (int v = 0; v < 4; ++v) {
            lam[v] = l[v];
            vid[v] = pidx(cell[0] + (cs[v] & 1), cell[1] + ((cs[v] >> 1) & 1), cell[2] + (cs[v] >> 2));
          }
        }
      }
      if (found) break;
      // local coordinates in the cell's mean affine frame (mean edge vector
      // per axis, origin at the cell centroid)
      double J[3][3], o[3];
      for (int d = 0; d < 3; ++d) {
        o[d] = 0.0;
        for (int cc = 0; cc < 8; ++cc) o[d] += V[cc][d];
        o[d] *= 0.125;
      }
      for (int ax = 0; ax < 3; ++ax)
        for (int d = 0; d < 3; ++d) {
          double e = 0.0;
          for (int cc = 0; cc < 8; ++cc)
            if (cc & (1 << ax)) e += V[cc][d] - V[cc ^ (1 << ax)][d];
          J[d][ax] = 0.25 * e;
        }
      const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                         J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                         J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
      if (!(fabs(det) > 1e-300)) break;
      const double r[3] = {q[0] - o[0], q[1] - o[1], q[2] - o[2]};
      double tl[3];
      tl[0] = (r[0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (r[1] * J[2][2] - J[1][2] * r[2]) +
               J[0][2] * (r[1] * J[2][1] - J[1][1] * r[2])) / det;
      tl[1] = (J[0][0] * (r[1] * J[2][2] - J[1][2] * r[2]) - r[0] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
               J[0][2] * (J[1][0] * r[2] - r[1] * J[2][0])) / det;
      tl[2] = (J[0][0] * (J[1][1] * r[2] - r[1] * J[2][1]) - J[0][1] * (J[1][0] * r[2] - r[1] * J[2][0]) +
               r[0] * (J[1][0] * J[2][1] - J[1][1] * J[2][0])) / det;
      bool moved = false;
      for (int d = 0; d < 3; ++d) {
        const double f = floor(tl[d] + 0.5);  // whole cells beyond this one
        if (f != 0.0) {
          const double nc = fmin(fmax((double)cell[d] + f, 0.0), (double)(N[d] - 2));
          if ((int)nc != cell[d]) { cell[d] = (int)nc; moved = true; }
        }
      }
      if (!moved) break;
    }
  }
  for (int c = 0; c < C; ++c) {
    double v;
    if (found) {
      v = 0.0;
      for (int k = 0; k < 4; ++k) v += lam[k] * field[vid[k] * C + c];
    } else {
      const int b = nn_idx[qi];
      v = b >= 0 ? field[(long long)b * C + c] : __longlong_as_double(0x7ff8000000000000LL);
    }
    out[qi * C + c] = v;
  }
}

}  // namespace

namespace vdvc {

int launch_material_cells(int Nx, int Ny, int Nz, int h, const void* img, int itype, double glt,
                          uint8_t* out, hipStream_t st) {
  const long long Ne = (long long)Nx * Ny * Nz;
  const unsigned nb = (unsigned)((Ne * 64 + VDVC_BLOCK - 1) / VDVC_BLOCK);
  hipLaunchKernelGGL(material_cells_kernel, dim3(nb), dim3(VDVC_BLOCK), 0, st, Nx, Ny, Nz, h, img,
                     itype, glt, out);
  return (int)hipGetLastError();
}

int launch_erode_cells(int Nx, int Ny, int Nz, const uint8_t* in, uint8_t* out, hipStream_t st) {
  const long long Ne = (long long)Nx * Ny * Nz;
  const unsigned nb = (unsigned)((Ne + VDVC_BLOCK - 1) / VDVC_BLOCK);
  hipLaunchKernelGGL(erode_cells_kernel, dim3(nb), dim3(VDVC_BLOCK), 0, st, Nx, Ny, Nz, in, out);
  return (int)hipGetLastError();
}

int launch_remap_linear(int Nx, int Ny, int Nz, int h, const double* pts, const double* qry,
                        const double* field, int C, const int* nn_idx, double* out,
                        hipStream_t st) {
  const long long Ne = (long long)Nx * Ny * Nz;
  const unsigned nb = (unsigned)((Ne + VDVC_BLOCK - 1) / VDVC_BLOCK);
  hipLaunchKernelGGL(remap_linear_kernel, dim3(nb), dim3(VDVC_BLOCK), 0, st, Nx, Ny, Nz, (double)h,
                     pts, qry, field, C, nn_idx, out);
  return (int)hipGetLastError();
}

int launch_elem_count(const CellIn& C, hipStream_t st) {
  const long long Ne = (long long)C.Nx * C.Ny * C.Nz;
  const long long rows = (long long)C.Ny * C.Nz;
  if (C.Nx <= 4096 && rows <= 0x7fffffffLL) {
    const size_t lds = (size_t)3 * C.Nx * sizeof(int);
    if (C.b)
      hipLaunchKernelGGL(elem_row_count_kernel<true>, dim3((unsigned)rows), dim3(VDVC_BLOCK), lds,
                         st, C);
    else
      hipLaunchKernelGGL(elem_row_count_kernel<false>, dim3((unsigned)rows), dim3(VDVC_BLOCK), lds,
                         st, C);
    return (int)hipGetLastError();
  }
  const unsigned nb = (unsigned)((Ne + VDVC_WAVES_PER_BLOCK - 1) / VDVC_WAVES_PER_BLOCK);
  if (C.b)
    hipLaunchKernelGGL(elem_count_kernel<true>, dim3(nb), dim3(VDVC_BLOCK), 0, st, C);
  else
    hipLaunchKernelGGL(elem_count_kernel<false>, dim3(nb), dim3(VDVC_BLOCK), 0, st, C);
  return (int)hipGetLastError();
}

int launch_elem_flags(const CellFlagsIn& F, hipStream_t st) {
  const long long Ne = (long long)F.Nx * F.Ny * F.Nz;
  if (F.unsafe_elem || F.lost_elem) {
    const unsigned nb = (unsigned)((Ne + VDVC_BLOCK - 1) / VDVC_BLOCK);
    hipLaunchKernelGGL(elem_flags_kernel, dim3(nb), dim3(VDVC_BLOCK), 0, st, F);
    const int rc = (int)hipGetLastError();
    if (rc) return rc;
  }
  if (F.unsafe_node) {
    const long long Nn = (long long)(F.Nx + 1) * (F.Ny + 1) * (F.Nz + 1);
    const unsigned nb = (unsigned)((Nn + VDVC_BLOCK - 1) / VDVC_BLOCK);
    hipLaunchKernelGGL(node_flags_kernel, dim3(nb), dim3(VDVC_BLOCK), 0, st, F);
  }
  return (int)hipGetLastError();
}

int launch_unsafe_voxels(int Nx, int Ny, int Nz, int h, const uint8_t* unsafe_elem,
                         uint8_t* out, hipStream_t st) {
  const long long rows = (long long)(Ny * h + 1) * (Nz * h + 1);
  const long long nb = (rows + VDVC_WAVES_PER_BLOCK - 1) / VDVC_WAVES_PER_BLOCK;
  const size_t lds = (size_t)VDVC_WAVES_PER_BLOCK * Nx;
  if (nb > 0x7fffffffLL || lds > 65536) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(unsafe_voxel_kernel, dim3((unsigned)nb), dim3(VDVC_BLOCK), lds, st, Nx, Ny, Nz,
                     h, unsafe_elem, out);
  return (int)hipGetLastError();
}

int launch_strains(const StrainIn& S, hipStream_t st) {
  const long long Ne = (long long)S.Nx * S.Ny * S.Nz;
  const unsigned nb = (unsigned)((Ne + VDVC_BLOCK - 1) / VDVC_BLOCK);
  hipLaunchKernelGGL(strain_kernel, dim3(nb), dim3(VDVC_BLOCK), 0, st, S);
  return (int)hipGetLastError();
}

int launch_equiv_strain(long long n, const double* principal, double k_eq, double* out,
                        hipStream_t st) {
  const unsigned nb = (unsigned)((n + VDVC_BLOCK - 1) / VDVC_BLOCK);
  hipLaunchKernelGGL(equiv_kernel, dim3(nb), dim3(VDVC_BLOCK), 0, st, n, principal, k_eq, out);
  return (int)hipGetLastError();
}

int launch_gn_convergence(int Nx, int Ny, int Nz, const void* dU, int dcode, const void* U,
                          int ucode, double* out, hipStream_t st) {
  const long long Ne = (long long)Nx * Ny * Nz;
  const unsigned nb = (unsigned)((Ne + VDVC_BLOCK - 1) / VDVC_BLOCK);
  hipLaunchKernelGGL(gn_convergence_kernel, dim3(nb), dim3(VDVC_BLOCK), 0, st, Nx, Ny, Nz, dU,
                     dcode, U, ucode, out);
  return (int)hipGetLastError();
}

int launch_elem_centres(int Nx, int Ny, int Nz, int h, const void* U, int ucode,
                        const double* crop, double* pts, double* qry, hipStream_t st) {
  const long long Ne = (long long)Nx * Ny * Nz;
  const unsigned nb = (unsigned)((Ne + VDVC_BLOCK - 1) / VDVC_BLOCK);
  hipLaunchKernelGGL(elem_centres_kernel, dim3(nb), dim3(VDVC_BLOCK), 0, st, Nx, Ny, Nz, h, U,
                     ucode, crop[0], crop[1], crop[2], pts, qry);
  return (int)hipGetLastError();
}

int launch_bbox(long long n, const double* pts, double* part, int nparts, double* out,
                hipStream_t st) {
  hipLaunchKernelGGL(bbox_part_kernel, dim3((unsigned)nparts), dim3(VDVC_BLOCK), 0, st, n, pts,
                     part);
  hipLaunchKernelGGL(bbox_final_kernel, dim3(1), dim3(VDVC_BLOCK), 0, st, nparts, part, out);
  return (int)hipGetLastError();
}

int nn_scan_chunks(long long ncell) { return (int)((ncell + kScanChunk - 1) / kScanChunk); }

int launch_nn_build(long long n, const double* pts, const NNGrid& g, int* cnt, int* off,
                    int* csum, int* list, hipStream_t st) {
  const long long ncell = (long long)g.G[0] * g.G[1] * g.G[2];
  const unsigned nb = (unsigned)((n + VDVC_BLOCK - 1) / VDVC_BLOCK);
  const int nchunk = nn_scan_chunks(ncell);
  hipLaunchKernelGGL(nn_count_kernel, dim3(nb), dim3(VDVC_BLOCK), 0, st, n, pts, g, cnt);
  hipLaunchKernelGGL(nn_chunk_sum_kernel, dim3((unsigned)nchunk), dim3(VDVC_BLOCK), 0, st, ncell,
                     cnt, csum);
  hipLaunchKernelGGL(nn_chunk_scan_kernel, dim3(1), dim3(1024), 0, st, nchunk, csum, ncell, off);
  hipLaunchKernelGGL(nn_chunk_apply_kernel, dim3((unsigned)nchunk), dim3(VDVC_BLOCK), 0, st, ncell,
                     cnt, csum, off);
  hipLaunchKernelGGL(nn_fill_kernel, dim3(nb), dim3(VDVC_BLOCK), 0, st, n, pts, g, cnt, list);
  return (int)hipGetLastError();
}

int launch_nn_query(long long n, const double* pts, const double* qry, const NNGrid& g,
                    const int* off, const int* list, const uint8_t* vals, uint8_t* out,
                    int* idx_out, hipStream_t st) {
  const unsigned nb = (unsigned)((n + VDVC_BLOCK - 1) / VDVC_BLOCK);
  hipLaunchKernelGGL(nn_query_kernel, dim3(nb), dim3(VDVC_BLOCK), 0, st, n, pts, qry, g, off,
                     list, vals, out, idx_out);
  return (int)hipGetLastError();
}

}  // namespace vdvc
