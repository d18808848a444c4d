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
//   * elem_centres_kernel + nearest_remap_kernel: the nearest-neighbour
//     remap of an element mask to the deformed configuration
//     (write_output.py:494-530): element centres moved by the mean corner
//     displacement, and for every regular centre (+ crop offset) the value of
//     the nearest moved centre, by exhaustive search over LDS tiles of
//     points (ties: the lowest element index).
// Elements x fastest (e = i + Nx*(j + Ny*k)), nodes n = i + (Nx+1)*(j +
// (Ny+1)*k), voxels x fastest on the (Nx*h+1) x (Ny*h+1) x (Nz*h+1) grid.
#include "vdvc_common.h"
#include "vdvc_kernels.h"
#include "vdvc_elem.h"
#include "vdvc_sample01.h"

namespace {

using vdvc_img::ld_img;

__global__ __launch_bounds__(VDVC_BLOCK) void elem_count_kernel(CellIn C) {
  const int lane = threadIdx.x & 63;
  const long long e = (long long)blockIdx.x * VDVC_WAVES_PER_BLOCK + (threadIdx.x >> 6);
  const long long Ne = (long long)C.Nx * C.Ny * C.Nz;
  if (e >= Ne) return;  // whole wavefronts
  const long long base = vdvc_elem::elem_base(e, C.Nx, C.Ny, C.h, C.Nvx, C.Nvy);
  int n0 = 0, n1 = 0, n2 = 0;
  vdvc_elem::walk(C.h, C.Nvx, C.Nvy, base, lane, [&](long long vi) {
    const bool a0 = C.a[vi] == 0;
    const bool b1 = C.b ? C.b[vi] != 0 : true;
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

__global__ __launch_bounds__(VDVC_BLOCK) void unsafe_voxel_kernel(int Nx, int Ny, int Nz, int h,
                                                                  const uint8_t* __restrict__ ue,
                                                                  uint8_t* __restrict__ out) {
  const int Nvx = Nx * h + 1, Nvy = Ny * h + 1;
  const int x = blockIdx.x * VDVC_BLOCK + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x >= Nvx) return;
  int x0, x1, y0, y1, z0, z1;
  axis_elems(x, h, Nx, x0, x1);
  axis_elems(y, h, Ny, y0, y1);
  axis_elems(z, h, Nz, z0, z1);
  bool u = false;
  for (int k = z0; k <= z1; ++k)
    for (int j = y0; j <= y1; ++j)
      for (int i = x0; i <= x1; ++i) u = u || ue[i + (long long)Nx * (j + (long long)Ny * k)] != 0;
  out[(long long)x + (long long)Nvx * ((long long)y + (long long)Nvy * z)] = u ? 1 : 0;
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

// exhaustive nearest point per query: squared distance ((d0^2 + d1^2) +
// d2^2), strict < so the lowest index wins ties
__global__ __launch_bounds__(VDVC_BLOCK) void nearest_remap_kernel(long long n,
                                                                   const double* __restrict__ pts,
                                                                   const double* __restrict__ qry,
                                                                   const uint8_t* __restrict__ vals,
                                                                   uint8_t* __restrict__ out) {
  __shared__ double tile[VDVC_BLOCK * 3];
  const long long qi = (long long)blockIdx.x * VDVC_BLOCK + threadIdx.x;
  const bool live = qi < n;
  const double q0 = live ? qry[3 * qi] : 0.0, q1 = live ? qry[3 * qi + 1] : 0.0,
               q2 = live ? qry[3 * qi + 2] : 0.0;
  double best = __longlong_as_double(0x7ff0000000000000LL);  // +inf
  long long bi = 0;
  for (long long t0 = 0; t0 < n; t0 += VDVC_BLOCK) {
    const long long pi = t0 + threadIdx.x;
    if (pi < n) {
      tile[3 * threadIdx.x] = pts[3 * pi];
      tile[3 * threadIdx.x + 1] = pts[3 * pi + 1];
      tile[3 * threadIdx.x + 2] = pts[3 * pi + 2];
    }
    __syncthreads();
    const int m = (int)(n - t0 < VDVC_BLOCK ? n - t0 : VDVC_BLOCK);
    for (int s = 0; s < m; ++s) {
      const double d0 = sub_rn(q0, tile[3 * s]), d1 = sub_rn(q1, tile[3 * s + 1]),
                   d2 = sub_rn(q2, tile[3 * s + 2]);
      const double dd = add_rn(add_rn(mul_rn(d0, d0), mul_rn(d1, d1)), mul_rn(d2, d2));
      if (dd < best) {
        best = dd;
        bi = t0 + s;
      }
    }
    __syncthreads();
  }
  if (live) out[qi] = vals[bi];
}

}  // namespace

namespace vdvc {

int launch_elem_count(const CellIn& C, hipStream_t st) {
  const long long Ne = (long long)C.Nx * C.Ny * C.Nz;
  const unsigned nb = (unsigned)((Ne + VDVC_WAVES_PER_BLOCK - 1) / VDVC_WAVES_PER_BLOCK);
  hipLaunchKernelGGL(elem_count_kernel, dim3(nb), dim3(VDVC_BLOCK), 0, st, C);
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
  const int Nvx = Nx * h + 1;
  dim3 g((unsigned)((Nvx + VDVC_BLOCK - 1) / VDVC_BLOCK), (unsigned)(Ny * h + 1),
         (unsigned)(Nz * h + 1));
  hipLaunchKernelGGL(unsafe_voxel_kernel, g, dim3(VDVC_BLOCK), 0, st, Nx, Ny, Nz, h, unsafe_elem,
                     out);
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

int launch_nearest_remap(long long n, const double* pts, const double* qry, const uint8_t* vals,
                         uint8_t* out, hipStream_t st) {
  const unsigned nb = (unsigned)((n + VDVC_BLOCK - 1) / VDVC_BLOCK);
  hipLaunchKernelGGL(nearest_remap_kernel, dim3(nb), dim3(VDVC_BLOCK), 0, st, n, pts, qry, vals,
                     out);
  return (int)hipGetLastError();
}

}  // namespace vdvc
