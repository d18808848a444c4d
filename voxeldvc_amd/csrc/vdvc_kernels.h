// Kernel parameter blocks and launchers shared by the C-ABI layer.
#pragma once
#include "vdvc_common.h"

enum { MODE_MATVEC = 0, MODE_RHS = 1, MODE_DIAG = 2 };

struct VoxIn {
  const void* g[3];          // gradient planes (GT)
  const uint8_t* mask;       // optional voxel mask (HASMASK)
  // RHS mode
  const void* fimg;          // reference image (FT), raw values
  const float* gvals;        // sampled deformed image
  const uint8_t* live;       // live mask (display statistics)
  float mean0, gmean, kscale;
  float* res;                // residual out
  uint16_t* fold_out[3];     // optional bf16 gradient * mask out
  double* stats;             // per-wave partial statistics [nwaves][4]
  const int* done;           // early-exit flag (device-resident PCG), may be null
};

struct SampleIn {
  const void* gimg;      // deformed image, x fastest, (gnx, gny, gnz)
  int gnx, gny, gnz;
  float gox, goy, goz;   // g origin
  float glt;
  const uint8_t* mask0;  // reference mask (may be null: all true)
  const uint8_t* bmask;  // frozen b_mask for the statistics (null: live mask)
  float* gvals;
  uint8_t* mask_out;
  double* stats;         // per-wave partials [nwaves][3]
};

// element stiffness operator: E_e * K (24x24, node-major DOFs) summed over
// the elements adjacent to node (i,j,k).  blockdiag=1 means K[3a+c][3b+d] =
// L8[a][b] delta_cd with L8 = K[3a][3b] (vector Laplacian), cheaper path.
struct KOp {
  const float* K;      // 24*24 row-major (device)
  int blockdiag;
  float E;             // scalar element scale
  const float* Earr;   // optional per-element scale (Nx*Ny*Nz)
};

struct NodeOp {
  const float* P;        // voxel partials (null: none)
  TilesDev T;
  float cH;              // scale of the partial sum
  KOp k1; const float* v1; float c1; int k1_on;
  KOp k2; const float* v2; float c2; int k2_on;
  const float* addin; float cadd;  // out += cadd * addin (may alias out)
  int interior;                    // if set: result *= interior-node mask (eq-gap P_i)
  float* out;
  const float* dotv;               // partial dot(dotv, out) in FP64 (null: none)
  double* dot_part;
  const int* done;                 // early-exit flag (PCG), may be null
};

// diag of the assembled H (from DIAG-mode partials) plus c1 * diag(K1)
// plus c2 * extra (e.g. a precomputed regularizer diagonal), and per-block
// minimum for the Jacobi shift
struct DiagOp {
  const float* P; TilesDev T;
  KOp k1; float c1; int k1_on;
  const float* extra; float cextra;
  float* out;
  float* min_part;
};

struct PcgState {
  double rz[2];     // r.z of the current / next iteration (parity buffers)
  double r0;        // b.b + 1e-30
  double relres;    // last relative residual
  int done;
  int iters;
};

// affine warp parameters (row-major A)
struct AffineIn {
  double A[9], t[3], c[3];
};

namespace vdvc {
int vec_width(int h, int gtype);
int launch_restrict_vec(int mode, int gtype, bool hasmask, int ftype, bool fold, const MeshDev& M,
                        const TilesDev& T, const VoxIn& in, const float* v, float* P,
                        hipStream_t st);
int launch_restrict(int mode, int gtype, bool hasmask, int ftype, bool fold, const MeshDev& M,
                    const TilesDev& T, const VoxIn& in, const float* v, float* P, hipStream_t st);
int launch_sample(int gitype, const MeshDev& M, const TilesDev& T, const SampleIn& S,
                  const float* U, hipStream_t st);
int launch_setup(int ftype, int gtype, const MeshDev& M, const void* f, float glt, void* gx,
                 void* gy, void* gz, uint8_t* mask0, hipStream_t st);
int launch_stats(int ftype, const MeshDev& M, const void* fld, const uint8_t* mask, int use_thresh,
                 float glt, double* part, int nblocks, hipStream_t st);
int launch_reduce(const double* part, long long n, int nv, double* out, hipStream_t st);
int launch_fold(long long n, const uint16_t* gx, const uint16_t* gy, const uint16_t* gz,
                const uint8_t* m, uint16_t* ox, uint16_t* oy, uint16_t* oz, hipStream_t st);
int launch_erode6(int nx, int ny, int nz, const uint8_t* in, uint8_t* out, hipStream_t st);
int launch_mul(long long n, const float* a, const float* b, float* out, hipStream_t st);
int launch_node_op(const MeshDev& M, const NodeOp& op, hipStream_t st);
int node_blocks(const MeshDev& M);
int vec_blocks(long long n);
int launch_diag_op(const MeshDev& M, const DiagOp& op, hipStream_t st);
int launch_jacobi(const float* diag, const float* min_part, int nparts, long long ndof,
                  float* minv, hipStream_t st);
int launch_pcg_init(long long ndof, const float* b, const float* minv, float* x, float* r,
                    float* z, float* p, double* part, PcgState* S, hipStream_t st);
int launch_pcg_update(long long ndof, int it, const double* pap_part, int npap, PcgState* S,
                      const float* minv, const float* p, const float* Ap, float* x, float* r,
                      float* z, double* part, hipStream_t st);
int launch_pcg_pupdate(long long ndof, int it, float tol, const double* part, PcgState* S,
                       const float* z, float* p, hipStream_t st);
int launch_axpby(long long n, float a, const float* x, float b, float* y, hipStream_t st);
int launch_dot(long long n, const float* a, const float* b, double* part, int nparts,
               hipStream_t st);
int launch_hutch_acc(long long ndof, const int8_t* z, const float* Rz, double* acc, hipStream_t st);
int launch_expand_i8(long long n, const int8_t* z, float* out, hipStream_t st);
int launch_hutch_final(long long ndof, const double* acc, int nprobes, float* out, double* scratch,
                       hipStream_t st);
int launch_stride(int type, const void* src, int nx, int ny, int nz, int s, void* dst, int mx,
                  int my, int mz, hipStream_t st);
int launch_phase_corr(const float* f, const float* g, int nx, int ny, int nz, double2* work,
                      long long* out_idx, hipStream_t st);
int launch_affine_cost(const float* f, int nx, int ny, int nz, const float* g, int gnx, int gny,
                       int gnz, const AffineIn& P, double* part, int nblocks, hipStream_t st);
}  // namespace vdvc
