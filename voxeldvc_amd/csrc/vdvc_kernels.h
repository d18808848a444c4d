// Kernel parameter blocks and launchers shared by the C-ABI layer.
#pragma once
#include "vdvc_common.h"

enum { MODE_MATVEC = 0, MODE_RHS = 1, MODE_DIAG = 2 };

struct VoxIn {
  const void* g[3];          // gradient planes (GT)
  const uint8_t* mask;       // optional voxel mask (HASMASK)
  // RHS mode
  const void* fimg;          // reference image (FT), raw values
  const void* gvals;         // sampled deformed image (R = RealOf<GT>)
  const uint8_t* live;       // live mask (display statistics)
  double mean0, gmean, kscale;  // converted to R in the kernel
  void* res;                 // residual out (R)
  uint16_t* fold_out[3];     // optional bf16 gradient * mask out
  double* stats;             // per-wave partial statistics [nwaves][4]
  const int* done;           // early-exit flag (device-resident PCG), may be null
  // products of the vectorised kernel: an element-stiffness term added to
  // the node partials, kc * sum_b kreg[a * 8 + b] v_b with kreg[a * 8 + b] =
  // K[3a][3b] of a block-diagonal element matrix (the Laplacian; device
  // array of 64) and kc = c E, each element once (kreg null: none)
  const float* kreg;
  float kc;
  // residual pass: node partials of the Hessian diagonal (same slot layout
  // as the product's), taken in the same pass when non-null
  float* pdiag;
};

struct SampleIn {
  const void* gimg;      // deformed image, x fastest, (gnx, gny, gnz)
  int gnx, gny, gnz;
  double gox, goy, goz;  // g origin (converted to R)
  double glt;            // threshold (converted to R)
  const uint8_t* mask0;  // reference mask (may be null: all true)
  const uint8_t* bmask;  // frozen b_mask for the statistics (null: live mask)
  void* gvals;           // R
  uint8_t* mask_out;
  double* stats;         // per-wave partials [nwaves][3]
  int tail_fold;         // the last strip's wavefronts also take column Nvx - 1
};

// post-solve voxel stages on dense voxel arrays (flat index x + Nvx*(y + Nvy*z))
enum { POST_INTERP = 0, POST_PUSH = 1, POST_MASK = 2 };
constexpr int kPostFzMax = 64;

struct PostIn {
  int Nx, Ny, Nz, h;
  int Nvx, Nvy, Nvz;
  int nnx, nny;
  long long Nvox;
  const void* U;          // nodal vector, R
  const void* img;        // sampled image, x fastest, (gnx, gny, gnz), type itype
  int itype;
  int gnx, gny, gnz;
  double ox, oy, oz;      // mask: image origin in the voxel frame
  int order, mode;        // push: spline order 0/1, mode 0 nearest / 1 constant
  double cval, glt;
  const uint8_t* mask0;   // mask: optional reference mask
  void* out;              // interp: R; push: itype; mask: uint8
  int channels_first;     // interp: (3, Nvox) instead of (Nvox, 3)
  unsigned img_bytes;     // bytes of img when below 2^32 (32-bit offsets), else 0
  // z factors (1 - xi, 1 + xi) of the Hex8 weights at local plane k, xi =
  // 2k/h - 1 formed in double as the kernels form it; filled for h <= 64
  int fz_tab;
  double fz[kPostFzMax + 1][2];
};

struct QualIn {
  int Nx, Ny, Nz, h, Nvx, Nvy;
  const void* res; int rtype;   // per-voxel residual (F32 / F64)
  const void* ref; int ftype;   // reference image (x fastest)
  double* std_out;              // (Ne,) element order x fastest
  double* zncc_out;
  int min_active;
};

struct ZnccIn {
  long long n;
  const void* f; int ftype;
  const void* g; int gtype;
  int g_is_residual;            // g := f - g
  const uint8_t* mask;          // optional
  long long chunk;              // voxels per block (a multiple of 1024 when vec)
  int vec;                      // 4 voxels per thread step (16-byte aligned arrays)
};

// preprocess affine stage on dense float32 volumes (x fastest)
#define VDVC_GAUSS_MAX_R 64
struct GaussIn {
  int nx, ny, nz;
  int axis;                  // 0, 1, 2
  int mode;                  // 0 'nearest', 1 'constant'
  int r;                     // radius (<= VDVC_GAUSS_MAX_R)
  double cval;
  double w[VDVC_GAUSS_MAX_R + 1];   // centre tap, then taps 1..r
};

struct AffIn {
  int nx, ny, nz;
  float M[9], tr[3], c[3];   // generative affine (float32 as the reference rounds it)
  double cval;
  const float* f;            // (smoothed) reference
  const float* g;            // (smoothed) deformed
  const float* gx;           // gradient planes of f (normal equations)
  const float* gy;
  const float* gz;
  const uint8_t* valid;      // g != cval
};

struct WarpIn {
  int nx, ny, nz;            // output (reference) grid
  int gnx, gny, gnz;         // sampled image
  float A[9], t[3], c[3];
};

// Element-scale write-output fields.  Per element (x fastest) over its
// (h+1)^3 voxels: c0 = #(a == 0), c1 = #(b != 0), c2 = #(b != 0 && a == 0)
// (b NULL counts as all non-zero); all_a0 (optional) = (c0 == (h+1)^3).
struct CellIn {
  int Nx, Ny, Nz, h, Nvx, Nvy;
  const uint8_t* a;
  const uint8_t* b;
  int* c0;
  int* c1;
  int* c2;
  uint8_t* all_a0;
};

// per-element flags from the counts: unsafe = boundary shell | an element
// within +-1 (sharing a node) has c0 > 0; lost = an element within +-1 has
// material (c1 > 0) and c2 / c1 > frac_threshold, | c1 / (h+1)^3 <
// min_material_frac | boundary shell; node = an adjacent element has
// c0 > 0 or is on the shell.  Outputs optional.
struct CellFlagsIn {
  int Nx, Ny, Nz, h;
  const int* c0;
  const int* c1;
  const int* c2;
  double frac_threshold, min_material_frac;
  uint8_t* unsafe_elem;
  uint8_t* lost_elem;
  uint8_t* unsafe_node;
};

// principal strains (descending) and equivalent strain at element centres
struct StrainIn {
  int Nx, Ny, Nz;
  const void* U; int ucode;  // flat nodal vector, node-major (F32 / F64)
  double g;                  // 0.25 / h: |dN_a / dx_d| at the centre
  double k_eq;               // sqrt(2) / 3
  double* principal;         // (Ne, 3)
  double* equiv;             // (Ne,) optional
};

// uniform grid of cells for the nearest-neighbour remap: origin o, cell
// edge cs (inv_cs = 1 / cs), G cells per axis
struct NNGrid {
  double o[3];
  double cs, inv_cs;
  int G[3];
};

// element stiffness operator: E_e * K (24x24, node-major DOFs) summed over
// the elements adjacent to node (i,j,k).  blockdiag=1 means K[3a+c][3b+d] =
// L8[a][b] delta_cd with L8 = K[3a][3b] (vector Laplacian), cheaper path.
template <typename R>
struct KOp {
  const R* K;          // 24*24 row-major (device)
  int blockdiag;
  R E;                 // scalar element scale
  const R* Earr;       // optional per-element scale (Nx*Ny*Nz)
  R kc[64];            // block-diagonal form: kc[a * 8 + b] = K[3a][3b] (launch argument)
};

template <typename R>
struct NodeOp {
  const R* P;            // voxel partials (null: none)
  TilesDev T;
  R cH;                  // scale of the partial sum
  KOp<R> k1; const R* v1; R c1; int k1_on;
  KOp<R> k2; const R* v2; R c2; int k2_on;
  const R* addin; R cadd;          // out += cadd * addin (may alias out)
  int interior;                    // if set: result *= interior-node mask (eq-gap P_i)
  R* out;
  const R* dotv;                   // partial dot(dotv, out) in FP64 (null: none; not out)
  double* dot_part;
  const int* done;                 // early-exit flag (PCG), may be null
};

// diag of the assembled H (from DIAG-mode partials) plus c1 * diag(K1)
// plus c2 * extra (e.g. a precomputed regularizer diagonal), and per-block
// minimum for the Jacobi shift
template <typename R>
struct DiagOp {
  const R* P; TilesDev T;
  KOp<R> k1; R c1; int k1_on;
  const R* extra; R cextra;
  R* out;
  R* min_part;
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

// strided volume copy with element-type conversion (vdvc_util.hip):
// dst[x + dp (y + ny z)] = src[x sx + y sy + z sz] for x < nx, 0 for
// nx <= x < dp
struct VolCopy {
  int nx, ny, nz;
  long long sx, sy, sz;   // source strides (elements)
  const void* src; int st;
  void* dst; int dt;
  int dp;                 // destination row pitch (>= nx)
};

// nodal field of one structured mesh interpolated onto another covering the
// same voxels (vdvc_util.hip)
struct NodeResample {
  int Mx, My, Mz, hs;     // source mesh (elements) and element size
  int Nx, Ny, Nz, hd;     // destination mesh and element size
  const void* src;
  void* dst;
  int rt;                 // working type of both (GT_F32 / GT_F64)
};

namespace vdvc {
int launch_resample_nodes(const NodeResample& R, hipStream_t s);
int launch_convert(long long n, const void* src, int st, void* dst, int dt, hipStream_t s);
int launch_copy(long long nbytes, const void* src, void* dst, hipStream_t s);
// read-only streaming probe over nbytes (a multiple of 16, 16-byte aligned)
int launch_stream_read(long long nbytes, const void* src, int grid, uint32_t key, uint32_t* out,
                       hipStream_t s);
int launch_fill(long long n, void* dst, int dt, double v, hipStream_t s);
int launch_copy_volume(const VolCopy& C, hipStream_t s);
int launch_axpby_x(long long n, double a, const void* x, int xt, double b, void* y, int yt,
                   hipStream_t s);
int vec_width(int h, int gtype);
bool rhs_diag_fused(int gtype, int ftype, bool fold, const TilesDev& T);
int launch_restrict_vec(int mode, int gtype, bool hasmask, int ftype, bool fold, const MeshDev& M,
                        const TilesDev& T, const VoxIn& in, const float* v, float* P,
                        hipStream_t st);
// generic voxel kernel; v and P are R = double for gtype GT_F64, else float
// the dense residual (g - g_mean) kscale + (f - mean0) over the b-mask from
// a residual pass's inputs (working type rtype, image type ftype)
int launch_residual_out(int rtype, int ftype, const MeshDev& M, const void* f, const void* gv,
                        const uint8_t* bm, double mean0, double gmean, double kscale, void* out,
                        hipStream_t st);
int launch_restrict(int mode, int gtype, bool hasmask, int ftype, bool fold, const MeshDev& M,
                    const TilesDev& T, const VoxIn& in, const void* v, void* P, hipStream_t st);
// U and gvals are double for an F64 image, else float
// f32w: float32 interpolation weights (float and half images)
int launch_sample(int gitype, const MeshDev& M, const TilesDev& T, const SampleIn& S,
                  const void* U, bool f32w, hipStream_t st);
int launch_setup(int ftype, int gtype, const MeshDev& M, const void* f, double glt, void* gx,
                 void* gy, void* gz, uint8_t* mask0, hipStream_t st);
int launch_level_image(int stype, int dtype, const MeshDev& M, const void* src, void* dst,
                       double glt, double* part, int nblocks, hipStream_t st);
int launch_stats(int ftype, const MeshDev& M, const void* fld, const uint8_t* mask, int use_thresh,
                 double glt, double* part, int nblocks, hipStream_t st);
int launch_reduce(const double* part, long long n, int nv, double* out, hipStream_t st);
int launch_fold(long long n, const uint16_t* gx, const uint16_t* gy, const uint16_t* gz,
                const uint8_t* m, uint16_t* ox, uint16_t* oy, uint16_t* oz, hipStream_t st);
int launch_erode6(int nx, int ny, int nz, const uint8_t* in, uint8_t* out, hipStream_t st);
template <typename R>
int launch_mul(long long n, const R* a, const R* b, R* out, hipStream_t st);
template <typename R>
int launch_node_op(const MeshDev& M, const NodeOp<R>& op, hipStream_t st);
int node_blocks(const MeshDev& M);
int vec_blocks(long long n);
template <typename R>
int launch_diag_op(const MeshDev& M, const DiagOp<R>& op, hipStream_t st);
template <typename R>
int launch_jacobi(const R* diag, const R* min_part, int nparts, long long ndof, R* minv,
                  hipStream_t st);
template <typename R>
int launch_pcg_init(long long ndof, const R* b, const R* minv, R* x, R* r, R* z, R* p,
                    double* part, PcgState* S, hipStream_t st);
// vec: the vectorised update kernels (every vector 16-byte aligned and
// pcg_vec_ok(ndof)); both updates of one solve take the same form
bool pcg_vec_ok(long long ndof);
template <typename R>
int launch_pcg_update(long long ndof, int it, const double* pap_part, int npap, PcgState* S,
                      const R* minv, const R* p, const R* Ap, R* x, R* r, R* z, double* part,
                      bool vec, hipStream_t st);

template <typename R>
int launch_pcg_pupdate(long long ndof, int it, double tol, const double* part, PcgState* S,
                       const R* z, R* p, bool vec, hipStream_t st);
template <typename R>
int launch_axpby(long long n, double a, const R* x, double b, R* y, hipStream_t st);
template <typename R>
int launch_dot(long long n, const R* a, const R* b, double* part, int nparts, hipStream_t st);
template <typename R>
int launch_dot2(long long n, const R* a, const R* b, const R* c, double* part, int nparts,
                hipStream_t st);
template <typename R>
int launch_gn_update(long long n, const R* dU, R* U, double* part, int nparts, hipStream_t st);
template <typename R>
int launch_calib_mode(int Nx, int Ny, int Nz, int h, R* out, hipStream_t st);
template <typename R>
int launch_hutch_acc(long long ndof, const int8_t* z, const R* Rz, double* acc, hipStream_t st);
template <typename R>
int launch_expand_i8(long long n, const int8_t* z, R* out, hipStream_t st);
template <typename R>
int launch_hutch_final(long long ndof, const double* acc, int nprobes, R* out, double* scratch,
                       hipStream_t st);
int launch_stride(int type, const void* src, int nx, int ny, int nz, int s, void* dst, int mx,
                  int my, int mz, hipStream_t st);
// (value, index) partials of the phase-correlation argmax: 16 bytes each
#define VDVC_ARG_PARTS 512
int launch_phase_corr(const float* f, const float* g, int nx, int ny, int nz, double2* work,
                      void* part, long long* out_idx, hipStream_t st);
int launch_affine_cost(const float* f, int nx, int ny, int nz, const float* g, int gnx, int gny,
                       int gnz, const AffineIn& P, double* part, int nblocks, hipStream_t st);
// learned-seed network (vdvc_seednet.hip)
int launch_conv_fwd(int N, int D, int H, int W, int C, int O, int stride, const float* x,
                    const float* w, const float* b, int relu, float* y, hipStream_t st);
int launch_conv_bwd_data(int N, int D, int H, int W, int C, int O, int stride, const float* dy,
                         const float* w, const float* act, float* dx, hipStream_t st);
int conv_wgrad_parts(int N, int D, int H, int W, int C, int O, int stride);
int launch_conv_bwd_weight(int N, int D, int H, int W, int C, int O, int stride, const float* x,
                           const float* dy, float* part, float* dw, float* db, hipStream_t st);
int launch_adam(long long n, float* p, const float* g, float* m, float* v, float lr, float b1,
                float b2, float eps, float c1, float c2, hipStream_t st);
int launch_mse(long long n, const float* pred, const float* label, float* grad, double* part,
               hipStream_t st);
int launch_post_voxel(int op, int rtype, const PostIn& P, hipStream_t st);
int launch_quality_cell(const QualIn& Q, hipStream_t st);
// pass 0: per-block (count, sum f, sum g); pass 1: per-block centred sums
// (f g, f f, g g) around the means that `red` (count, sum f, sum g) defines
int launch_zncc_pass(int pass, const ZnccIn& Z, const double* red, double* part, int nblocks,
                     hipStream_t st);
int launch_gauss_axis(const GaussIn& G, const float* in, float* out, hipStream_t st);
int launch_mask_prep(long long n, const float* g, float cval, float* gv, float* vf,
                     uint8_t* valid, hipStream_t st);
int launch_masked_ratio(long long n, const float* num, const float* den, const float* g,
                        float* out, hipStream_t st);
int launch_gradient(int nx, int ny, int nz, const float* f, float* gx, float* gy, float* gz,
                    hipStream_t st);
int launch_affine_normal(const AffIn& A, double* part, int nblocks, hipStream_t st);
int launch_affine_rms(const AffIn& A, double* part, int nblocks, hipStream_t st);
int launch_warp_affine(const WarpIn& W, const float* g, float* out, hipStream_t st);
int launch_overlap_proj(int nx, int ny, int nz, const void* f, int ftype, const float* g,
                        double glt, uint8_t* flags, hipStream_t st);
int launch_elem_count(const CellIn& C, hipStream_t st);
int launch_elem_flags(const CellFlagsIn& F, hipStream_t st);
int launch_unsafe_voxels(int Nx, int Ny, int Nz, int h, const uint8_t* unsafe_elem,
                         uint8_t* out, hipStream_t st);
int launch_strains(const StrainIn& S, hipStream_t st);
int launch_equiv_strain(long long n, const double* principal, double k_eq, double* out,
                        hipStream_t st);
int launch_gn_convergence(int Nx, int Ny, int Nz, const void* dU, int dcode, const void* U,
                          int ucode, double* out, hipStream_t st);
int launch_elem_centres(int Nx, int Ny, int Nz, int h, const void* U, int ucode,
                        const double* crop, double* pts, double* qry, hipStream_t st);
// bounding box of n points (x, y, z): out = (min x, min y, min z, max x,
// max y, max z); part holds 6 * nparts doubles
int launch_bbox(long long n, const double* pts, double* part, int nparts, double* out,
                hipStream_t st);
// cnt (zeroed, ncell ints) -> cell offsets off (ncell + 1) and the point
// list (n ints) sorted by cell; csum holds nn_scan_chunks(ncell) ints;
// nparts of launch_bbox <= 256
int nn_scan_chunks(long long ncell);
int launch_nn_build(long long n, const double* pts, const NNGrid& g, int* cnt, int* off,
                    int* csum, int* list, hipStream_t st);
// vals/out: the nearest point's value (optional); idx_out: its index (optional)
int launch_nn_query(long long n, const double* pts, const double* qry, const NNGrid& g,
                    const int* off, const int* list, const uint8_t* vals, uint8_t* out,
                    int* idx_out, hipStream_t st);
int launch_material_cells(int Nx, int Ny, int Nz, int h, const void* img, int itype, double glt,
                          uint8_t* out, hipStream_t st);
int launch_erode_cells(int Nx, int Ny, int Nz, const uint8_t* in, uint8_t* out, hipStream_t st);
int launch_remap_linear(int Nx, int Ny, int Nz, int h, const double* pts, const double* qry,
                        const double* field, int C, const int* nn_idx, double* out,
                        hipStream_t st);
int launch_seed_input(int itype, const void* f, const void* g, int nx, int ny, int nz, int s,
                      double* part, int nparts, double* stats6, float* out, hipStream_t st);
}  // namespace vdvc
