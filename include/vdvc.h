/*
 * libvdvc -- gfx950 kernels for the global Gauss-Newton DVC solve.
 *
 * C ABI: plain pointers and sizes, `int` status (0 = ok, < 0 = error, message
 * from vdvc_last_error()).  Every pointer named d_* is device memory on the
 * context's device; h_* pointers are host memory.  All work is enqueued on
 * the stream given at context creation; functions returning host values
 * synchronise that stream.
 *
 * Layouts (the reference's conventions, src/voxeldvc/engine/kernels_dvc.py:9-21,
 * with a padded row pitch):
 *   voxel volumes of a mesh level (images f, gradients, masks, sampled g,
 *                  residual): x fastest, index x + pitch*(y + Nvy*z),
 *                  Nv* = N*·h + 1, pitch = vdvc_voxel_pitch(Nvx) (a multiple
 *                  of 16 >= Nvx); the voxel order is the reference's
 *                  ravel(order='F') of an [x][y][z] array; padding voxels are
 *                  zero and never read as data;
 *   gradients:     three planes (gx, gy, gz), each a voxel volume;
 *   alignment:     voxel volumes should start on a 16-byte boundary (any
 *                  allocation does); the Hessian, diagonal, residual / RHS,
 *                  Jacobi and PCG entry points accept any element-aligned
 *                  base, but one that is not 16-byte aligned (an offset
 *                  view) runs the slower one-voxel-per-lane kernels;
 *   nodal vectors: DOF 3*(i + j*(Nx+1) + k*(Nx+1)*(Ny+1)) + c;
 *   deformed image g for sampling: x fastest, dims (gnx, gny, gnz).
 * Element types: VDVC_F32 float32, VDVC_BF16 bfloat16, VDVC_F16 float16,
 * VDVC_F64 float64; the post-solve stages also read the images' integer
 * dtypes VDVC_U8, VDVC_U16, VDVC_I16 and VDVC_BOOL.
 *
 * Post-solve stages (vdvc_interp_field ... vdvc_quality_cell) take dense
 * voxel arrays, no row pitch: index x + Nvx*(y + Nvy*z), the reference's
 * ravel(order='F').
 *
 * Working type.  Nodal vectors (v, U, b, x, diagonals, M_inv), sampled
 * images, residuals and the node partials are float64 on the FP64 path and
 * float32 otherwise.  The path is selected by the storage type: gradients
 * (gtype) or deformed image (gitype) VDVC_F64 select float64 (the
 * reference's dtype=float64 solve, engine/correlate_gpu.py:910-914), every
 * other storage type float32; node-only functions take it as `rtype`
 * (VDVC_F32 or VDVC_F64).  Scalars (E, eps_reg, ll, tol, means) are passed
 * as double and rounded once to the working type.
 *
 * The reference has no FFI for this path; its boundary is the Python
 * function API of voxeldvc.engine with array-module injection.  Each entry
 * below names the reference function it replaces; voxeldvc_amd.engine binds
 * them through ctypes with the reference signatures.
 */
#ifndef VDVC_H
#define VDVC_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VDVC_ABI_VERSION 12

enum { VDVC_F32 = 0, VDVC_BF16 = 1, VDVC_F16 = 2, VDVC_F64 = 3 };
/* image element types of the post-solve stages (the images' own dtypes;
 * VDVC_BOOL is one byte, 0 or 1) */
enum { VDVC_U8 = 4, VDVC_U16 = 5, VDVC_I16 = 6, VDVC_BOOL = 7 };
enum { VDVC_REG_NONE = 0, VDVC_REG_K = 1, VDVC_REG_EQGAP = 2 };

typedef struct vdvc_ctx vdvc_ctx;

int vdvc_abi_version(void);
const char* vdvc_last_error(void);

/* Row pitch (elements) of the voxel volumes of a mesh level with Nvx voxels
 * per row. */
int vdvc_voxel_pitch(int nvx);

/* Context: one device, one stream (a hipStream_t passed as an integer; 0 is
 * the null stream), library-owned scratch.  One host thread per context.
 * (The reference runs on CuPy's null stream, run_dvc.py:134.) */
int vdvc_ctx_create(int device, uintptr_t stream, vdvc_ctx** out);
int vdvc_ctx_destroy(vdvc_ctx* ctx);
int vdvc_synchronize(vdvc_ctx* ctx);

/* Measurement hook (no reference counterpart): between begin and end, every
 * launch of the voxel-streaming kernel (Hessian matvec / diagonal / RHS,
 * including those inside vdvc_pcg), of vdvc_copy and of vdvc_stream_read is
 * bracketed by HIP events on the context stream; end synchronises and returns
 * the summed kernel time and the number of launches. */
int vdvc_timing_begin(vdvc_ctx* ctx);
int vdvc_timing_end(vdvc_ctx* ctx, double* h_total_ms, int* h_launches);

/* Streaming-read probe (no reference counterpart: it measures the HBM read
 * ceiling the voxel kernels are compared against).  `grid` blocks of 256
 * lanes read nbytes (a positive multiple of 16, d_src 16-byte aligned) with
 * 16-byte nontemporal loads, four in flight per lane, and fold them with
 * xor; block b writes its fold to d_out[b] only if it equals `key` (d_out
 * holds `grid` words), so for a key that is not a fold nothing is written. */
int vdvc_stream_read(vdvc_ctx* ctx, int64_t nbytes, const void* d_src, int grid, uint32_t key,
                     uint32_t* d_out);

/* Reference-image setup: np.gradient per axis (edge_order=1), mask0 = f > glt,
 * gradient zeroed where the 6-cross erosion of mask0 is false.
 * Replaces image_grad_and_values + erode_mask0_vox + the zeroing in
 * correlate_gpu (engine/correlate_gpu.py:49-94, 951-953).
 * d_f: ftype (F32/F16, or F64 with F64 gradients); gradient planes: gtype
 * (F32/BF16/F64); d_mask0: uint8. */
int vdvc_image_setup(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_f, int ftype,
                     double glt, void* d_gx, void* d_gy, void* d_gz, int gtype,
                     uint8_t* d_mask0);

/* Multiplicity-weighted masked (count, sum, sum of squares) of a voxel field
 * in FP64.  Mask: d_mask (uint8) or, if use_thresh, field > glt; d_mask NULL
 * and !use_thresh means all voxels.  Replaces the reductions of
 * masked_quadpoint_mean_std (engine/correlate_gpu.py:113-157). */
int vdvc_masked_stats(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_field,
                      int ftype, const uint8_t* d_mask, int use_thresh, double glt,
                      double* h_out3);

/* The reference image's upload and mask0 statistics in one pass: d_src, a
 * dense x-fastest volume of (Nx*h+1, Ny*h+1, Nz*h+1) voxels of stype, is
 * rounded once to dtype and written to d_dst in the pitched level layout
 * (padding columns zero), and h_out3 receives the multiplicity-weighted
 * (count, sum, sum of squares) of the stored values above glt, as
 * vdvc_masked_stats with use_thresh on d_dst would (the summation order
 * differs).  Replaces the conversion of f_pix to the working dtype and the
 * mean0 / std0 reduction of correlate_gpu (engine/correlate_gpu.py:938,
 * 952, 961; masked_quadpoint_mean_std at :113-157).  Type pairs (stype,
 * dtype): (F32, F16), (F32, F32), (F64, F64), (F64, F32). */
int vdvc_level_image(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_src, int stype,
                     void* d_dst, int dtype, double glt, double* h_out3);

/* g(x + u(x)) at every voxel, active mask = in_bounds & mask0 & (g > glt), and
 * the multiplicity-weighted (count, sum, sum sq) of g over d_bmask (or over
 * the new mask when d_bmask is NULL).  Replaces sample_deformed_field
 * (engine/correlate_gpu.py:201-281) fused with the b-mask statistics of
 * correlate_gpu (:1091).  d_g gitype (F32/F16/F64); d_U and d_gvals in the
 * working type of gitype.  Interpolation weights: f32_weights = 0 forms them
 * and the trilinear combination in float64 in scipy's map_coordinates order
 * (the reference's values bit for bit); 1 in float32 (the mixed-precision
 * form; float32 / float16 images only). */
int vdvc_sample(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_U, const void* d_g,
                int gitype, int gnx, int gny, int gnz, double gox, double goy, double goz,
                double glt, const uint8_t* d_mask0, const uint8_t* d_bmask, void* d_gvals,
                uint8_t* d_mask_out, double* h_stats3, int f32_weights);

/* Residual res = ((g - gmean) * kscale + (f - mean0)) zeroed outside d_bmask
 * (working type, the reference's operation order; stored into d_res in the
 * level layout unless d_res is NULL), RHS b = sum W bmask res v_eq,
 * and (weighted count, sum, sum sq, plain voxel count) of res over d_live,
 * multiplicity-weighted except the last, into h_stats4 -- only when
 * h_stats4 is non-NULL (the reference forms them only to print them,
 * engine/correlate_gpu.py:1132-1136; d_live is then not read).  If d_fold_x is non-NULL
 * also writes bfloat16 planes g * bmask.  Replaces correlate_gpu.py:1092-1109
 * + compute_b_gpu (:341-378) + the display statistics (:1133). */
int vdvc_residual_rhs(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_gx,
                      const void* d_gy, const void* d_gz, int gtype, const void* d_f, int ftype,
                      double mean0, const void* d_gvals, double gmean, double kscale,
                      const uint8_t* d_bmask, const uint8_t* d_live, void* d_res, void* d_b,
                      uint16_t* d_fold_x, uint16_t* d_fold_y, uint16_t* d_fold_z,
                      double* h_stats4);

/* vdvc_residual_rhs followed by vdvc_jacobi for the H of a dynamic mask (the
 * refolded planes d_fold_* when given, else the gradients with d_bmask), in
 * one voxel pass where the library has it (float32 image and gradients
 * without refold, or float16 image and bfloat16 gradients with refold, on
 * its vectorised plan): the diagonal's voxel sums are taken beside b's from
 * the same gradient loads.  Outputs as the two calls; the diagonal equals
 * vdvc_jacobi's to float32 rounding (its partial sums follow the residual
 * pass's tiling).  Replaces one iteration's b, diag(H) and M_inv of
 * correlate_gpu (engine/correlate_gpu.py:1092-1120). */
int vdvc_residual_rhs_jacobi(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_gx,
                             const void* d_gy, const void* d_gz, int gtype, const void* d_f,
                             int ftype, double mean0, const void* d_gvals, double gmean,
                             double kscale, const uint8_t* d_bmask, const uint8_t* d_live,
                             void* d_res, void* d_b, uint16_t* d_fold_x, uint16_t* d_fold_y,
                             uint16_t* d_fold_z, double* h_stats4, int reg_kind,
                             const double* h_K1, const double* h_K2, double E, double eps_reg,
                             double ll, const void* d_extra, void* d_diag, void* d_minv);

/* The residual of vdvc_residual_rhs, (g - g_mean) * kscale + (f - mean0) in
 * the working type rtype (the type of d_gvals) and zero outside d_bmask,
 * formed from the same level-layout inputs with the same single-rounded
 * steps (bit for bit the d_res of that call) and written in the dense layout
 * x + Nvx (y + Nvy z) (d_res of Nvox entries).  d_res of vdvc_residual_rhs
 * may be NULL: a GN loop then stores no residual per iteration and forms the
 * returned one (the last iteration's) once with this call.  The residual of
 * correlate_gpu.py:1092-1109 as the solve returns it (:1138). */
int vdvc_residual_out(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_f, int ftype,
                      double mean0, const void* d_gvals, int rtype, double gmean, double kscale,
                      const uint8_t* d_bmask, void* d_res);

/* out = H v, H = sum_e sum_q w_q m_q v_eq v_eq^T.  d_mask may be NULL (mask
 * folded into the gradients).  Replaces matvec_H (engine/kernels_dvc.py:452-506). */
int vdvc_matvec_H(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_gx,
                  const void* d_gy, const void* d_gz, int gtype, const uint8_t* d_mask,
                  const void* d_v, void* d_out);

/* out = H v exactly as vdvc_matvec_H computes it (same kernels, same
 * result bit for bit), without allocating: the node partials live in a
 * library-static device array of 64 MiB per GPU (taken when the first
 * context on that GPU is created and held until the process ends) instead
 * of context scratch, and no timing events are recorded.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * Returns -1 when the mesh needs more than 64 MiB of partials (every
 * 513^3-voxel mesh at h >= 4 fits).  Nothing in the array has to start at
 * zero; every call on a GPU shares that GPU's array, so calls that overlap
 * on one GPU (any context or stream) race. */
int vdvc_matvec_H_noalloc(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_gx,
                          const void* d_gy, const void* d_gz, int gtype, const uint8_t* d_mask,
                          const void* d_v, void* d_out);

/* out = diag(H).  Replaces build_diagonal_H (engine/kernels_dvc.py:509-543). */
int vdvc_diag_H(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_gx, const void* d_gy,
                const void* d_gz, int gtype, const uint8_t* d_mask, void* d_out);

/* out = sum_e E_e P_e^T K P_e v with a 24x24 element matrix (h_K, float64
 * row-major, applied in the working type rtype) and E_e = E or d_E[e].
 * Replaces matvec_K_ref (engine/kernels_dvc.py:349-374). */
int vdvc_matvec_K(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int rtype, const double* h_K, double E,
                  const void* d_E, const void* d_v, void* d_out);

/* out = diagonal of the assembled K.  Replaces build_diagonal_K_ref
 * (engine/kernels_dvc.py:377-392). */
int vdvc_diag_K(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int rtype, const double* h_K, double E,
                const void* d_E, void* d_out);

/* out = K (P_i (K v)), P_i the interior-node DOF projector.  Replaces
 * matvec_equilibrium_gap (engine/kernels_dvc.py:402-417). */
int vdvc_matvec_eqgap(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int rtype, const double* h_Kel,
                      double E, const void* d_E, const void* d_v, void* d_out);

/* Hutchinson estimate diag(R_m) ~ mean_p z_p * (R_m z_p) over n_probes
 * Rademacher probes d_probes (int8, n_probes x Ndof), accumulated in FP64,
 * clipped below at 1e-12 * max|d|.  Replaces build_diagonal_equilibrium_gap
 * (engine/kernels_dvc.py:420-437); the caller supplies the reference's probe
 * stream. */
int vdvc_hutchinson_eqgap(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int rtype, const double* h_Kel,
                          double E, const int8_t* d_probes, int n_probes, void* d_out);

/* Regularizer used by vdvc_jacobi / vdvc_pcg:
 *   VDVC_REG_NONE:  R = 0
 *   VDVC_REG_K:     R = assembled E*K1            (Laplacian: K1 = K_ref_laplacian)
 *   VDVC_REG_EQGAP: R = K1 P_i K1 + eps_reg*K2    (K1 elastic, K2 Laplacian)
 * (engine/correlate_gpu.py:974-988). */

/* diag = diag(H) + ll*diag(R) (+ ll*d_extra for REG_EQGAP, the precomputed
 * Hutchinson diagonal of K1 P_i K1) and M_inv = 1/(diag + 1e-5*min diag), on
 * device.  Replaces the Jacobi set-up of correlate_gpu.py:1042-1047,
 * 1113-1120. */
int vdvc_jacobi(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_gx, const void* d_gy,
                const void* d_gz, int gtype, const uint8_t* d_mask, int reg_kind,
                const double* h_K1, const double* h_K2, double E, double eps_reg, double ll,
                const void* d_extra, void* d_diag, void* d_minv);

/* Jacobi-PCG for (H + ll R) x = b from x0 = 0, FP64 dots, stop when
 * sqrt(r.r/(b.b + 1e-30)) < tol, device-resident loop (the host polls the
 * stop flag first after check_every iterations, then after the iterations
 * the residual's rate since the last poll needs, at most 4 * check_every;
 * the schedule changes no result).  Replaces pcg
 * (engine/kernels_dvc.py:550-583) with the matvec_full closure of
 * correlate_gpu.py:1122-1126. */
int vdvc_pcg(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_gx, const void* d_gy,
             const void* d_gz, int gtype, const uint8_t* d_mask, int reg_kind,
             const double* h_K1, const double* h_K2, double E, double eps_reg, double ll,
             const void* d_b, const void* d_minv, void* d_x, double tol, int max_iter,
             int check_every, int* h_iters, double* h_relres);

/* out = H v + ll R v (the PCG operator).  Replaces matvec_full
 * (engine/correlate_gpu.py:1122-1126). */
int vdvc_apply_A(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_gx, const void* d_gy,
                 const void* d_gz, int gtype, const uint8_t* d_mask, int reg_kind,
                 const double* h_K1, const double* h_K2, double E, double eps_reg, double ll,
                 const void* d_v, void* d_out);

/* FP64 dot product of two vectors of type rtype.  Replaces xp.dot /
 * linalg.norm (engine/kernels_dvc.py:558-578, engine/correlate_gpu.py:1131). */
int vdvc_dot(vdvc_ctx* ctx, int64_t n, int rtype, const void* d_a, const void* d_b,
             double* h_out);

/* The plane-wave calibration mode of the regularization weight, in rtype:
 * out[3n] = cos(j*h/T*2*pi) at node n = i + (Nx+1)*(j + (Ny+1)*k), T = 10*h,
 * evaluated one rounded operation at a time from the left; out[3n+1] =
 * out[3n+2] = 0.  Replaces the V construction of correlate_gpu.py:1016-1020. */
int vdvc_calibration_mode(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, int rtype, void* d_out);

/* h_out2 = (a.b, a.c) in FP64, each equal bit for bit to vdvc_dot of the
 * pair, in one pass and one host round trip.  Replaces the two dot products
 * of the ll calibration, H0 = V.HV and L0 = V.LV (engine/correlate_gpu.py:
 * 1022-1025). */
int vdvc_dot2(vdvc_ctx* ctx, int64_t n, int rtype, const void* d_a, const void* d_b,
              const void* d_c, double* h_out2);

/* The GN update and its convergence norms: U <- U + dU in rtype, then
 * h_out2 = (dU.dU, U.U) in FP64, equal bit for bit to vdvc_axpby(1, dU, 1, U)
 * followed by vdvc_dot(dU, dU) and vdvc_dot(U, U).  Replaces U = U + dU and
 * the two norms of err = ||dU|| / ||U|| (engine/correlate_gpu.py:1129-1131). */
int vdvc_gn_update(vdvc_ctx* ctx, int64_t n, int rtype, const void* d_dU, void* d_U,
                   double* h_out2);

/* y = a*x + b*y in rtype. */
int vdvc_axpby(vdvc_ctx* ctx, int64_t n, int rtype, double a, const void* d_x, double b,
               void* d_y);

/* out = a * b elementwise in rtype; the Jacobi application z = M_inv r of
 * pcg (engine/kernels_dvc.py:556, 577). */
int vdvc_mul(vdvc_ctx* ctx, int64_t n, int rtype, const void* d_a, const void* d_b, void* d_out);

/* Folded bfloat16 gradient planes o = g * mask (raw bf16 bits), the storage
 * the Hessian kernels read when d_mask is NULL.  No reference counterpart:
 * the reference keeps gradient and mask apart (engine/kernels_dvc.py:490-500);
 * H uses m g g^T with m in {0,1}, so folding is exact. */
int vdvc_fold_grad(vdvc_ctx* ctx, int64_t nvox, const uint16_t* d_gx, const uint16_t* d_gy,
                   const uint16_t* d_gz, const uint8_t* d_mask, uint16_t* d_ox, uint16_t* d_oy,
                   uint16_t* d_oz);

/* 6-cross erosion of a uint8 volume (x fastest) with edge replication.
 * Replaces erode_mask0_vox (engine/correlate_gpu.py:75-94). */
int vdvc_erode6(vdvc_ctx* ctx, int nx, int ny, int nz, const uint8_t* d_in, uint8_t* d_out);

/* Stride-s decimation of a volume, dst dims ((n-1)/s + 1 per axis).
 * Replaces f_pix[::s, ::s, ::s] (engine/correlate_gpu.py:1257-1258). */
int vdvc_stride(vdvc_ctx* ctx, const void* d_src, int type, int nx, int ny, int nz, int s,
                void* d_dst);

/* Phase correlation of two float32 volumes of equal dims (x fastest):
 * returns in *h_cidx the argmax of ifftn(F conj(G)/(|.|+1e-12)).real as a
 * C-order index ((x*ny + y)*nz + z); each axis <= 4096, at most 2^26
 * voxels.  Replaces the FFT part of fft_rigid_shift
 * (engine/correlate_gpu.py:393-405). */
int vdvc_phase_corr(vdvc_ctx* ctx, const float* d_f, const float* d_g, int nx, int ny, int nz,
                    int64_t* h_cidx);

/* Affine warp A(x - c) + c + t of every voxel of f, order-1 sample of g, and
 * the FP64 sums (n, sum f, sum g, sum f^2, sum g^2, sum f g) over in-bounds
 * samples.  Replaces warp_and_cost / _zncc_masked
 * (engine/correlate_gpu.py:417-433, 523-536). */
int vdvc_affine_sums(vdvc_ctx* ctx, const float* d_f, int nx, int ny, int nz, const float* d_g,
                     int gnx, int gny, int gnz, const double* h_A9, const double* h_t3,
                     const double* h_c3, double* h_out6);

/* ---- learned seed (no reference counterpart) ------------------------------
 * A small 3-D convolutional network predicts coarse nodal displacements that
 * seed the finest-scale GN solve in place of the coarse pyramid levels of
 * multiscale_correlate_gpu (engine/correlate_gpu.py:1143-1305).  Tensors are
 * float32, channels last: activations [n][z][y][x][c], weights [o][tap][c]
 * with tap = (dz*3 + dy)*3 + dx; 3x3x3 kernels, zero padding 1, stride 1 or
 * 2 (output extent (n - 1)/stride + 1).  Products run on the matrix cores
 * with bfloat16 operands and float32 accumulation.  C, O <= 32. */

/* y = conv(x, w) + b, optional ReLU. */
int vdvc_conv3d_forward(vdvc_ctx* ctx, int N, int D, int H, int W, int C, int O, int stride,
                        const float* d_x, const float* d_w, const float* d_b, int relu,
                        float* d_y);

/* dx = conv^T(dy, w), times (d_act > 0) when d_act (the layer input, a ReLU
 * output) is given. */
int vdvc_conv3d_backward_data(vdvc_ctx* ctx, int N, int D, int H, int W, int C, int O,
                              int stride, const float* d_dy, const float* d_w,
                              const float* d_act, float* d_dx);

/* dw = sum_p dy (x) x-patch, db = sum_p dy (fixed-order reductions). */
int vdvc_conv3d_backward_weight(vdvc_ctx* ctx, int N, int D, int H, int W, int C, int O,
                                int stride, const float* d_x, const float* d_dy, float* d_dw,
                                float* d_db);

/* One bias-corrected Adam step (step >= 1) on n float32 parameters. */
int vdvc_adam_step(vdvc_ctx* ctx, int64_t n, float* d_p, const float* d_g, float* d_m,
                   float* d_v, double lr, double beta1, double beta2, double eps, int step);

/* Mean squared error over n entries: *h_loss (FP64 sum / n) and
 * d_grad = 2 (pred - label) / n. */
int vdvc_mse(vdvc_ctx* ctx, int64_t n, const float* d_pred, const float* d_label, float* d_grad,
             double* h_loss);

/* Network input [z][y][x][2] from the stride-s decimation of f and g
 * (itype F32 or F16, x-fastest dense volumes of equal dims), each channel
 * standardised by its own mean and standard deviation. */
int vdvc_seed_input(vdvc_ctx* ctx, const void* d_f, const void* d_g, int itype, int nx, int ny,
                    int nz, int s, float* d_out);

/* ---- post-solve voxel stages (dense voxel arrays) ---- */

/* Nodal field U at every voxel with the voxel-centre Hex8 stencil; d_out is
 * (Nvox, 3) or, with channels_first, (3, Nvox), in U's type (rtype F32 /
 * F64).  Replaces interp_field_vox (engine/correlate_gpu.py:160-198). */
int vdvc_interp_field(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_U, int rtype,
                      void* d_out, int channels_first);

/* Active mask mask0 & g.InBounds(x + u - origin) & g(x + u - origin) > glt,
 * g sampled trilinearly in its own dtype (gtype).  d_mask0 may be NULL (all
 * true); d_mask_out is one byte per voxel.  Replaces compute_active_mask
 * (engine/correlate_gpu.py:284-338). */
int vdvc_active_mask(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_U, int rtype,
                     const void* d_g, int gtype, int gnx, int gny, int gnz, double gox, double goy,
                     double goz, const uint8_t* d_mask0, double glt, uint8_t* d_mask_out);

/* d_out(x) = field(x - u(x)) on the (Nvx, Nvy, Nvz) grid, in the field's type
 * (ftype), with scipy.ndimage.map_coordinates semantics for order 0 / 1 and
 * mode 0 'nearest' / 1 'constant' (cval).  Replaces push_field_forward /
 * push_reference_forward (engine/write_output.py:571-622). */
int vdvc_push_forward(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_U, int rtype,
                      const void* d_field, int ftype, int order, int mode, double cval,
                      void* d_out);

/* Zero-normalised cross-correlation of f and g (or of f and f - g when
 * g_is_residual) over d_mask (NULL: all n entries), FP64 two-pass; *h_out is
 * NaN for an empty or constant selection.  Replaces compute_zncc
 * (engine/write_output.py:658-690) and its caller's g_warped = f - res
 * (:786-787). */
int vdvc_zncc(vdvc_ctx* ctx, int64_t n, const void* d_f, int ftype, const void* d_g, int gtype,
              int g_is_residual, const uint8_t* d_mask, double* h_out);

/* Per-element population std of the residual and ZNCC of (ref, ref - res)
 * over the element's (h+1)^3 voxels with res != 0; NaN below 8 active voxels
 * (and ZNCC for zero variance).  d_res rtype (F32 / F64), d_ref ftype; d_std
 * and d_zncc are (Nx*Ny*Nz,) float64 in element order i + Nx*(j + Ny*k).
 * Replaces compute_correlation_quality_cell (engine/write_output.py:380-447). */
int vdvc_quality_cell(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_res, int rtype,
                      const void* d_ref, int ftype, double* d_std, double* d_zncc);

/* ---- element-scale write-output fields (dense voxel arrays, elements
 * i + Nx*(j + Ny*k), nodes i + (Nx+1)*(j + (Ny+1)*k); uint8 masks 0 / 1) ----
 *
 * Unsafe and overlap-lost masks from the per-voxel active mask d_active and
 * material mask d_mask0 (ref > glt).  Outputs are optional (NULL: skipped):
 *   d_unsafe_elem (Ne,): shares a node with an element holding an inactive
 *     voxel, or lies on the outer element shell -- compute_unsafe_element_mask
 *     (engine/kernels_dvc.py:196-213, via _unsafe_elements :118-163);
 *   d_unsafe_node (Nn,): belongs to an element holding an inactive voxel or
 *     to a shell element -- compute_unsafe_node_mask (:216-233);
 *   d_unsafe_vox (Nvox,): voxel of an unsafe element --
 *     compute_unsafe_voxel_mask (:166-193);
 *   d_overlap_lost (Ne,): shares a node with an element whose material lost
 *     overlap (lost / material > frac_threshold), or has material fraction
 *     below min_material_frac, or lies on the shell -- needs d_mask0;
 *     compute_overlap_lost_element_mask (:236-340). */
int vdvc_cell_masks(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const uint8_t* d_active,
                    const uint8_t* d_mask0, double frac_threshold, double min_material_frac,
                    uint8_t* d_unsafe_elem, uint8_t* d_unsafe_node, uint8_t* d_unsafe_vox,
                    uint8_t* d_overlap_lost);

/* d_safe (Ne,) = 1 where none of the element's (h+1)^3 voxels is set in
 * d_unsafe_vox.  Replaces compute_safe_cell_mask
 * (engine/write_output.py:352-377). */
int vdvc_safe_cells(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const uint8_t* d_unsafe_vox,
                    uint8_t* d_safe);

/* Principal small strains e1 >= e2 >= e3 at each element centre from the
 * nodal displacement d_U (utype F32 / F64): Hex8 gradients at the centre,
 * eps = (grad u + grad u^T) / 2, eigenvalues in float64; d_principal
 * (Ne, 3) float64, d_equiv (Ne,) float64 von Mises equivalent strain
 * (optional).  Replaces compute_principal_strains_cell and
 * compute_equivalent_strain (engine/write_output.py:256-306). */
int vdvc_principal_strains(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_U,
                           int utype, double* d_principal, double* d_equiv);

/* Equivalent strain of n principal-strain triples (float64).  Replaces
 * compute_equivalent_strain (engine/write_output.py:284-306). */
int vdvc_equivalent_strain(vdvc_ctx* ctx, long long n, const double* d_principal,
                           double* d_out);

/* d_out (Ne,) float64 = ||dU|| / ||U|| over each element's 8 corner nodes'
 * 24 DOFs, NaN where ||U|| = 0.  Replaces compute_gn_convergence_cell
 * (engine/write_output.py:450-491). */
int vdvc_gn_convergence(vdvc_ctx* ctx, int Nx, int Ny, int Nz, const void* d_dU, int dtype,
                        const void* d_U, int utype, double* d_out);

/* Element mask d_mask (Ne,) carried to the deformed configuration: element
 * centres (i + 0.5) h moved by the mean of their 8 corner displacements;
 * d_out at each regular centre + h_crop[3] takes the value of the nearest
 * moved centre (exact search over a uniform grid of cells, float64 squared
 * distances; ties: lowest element index).  Replaces remap_element_mask_to_deformed
 * (engine/write_output.py:494-530, scipy griddata 'nearest'). */
int vdvc_remap_mask_deformed(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_U,
                             int utype, const double* h_crop, const uint8_t* d_mask,
                             uint8_t* d_out);

/* Per-element field d_field (Ne, C) float64, element-major (e*C + c), carried
 * to the deformed configuration: element centres moved as in
 * vdvc_remap_mask_deformed; each regular centre + h_crop takes the linear
 * interpolation of the field over the tetrahedron of the moved-centre
 * lattice that contains it (each lattice cell split into the 6 tetrahedra
 * around its (0,0,0)-(1,1,1) diagonal, a conforming piecewise-linear
 * interpolant), or, outside the moved lattice, the value of the nearest
 * moved centre.  Replaces remap_cell_field_to_deformed
 * (engine/write_output.py:533-568: scipy griddata 'linear' with a
 * 'nearest' fill), whose Qhull Delaunay tetrahedra can differ from these
 * inside a cell: both reproduce linear fields exactly and agree to the
 * field's second differences otherwise; 1 <= C <= 64. */
int vdvc_remap_field_deformed(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_U,
                              int utype, const double* h_crop, const double* d_field, int C,
                              double* d_out);

/* d_out (Ne,) = 1 where the mean of the reference image (dense voxel array,
 * its own dtype itype) over the element's (h+1)^3 voxels exceeds glt (float64
 * sums: exact for integer images).  Replaces compute_material_cell_mask
 * (engine/write_output.py:309-338). */
int vdvc_material_cells(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_ref,
                        int itype, double glt, uint8_t* d_out);

/* One-element erosion of an element mask (uint8, elements x fastest) over
 * the 26-neighbourhood, the mesh border counting as set.  Replaces
 * compute_material_cell_mask_eroded (engine/write_output.py:341-349, scipy
 * binary_erosion with a 3x3x3 structure and border_value=1). */
int vdvc_erode_cells(vdvc_ctx* ctx, int Nx, int Ny, int Nz, const uint8_t* d_in, uint8_t* d_out);

/* ---- preprocess affine stage (dense float32 volumes, x fastest) ---- */

/* scipy.ndimage.gaussian_filter of a float32 volume: one symmetric
 * correlation per axis (0, 1, 2) in float64, SciPy's order, each pass stored
 * float32; h_w = the r + 1 weights centre-first (the normalised Gaussian of
 * radius r = int(4 sigma + 0.5)), mode 0 'nearest' / 1 'constant' (cval).
 * Replaces the gaussian_filter calls of _smooth_plain / _smooth_masked
 * (engine/correlate_gpu.py:570-598). */
int vdvc_gaussian_filter(vdvc_ctx* ctx, int nx, int ny, int nz, const float* d_in, float* d_out,
                         const double* h_w, int r, int mode, double cval);

/* Normalised-convolution smoothing of g: smooth(g valid) / smooth(valid)
 * ('constant' 0) where the denominator exceeds 1e-6, g elsewhere, with
 * valid = (g != cval) also written to d_valid (optional); r < 0 writes only
 * d_valid.  Replaces _smooth_masked (engine/correlate_gpu.py:586-598) and the
 * fit mask valid_g (:697). */
int vdvc_smooth_masked(vdvc_ctx* ctx, int nx, int ny, int nz, const float* d_g, double cval,
                       const double* h_w, int r, float* d_out, uint8_t* d_valid);

/* np.gradient of a float32 volume (unit spacing, edge order 1, dims >= 2).
 * Replaces xp.gradient(f_s) of affine_refine_gn (engine/correlate_gpu.py:749). */
int vdvc_gradient3(vdvc_ctx* ctx, int nx, int ny, int nz, const float* d_f, float* d_gx,
                   float* d_gy, float* d_gz);

/* One Gauss-Newton step's normal equations of the affine refinement:
 * H = sum J J^T (12 x 12, row-major) and b = sum J e in float64 over all
 * voxels, J and e in float32 at s = x - M (x - c) - t_r (M, t_r rounded to
 * float32), e = (g - f(s)) mask, mask = in-bounds & valid.  Replaces the slab
 * loop of affine_refine_gn (engine/correlate_gpu.py:755-793). */
int vdvc_affine_normal(vdvc_ctx* ctx, int nx, int ny, int nz, const float* d_f, const float* d_g,
                       const float* d_gx, const float* d_gy, const float* d_gz,
                       const uint8_t* d_valid, const double* h_M, const double* h_tr,
                       const double* h_c, double cval, double* h_H, double* h_b);

/* (sum e^2, count) of the masked residual e = g - f(s) over in-bounds voxels
 * with valid g.  Replaces _rms of affine_refine_gn
 * (engine/correlate_gpu.py:725-743). */
int vdvc_affine_rms(vdvc_ctx* ctx, int nx, int ny, int nz, const float* d_f, const float* d_g,
                    const uint8_t* d_valid, const double* h_M, const double* h_tr,
                    const double* h_c, double cval, double* h_out2);

/* d_out (nx, ny, nz) = g (gnx, gny, gnz) sampled at A (x - c) + c + t
 * (float32; A, t, c rounded to float32), trilinear, 0 outside.  Replaces
 * warp_onto_ref_grid (preprocess.py:233-285). */
int vdvc_warp_affine(vdvc_ctx* ctx, int nx, int ny, int nz, const float* d_g, int gnx, int gny,
                     int gnz, const double* h_A, const double* h_t, const double* h_c,
                     float* d_out);

/* Bounding box (first and last index per axis, -1 when empty) of
 * (f > glt) & (g_warped > glt), f in its own dtype (ftype).  Replaces the
 * axis projections of compute_crop (preprocess.py:315-346). */
int vdvc_overlap_bounds(vdvc_ctx* ctx, int nx, int ny, int nz, const void* d_f, int ftype,
                        const float* d_gw, double glt, int* h_lo3, int* h_hi3);

/* ---- data movement at the solve's boundary --------------------------------
 * Element-type conversions and copies the engine needs between the caller's
 * arrays and the level layout.  They replace the xp.asarray(..., dtype)
 * conversions of the reference's engine boundary (engine/correlate_gpu.py:
 * 938-939), its per-level copies and the float64 promotion of its solve
 * vectors (:1042-1047); no arithmetic beyond the conversions.  Types: any
 * VDVC_* element type; conversions round to nearest even, float16 and
 * bfloat16 from float64 without double rounding; integer targets truncate. */

/* dst[i] = src[i] converted, i < n. */
int vdvc_convert(vdvc_ctx* ctx, int64_t n, const void* d_src, int stype, void* d_dst,
                 int dtype);

/* Byte copy of nbytes (non-overlapping buffers). */
int vdvc_copy(vdvc_ctx* ctx, int64_t nbytes, const void* d_src, void* d_dst);

/* dst[i] = value converted to dtype, i < n. */
int vdvc_fill(vdvc_ctx* ctx, int64_t n, int dtype, double value, void* d_dst);

/* Strided volume copy with conversion: dst[x + dpitch*(y + ny*z)] =
 * src[x*sx + y*sy + z*sz] (strides in elements) for x < nx, and 0 in the
 * row padding nx <= x < dpitch.  With the source strides of a C-order
 * [x][y][z] array (ny*nz, nz, 1) it is the reference's ravel(order='F')
 * into the level layout (engine/correlate_gpu.py:50-52, 60); ny, nz <= 65535. */
int vdvc_copy_volume(vdvc_ctx* ctx, int nx, int ny, int nz, const void* d_src, int stype,
                     int64_t sx, int64_t sy, int64_t sz, void* d_dst, int dtype, int dpitch);

/* Nodal field of an (Mx, My, Mz)-element mesh of element size hs resampled
 * onto an (Nx, Ny, Nz)-element mesh of element size hd covering the same
 * voxels (M*hs == N*hd per axis): each destination node takes the Hex8
 * (trilinear) interpolation of the source field at its voxel position
 * (float64 weights; node picks when hd is a multiple of hs).  Both vectors
 * in rtype, the reference's DOF order.  The U <- U * (prev_s / s) carry-over
 * between pyramid levels (engine/correlate_gpu.py:1263) keeps one mesh; this
 * is its counterpart between element sizes, used to seed a mesh from the
 * learned network's h = 8 prediction. */
int vdvc_resample_nodes(vdvc_ctx* ctx, int Mx, int My, int Mz, int hs, const void* d_src,
                        int Nx, int Ny, int Nz, int hd, void* d_dst, int rtype);

/* y = a*x + b*y with x of xtype (F32/F16/BF16/F64) and y of ytype (F32/F64),
 * evaluated in float64 and rounded once to ytype (a float32 correction
 * accumulated into a float64 vector is exact before that rounding). */
int vdvc_axpby_x(vdvc_ctx* ctx, int64_t n, double a, const void* d_x, int xtype, double b,
                 void* d_y, int ytype);

#ifdef __cplusplus
}
#endif
#endif /* VDVC_H */
