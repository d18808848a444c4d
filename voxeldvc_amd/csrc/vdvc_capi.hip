// C-ABI entry points (declared in include/vdvc.h): argument checks, scratch
// management, tile planning and kernel sequencing.  No compute happens on the
// host; every numeric step is one of the gfx950 kernels in this directory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/vdvc.h"
#include "vdvc_kernels.h"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define VDVC_HIP(call)                                                                    \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) return fail(-2, "%s: %s", #call, hipGetErrorString(e_));         \
  } while (0)
#define VDVC_LAUNCH(call)                                                                 \
  do {                                                                                    \
    int e_ = (call);                                                                      \
    if (e_ != 0) return fail(-2, "%s: %s", #call, hipGetErrorString((hipError_t)e_));     \
  } while (0)

// doubles of pinned host staging per context (largest host read-back: the
// 90 normal-equation sums of the affine refinement)
#define VDVC_HPIN 256

// grow-only device scratch buffer
struct Buf {
  void* p = nullptr;
  size_t n = 0;
  int ensure(size_t bytes) {
    if (bytes <= n) return 0;
    if (p) {
      hipError_t e = hipFree(p);
      p = nullptr;
      n = 0;
      if (e != hipSuccess) return fail(-2, "hipFree: %s", hipGetErrorString(e));
    }
    size_t want = bytes + bytes / 8 + 256;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(-3, "hipMalloc(%zu): %s", want, hipGetErrorString(e));
    }
    n = want;
    return 0;
  }
  template <typename T> T* as() const { return reinterpret_cast<T*>(p); }
  ~Buf() { if (p) (void)hipFree(p); }
};

}  // namespace

// The no-allocation entry points keep their scratch (the node partials of
// the voxel kernels) in this file-scope device array, one copy per GPU,
// held until the process ends.  Host code never passes `scratch` itself to
// a kernel (in host code the name is a host-side stand-in); it passes the
// device address hipGetSymbolAddress returned for the current device.
#define VDVC_NOALLOC_SCRATCH_BYTES (64u << 20)
static __device__ __align__(256) float scratch[VDVC_NOALLOC_SCRATCH_BYTES / sizeof(float)];

namespace {
#define VDVC_MAX_DEVICES 64
std::mutex g_scratch_mu;
void* g_scratch_ptr[VDVC_MAX_DEVICES] = {nullptr};

// device address of `scratch` on `device` (current), taken once per device
int scratch_address(int device, void** out) {
  if (device < 0 || device >= VDVC_MAX_DEVICES) return fail(-1, "device %d out of range", device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(-2, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return 0;
}
}  // namespace

struct vdvc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  void* scratch = nullptr;  // this device's address of the static scratch array
  Buf P;          // voxel-kernel node partials
  Buf Pd;         // the residual pass's Hessian-diagonal partials
  Buf stats;      // per-wave / per-block FP64 partials
  Buf red;        // reduced FP64 results
  Buf nodeA, nodeB, nodeC;  // node-vector temporaries
  Buf pr, pz, pp, pAp;      // PCG vectors
  Buf dparts, dparts2;      // FP64 dot partials
  Buf pstate;               // PcgState
  Buf fmin;                 // per-block minima
  Buf K1[2], K2[2];         // element matrices (24x24) in float / double
  Buf hacc;                 // Hutchinson FP64 accumulator
  Buf dft;                  // DFT work volumes
  Buf wpart;                // conv weight-gradient partials
  Buf vtmp[4];              // dense float32 volume temporaries (affine stage)
  double* hpin = nullptr;   // pinned host staging (VDVC_HPIN doubles)
  std::vector<double> k1_cache[2], k2_cache[2];
  // 8 x 8 float form of a block-diagonal K1 for the voxel kernels, and the
  // matrix it was formed from
  Buf kc32;
  std::vector<double> kc32_src;
  // optional per-launch timing of the voxel-streaming kernels (HIP events
  // recorded on the context stream around each launch)
  bool timing = false;
  std::vector<hipEvent_t> ev;
  size_t ev_used = 0;
};

namespace {

int check_mesh(int Nx, int Ny, int Nz, int h) {
  if (Nx < 1 || Ny < 1 || Nz < 1 || h < 1)
    return fail(-1, "invalid mesh Nx=%d Ny=%d Nz=%d h=%d", Nx, Ny, Nz, h);
  const long long nvox = (long long)(Nx * (long long)h + 1) * (Ny * (long long)h + 1) *
                         (Nz * (long long)h + 1);
  const long long ndof = 3LL * (Nx + 1) * (Ny + 1) * (Nz + 1);
  if (nvox > 0x7fffffffLL || ndof > 0x7fffffffLL)
    return fail(-1, "Nvox=%lld / Ndof=%lld exceed int32 range; matrix-free index arrays "
                    "assume int32 indices", nvox, ndof);
  return 0;
}

MeshDev make_mesh(int Nx, int Ny, int Nz, int h) {
  MeshDev M;
  M.Nx = Nx; M.Ny = Ny; M.Nz = Nz; M.h = h;
  M.Nvx = Nx * h + 1; M.Nvy = Ny * h + 1; M.Nvz = Nz * h + 1;
  M.pitch = vdvc_pitch_of(M.Nvx);
  M.nnx = Nx + 1; M.nny = Ny + 1; M.nnz = Nz + 1;
  M.Nvox = (long long)M.Nvx * M.Nvy * M.Nvz;
  M.Nvox_p = (long long)M.pitch * M.Nvy * M.Nvz;
  M.Nnodes = (long long)M.nnx * M.nny * M.nnz;
  M.inv_h = 1.0f / (float)h;
  return M;
}

// z tiling shared by both decompositions: enough wavefronts to fill 256 CUs
void plan_z(const MeshDev& M, TilesDev& T, long long target) {
  const long long base = (long long)T.nsx * M.Ny * T.nys;
  long long nzt_t = (target + base - 1) / base;
  if (nzt_t < 1) nzt_t = 1;
  int ez = (int)((M.Nz + nzt_t - 1) / nzt_t);
  if (ez < 1) ez = 1;
  T.ez = ez;
  T.nzt = (M.Nz + ez - 1) / ez;
  T.wave_stride = (long long)(ez + 1) * 2 * T.px * 3;
  T.nwaves = base * T.nzt;
}

// generic decomposition: 64-voxel x strips over [0, Nvx), one voxel per lane
TilesDev make_tiles_generic(const MeshDev& M) {
  TilesDev T;
  T.vec = 0;
  T.epl = 1;
  T.sw = 64;
  T.nys = 1;
  T.nsx = (M.Nvx + 63) / 64;
  T.px = 63 / M.h + 3;
  plan_z(M, T, 16384);
  return T;
}

// decomposition of the voxel-streaming Hessian kernels: vectorised strips of
// 64*V voxels over [0, Nx*h) when h is a power of two, else generic.
// `pairs`: the caller's kernel mode takes paired-element lanes (the products
// and the diagonal; not the residual pass).
TilesDev make_tiles(const MeshDev& M, int gtype, bool pairs) {
  int V = vdvc::vec_width(M.h, gtype);
  if (V == 0) return make_tiles_generic(M);
  // a row narrower than a 64 V-voxel strip would leave lanes idle: narrow
  // the lanes (V stays a power of two dividing h) until the strip fits
  while (V > 1 && 64 * V > M.Nx * M.h && getenv("VDVC_VEC_BF16") == nullptr &&
         getenv("VDVC_VEC_F32") == nullptr)
    V >>= 1;
  // paired elements: where a lane of one element holds h = 2 or 4 voxels and
  // a 16-byte lane holds two elements, a lane may take two (V = 2h) if every
  // lane owns two whole elements (Nx even).  Planned for bfloat16 at h = 4
  // when a strip fits the row (measured at 513^3: 191 against 203-210 us);
  // for float32 at h = 2 the paired lanes' registers cost more than they
  // save (257^3: 80 against 75 us).  Tuning override, read at every plan:
  // VDVC_EPL = 1 (never) or 2 (whenever the element count allows).
  const char* ep = getenv("VDVC_EPL");
  const int po = ep ? atoi(ep) : 0;
  int E = 1;
  if (pairs && po != 1 && V == M.h && (M.h == 2 || M.h == 4) &&
      vdvc::vec_width(2 * M.h, gtype) == 2 * M.h && M.Nx % 2 == 0 &&
      (po == 2 || (gtype == VDVC_BF16 && M.h == 4 && 64 * 2 * M.h <= M.Nx * M.h))) {
    V = 2 * M.h;
    E = 2;
  }
  TilesDev T;
  T.vec = V;
  T.epl = E;
  T.sw = 64 * V;
  T.nsx = (M.Nx * M.h + T.sw - 1) / T.sw;
  T.px = 64 * V / M.h + 2;
  T.nys = 1;
  plan_z(M, T, 8192);
  // too few wavefronts for the chip even with one element layer per wave
  // (fewer than one per SIMD slot at the kernel's occupancy of 3-4): split
  // each element row's rows between two wavefronts.  The split doubles the
  // node-partial traffic, so it is not used beyond that point.
  // tuning override of the row split, read at every plan: VDVC_NYS = 1 or 2
  const char* e = getenv("VDVC_NYS");
  const int ov = e ? atoi(e) : 0;
  const bool split = ov == 1 ? false : ov == 2 ? true : T.nwaves < 4096;
  if (split && M.h >= 4 && (M.h % 2) == 0) {
    T.nys = 2;
    plan_z(M, T, 8192);
  } else if (T.ez == 1 && T.nwaves >= 4096 && M.Nz >= 2 && (gtype == VDVC_BF16 || M.h <= 4)) {
    // two element layers per wavefront where one layer already gives at
    // least 4096 wavefronts: each interior node plane's partials are written
    // once instead of twice and the waves stay at >= 2 per SIMD.  Measured
    // at 64^3 elements: bf16 h = 8 161 -> 155 us, float32 h = 4 55.9 ->
    // 53.1 us, h = 2 20.8 -> 20.3 us per product; float32 at h = 8 was
    // 1 % slower with two layers (three waves per SIMD at its occupancy
    // want the larger grid), so it keeps one
    T.ez = 2;
    T.nzt = (M.Nz + 1) / 2;
    T.wave_stride = (long long)(T.ez + 1) * 2 * T.px * 3;
    T.nwaves = (long long)T.nsx * M.Ny * T.nys * T.nzt;
  }
  // tuning override of the element layers per wavefront, read at every
  // plan: VDVC_EZ = layers (>= 1)
  const char* ez = getenv("VDVC_EZ");
  const int eo = ez ? atoi(ez) : 0;
  if (eo >= 1) {
    T.ez = std::min(eo, M.Nz);
    T.nzt = (M.Nz + T.ez - 1) / T.ez;
    T.wave_stride = (long long)(T.ez + 1) * 2 * T.px * 3;
    T.nwaves = (long long)T.nsx * M.Ny * T.nys * T.nzt;
  }
  return T;
}

// the vectorised voxel kernels load (and, for the residual and the folded
// planes, store) 16-byte words at row starts, which sit on 16-element pitch
// boundaries; a streamed array whose base is not 16-byte aligned (an offset
// view) takes the generic one-voxel-per-lane decomposition instead
bool streams_aligned(const VoxIn& in) {
  auto ok = [](const void* p) { return ((uintptr_t)p & 15u) == 0; };
  return ok(in.g[0]) && ok(in.g[1]) && ok(in.g[2]) && ok(in.mask) && ok(in.fimg) &&
         ok(in.gvals) && ok(in.live) && ok(in.res) && ok(in.fold_out[0]) &&
         ok(in.fold_out[1]) && ok(in.fold_out[2]);
}

TilesDev tiles_for(const MeshDev& M, int gtype, const VoxIn& in, bool pairs) {
  TilesDev T = make_tiles(M, gtype, pairs);
  if (T.vec && !streams_aligned(in)) T = make_tiles_generic(M);
  return T;
}

int gtype_ok(int t) { return t == VDVC_F32 || t == VDVC_BF16 || t == VDVC_F64; }
int ftype_ok(int t) { return t == VDVC_F32 || t == VDVC_F16 || t == VDVC_F64; }
int rtype_ok(int t) { return t == VDVC_F32 || t == VDVC_F64; }

// index of the working type in per-type caches
template <typename R> constexpr int rix() { return sizeof(R) == 8 ? 1 : 0; }

// element matrix in the working type R, uploaded once per distinct matrix
template <typename R>
int upload_K(vdvc_ctx* ctx, Buf& dst, std::vector<double>& cache, const double* hK,
             int* blockdiag) {
  if (!hK) return fail(-1, "element matrix is NULL");
  bool same = cache.size() == 576 && std::memcmp(cache.data(), hK, 576 * sizeof(double)) == 0;
  bool bd = true;
  for (int a = 0; a < 8 && bd; ++a)
    for (int b = 0; b < 8 && bd; ++b)
      for (int c = 0; c < 3 && bd; ++c)
        for (int d = 0; d < 3 && bd; ++d) {
          const double v = hK[(3 * a + c) * 24 + 3 * b + d];
          if (c != d && v != 0.0) bd = false;
          if (c == d && v != hK[(3 * a) * 24 + 3 * b]) bd = false;
        }
  *blockdiag = bd ? 1 : 0;
  if (same && dst.p) return 0;
  int rc = dst.ensure(576 * sizeof(R));
  if (rc) return rc;
  R tmp[576];
  for (int q = 0; q < 576; ++q) tmp[q] = (R)hK[q];
  VDVC_HIP(hipMemcpyAsync(dst.p, tmp, sizeof(tmp), hipMemcpyHostToDevice, ctx->stream));
  VDVC_HIP(hipStreamSynchronize(ctx->stream));
  cache.assign(hK, hK + 576);
  return 0;
}

template <typename R>
int upload_K1(vdvc_ctx* ctx, const double* hK, int* bd) {
  return upload_K<R>(ctx, ctx->K1[rix<R>()], ctx->k1_cache[rix<R>()], hK, bd);
}
template <typename R>
int upload_K2(vdvc_ctx* ctx, const double* hK, int* bd) {
  return upload_K<R>(ctx, ctx->K2[rix<R>()], ctx->k2_cache[rix<R>()], hK, bd);
}

int fetch(vdvc_ctx* ctx, const void* d, size_t bytes, void* h) {
  if (bytes > VDVC_HPIN * sizeof(double)) return fail(-1, "fetch too large");
  VDVC_HIP(hipMemcpyAsync(ctx->hpin, d, bytes, hipMemcpyDeviceToHost, ctx->stream));
  VDVC_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(h, ctx->hpin, bytes);
  return 0;
}

VoxIn vox_in(const void* gx, const void* gy, const void* gz, const uint8_t* mask) {
  VoxIn in;
  std::memset(&in, 0, sizeof(in));
  in.g[0] = gx; in.g[1] = gy; in.g[2] = gz;
  in.mask = mask;
  return in;
}

int timed_event(vdvc_ctx* ctx) {
  if (!ctx->timing) return 0;
  if (ctx->ev_used == ctx->ev.size()) {
    hipEvent_t e;
    VDVC_HIP(hipEventCreate(&e));
    ctx->ev.push_back(e);
  }
  VDVC_HIP(hipEventRecord(ctx->ev[ctx->ev_used++], ctx->stream));
  return 0;
}

int launch_restrict_timed(vdvc_ctx* ctx, int mode, int gtype, const MeshDev& M, const TilesDev& T,
                          const VoxIn& in, int ftype, bool fold, const void* v) {
  int rc = timed_event(ctx);
  if (rc) return rc;
  if (T.vec)
    VDVC_LAUNCH(vdvc::launch_restrict_vec(mode, gtype, in.mask != nullptr, ftype, fold, M, T, in,
                                          static_cast<const float*>(v), ctx->P.as<float>(),
                                          ctx->stream));
  else
    VDVC_LAUNCH(vdvc::launch_restrict(mode, gtype, in.mask != nullptr, ftype, fold, M, T, in, v,
                                      ctx->P.p, ctx->stream));
  return timed_event(ctx);
}

// element types of the post-solve stages' images (the images' own dtypes)
int itype_ok(int t) {
  return t == VDVC_F32 || t == VDVC_F64 || t == VDVC_U8 || t == VDVC_U16 || t == VDVC_I16 ||
         t == VDVC_BOOL;
}

// dense voxel grid of a mesh for the post-solve kernels (grid y / z limits)
int post_grid_ok(const MeshDev& M) {
  if ((M.Nvy + 3) / 4 > 65535 || M.Nz > 65535)
    return fail(-1, "voxel grid %d x %d x %d exceeds the launch grid (Nvy <= 262140, Nz <= 65535)",
                M.Nvx, M.Nvy, M.Nvz);
  return 0;
}

PostIn post_in(const MeshDev& M, const void* U) {
  PostIn P;
  std::memset(&P, 0, sizeof(P));
  P.Nx = M.Nx; P.Ny = M.Ny; P.Nz = M.Nz; P.h = M.h;
  P.Nvx = M.Nvx; P.Nvy = M.Nvy; P.Nvz = M.Nvz;
  P.nnx = M.nnx; P.nny = M.nny;
  P.Nvox = M.Nvox;
  P.U = U;
  // z factors of the Hex8 voxel weights per local plane k = 0..h: the same
  // IEEE double operations as the kernels' (2k / h) - 1, 1 - xi, 1 + xi
  if (M.h <= kPostFzMax) {
    P.fz_tab = 1;
    for (int k = 0; k <= M.h; ++k) {
      const double q = (2.0 * k) / (double)M.h;
      const double xi = q - 1.0;
      P.fz[k][0] = 1.0 - xi;
      P.fz[k][1] = 1.0 + xi;
    }
  }
  return P;
}

// the sampled image of a post-solve stage; images below 2^32 bytes are read
// with 32-bit offsets
void post_img(PostIn& P, const void* img, int itype, int gnx, int gny, int gnz) {
  P.img = img; P.itype = itype; P.gnx = gnx; P.gny = gny; P.gnz = gnz;
  const size_t es = itype == VDVC_F64 ? 8 : itype == VDVC_F32 ? 4
                    : (itype == VDVC_U16 || itype == VDVC_I16) ? 2 : 1;
  const unsigned long long bytes =
      (unsigned long long)gnx * (unsigned long long)gny * (unsigned long long)gnz * es;
  P.img_bytes = bytes < (1ull << 32) ? (unsigned)bytes : 0u;
}

int run_post_timed(vdvc_ctx* ctx, int op, int rtype, const PostIn& P) {
  int rc = timed_event(ctx);
  if (rc) return rc;
  VDVC_LAUNCH(vdvc::launch_post_voxel(op, rtype, P, ctx->stream));
  return timed_event(ctx);
}

size_t real_size(int gtype) { return gtype == VDVC_F64 ? sizeof(double) : sizeof(float); }

int run_restrict(vdvc_ctx* ctx, int mode, int gtype, const MeshDev& M, const TilesDev& T,
                 const VoxIn& in, int ftype, bool fold, const void* v) {
  int rc = ctx->P.ensure((size_t)T.nwaves * T.wave_stride * real_size(gtype));
  if (rc) return rc;
  return launch_restrict_timed(ctx, mode, gtype, M, T, in, ftype, fold, v);
}

template <typename R>
NodeOp<R> node_op0() {
  NodeOp<R> op;
  std::memset(&op, 0, sizeof(op));
  op.cH = R(1);
  return op;
}

// element-matrix operand of a node kernel: the device copy, and for a
// block-diagonal matrix its 8 x 8 scalar form as launch arguments (from the
// host copy the device copy was converted from, so the values are the same)
template <typename R>
KOp<R> kop(const Buf& b, const std::vector<double>& hk, int blockdiag, double E,
           const R* Earr) {
  KOp<R> k;
  std::memset(&k, 0, sizeof(k));
  k.K = b.as<const R>();
  k.blockdiag = blockdiag;
  k.E = (R)E;
  k.Earr = Earr;
  if (blockdiag && hk.size() == 576)
    for (int q = 0; q < 64; ++q) k.kc[q] = (R)hk[(3 * (q >> 3)) * 24 + 3 * (q & 7)];
  return k;
}

template <typename R>
KOp<R> kop1(vdvc_ctx* ctx, int blockdiag, double E, const R* Earr) {
  return kop<R>(ctx->K1[rix<R>()], ctx->k1_cache[rix<R>()], blockdiag, E, Earr);
}

template <typename R>
KOp<R> kop2(vdvc_ctx* ctx, int blockdiag, double E, const R* Earr) {
  return kop<R>(ctx->K2[rix<R>()], ctx->k2_cache[rix<R>()], blockdiag, E, Earr);
}

// out = Hpartials (if P) + ll * R v, with R per reg_kind; eq-gap uses a
// temporary for the inner P_i K v; optional FP64 dot partials of (dotv . out)
template <typename R>
int apply_reg_sum(vdvc_ctx* ctx, const MeshDev& M, const TilesDev* T, bool useP, int reg_kind,
                  int bd1, int bd2, double E, double eps_reg, double ll, const R* v, R* out,
                  const R* dotv, double* dot_part, const int* done, bool reg_fused = false) {
  NodeOp<R> op = node_op0<R>();
  if (useP) { op.P = ctx->P.as<const R>(); op.T = *T; }
  op.done = done;
  // reg_fused: the voxel kernel already added ll R v to the partials
  if (reg_kind == VDVC_REG_K && ll != 0.0 && !reg_fused) {
    op.k1 = kop1<R>(ctx, bd1, E, nullptr); op.v1 = v; op.c1 = (R)ll; op.k1_on = 1;
  }
  if (reg_kind != VDVC_REG_EQGAP || ll == 0.0) {
    op.out = out; op.dotv = dotv; op.dot_part = dot_part;
    VDVC_LAUNCH(vdvc::launch_node_op(M, op, ctx->stream));
    return 0;
  }
  // equilibrium gap: out = H v (partials); w = P_i (K1 v); out += ll (K1 w + eps K2 v)
  int rc = ctx->nodeC.ensure((size_t)M.Nnodes * 3 * sizeof(R));
  if (rc) return rc;
  R* w = ctx->nodeC.as<R>();
  op.out = out;
  VDVC_LAUNCH(vdvc::launch_node_op(M, op, ctx->stream));
  NodeOp<R> o2 = node_op0<R>();
  o2.done = done;
  o2.k1 = kop1<R>(ctx, bd1, E, nullptr); o2.v1 = v; o2.c1 = R(1); o2.k1_on = 1;
  o2.interior = 1;
  o2.out = w;
  VDVC_LAUNCH(vdvc::launch_node_op(M, o2, ctx->stream));
  NodeOp<R> o3 = node_op0<R>();
  o3.done = done;
  o3.addin = out; o3.cadd = R(1);
  o3.k1 = kop1<R>(ctx, bd1, E, nullptr); o3.v1 = w; o3.c1 = (R)ll; o3.k1_on = 1;
  o3.k2 = kop2<R>(ctx, bd2, E, nullptr); o3.v2 = v; o3.c2 = (R)(ll * eps_reg); o3.k2_on = 1;
  o3.out = out; o3.dotv = dotv; o3.dot_part = dot_part;
  VDVC_LAUNCH(vdvc::launch_node_op(M, o3, ctx->stream));
  return 0;
}

template <typename R>
int prep_reg(vdvc_ctx* ctx, int reg_kind, const double* hK1, const double* hK2, int* bd1,
             int* bd2) {
  *bd1 = *bd2 = 0;
  if (reg_kind == VDVC_REG_NONE) return 0;
  if (reg_kind != VDVC_REG_K && reg_kind != VDVC_REG_EQGAP)
    return fail(-1, "reg_kind must be VDVC_REG_NONE, VDVC_REG_K or VDVC_REG_EQGAP");
  int rc = upload_K1<R>(ctx, hK1, bd1);
  if (rc) return rc;
  if (reg_kind == VDVC_REG_EQGAP) return upload_K2<R>(ctx, hK2, bd2);
  return 0;
}

// ---- working-type implementations of the entry points ----------------------

template <typename R>
int residual_rhs_impl(vdvc_ctx* ctx, const MeshDev& M, const TilesDev& T, int gtype,
                      const VoxIn& in, int ftype, bool fold, R* d_b, double* h_stats4) {
  int rc;
  if ((rc = run_restrict(ctx, MODE_RHS, gtype, M, T, in, ftype, fold, nullptr))) return rc;
  NodeOp<R> op = node_op0<R>();
  op.P = ctx->P.as<const R>(); op.T = T; op.out = d_b;
  VDVC_LAUNCH(vdvc::launch_node_op(M, op, ctx->stream));
  if (!h_stats4) return 0;
  VDVC_LAUNCH(vdvc::launch_reduce(ctx->stats.as<double>(), T.nwaves, 4, ctx->red.as<double>(),
                                  ctx->stream));
  return fetch(ctx, ctx->red.p, 4 * sizeof(double), h_stats4);
}

template <typename R>
int matvec_H_impl(vdvc_ctx* ctx, const MeshDev& M, const TilesDev& T, int gtype, const VoxIn& in,
                  const R* v, R* out) {
  int rc;
  if ((rc = run_restrict(ctx, MODE_MATVEC, gtype, M, T, in, VDVC_F32, false, v))) return rc;
  NodeOp<R> op = node_op0<R>();
  op.P = ctx->P.as<const R>(); op.T = T; op.out = out;
  VDVC_LAUNCH(vdvc::launch_node_op(M, op, ctx->stream));
  return 0;
}

// H v with the node partials in the static scratch array: the same kernels
// and decomposition as matvec_H_impl, no allocation, no timing events
template <typename R>
int matvec_H_noalloc_impl(vdvc_ctx* ctx, const MeshDev& M, const TilesDev& T, int gtype,
                          const VoxIn& in, const R* v, R* out) {
  const size_t need = (size_t)T.nwaves * T.wave_stride * sizeof(R);
  if (need > VDVC_NOALLOC_SCRATCH_BYTES)
    return fail(-1, "mesh %d x %d x %d (h=%d) needs %zu bytes of node partials; the "
                    "no-allocation scratch holds %u", M.Nx, M.Ny, M.Nz, M.h, need,
                VDVC_NOALLOC_SCRATCH_BYTES);
  if (!ctx->scratch) return fail(-2, "no scratch address for device %d", ctx->device);
  R* P = static_cast<R*>(ctx->scratch);
  if (T.vec)
    VDVC_LAUNCH(vdvc::launch_restrict_vec(MODE_MATVEC, gtype, in.mask != nullptr, VDVC_F32, false,
                                          M, T, in, reinterpret_cast<const float*>(v),
                                          reinterpret_cast<float*>(P), ctx->stream));
  else
    VDVC_LAUNCH(vdvc::launch_restrict(MODE_MATVEC, gtype, in.mask != nullptr, VDVC_F32, false, M,
                                      T, in, v, P, ctx->stream));
  NodeOp<R> op = node_op0<R>();
  op.P = P; op.T = T; op.out = out;
  VDVC_LAUNCH(vdvc::launch_node_op(M, op, ctx->stream));
  return 0;
}

template <typename R>
int diag_H_impl(vdvc_ctx* ctx, const MeshDev& M, const TilesDev& T, int gtype, const VoxIn& in,
                R* out) {
  int rc;
  if ((rc = run_restrict(ctx, MODE_DIAG, gtype, M, T, in, VDVC_F32, false, nullptr))) return rc;
  DiagOp<R> op;
  std::memset(&op, 0, sizeof(op));
  op.P = ctx->P.as<const R>(); op.T = T; op.out = out;
  VDVC_LAUNCH(vdvc::launch_diag_op(M, op, ctx->stream));
  return 0;
}

template <typename R>
int matvec_K_impl(vdvc_ctx* ctx, const MeshDev& M, const double* h_K, double E, const R* d_E,
                  const R* v, R* out) {
  int bd = 0, rc;
  if ((rc = upload_K1<R>(ctx, h_K, &bd))) return rc;
  NodeOp<R> op = node_op0<R>();
  op.k1 = kop1<R>(ctx, bd, E, d_E); op.v1 = v; op.c1 = R(1); op.k1_on = 1;
  op.out = out;
  VDVC_LAUNCH(vdvc::launch_node_op(M, op, ctx->stream));
  return 0;
}

template <typename R>
int diag_K_impl(vdvc_ctx* ctx, const MeshDev& M, const double* h_K, double E, const R* d_E,
                R* out) {
  int bd = 0, rc;
  if ((rc = upload_K1<R>(ctx, h_K, &bd))) return rc;
  DiagOp<R> op;
  std::memset(&op, 0, sizeof(op));
  op.k1 = kop1<R>(ctx, bd, E, d_E); op.c1 = R(1); op.k1_on = 1;
  op.out = out;
  VDVC_LAUNCH(vdvc::launch_diag_op(M, op, ctx->stream));
  return 0;
}

template <typename R>
int eqgap_impl(vdvc_ctx* ctx, const MeshDev& M, const double* h_Kel, double E, const R* d_E,
               const R* v, R* out) {
  int bd = 0, rc;
  if ((rc = upload_K1<R>(ctx, h_Kel, &bd))) return rc;
  if ((rc = ctx->nodeC.ensure((size_t)M.Nnodes * 3 * sizeof(R)))) return rc;
  R* w = ctx->nodeC.as<R>();
  const Buf& K1 = ctx->K1[rix<R>()];
  NodeOp<R> o1 = node_op0<R>();
  o1.k1 = kop1<R>(ctx, bd, E, d_E); o1.v1 = v; o1.c1 = R(1); o1.k1_on = 1;
  o1.interior = 1; o1.out = w;
  VDVC_LAUNCH(vdvc::launch_node_op(M, o1, ctx->stream));
  NodeOp<R> o2 = node_op0<R>();
  o2.k1 = kop1<R>(ctx, bd, E, d_E); o2.v1 = w; o2.c1 = R(1); o2.k1_on = 1;
  o2.out = out;
  VDVC_LAUNCH(vdvc::launch_node_op(M, o2, ctx->stream));
  return 0;
}

template <typename R>
int hutch_impl(vdvc_ctx* ctx, const MeshDev& M, const double* h_Kel, double E,
               const int8_t* d_probes, int n_probes, R* out) {
  const long long ndof = 3 * M.Nnodes;
  int rc;
  if ((rc = ctx->hacc.ensure((size_t)ndof * sizeof(double)))) return rc;
  if ((rc = ctx->nodeA.ensure((size_t)ndof * sizeof(R)))) return rc;
  if ((rc = ctx->nodeB.ensure((size_t)ndof * sizeof(R)))) return rc;
  if ((rc = ctx->dparts.ensure(1024 * sizeof(double)))) return rc;
  VDVC_HIP(hipMemsetAsync(ctx->hacc.p, 0, (size_t)ndof * sizeof(double), ctx->stream));
  for (int p = 0; p < n_probes; ++p) {
    const int8_t* z = d_probes + (size_t)p * ndof;
    VDVC_LAUNCH(vdvc::launch_expand_i8(ndof, z, ctx->nodeA.as<R>(), ctx->stream));
    if ((rc = eqgap_impl<R>(ctx, M, h_Kel, E, nullptr, ctx->nodeA.as<R>(), ctx->nodeB.as<R>())))
      return rc;
    VDVC_LAUNCH(vdvc::launch_hutch_acc(ndof, z, ctx->nodeB.as<const R>(), ctx->hacc.as<double>(),
                                       ctx->stream));
  }
  VDVC_LAUNCH(vdvc::launch_hutch_final(ndof, ctx->hacc.as<double>(), n_probes, out,
                                       ctx->dparts.as<double>(), ctx->stream));
  return 0;
}

template <typename R>
int jacobi_finish(vdvc_ctx* ctx, const MeshDev& M, const TilesDev& T, const R* P, int reg_kind,
                  const double* h_K1, const double* h_K2, double E, double eps_reg, double ll,
                  const R* d_extra, R* d_diag, R* d_minv);

template <typename R>
int jacobi_impl(vdvc_ctx* ctx, const MeshDev& M, const TilesDev& T, int gtype, const VoxIn& in,
                int reg_kind, const double* h_K1, const double* h_K2, double E, double eps_reg,
                double ll, const R* d_extra, R* d_diag, R* d_minv) {
  int rc;
  if ((rc = run_restrict(ctx, MODE_DIAG, gtype, M, T, in, VDVC_F32, false, nullptr))) return rc;
  return jacobi_finish<R>(ctx, M, T, ctx->P.as<const R>(), reg_kind, h_K1, h_K2, E, eps_reg, ll,
                          d_extra, d_diag, d_minv);
}

// the node side of the Jacobi preconditioner from the diagonal's voxel
// partials P (plan T): gather + regularizer diagonal, then the inverse
template <typename R>
int jacobi_finish(vdvc_ctx* ctx, const MeshDev& M, const TilesDev& T, const R* P, int reg_kind,
                  const double* h_K1, const double* h_K2, double E, double eps_reg, double ll,
                  const R* d_extra, R* d_diag, R* d_minv) {
  int bd1 = 0, bd2 = 0, rc;
  if ((rc = prep_reg<R>(ctx, reg_kind, h_K1, h_K2, &bd1, &bd2))) return rc;
  const int nb = vdvc::node_blocks(M);
  if ((rc = ctx->fmin.ensure((size_t)nb * sizeof(R)))) return rc;
  DiagOp<R> op;
  std::memset(&op, 0, sizeof(op));
  op.P = P; op.T = T; op.out = d_diag; op.min_part = ctx->fmin.as<R>();
  if (reg_kind == VDVC_REG_K) {
    op.k1 = kop1<R>(ctx, bd1, E, nullptr); op.c1 = (R)ll; op.k1_on = 1;
  } else if (reg_kind == VDVC_REG_EQGAP) {
    // diag R = hutchinson(K1 P_i K1) + eps * diag(K2): K2 part here, the
    // Hutchinson part via d_extra
    if (!d_extra) return fail(-1, "REG_EQGAP needs the Hutchinson diagonal in d_extra");
    op.k1 = kop2<R>(ctx, bd2, E, nullptr); op.c1 = (R)(ll * eps_reg);
    op.k1_on = 1;
    op.extra = d_extra; op.cextra = (R)ll;
  }
  VDVC_LAUNCH(vdvc::launch_diag_op(M, op, ctx->stream));
  VDVC_LAUNCH(vdvc::launch_jacobi(d_diag, ctx->fmin.as<const R>(), nb, 3 * M.Nnodes, d_minv,
                                  ctx->stream));
  return 0;
}

// The Laplacian term of a float32 product is added by the vectorised voxel
// kernel itself (single-element lanes), element by element, as c E K (the
// block-diagonal matrix's 8 x 8 form, c = ll): the node pass is then the
// partials' gather alone.  Returns whether `in` now carries the term.
template <typename R>
int fuse_reg(vdvc_ctx* ctx, const TilesDev& T, int reg_kind, int bd1, double E, double ll,
             VoxIn& in, bool* fused) {
  *fused = false;
  in.kreg = nullptr;
  in.kc = 0.0f;
  const std::vector<double>& hk = ctx->k1_cache[rix<R>()];
  // the voxel kernels that carry the term: the vectorised kernel with
  // single-element lanes (paired-element lanes would spill registers)
  const bool kernel_has = T.vec && T.epl == 1;
  if (sizeof(R) != 4 || !kernel_has || reg_kind != VDVC_REG_K || ll == 0.0 || !bd1 ||
      hk.size() != 576 || getenv("VDVC_NO_REG_FUSE"))
    return 0;
  if (ctx->kc32_src != hk || !ctx->kc32.p) {
    int rc = ctx->kc32.ensure(64 * sizeof(float));
    if (rc) return rc;
    float tmp[64];
    for (int q = 0; q < 64; ++q) tmp[q] = (float)hk[(3 * (q >> 3)) * 24 + 3 * (q & 7)];
    VDVC_HIP(hipMemcpyAsync(ctx->kc32.p, tmp, sizeof(tmp), hipMemcpyHostToDevice, ctx->stream));
    VDVC_HIP(hipStreamSynchronize(ctx->stream));
    ctx->kc32_src = hk;
  }
  in.kreg = ctx->kc32.as<const float>();
  in.kc = (float)(ll * E);
  *fused = true;
  return 0;
}

template <typename R>
int apply_A_impl(vdvc_ctx* ctx, const MeshDev& M, const TilesDev& T, int gtype, const VoxIn& in,
                 int reg_kind, const double* h_K1, const double* h_K2, double E, double eps_reg,
                 double ll, const R* v, R* out) {
  int bd1 = 0, bd2 = 0, rc;
  if ((rc = prep_reg<R>(ctx, reg_kind, h_K1, h_K2, &bd1, &bd2))) return rc;
  VoxIn vin = in;
  bool fused = false;
  if ((rc = fuse_reg<R>(ctx, T, reg_kind, bd1, E, ll, vin, &fused))) return rc;
  if ((rc = run_restrict(ctx, MODE_MATVEC, gtype, M, T, vin, VDVC_F32, false, v))) return rc;
  return apply_reg_sum<R>(ctx, M, &T, true, reg_kind, bd1, bd2, E, eps_reg, ll, v, out, nullptr,
                          nullptr, nullptr, fused);
}

template <typename R>
int pcg_impl(vdvc_ctx* ctx, const MeshDev& M, const TilesDev& T, int gtype, VoxIn in,
             int reg_kind, const double* h_K1, const double* h_K2, double E, double eps_reg,
             double ll, const R* d_b, const R* d_minv, R* d_x, double tol, int max_iter,
             int check_every, int* h_iters, double* h_relres) {
  int bd1 = 0, bd2 = 0, rc;
  if ((rc = prep_reg<R>(ctx, reg_kind, h_K1, h_K2, &bd1, &bd2))) return rc;
  const long long ndof = 3 * M.Nnodes;
  const size_t vb = (size_t)ndof * sizeof(R);
  if ((rc = ctx->pr.ensure(vb)) || (rc = ctx->pz.ensure(vb)) || (rc = ctx->pp.ensure(vb)) ||
      (rc = ctx->pAp.ensure(vb)))
    return rc;
  const int nvb = vdvc::vec_blocks(ndof);
  const int nnb = vdvc::node_blocks(M);
  if ((rc = ctx->dparts.ensure((size_t)2 * nvb * sizeof(double)))) return rc;
  if ((rc = ctx->dparts2.ensure((size_t)nnb * sizeof(double)))) return rc;
  if ((rc = ctx->pstate.ensure(sizeof(PcgState)))) return rc;
  if ((rc = ctx->P.ensure((size_t)T.nwaves * T.wave_stride * sizeof(R)))) return rc;
  R* r = ctx->pr.as<R>();
  R* z = ctx->pz.as<R>();
  R* p = ctx->pp.as<R>();
  R* Ap = ctx->pAp.as<R>();
  PcgState* S = ctx->pstate.as<PcgState>();
  const int* done = &S->done;
  VDVC_LAUNCH(vdvc::launch_pcg_init(ndof, d_b, d_minv, d_x, r, z, p, ctx->dparts.as<double>(), S,
                                    ctx->stream));
  in.done = done;
  bool fused = false;
  if ((rc = fuse_reg<R>(ctx, T, reg_kind, bd1, E, ll, in, &fused))) return rc;
  // the library's own vectors are 256-byte aligned; the caller's x and M_inv
  // decide whether the vectorised updates apply
  const bool vec = vdvc::pcg_vec_ok(ndof) && (((uintptr_t)d_x | (uintptr_t)d_minv) & 15u) == 0;
  // Host polls: the first after check_every iterations, each later one after
  // the iterations the relative residual's rate since the previous poll
  // needs to reach tol (at least one, at most four check intervals).  The
  // device stops by itself at convergence -- launches after it exit at once
  // -- so the poll schedule changes no result, only how many early-exit
  // launches and host round trips a solve spends.
  int it = 0, batch = check_every, prev_it = 0;
  double prev_rel = 1.0;
  PcgState hs;
  std::memset(&hs, 0, sizeof(hs));
  bool fetched = false;
  while (it < max_iter) {
    const int nrun = (max_iter - it) < batch ? (max_iter - it) : batch;
    for (int q = 0; q < nrun; ++q, ++it) {
      if ((rc = launch_restrict_timed(ctx, MODE_MATVEC, gtype, M, T, in, VDVC_F32, false, p)))
        return rc;
      if ((rc = apply_reg_sum<R>(ctx, M, &T, true, reg_kind, bd1, bd2, E, eps_reg, ll, p, Ap, p,
                                 ctx->dparts2.as<double>(), done, fused)))
        return rc;
      VDVC_LAUNCH(vdvc::launch_pcg_update(ndof, it, ctx->dparts2.as<double>(), nnb, S, d_minv,
                                          (const R*)p, (const R*)Ap, d_x, r, z,
                                          ctx->dparts.as<double>(), vec, ctx->stream));
      VDVC_LAUNCH(vdvc::launch_pcg_pupdate(ndof, it, tol, ctx->dparts.as<double>(), S,
                                           (const R*)z, p, vec, ctx->stream));
    }
    if ((rc = fetch(ctx, S, sizeof(PcgState), &hs))) return rc;
    fetched = true;
    if (hs.done) break;
    batch = check_every;
    const double rel = hs.relres;
    if (tol > 0.0 && rel > tol && rel < prev_rel && it > prev_it) {
      const double rate = std::log(rel / prev_rel) / (double)(it - prev_it);
      const double need = std::ceil(std::log(tol / rel) / rate);
      batch = (int)std::min(4.0 * check_every, std::max(1.0, need));
    }
    prev_rel = rel;
    prev_it = it;
  }
  if (!fetched && (rc = fetch(ctx, S, sizeof(PcgState), &hs))) return rc;
  if (h_iters) *h_iters = hs.iters;
  if (h_relres) *h_relres = hs.relres;
  return 0;
}

// call F<float> or F<double> by working type
#define VDVC_BY_RTYPE(rt, call_f, call_d) ((rt) == VDVC_F64 ? (call_d) : (call_f))

}  // namespace

extern "C" {

int vdvc_abi_version(void) { return VDVC_ABI_VERSION; }

int vdvc_voxel_pitch(int nvx) { return nvx < 1 ? -1 : vdvc_pitch_of(nvx); }

const char* vdvc_last_error(void) { return g_err.c_str(); }

int vdvc_ctx_create(int device, uintptr_t stream, vdvc_ctx** out) {
  if (!out) return fail(-1, "out is NULL");
  int ndev = 0;
  VDVC_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(-1, "device %d not present (%d devices)", device, ndev);
  VDVC_HIP(hipSetDevice(device));
  void* sp = nullptr;
  int rc = scratch_address(device, &sp);
  if (rc) return rc;
  vdvc_ctx* c = new vdvc_ctx();
  c->device = device;
  c->stream = reinterpret_cast<hipStream_t>(stream);
  c->scratch = sp;
  hipError_t e = hipHostMalloc((void**)&c->hpin, VDVC_HPIN * sizeof(double), hipHostMallocDefault);
  if (e != hipSuccess) {
    delete c;
    return fail(-3, "hipHostMalloc: %s", hipGetErrorString(e));
  }
  *out = c;
  return 0;
}

int vdvc_ctx_destroy(vdvc_ctx* ctx) {
  if (!ctx) return 0;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (hipEvent_t e : ctx->ev) (void)hipEventDestroy(e);
  if (ctx->hpin) (void)hipHostFree(ctx->hpin);
  delete ctx;
  return 0;
}

int vdvc_timing_begin(vdvc_ctx* ctx) {
  if (!ctx) return fail(-1, "ctx is NULL");
  ctx->timing = true;
  ctx->ev_used = 0;
  return 0;
}

int vdvc_timing_end(vdvc_ctx* ctx, double* h_total_ms, int* h_launches) {
  if (!ctx || !h_total_ms || !h_launches) return fail(-1, "NULL argument");
  VDVC_HIP(hipStreamSynchronize(ctx->stream));
  double tot = 0.0;
  for (size_t q = 0; q + 1 < ctx->ev_used; q += 2) {
    float ms = 0.0f;
    VDVC_HIP(hipEventElapsedTime(&ms, ctx->ev[q], ctx->ev[q + 1]));
    tot += ms;
  }
  *h_total_ms = tot;
  *h_launches = (int)(ctx->ev_used / 2);
  ctx->timing = false;
  ctx->ev_used = 0;
  return 0;
}

int vdvc_synchronize(vdvc_ctx* ctx) {
  if (!ctx) return fail(-1, "ctx is NULL");
  VDVC_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int vdvc_image_setup(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_f, int ftype,
                     double glt, void* d_gx, void* d_gy, void* d_gz, int gtype,
                     uint8_t* d_mask0) {
  if (!ctx) return fail(-1, "ctx is NULL");
  int rc = check_mesh(Nx, Ny, Nz, h);
  if (rc) return rc;
  if (!ftype_ok(ftype) || !gtype_ok(gtype)) return fail(-1, "unsupported element type");
  if ((ftype == VDVC_F64) != (gtype == VDVC_F64))
    return fail(-1, "the FP64 path needs a float64 image and float64 gradients");
  if (!d_f || !d_gx || !d_gy || !d_gz || !d_mask0) return fail(-1, "NULL buffer");
  MeshDev M = make_mesh(Nx, Ny, Nz, h);
  VDVC_LAUNCH(vdvc::launch_setup(ftype, gtype, M, d_f, glt, d_gx, d_gy, d_gz, d_mask0, ctx->stream));
  return 0;
}

int vdvc_masked_stats(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_field,
                      int ftype, const uint8_t* d_mask, int use_thresh, double glt,
                      double* h_out3) {
  if (!ctx || !h_out3 || !d_field) return fail(-1, "NULL argument");
  int rc = check_mesh(Nx, Ny, Nz, h);
  if (rc) return rc;
  if (!ftype_ok(ftype)) return fail(-1, "unsupported field type");
  MeshDev M = make_mesh(Nx, Ny, Nz, h);
  const int nb = 2048;
  if ((rc = ctx->stats.ensure((size_t)nb * 3 * sizeof(double)))) return rc;
  if ((rc = ctx->red.ensure(8 * sizeof(double)))) return rc;
  VDVC_LAUNCH(vdvc::launch_stats(ftype, M, d_field, d_mask, use_thresh, glt,
                                 ctx->stats.as<double>(), nb, ctx->stream));
  VDVC_LAUNCH(vdvc::launch_reduce(ctx->stats.as<double>(), nb, 3, ctx->red.as<double>(), ctx->stream));
  return fetch(ctx, ctx->red.p, 3 * sizeof(double), h_out3);
}

int vdvc_level_image(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_src, int stype,
                     void* d_dst, int dtype, double glt, double* h_out3) {
  if (!ctx || !h_out3 || !d_src || !d_dst) return fail(-1, "NULL argument");
  int rc = check_mesh(Nx, Ny, Nz, h);
  if (rc) return rc;
  if (!((stype == VDVC_F32 && (dtype == VDVC_F16 || dtype == VDVC_F32)) ||
        (stype == VDVC_F64 && (dtype == VDVC_F64 || dtype == VDVC_F32))))
    return fail(-1, "unsupported (source, level) type pair");
  auto esz = [](int t) -> uintptr_t { return t == VDVC_F64 ? 8 : (t == VDVC_F16 ? 2 : 4); };
  if (((uintptr_t)d_src & (esz(stype) - 1)) || ((uintptr_t)d_dst & (esz(dtype) - 1)))
    return fail(-1, "volume not aligned to its element size");
  MeshDev M = make_mesh(Nx, Ny, Nz, h);
  const int nb = 2048;
  if ((rc = ctx->stats.ensure((size_t)nb * 3 * sizeof(double)))) return rc;
  if ((rc = ctx->red.ensure(8 * sizeof(double)))) return rc;
  VDVC_LAUNCH(vdvc::launch_level_image(stype, dtype, M, d_src, d_dst, glt,
                                       ctx->stats.as<double>(), nb, ctx->stream));
  VDVC_LAUNCH(vdvc::launch_reduce(ctx->stats.as<double>(), nb, 3, ctx->red.as<double>(), ctx->stream));
  return fetch(ctx, ctx->red.p, 3 * sizeof(double), h_out3);
}

int vdvc_sample(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_U, const void* d_g,
                int gitype, int gnx, int gny, int gnz, double gox, double goy, double goz,
                double glt, const uint8_t* d_mask0, const uint8_t* d_bmask, void* d_gvals,
                uint8_t* d_mask_out, double* h_stats3, int f32_weights) {
  if (!ctx) return fail(-1, "ctx is NULL");
  if (f32_weights && gitype == VDVC_F64)
    return fail(-1, "float32 interpolation weights need a float32 or float16 image");
  int rc = check_mesh(Nx, Ny, Nz, h);
  if (rc) return rc;
  if (!ftype_ok(gitype)) return fail(-1, "unsupported image type");
  if (gnx < 1 || gny < 1 || gnz < 1) return fail(-1, "invalid g dims");
  if (!d_U || !d_g || !d_gvals || !d_mask_out) return fail(-1, "NULL buffer");
  MeshDev M = make_mesh(Nx, Ny, Nz, h);
  TilesDev T = make_tiles_generic(M);
  // Nvx = 64 k + 1 (e.g. 513): the one-column last strip is folded into the
  // strip before it rather than given wavefronts with one active lane
  const int tail_fold = (M.Nvx > 64 && (M.Nvx - 1) % 64 == 0) ? 1 : 0;
  if (tail_fold) {
    T.nsx = (M.Nvx - 1) / 64;
    plan_z(M, T, 16384);
  }
  if ((rc = ctx->stats.ensure((size_t)T.nwaves * 3 * sizeof(double)))) return rc;
  if ((rc = ctx->red.ensure(8 * sizeof(double)))) return rc;
  SampleIn S;
  S.tail_fold = tail_fold;
  S.gimg = d_g; S.gnx = gnx; S.gny = gny; S.gnz = gnz;
  S.gox = gox; S.goy = goy; S.goz = goz; S.glt = glt;
  S.mask0 = d_mask0; S.bmask = d_bmask; S.gvals = d_gvals; S.mask_out = d_mask_out;
  S.stats = ctx->stats.as<double>();
  VDVC_LAUNCH(vdvc::launch_sample(gitype, M, T, S, d_U, f32_weights != 0, ctx->stream));
  VDVC_LAUNCH(vdvc::launch_reduce(ctx->stats.as<double>(), T.nwaves, 3, ctx->red.as<double>(),
                                  ctx->stream));
  if (h_stats3) return fetch(ctx, ctx->red.p, 3 * sizeof(double), h_stats3);
  return 0;
}

int vdvc_residual_rhs(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_gx,
                      const void* d_gy, const void* d_gz, int gtype, const void* d_f, int ftype,
                      double mean0, const void* d_gvals, double gmean, double kscale,
                      const uint8_t* d_bmask, const uint8_t* d_live, void* d_res, void* d_b,
                      uint16_t* d_fold_x, uint16_t* d_fold_y, uint16_t* d_fold_z,
                      double* h_stats4) {
  if (!ctx) return fail(-1, "ctx is NULL");
  int rc = check_mesh(Nx, Ny, Nz, h);
  if (rc) return rc;
  if (!gtype_ok(gtype) || !ftype_ok(ftype)) return fail(-1, "unsupported element type");
  if ((ftype == VDVC_F64) != (gtype == VDVC_F64))
    return fail(-1, "the FP64 path needs a float64 image and float64 gradients");
  if (!d_gx || !d_gy || !d_gz || !d_f || !d_gvals || !d_bmask || !d_live || !d_b)
    return fail(-1, "NULL buffer");
  const bool fold = d_fold_x != nullptr;
  if (fold && (!d_fold_y || !d_fold_z)) return fail(-1, "fold planes must all be given");
  if (fold && gtype == VDVC_F64) return fail(-1, "folded bfloat16 planes need float32 input");
  MeshDev M = make_mesh(Nx, Ny, Nz, h);
  VoxIn in = vox_in(d_gx, d_gy, d_gz, d_bmask);
  in.fimg = d_f; in.gvals = d_gvals; in.live = d_live;
  in.mean0 = mean0; in.gmean = gmean; in.kscale = kscale;
  in.res = d_res;
  in.fold_out[0] = d_fold_x; in.fold_out[1] = d_fold_y; in.fold_out[2] = d_fold_z;
  TilesDev T = tiles_for(M, gtype, in, false);
  if ((rc = ctx->stats.ensure((size_t)T.nwaves * 4 * sizeof(double)))) return rc;
  if ((rc = ctx->red.ensure(8 * sizeof(double)))) return rc;
  // the display statistics only when the caller asks for them
  in.stats = h_stats4 ? ctx->stats.as<double>() : nullptr;
  return VDVC_BY_RTYPE(gtype,
                       residual_rhs_impl<float>(ctx, M, T, gtype, in, ftype, fold,
                                                static_cast<float*>(d_b), h_stats4),
                       residual_rhs_impl<double>(ctx, M, T, gtype, in, ftype, fold,
                                                 static_cast<double*>(d_b), h_stats4));
}

int vdvc_residual_rhs_jacobi(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_gx,
                             const void* d_gy, const void* d_gz, int gtype, const void* d_f,
                             int ftype, double mean0, const void* d_gvals, double gmean,
                             double kscale, const uint8_t* d_bmask, const uint8_t* d_live,
                             void* d_res, void* d_b, uint16_t* d_fold_x, uint16_t* d_fold_y,
                             uint16_t* d_fold_z, double* h_stats4, int reg_kind,
                             const double* h_K1, const double* h_K2, double E, double eps_reg,
                             double ll, const void* d_extra, void* d_diag, void* d_minv) {
  if (!ctx) return fail(-1, "ctx is NULL");
  int rc = check_mesh(Nx, Ny, Nz, h);
  if (rc) return rc;
  if (!gtype_ok(gtype) || !ftype_ok(ftype)) return fail(-1, "unsupported element type");
  if (!d_diag || !d_minv) return fail(-1, "NULL buffer");
  const bool fold = d_fold_x != nullptr;
  // the H whose diagonal is taken: the refolded planes, or the gradients
  // with the b-mask
  const void* hg[3] = {fold ? d_fold_x : d_gx, fold ? d_fold_y : d_gy, fold ? d_fold_z : d_gz};
  const int hgt = fold ? VDVC_BF16 : gtype;
  const uint8_t* hm = fold ? nullptr : d_bmask;
  MeshDev M = make_mesh(Nx, Ny, Nz, h);
  VoxIn in = vox_in(d_gx, d_gy, d_gz, d_bmask);
  in.fimg = d_f; in.gvals = d_gvals; in.live = d_live;
  in.mean0 = mean0; in.gmean = gmean; in.kscale = kscale;
  in.res = d_res;
  in.fold_out[0] = d_fold_x; in.fold_out[1] = d_fold_y; in.fold_out[2] = d_fold_z;
  const TilesDev T = tiles_for(M, gtype, in, false);
  if (gtype == VDVC_F64 || !vdvc::rhs_diag_fused(gtype, ftype, fold, T) ||
      getenv("VDVC_NO_RHS_DIAG")) {
    // no fused kernel for this form or plan: the two passes
    rc = vdvc_residual_rhs(ctx, Nx, Ny, Nz, h, d_gx, d_gy, d_gz, gtype, d_f, ftype, mean0, d_gvals,
                           gmean, kscale, d_bmask, d_live, d_res, d_b, d_fold_x, d_fold_y,
                           d_fold_z, h_stats4);
    if (rc) return rc;
    return vdvc_jacobi(ctx, Nx, Ny, Nz, h, hg[0], hg[1], hg[2], hgt, hm, reg_kind, h_K1, h_K2, E,
                       eps_reg, ll, d_extra, d_diag, d_minv);
  }
  if (!d_gx || !d_gy || !d_gz || !d_f || !d_gvals || !d_bmask || !d_live || !d_b)
    return fail(-1, "NULL buffer");
  if (fold && (!d_fold_y || !d_fold_z)) return fail(-1, "fold planes must all be given");
  if ((rc = ctx->Pd.ensure((size_t)T.nwaves * T.wave_stride * sizeof(float)))) return rc;
  if ((rc = ctx->stats.ensure((size_t)T.nwaves * 4 * sizeof(double)))) return rc;
  if ((rc = ctx->red.ensure(8 * sizeof(double)))) return rc;
  in.stats = h_stats4 ? ctx->stats.as<double>() : nullptr;
  in.pdiag = ctx->Pd.as<float>();
  // the diagonal's node pass before the statistics' fetch, so that the host
  // round trip comes last
  if ((rc = residual_rhs_impl<float>(ctx, M, T, gtype, in, ftype, fold, static_cast<float*>(d_b),
                                     nullptr)))
    return rc;
  if ((rc = jacobi_finish<float>(ctx, M, T, ctx->Pd.as<const float>(), reg_kind, h_K1, h_K2, E,
                                 eps_reg, ll, static_cast<const float*>(d_extra),
                                 static_cast<float*>(d_diag), static_cast<float*>(d_minv))))
    return rc;
  if (!h_stats4) return 0;
  VDVC_LAUNCH(vdvc::launch_reduce(ctx->stats.as<double>(), T.nwaves, 4, ctx->red.as<double>(),
                                  ctx->stream));
  return fetch(ctx, ctx->red.p, 4 * sizeof(double), h_stats4);
}

int vdvc_residual_out(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_f, int ftype,
                      double mean0, const void* d_gvals, int rtype, double gmean, double kscale,
                      const uint8_t* d_bmask, void* d_res) {
  if (!ctx) return fail(-1, "ctx is NULL");
  int rc = check_mesh(Nx, Ny, Nz, h);
  if (rc) return rc;
  if (!d_f || !d_gvals || !d_bmask || !d_res) return fail(-1, "NULL buffer");
  if (!rtype_ok(rtype)) return fail(-1, "rtype must be VDVC_F32 or VDVC_F64");
  if (rtype == VDVC_F64 ? ftype != VDVC_F64 : (ftype != VDVC_F32 && ftype != VDVC_F16))
    return fail(-1, "image type %d does not match working type %d", ftype, rtype);
  const MeshDev M = make_mesh(Nx, Ny, Nz, h);
  VDVC_LAUNCH(vdvc::launch_residual_out(rtype, ftype, M, d_f, d_gvals, d_bmask, mean0, gmean,
                                        kscale, d_res, ctx->stream));
  return 0;
}

int vdvc_matvec_H(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_gx,
                  const void* d_gy, const void* d_gz, int gtype, const uint8_t* d_mask,
                  const void* d_v, void* d_out) {
  if (!ctx) return fail(-1, "ctx is NULL");
  int rc = check_mesh(Nx, Ny, Nz, h);
  if (rc) return rc;
  if (!gtype_ok(gtype)) return fail(-1, "unsupported gradient type");
  if (!d_gx || !d_gy || !d_gz || !d_v || !d_out) return fail(-1, "NULL buffer");
  MeshDev M = make_mesh(Nx, Ny, Nz, h);
  const VoxIn in = vox_in(d_gx, d_gy, d_gz, d_mask);
  const TilesDev T = tiles_for(M, gtype, in, true);
  return VDVC_BY_RTYPE(gtype,
                       matvec_H_impl<float>(ctx, M, T, gtype, in, static_cast<const float*>(d_v),
                                            static_cast<float*>(d_out)),
                       matvec_H_impl<double>(ctx, M, T, gtype, in,
                                             static_cast<const double*>(d_v),
                                             static_cast<double*>(d_out)));
}

int vdvc_matvec_H_noalloc(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_gx,
                          const void* d_gy, const void* d_gz, int gtype, const uint8_t* d_mask,
                          const void* d_v, void* d_out) {
  if (!ctx) return fail(-1, "ctx is NULL");
  int rc = check_mesh(Nx, Ny, Nz, h);
  if (rc) return rc;
  if (!gtype_ok(gtype)) return fail(-1, "unsupported gradient type");
  if (!d_gx || !d_gy || !d_gz || !d_v || !d_out) return fail(-1, "NULL buffer");
  MeshDev M = make_mesh(Nx, Ny, Nz, h);
  const VoxIn in = vox_in(d_gx, d_gy, d_gz, d_mask);
  TilesDev T = tiles_for(M, gtype, in, true);
  // paired-element strips write more node partials; single-element lanes
  // when the pairs' would not fit the static scratch
  const size_t rsz = gtype == VDVC_F64 ? sizeof(double) : sizeof(float);
  if (T.epl > 1 && (size_t)T.nwaves * T.wave_stride * rsz > VDVC_NOALLOC_SCRATCH_BYTES)
    T = tiles_for(M, gtype, in, false);
  return VDVC_BY_RTYPE(gtype,
                       matvec_H_noalloc_impl<float>(ctx, M, T, gtype, in,
                                                    static_cast<const float*>(d_v),
                                                    static_cast<float*>(d_out)),
                       matvec_H_noalloc_impl<double>(ctx, M, T, gtype, in,
                                                     static_cast<const double*>(d_v),
                                                     static_cast<double*>(d_out)));
}

int vdvc_diag_H(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_gx, const void* d_gy,
                const void* d_gz, int gtype, const uint8_t* d_mask, void* d_out) {
  if (!ctx) return fail(-1, "ctx is NULL");
  int rc = check_mesh(Nx, Ny, Nz, h);
  if (rc) return rc;
  if (!gtype_ok(gtype)) return fail(-1, "unsupported gradient type");
  if (!d_gx || !d_gy || !d_gz || !d_out) return fail(-1, "NULL buffer");
  MeshDev M = make_mesh(Nx, Ny, Nz, h);
  const VoxIn in = vox_in(d_gx, d_gy, d_gz, d_mask);
  const TilesDev T = tiles_for(M, gtype, in, true);
  return VDVC_BY_RTYPE(gtype, diag_H_impl<float>(ctx, M, T, gtype, in, static_cast<float*>(d_out)),
                       diag_H_impl<double>(ctx, M, T, gtype, in, static_cast<double*>(d_out)));
}

int vdvc_matvec_K(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int rtype, const double* h_K, double E,
                  const void* d_E, const void* d_v, void* d_out) {
  if (!ctx) return fail(-1, "ctx is NULL");
  int rc = check_mesh(Nx, Ny, Nz, 1);
  if (rc) return rc;
  if (!rtype_ok(rtype)) return fail(-1, "rtype must be VDVC_F32 or VDVC_F64");
  if (!d_v || !d_out) return fail(-1, "NULL buffer");
  MeshDev M = make_mesh(Nx, Ny, Nz, 1);
  return VDVC_BY_RTYPE(rtype,
                       matvec_K_impl<float>(ctx, M, h_K, E, static_cast<const float*>(d_E),
                                            static_cast<const float*>(d_v),
                                            static_cast<float*>(d_out)),
                       matvec_K_impl<double>(ctx, M, h_K, E, static_cast<const double*>(d_E),
                                             static_cast<const double*>(d_v),
                                             static_cast<double*>(d_out)));
}

int vdvc_diag_K(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int rtype, const double* h_K, double E,
                const void* d_E, void* d_out) {
  if (!ctx) return fail(-1, "ctx is NULL");
  int rc = check_mesh(Nx, Ny, Nz, 1);
  if (rc) return rc;
  if (!rtype_ok(rtype)) return fail(-1, "rtype must be VDVC_F32 or VDVC_F64");
  if (!d_out) return fail(-1, "NULL buffer");
  MeshDev M = make_mesh(Nx, Ny, Nz, 1);
  return VDVC_BY_RTYPE(rtype,
                       diag_K_impl<float>(ctx, M, h_K, E, static_cast<const float*>(d_E),
                                          static_cast<float*>(d_out)),
                       diag_K_impl<double>(ctx, M, h_K, E, static_cast<const double*>(d_E),
                                           static_cast<double*>(d_out)));
}

int vdvc_matvec_eqgap(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int rtype, const double* h_Kel,
                      double E, const void* d_E, const void* d_v, void* d_out) {
  if (!ctx) return fail(-1, "ctx is NULL");
  int rc = check_mesh(Nx, Ny, Nz, 1);
  if (rc) return rc;
  if (!rtype_ok(rtype)) return fail(-1, "rtype must be VDVC_F32 or VDVC_F64");
  if (!d_v || !d_out) return fail(-1, "NULL buffer");
  MeshDev M = make_mesh(Nx, Ny, Nz, 1);
  return VDVC_BY_RTYPE(rtype,
                       eqgap_impl<float>(ctx, M, h_Kel, E, static_cast<const float*>(d_E),
                                         static_cast<const float*>(d_v),
                                         static_cast<float*>(d_out)),
                       eqgap_impl<double>(ctx, M, h_Kel, E, static_cast<const double*>(d_E),
                                          static_cast<const double*>(d_v),
                                          static_cast<double*>(d_out)));
}

int vdvc_hutchinson_eqgap(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int rtype, const double* h_Kel,
                          double E, const int8_t* d_probes, int n_probes, void* d_out) {
  if (!ctx) return fail(-1, "ctx is NULL");
  int rc = check_mesh(Nx, Ny, Nz, 1);
  if (rc) return rc;
  if (!rtype_ok(rtype)) return fail(-1, "rtype must be VDVC_F32 or VDVC_F64");
  if (!d_probes || !d_out || n_probes < 1) return fail(-1, "invalid probes / output");
  MeshDev M = make_mesh(Nx, Ny, Nz, 1);
  return VDVC_BY_RTYPE(rtype,
                       hutch_impl<float>(ctx, M, h_Kel, E, d_probes, n_probes,
                                         static_cast<float*>(d_out)),
                       hutch_impl<double>(ctx, M, h_Kel, E, d_probes, n_probes,
                                          static_cast<double*>(d_out)));
}

int vdvc_jacobi(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_gx, const void* d_gy,
                const void* d_gz, int gtype, const uint8_t* d_mask, int reg_kind,
                const double* h_K1, const double* h_K2, double E, double eps_reg, double ll,
                const void* d_extra, void* d_diag, void* d_minv) {
  if (!ctx) return fail(-1, "ctx is NULL");
  int rc = check_mesh(Nx, Ny, Nz, h);
  if (rc) return rc;
  if (!gtype_ok(gtype)) return fail(-1, "unsupported gradient type");
  if (!d_gx || !d_gy || !d_gz || !d_diag || !d_minv) return fail(-1, "NULL buffer");
  MeshDev M = make_mesh(Nx, Ny, Nz, h);
  const VoxIn in = vox_in(d_gx, d_gy, d_gz, d_mask);
  const TilesDev T = tiles_for(M, gtype, in, true);
  return VDVC_BY_RTYPE(
      gtype,
      jacobi_impl<float>(ctx, M, T, gtype, in, reg_kind, h_K1, h_K2, E, eps_reg, ll,
                         static_cast<const float*>(d_extra), static_cast<float*>(d_diag),
                         static_cast<float*>(d_minv)),
      jacobi_impl<double>(ctx, M, T, gtype, in, reg_kind, h_K1, h_K2, E, eps_reg, ll,
                          static_cast<const double*>(d_extra), static_cast<double*>(d_diag),
                          static_cast<double*>(d_minv)));
}

int vdvc_apply_A(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_gx, const void* d_gy,
                 const void* d_gz, int gtype, const uint8_t* d_mask, int reg_kind,
                 const double* h_K1, const double* h_K2, double E, double eps_reg, double ll,
                 const void* d_v, void* d_out) {
  if (!ctx) return fail(-1, "ctx is NULL");
  int rc = check_mesh(Nx, Ny, Nz, h);
  if (rc) return rc;
  if (!gtype_ok(gtype)) return fail(-1, "unsupported gradient type");
  if (!d_gx || !d_gy || !d_gz || !d_v || !d_out) return fail(-1, "NULL buffer");
  MeshDev M = make_mesh(Nx, Ny, Nz, h);
  const VoxIn in = vox_in(d_gx, d_gy, d_gz, d_mask);
  const TilesDev T = tiles_for(M, gtype, in, true);
  return VDVC_BY_RTYPE(
      gtype,
      apply_A_impl<float>(ctx, M, T, gtype, in, reg_kind, h_K1, h_K2, E, eps_reg, ll,
                          static_cast<const float*>(d_v), static_cast<float*>(d_out)),
      apply_A_impl<double>(ctx, M, T, gtype, in, reg_kind, h_K1, h_K2, E, eps_reg, ll,
                           static_cast<const double*>(d_v), static_cast<double*>(d_out)));
}

int vdvc_pcg(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_gx, const void* d_gy,
             const void* d_gz, int gtype, const uint8_t* d_mask, int reg_kind,
             const double* h_K1, const double* h_K2, double E, double eps_reg, double ll,
             const void* d_b, const void* d_minv, void* d_x, double tol, int max_iter,
             int check_every, int* h_iters, double* h_relres) {
  if (!ctx) return fail(-1, "ctx is NULL");
  int rc = check_mesh(Nx, Ny, Nz, h);
  if (rc) return rc;
  if (!gtype_ok(gtype)) return fail(-1, "unsupported gradient type");
  if (!d_gx || !d_gy || !d_gz || !d_b || !d_minv || !d_x) return fail(-1, "NULL buffer");
  if (max_iter < 0) return fail(-1, "max_iter < 0");
  if (check_every < 1) check_every = 1;
  MeshDev M = make_mesh(Nx, Ny, Nz, h);
  const VoxIn in = vox_in(d_gx, d_gy, d_gz, d_mask);
  const TilesDev T = tiles_for(M, gtype, in, true);
  return VDVC_BY_RTYPE(
      gtype,
      pcg_impl<float>(ctx, M, T, gtype, in, reg_kind, h_K1, h_K2, E, eps_reg, ll,
                      static_cast<const float*>(d_b), static_cast<const float*>(d_minv),
                      static_cast<float*>(d_x), tol, max_iter, check_every, h_iters, h_relres),
      pcg_impl<double>(ctx, M, T, gtype, in, reg_kind, h_K1, h_K2, E, eps_reg, ll,
                       static_cast<const double*>(d_b), static_cast<const double*>(d_minv),
                       static_cast<double*>(d_x), tol, max_iter, check_every, h_iters,
                       h_relres));
}

int vdvc_dot(vdvc_ctx* ctx, int64_t n, int rtype, const void* d_a, const void* d_b,
             double* h_out) {
  if (!ctx || !d_a || !d_b || !h_out) return fail(-1, "NULL argument");
  if (n < 0) return fail(-1, "n < 0");
  if (!rtype_ok(rtype)) return fail(-1, "rtype must be VDVC_F32 or VDVC_F64");
  int rc;
  const int nb = 1024;
  if ((rc = ctx->stats.ensure((size_t)nb * sizeof(double)))) return rc;
  if ((rc = ctx->red.ensure(8 * sizeof(double)))) return rc;
  if (rtype == VDVC_F64)
    VDVC_LAUNCH(vdvc::launch_dot(n, static_cast<const double*>(d_a),
                                 static_cast<const double*>(d_b), ctx->stats.as<double>(), nb,
                                 ctx->stream));
  else
    VDVC_LAUNCH(vdvc::launch_dot(n, static_cast<const float*>(d_a),
                                 static_cast<const float*>(d_b), ctx->stats.as<double>(), nb,
                                 ctx->stream));
  VDVC_LAUNCH(vdvc::launch_reduce(ctx->stats.as<double>(), nb, 1, ctx->red.as<double>(), ctx->stream));
  return fetch(ctx, ctx->red.p, sizeof(double), h_out);
}

int vdvc_dot2(vdvc_ctx* ctx, int64_t n, int rtype, const void* d_a, const void* d_b,
              const void* d_c, double* h_out2) {
  if (!ctx || !d_a || !d_b || !d_c || !h_out2) return fail(-1, "NULL argument");
  if (n < 0) return fail(-1, "n < 0");
  if (!rtype_ok(rtype)) return fail(-1, "rtype must be VDVC_F32 or VDVC_F64");
  int rc;
  const int nb = 1024;
  if ((rc = ctx->stats.ensure((size_t)nb * 2 * sizeof(double)))) return rc;
  if ((rc = ctx->red.ensure(8 * sizeof(double)))) return rc;
  if (rtype == VDVC_F64)
    VDVC_LAUNCH(vdvc::launch_dot2(n, static_cast<const double*>(d_a),
                                  static_cast<const double*>(d_b),
                                  static_cast<const double*>(d_c), ctx->stats.as<double>(), nb,
                                  ctx->stream));
  else
    VDVC_LAUNCH(vdvc::launch_dot2(n, static_cast<const float*>(d_a),
                                  static_cast<const float*>(d_b), static_cast<const float*>(d_c),
                                  ctx->stats.as<double>(), nb, ctx->stream));
  VDVC_LAUNCH(vdvc::launch_reduce(ctx->stats.as<double>(), nb, 2, ctx->red.as<double>(), ctx->stream));
  return fetch(ctx, ctx->red.p, 2 * sizeof(double), h_out2);
}

int vdvc_gn_update(vdvc_ctx* ctx, int64_t n, int rtype, const void* d_dU, void* d_U,
                   double* h_out2) {
  if (!ctx || !d_dU || !d_U || !h_out2) return fail(-1, "NULL argument");
  if (n < 0) return fail(-1, "n < 0");
  if (!rtype_ok(rtype)) return fail(-1, "rtype must be VDVC_F32 or VDVC_F64");
  int rc;
  const int nb = 1024;
  if ((rc = ctx->stats.ensure((size_t)nb * 2 * sizeof(double)))) return rc;
  if ((rc = ctx->red.ensure(8 * sizeof(double)))) return rc;
  if (rtype == VDVC_F64)
    VDVC_LAUNCH(vdvc::launch_gn_update(n, static_cast<const double*>(d_dU),
                                       static_cast<double*>(d_U), ctx->stats.as<double>(), nb,
                                       ctx->stream));
  else
    VDVC_LAUNCH(vdvc::launch_gn_update(n, static_cast<const float*>(d_dU),
                                       static_cast<float*>(d_U), ctx->stats.as<double>(), nb,
                                       ctx->stream));
  VDVC_LAUNCH(vdvc::launch_reduce(ctx->stats.as<double>(), nb, 2, ctx->red.as<double>(), ctx->stream));
  return fetch(ctx, ctx->red.p, 2 * sizeof(double), h_out2);
}

int vdvc_axpby(vdvc_ctx* ctx, int64_t n, int rtype, double a, const void* d_x, double b,
               void* d_y) {
  if (!ctx || !d_x || !d_y) return fail(-1, "NULL argument");
  if (n < 0) return fail(-1, "n < 0");
  if (!rtype_ok(rtype)) return fail(-1, "rtype must be VDVC_F32 or VDVC_F64");
  if (n == 0) return 0;
  if (rtype == VDVC_F64)
    VDVC_LAUNCH(vdvc::launch_axpby(n, a, static_cast<const double*>(d_x), b,
                                   static_cast<double*>(d_y), ctx->stream));
  else
    VDVC_LAUNCH(vdvc::launch_axpby(n, a, static_cast<const float*>(d_x), b,
                                   static_cast<float*>(d_y), ctx->stream));
  return 0;
}

int vdvc_calibration_mode(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, int rtype, void* d_out) {
  if (!ctx || !d_out) return fail(-1, "NULL argument");
  int rc = check_mesh(Nx, Ny, Nz, h);
  if (rc) return rc;
  if (!rtype_ok(rtype)) return fail(-1, "rtype must be VDVC_F32 or VDVC_F64");
  if (rtype == VDVC_F64)
    VDVC_LAUNCH(vdvc::launch_calib_mode(Nx, Ny, Nz, h, static_cast<double*>(d_out), ctx->stream));
  else
    VDVC_LAUNCH(vdvc::launch_calib_mode(Nx, Ny, Nz, h, static_cast<float*>(d_out), ctx->stream));
  return 0;
}

int vdvc_fold_grad(vdvc_ctx* ctx, int64_t nvox, const uint16_t* d_gx, const uint16_t* d_gy,
                   const uint16_t* d_gz, const uint8_t* d_mask, uint16_t* d_ox, uint16_t* d_oy,
                   uint16_t* d_oz) {
  if (!ctx || !d_gx || !d_gy || !d_gz || !d_mask || !d_ox || !d_oy || !d_oz)
    return fail(-1, "NULL argument");
  if (nvox < 0) return fail(-1, "nvox < 0");
  if (nvox == 0) return 0;
  VDVC_LAUNCH(vdvc::launch_fold(nvox, d_gx, d_gy, d_gz, d_mask, d_ox, d_oy, d_oz, ctx->stream));
  return 0;
}

int vdvc_erode6(vdvc_ctx* ctx, int nx, int ny, int nz, const uint8_t* d_in, uint8_t* d_out) {
  if (!ctx || !d_in || !d_out) return fail(-1, "NULL argument");
  if (nx < 1 || ny < 1 || nz < 1) return fail(-1, "invalid dims");
  VDVC_LAUNCH(vdvc::launch_erode6(nx, ny, nz, d_in, d_out, ctx->stream));
  return 0;
}

int vdvc_mul(vdvc_ctx* ctx, int64_t n, int rtype, const void* d_a, const void* d_b, void* d_out) {
  if (!ctx || !d_a || !d_b || !d_out) return fail(-1, "NULL argument");
  if (n < 0) return fail(-1, "n < 0");
  if (!rtype_ok(rtype)) return fail(-1, "rtype must be VDVC_F32 or VDVC_F64");
  if (n == 0) return 0;
  if (rtype == VDVC_F64)
    VDVC_LAUNCH(vdvc::launch_mul(n, static_cast<const double*>(d_a),
                                 static_cast<const double*>(d_b), static_cast<double*>(d_out),
                                 ctx->stream));
  else
    VDVC_LAUNCH(vdvc::launch_mul(n, static_cast<const float*>(d_a),
                                 static_cast<const float*>(d_b), static_cast<float*>(d_out),
                                 ctx->stream));
  return 0;
}

int vdvc_stride(vdvc_ctx* ctx, const void* d_src, int type, int nx, int ny, int nz, int s,
                void* d_dst) {
  if (!ctx || !d_src || !d_dst) return fail(-1, "NULL argument");
  if (!ftype_ok(type)) return fail(-1, "unsupported element type");
  if (nx < 1 || ny < 1 || nz < 1 || s < 1) return fail(-1, "invalid dims / stride");
  const int mx = (nx - 1) / s + 1, my = (ny - 1) / s + 1, mz = (nz - 1) / s + 1;
  VDVC_LAUNCH(vdvc::launch_stride(type, d_src, nx, ny, nz, s, d_dst, mx, my, mz, ctx->stream));
  return 0;
}

int vdvc_phase_corr(vdvc_ctx* ctx, const float* d_f, const float* d_g, int nx, int ny, int nz,
                    int64_t* h_cidx) {
  if (!ctx || !d_f || !d_g || !h_cidx) return fail(-1, "NULL argument");
  if (nx < 1 || ny < 1 || nz < 1) return fail(-1, "invalid dims");
  const long long n = (long long)nx * ny * nz;
  if (n > (1LL << 26)) return fail(-1, "phase correlation volume too large (%lld voxels)", n);
  if (nx > 4096 || ny > 4096 || nz > 4096)
    return fail(-1, "phase correlation axes must be <= 4096 (got %d %d %d)", nx, ny, nz);
  int rc;
  if ((rc = ctx->dft.ensure((size_t)3 * n * sizeof(double2)))) return rc;
  if ((rc = ctx->red.ensure(8 * sizeof(double)))) return rc;
  if ((rc = ctx->stats.ensure((size_t)VDVC_ARG_PARTS * 16))) return rc;
  VDVC_LAUNCH(vdvc::launch_phase_corr(d_f, d_g, nx, ny, nz, ctx->dft.as<double2>(),
                                      ctx->stats.p, ctx->red.as<long long>(), ctx->stream));
  long long idx = 0;
  if ((rc = fetch(ctx, ctx->red.p, sizeof(long long), &idx))) return rc;
  *h_cidx = idx;
  return 0;
}

int vdvc_affine_sums(vdvc_ctx* ctx, const float* d_f, int nx, int ny, int nz, const float* d_g,
                     int gnx, int gny, int gnz, const double* h_A9, const double* h_t3,
                     const double* h_c3, double* h_out6) {
  if (!ctx || !d_f || !d_g || !h_A9 || !h_t3 || !h_c3 || !h_out6) return fail(-1, "NULL argument");
  if (nx < 1 || ny < 1 || nz < 1 || gnx < 1 || gny < 1 || gnz < 1) return fail(-1, "invalid dims");
  AffineIn P;
  for (int q = 0; q < 9; ++q) P.A[q] = h_A9[q];
  for (int q = 0; q < 3; ++q) { P.t[q] = h_t3[q]; P.c[q] = h_c3[q]; }
  const int nb = 1024;
  int rc;
  if ((rc = ctx->stats.ensure((size_t)nb * 6 * sizeof(double)))) return rc;
  if ((rc = ctx->red.ensure(8 * sizeof(double)))) return rc;
  VDVC_LAUNCH(vdvc::launch_affine_cost(d_f, nx, ny, nz, d_g, gnx, gny, gnz, P,
                                       ctx->stats.as<double>(), nb, ctx->stream));
  VDVC_LAUNCH(vdvc::launch_reduce(ctx->stats.as<double>(), nb, 6, ctx->red.as<double>(), ctx->stream));
  return fetch(ctx, ctx->red.p, 6 * sizeof(double), h_out6);
}

static int conv_args(int N, int D, int H, int W, int C, int O, int stride) {
  if (N < 1 || D < 1 || H < 1 || W < 1 || C < 1 || O < 1)
    return fail(-1, "invalid conv shape N=%d D=%d H=%d W=%d C=%d O=%d", N, D, H, W, C, O);
  if (C > 32 || O > 32) return fail(-1, "conv channels must be <= 32 (C=%d O=%d)", C, O);
  if (stride != 1 && stride != 2) return fail(-1, "conv stride must be 1 or 2");
  if ((long long)N * D * H * W * (C > O ? C : O) > 0x7fffffffLL)
    return fail(-1, "conv tensor too large");
  return 0;
}

int vdvc_conv3d_forward(vdvc_ctx* ctx, int N, int D, int H, int W, int C, int O, int stride,
                        const float* d_x, const float* d_w, const float* d_b, int relu,
                        float* d_y) {
  if (!ctx || !d_x || !d_w || !d_y) return fail(-1, "NULL argument");
  int rc = conv_args(N, D, H, W, C, O, stride);
  if (rc) return rc;
  VDVC_LAUNCH(vdvc::launch_conv_fwd(N, D, H, W, C, O, stride, d_x, d_w, d_b, relu, d_y,
                                    ctx->stream));
  return 0;
}

int vdvc_conv3d_backward_data(vdvc_ctx* ctx, int N, int D, int H, int W, int C, int O,
                              int stride, const float* d_dy, const float* d_w,
                              const float* d_act, float* d_dx) {
  if (!ctx || !d_dy || !d_w || !d_dx) return fail(-1, "NULL argument");
  int rc = conv_args(N, D, H, W, C, O, stride);
  if (rc) return rc;
  VDVC_LAUNCH(vdvc::launch_conv_bwd_data(N, D, H, W, C, O, stride, d_dy, d_w, d_act, d_dx,
                                         ctx->stream));
  return 0;
}

int vdvc_conv3d_backward_weight(vdvc_ctx* ctx, int N, int D, int H, int W, int C, int O,
                                int stride, const float* d_x, const float* d_dy, float* d_dw,
                                float* d_db) {
  if (!ctx || !d_x || !d_dy || !d_dw || !d_db) return fail(-1, "NULL argument");
  int rc = conv_args(N, D, H, W, C, O, stride);
  if (rc) return rc;
  const int np = vdvc::conv_wgrad_parts(N, D, H, W, C, O, stride);
  if ((rc = ctx->wpart.ensure((size_t)np * O * 27 * C * sizeof(float)))) return rc;
  VDVC_LAUNCH(vdvc::launch_conv_bwd_weight(N, D, H, W, C, O, stride, d_x, d_dy,
                                           ctx->wpart.as<float>(), d_dw, d_db, ctx->stream));
  return 0;
}

int vdvc_adam_step(vdvc_ctx* ctx, int64_t n, float* d_p, const float* d_g, float* d_m,
                   float* d_v, double lr, double beta1, double beta2, double eps, int step) {
  if (!ctx || !d_p || !d_g || !d_m || !d_v) return fail(-1, "NULL argument");
  if (n < 0 || step < 1) return fail(-1, "n < 0 or step < 1");
  if (n == 0) return 0;
  const double c1 = 1.0 - std::pow(beta1, step), c2 = 1.0 - std::pow(beta2, step);
  VDVC_LAUNCH(vdvc::launch_adam(n, d_p, d_g, d_m, d_v, (float)lr, (float)beta1, (float)beta2,
                                (float)eps, (float)c1, (float)c2, ctx->stream));
  return 0;
}

int vdvc_mse(vdvc_ctx* ctx, int64_t n, const float* d_pred, const float* d_label, float* d_grad,
             double* h_loss) {
  if (!ctx || !d_pred || !d_label || !d_grad || !h_loss) return fail(-1, "NULL argument");
  if (n < 1) return fail(-1, "n < 1");
  const long long nb = (n + 255) / 256;
  int rc;
  if ((rc = ctx->stats.ensure((size_t)nb * sizeof(double)))) return rc;
  if ((rc = ctx->red.ensure(8 * sizeof(double)))) return rc;
  VDVC_LAUNCH(vdvc::launch_mse(n, d_pred, d_label, d_grad, ctx->stats.as<double>(), ctx->stream));
  VDVC_LAUNCH(vdvc::launch_reduce(ctx->stats.as<double>(), nb, 1, ctx->red.as<double>(),
                                  ctx->stream));
  double s = 0.0;
  if ((rc = fetch(ctx, ctx->red.p, sizeof(double), &s))) return rc;
  *h_loss = s / (double)n;
  return 0;
}

int vdvc_seed_input(vdvc_ctx* ctx, const void* d_f, const void* d_g, int itype, int nx, int ny,
                    int nz, int s, float* d_out) {
  if (!ctx || !d_f || !d_g || !d_out) return fail(-1, "NULL argument");
  if (itype != VDVC_F32 && itype != VDVC_F16) return fail(-1, "image type must be F32 or F16");
  if (nx < 1 || ny < 1 || nz < 1 || s < 1) return fail(-1, "invalid dims / stride");
  const int nparts = 512;
  int rc;
  if ((rc = ctx->stats.ensure((size_t)6 * nparts * sizeof(double)))) return rc;
  if ((rc = ctx->red.ensure(8 * sizeof(double)))) return rc;
  VDVC_LAUNCH(vdvc::launch_seed_input(itype, d_f, d_g, nx, ny, nz, s, ctx->stats.as<double>(),
                                      nparts, ctx->red.as<double>(), d_out, ctx->stream));
  return 0;
}

int vdvc_interp_field(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_U, int rtype,
                      void* d_out, int channels_first) {
  if (!ctx || !d_U || !d_out) return fail(-1, "NULL argument");
  int rc = check_mesh(Nx, Ny, Nz, h);
  if (rc) return rc;
  if (!rtype_ok(rtype)) return fail(-1, "U type must be F32 or F64");
  MeshDev M = make_mesh(Nx, Ny, Nz, h);
  if ((rc = post_grid_ok(M))) return rc;
  PostIn P = post_in(M, d_U);
  P.out = d_out;
  P.channels_first = channels_first ? 1 : 0;
  return run_post_timed(ctx, POST_INTERP, rtype, P);
}

int vdvc_active_mask(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_U, int rtype,
                     const void* d_g, int gtype, int gnx, int gny, int gnz, double gox, double goy,
                     double goz, const uint8_t* d_mask0, double glt, uint8_t* d_mask_out) {
  if (!ctx || !d_U || !d_g || !d_mask_out) return fail(-1, "NULL argument");
  int rc = check_mesh(Nx, Ny, Nz, h);
  if (rc) return rc;
  if (!rtype_ok(rtype)) return fail(-1, "U type must be F32 or F64");
  if (!itype_ok(gtype)) return fail(-1, "unsupported image type %d", gtype);
  if (gnx < 1 || gny < 1 || gnz < 1) return fail(-1, "invalid image dims");
  MeshDev M = make_mesh(Nx, Ny, Nz, h);
  if ((rc = post_grid_ok(M))) return rc;
  PostIn P = post_in(M, d_U);
  post_img(P, d_g, gtype, gnx, gny, gnz);
  P.ox = gox; P.oy = goy; P.oz = goz; P.glt = glt; P.mask0 = d_mask0;
  P.out = d_mask_out;
  return run_post_timed(ctx, POST_MASK, rtype, P);
}

int vdvc_push_forward(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_U, int rtype,
                      const void* d_field, int ftype, int order, int mode, double cval,
                      void* d_out) {
  if (!ctx || !d_U || !d_field || !d_out) return fail(-1, "NULL argument");
  int rc = check_mesh(Nx, Ny, Nz, h);
  if (rc) return rc;
  if (!rtype_ok(rtype)) return fail(-1, "U type must be F32 or F64");
  if (!itype_ok(ftype)) return fail(-1, "unsupported field type %d", ftype);
  if (order != 0 && order != 1) return fail(-1, "order must be 0 or 1");
  if (mode != 0 && mode != 1) return fail(-1, "mode must be 0 (nearest) or 1 (constant)");
  MeshDev M = make_mesh(Nx, Ny, Nz, h);
  if ((rc = post_grid_ok(M))) return rc;
  PostIn P = post_in(M, d_U);
  post_img(P, d_field, ftype, M.Nvx, M.Nvy, M.Nvz);
  P.order = order; P.mode = mode; P.cval = cval;
  P.out = d_out;
  return run_post_timed(ctx, POST_PUSH, rtype, P);
}

int vdvc_zncc(vdvc_ctx* ctx, int64_t n, const void* d_f, int ftype, const void* d_g, int gtype,
              int g_is_residual, const uint8_t* d_mask, double* h_out) {
  if (!ctx || !h_out) return fail(-1, "NULL argument");
  if (n < 0) return fail(-1, "n < 0");
  if (n > 0 && (!d_f || !d_g)) return fail(-1, "NULL argument");
  if (!itype_ok(ftype) || !itype_ok(gtype)) return fail(-1, "unsupported element type");
  const double nan = std::nan("");
  if (n == 0) {
    *h_out = nan;
    return 0;
  }
  const int nb = (int)std::min<int64_t>(1024, (n + VDVC_BLOCK - 1) / VDVC_BLOCK);
  int rc;
  if ((rc = ctx->stats.ensure((size_t)nb * 3 * sizeof(double)))) return rc;
  if ((rc = ctx->red.ensure(8 * sizeof(double)))) return rc;
  ZnccIn Z;
  Z.n = n; Z.f = d_f; Z.ftype = ftype; Z.g = d_g; Z.gtype = gtype;
  Z.g_is_residual = g_is_residual ? 1 : 0; Z.mask = d_mask;
  auto al16 = [](const void* p) { return p == nullptr || ((uintptr_t)p & 15) == 0; };
  Z.vec = al16(d_f) && al16(d_g) && al16(d_mask) ? 1 : 0;
  Z.chunk = ((n + nb - 1) / nb + 1023) / 1024 * 1024;
  double* part = ctx->stats.as<double>();
  double* red = ctx->red.as<double>();
  VDVC_LAUNCH(vdvc::launch_zncc_pass(0, Z, nullptr, part, nb, ctx->stream));
  VDVC_LAUNCH(vdvc::launch_reduce(part, nb, 3, red, ctx->stream));
  double s0[3];
  if ((rc = fetch(ctx, red, 3 * sizeof(double), s0))) return rc;
  if (!(s0[0] > 0.0)) {  // empty selection: the mean is nan
    *h_out = nan;
    return 0;
  }
  VDVC_LAUNCH(vdvc::launch_zncc_pass(1, Z, red, part, nb, ctx->stream));
  VDVC_LAUNCH(vdvc::launch_reduce(part, nb, 3, red + 4, ctx->stream));
  double s1[3];
  if ((rc = fetch(ctx, red + 4, 3 * sizeof(double), s1))) return rc;
  const double den = std::sqrt(s1[1] * s1[2]);
  *h_out = den > 0.0 ? s1[0] / den : nan;
  return 0;
}

int vdvc_quality_cell(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_res, int rtype,
                      const void* d_ref, int ftype, double* d_std, double* d_zncc) {
  if (!ctx || !d_res || !d_ref || !d_std || !d_zncc) return fail(-1, "NULL argument");
  int rc = check_mesh(Nx, Ny, Nz, h);
  if (rc) return rc;
  if (!rtype_ok(rtype)) return fail(-1, "residual type must be F32 or F64");
  if (!itype_ok(ftype)) return fail(-1, "unsupported image type %d", ftype);
  MeshDev M = make_mesh(Nx, Ny, Nz, h);
  QualIn Q;
  Q.Nx = Nx; Q.Ny = Ny; Q.Nz = Nz; Q.h = h; Q.Nvx = M.Nvx; Q.Nvy = M.Nvy;
  Q.res = d_res; Q.rtype = rtype; Q.ref = d_ref; Q.ftype = ftype;
  Q.std_out = d_std; Q.zncc_out = d_zncc; Q.min_active = 8;
  if ((rc = timed_event(ctx))) return rc;
  VDVC_LAUNCH(vdvc::launch_quality_cell(Q, ctx->stream));
  return timed_event(ctx);
}

// ---- element-scale write-output fields ----

static int cell_voxel_grid_ok(int Nx, int Ny, int Nz, int h) {
  if (Ny * (long long)h + 1 > 65535 || Nz * (long long)h + 1 > 65535)
    return fail(-1, "voxel grid y / z extent exceeds 65535");
  return 0;
}

int vdvc_cell_masks(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const uint8_t* d_active,
                    const uint8_t* d_mask0, double frac_threshold, double min_material_frac,
                    uint8_t* d_unsafe_elem, uint8_t* d_unsafe_node, uint8_t* d_unsafe_vox,
                    uint8_t* d_overlap_lost) {
  if (!ctx || !d_active) return fail(-1, "NULL argument");
  if (d_overlap_lost && !d_mask0) return fail(-1, "the overlap-lost mask needs mask0");
  int rc = check_mesh(Nx, Ny, Nz, h);
  if (rc) return rc;
  if (d_unsafe_vox && (rc = cell_voxel_grid_ok(Nx, Ny, Nz, h))) return rc;
  const MeshDev M = make_mesh(Nx, Ny, Nz, h);
  const size_t Ne = (size_t)Nx * Ny * Nz;
  for (int k = 0; k < 3; ++k)
    if ((rc = ctx->vtmp[k].ensure(Ne * sizeof(int)))) return rc;
  uint8_t* ue = d_unsafe_elem;
  if (!ue && d_unsafe_vox) {
    if ((rc = ctx->vtmp[3].ensure(Ne))) return rc;
    ue = ctx->vtmp[3].as<uint8_t>();
  }
  CellIn C;
  C.Nx = Nx; C.Ny = Ny; C.Nz = Nz; C.h = h; C.Nvx = M.Nvx; C.Nvy = M.Nvy;
  C.a = d_active; C.b = d_mask0;
  C.c0 = ctx->vtmp[0].as<int>(); C.c1 = ctx->vtmp[1].as<int>(); C.c2 = ctx->vtmp[2].as<int>();
  C.all_a0 = nullptr;
  CellFlagsIn F;
  F.Nx = Nx; F.Ny = Ny; F.Nz = Nz; F.h = h;
  F.c0 = C.c0; F.c1 = C.c1; F.c2 = C.c2;
  F.frac_threshold = frac_threshold; F.min_material_frac = min_material_frac;
  F.unsafe_elem = ue; F.lost_elem = d_overlap_lost; F.unsafe_node = d_unsafe_node;
  if ((rc = timed_event(ctx))) return rc;
  VDVC_LAUNCH(vdvc::launch_elem_count(C, ctx->stream));
  VDVC_LAUNCH(vdvc::launch_elem_flags(F, ctx->stream));
  if (d_unsafe_vox) VDVC_LAUNCH(vdvc::launch_unsafe_voxels(Nx, Ny, Nz, h, ue, d_unsafe_vox, ctx->stream));
  return timed_event(ctx);
}

int vdvc_safe_cells(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const uint8_t* d_unsafe_vox,
                    uint8_t* d_safe) {
  if (!ctx || !d_unsafe_vox || !d_safe) return fail(-1, "NULL argument");
  int rc = check_mesh(Nx, Ny, Nz, h);
  if (rc) return rc;
  const MeshDev M = make_mesh(Nx, Ny, Nz, h);
  CellIn C;
  C.Nx = Nx; C.Ny = Ny; C.Nz = Nz; C.h = h; C.Nvx = M.Nvx; C.Nvy = M.Nvy;
  C.a = d_unsafe_vox; C.b = nullptr;
  C.c0 = nullptr; C.c1 = nullptr; C.c2 = nullptr; C.all_a0 = d_safe;
  if ((rc = timed_event(ctx))) return rc;
  VDVC_LAUNCH(vdvc::launch_elem_count(C, ctx->stream));
  return timed_event(ctx);
}

static const double kEquivStrain = std::sqrt(2.0) / 3.0;

int vdvc_principal_strains(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_U,
                           int utype, double* d_principal, double* d_equiv) {
  if (!ctx || !d_U || !d_principal) return fail(-1, "NULL argument");
  int rc = check_mesh(Nx, Ny, Nz, h);
  if (rc) return rc;
  if (!rtype_ok(utype)) return fail(-1, "U type must be VDVC_F32 or VDVC_F64");
  StrainIn S;
  S.Nx = Nx; S.Ny = Ny; S.Nz = Nz; S.U = d_U; S.ucode = utype;
  S.g = 0.25 / (double)h; S.k_eq = kEquivStrain;
  S.principal = d_principal; S.equiv = d_equiv;
  if ((rc = timed_event(ctx))) return rc;
  VDVC_LAUNCH(vdvc::launch_strains(S, ctx->stream));
  return timed_event(ctx);
}

int vdvc_equivalent_strain(vdvc_ctx* ctx, long long n, const double* d_principal,
                           double* d_out) {
  if (!ctx || !d_principal || !d_out) return fail(-1, "NULL argument");
  if (n < 0 || n > 0x7fffffffLL) return fail(-1, "invalid count %lld", n);
  if (n == 0) return 0;
  VDVC_LAUNCH(vdvc::launch_equiv_strain(n, d_principal, kEquivStrain, d_out, ctx->stream));
  return 0;
}

int vdvc_gn_convergence(vdvc_ctx* ctx, int Nx, int Ny, int Nz, const void* d_dU, int dtype,
                        const void* d_U, int utype, double* d_out) {
  if (!ctx || !d_dU || !d_U || !d_out) return fail(-1, "NULL argument");
  int rc = check_mesh(Nx, Ny, Nz, 1);
  if (rc) return rc;
  if (!rtype_ok(dtype) || !rtype_ok(utype)) return fail(-1, "dU / U types must be VDVC_F32 or VDVC_F64");
  if ((rc = timed_event(ctx))) return rc;
  VDVC_LAUNCH(vdvc::launch_gn_convergence(Nx, Ny, Nz, d_dU, dtype, d_U, utype, d_out, ctx->stream));
  return timed_event(ctx);
}

}  // extern "C"

namespace {

// Moved element centres (pts) and regular query centres + crop (qry) of a
// mesh, and the uniform grid of cells over the moved centres that the exact
// nearest-centre search walks (counting sort by cell: off / list).  The
// grid box is the moved centres' bounding box clipped to the query box
// +- 64 h, so a few far outliers (a diverged element) cannot coarsen the
// cells for everyone else; points beyond the box are filed in its boundary
// cells, which keeps the search exact (they are visited with those cells
// and are farther than the ring bound the search stops on).
struct NNPrep {
  double* pts;
  double* qry;
  NNGrid g;
  int* off;
  int* list;
};

int nn_prepare(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_U, int utype,
               const double* h_crop, NNPrep* P) {
  const long long Ne = (long long)Nx * Ny * Nz;
  const int nparts = 256;
  int rc;
  if ((rc = ctx->vtmp[0].ensure((size_t)Ne * 3 * sizeof(double)))) return rc;
  if ((rc = ctx->vtmp[1].ensure((size_t)Ne * 3 * sizeof(double)))) return rc;
  if ((rc = ctx->stats.ensure((size_t)nparts * 6 * sizeof(double)))) return rc;
  if ((rc = ctx->red.ensure(8 * sizeof(double)))) return rc;
  P->pts = ctx->vtmp[0].as<double>();
  P->qry = ctx->vtmp[1].as<double>();
  VDVC_LAUNCH(vdvc::launch_elem_centres(Nx, Ny, Nz, h, d_U, utype, h_crop, P->pts, P->qry,
                                        ctx->stream));
  VDVC_LAUNCH(vdvc::launch_bbox(Ne, P->pts, ctx->stats.as<double>(), nparts,
                                ctx->red.as<double>(), ctx->stream));
  double bb[6];
  if ((rc = fetch(ctx, ctx->red.p, 6 * sizeof(double), bb))) return rc;
  for (int c = 0; c < 6; ++c)
    if (!std::isfinite(bb[c])) return fail(-1, "displacement field is not finite");
  const int N[3] = {Nx, Ny, Nz};
  for (int c = 0; c < 3; ++c) {
    const double qlo = 0.5 * h + h_crop[c], qhi = (N[c] - 0.5) * h + h_crop[c];
    const double lo = qlo - 64.0 * h, hi = qhi + 64.0 * h;
    bb[c] = std::min(std::max(bb[c], lo), hi);
    bb[3 + c] = std::min(std::max(bb[3 + c], lo), hi);
  }
  // cells of edge h over the box, coarsened while there are more than 4
  // cells per point
  NNGrid& g = P->g;
  double cs = (double)h;
  long long ncell = 0;
  for (int it = 0; it < 64; ++it) {
    ncell = 1;
    for (int c = 0; c < 3; ++c) {
      const double span = std::floor((bb[3 + c] - bb[c]) / cs) + 1.0;
      g.G[c] = span > 1e9 ? 1000000000 : (int)span;
      ncell *= g.G[c];
      if (ncell > 0x3fffffffLL) break;
    }
    if (ncell <= 4 * Ne + 64) break;
    cs *= 2.0;
  }
  if (ncell > 4 * Ne + 64) return fail(-1, "point spread too large for the remap grid");
  for (int c = 0; c < 3; ++c) g.o[c] = bb[c];
  g.cs = cs;
  g.inv_cs = 1.0 / cs;
  const long long nchunk = vdvc::nn_scan_chunks(ncell);
  if ((rc = ctx->vtmp[3].ensure((size_t)(ncell + nchunk) * sizeof(int)))) return rc;
  if ((rc = ctx->vtmp[2].ensure((size_t)(ncell + 1 + Ne) * sizeof(int)))) return rc;
  int* cnt = ctx->vtmp[3].as<int>();
  int* csum = cnt + ncell;
  P->off = ctx->vtmp[2].as<int>();
  P->list = P->off + ncell + 1;
  // zero the int counters (float 0.0 has all-zero bits)
  VDVC_LAUNCH(vdvc::launch_fill(ncell, cnt, VDVC_F32, 0.0, ctx->stream));
  VDVC_LAUNCH(vdvc::launch_nn_build(Ne, P->pts, g, cnt, P->off, csum, P->list, ctx->stream));
  return 0;
}

int remap_args_ok(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_U, int utype,
                  const double* h_crop) {
  if (!ctx || !d_U || !h_crop) return fail(-1, "NULL argument");
  int rc = check_mesh(Nx, Ny, Nz, h);
  if (rc) return rc;
  if (!rtype_ok(utype)) return fail(-1, "U type must be VDVC_F32 or VDVC_F64");
  for (int c = 0; c < 3; ++c)
    if (!std::isfinite(h_crop[c])) return fail(-1, "crop offset must be finite");
  return 0;
}

}  // namespace

extern "C" {

int vdvc_remap_mask_deformed(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_U,
                             int utype, const double* h_crop, const uint8_t* d_mask,
                             uint8_t* d_out) {
  int rc = remap_args_ok(ctx, Nx, Ny, Nz, h, d_U, utype, h_crop);
  if (rc) return rc;
  if (!d_mask || !d_out) return fail(-1, "NULL argument");
  const long long Ne = (long long)Nx * Ny * Nz;
  if ((rc = timed_event(ctx))) return rc;
  NNPrep P;
  if ((rc = nn_prepare(ctx, Nx, Ny, Nz, h, d_U, utype, h_crop, &P))) return rc;
  VDVC_LAUNCH(vdvc::launch_nn_query(Ne, P.pts, P.qry, P.g, P.off, P.list, d_mask, d_out, nullptr,
                                    ctx->stream));
  return timed_event(ctx);
}

int vdvc_remap_field_deformed(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_U,
                              int utype, const double* h_crop, const double* d_field, int C,
                              double* d_out) {
  int rc = remap_args_ok(ctx, Nx, Ny, Nz, h, d_U, utype, h_crop);
  if (rc) return rc;
  if (!d_field || !d_out) return fail(-1, "NULL argument");
  if (C < 1 || C > 64) return fail(-1, "channels must be 1..64 (got %d)", C);
  const long long Ne = (long long)Nx * Ny * Nz;
  if ((rc = timed_event(ctx))) return rc;
  NNPrep P;
  if ((rc = nn_prepare(ctx, Nx, Ny, Nz, h, d_U, utype, h_crop, &P))) return rc;
  if ((rc = ctx->nodeA.ensure((size_t)Ne * sizeof(int)))) return rc;
  int* idx = ctx->nodeA.as<int>();
  VDVC_LAUNCH(vdvc::launch_nn_query(Ne, P.pts, P.qry, P.g, P.off, P.list, nullptr, nullptr, idx,
                                    ctx->stream));
  VDVC_LAUNCH(vdvc::launch_remap_linear(Nx, Ny, Nz, h, P.pts, P.qry, d_field, C, idx, d_out,
                                        ctx->stream));
  return timed_event(ctx);
}

int vdvc_material_cells(vdvc_ctx* ctx, int Nx, int Ny, int Nz, int h, const void* d_ref,
                        int itype, double glt, uint8_t* d_out) {
  if (!ctx || !d_ref || !d_out) return fail(-1, "NULL argument");
  int rc = check_mesh(Nx, Ny, Nz, h);
  if (rc) return rc;
  if (!itype_ok(itype)) return fail(-1, "unsupported image type");
  VDVC_LAUNCH(vdvc::launch_material_cells(Nx, Ny, Nz, h, d_ref, itype, glt, d_out, ctx->stream));
  return 0;
}

int vdvc_erode_cells(vdvc_ctx* ctx, int Nx, int Ny, int Nz, const uint8_t* d_in, uint8_t* d_out) {
  if (!ctx || !d_in || !d_out) return fail(-1, "NULL argument");
  if (Nx < 1 || Ny < 1 || Nz < 1) return fail(-1, "invalid mesh");
  if ((long long)Nx * Ny * Nz > 0x7fffffffLL) return fail(-1, "mesh too large");
  VDVC_LAUNCH(vdvc::launch_erode_cells(Nx, Ny, Nz, d_in, d_out, ctx->stream));
  return 0;
}

}  // extern "C"

// ---- preprocess affine stage ----

static int dims_ok(int nx, int ny, int nz) {
  if (nx < 2 || ny < 2 || nz < 2) return fail(-1, "volume dims must be >= 2 (got %d %d %d)", nx, ny, nz);
  if (ny > 65535 || nz > 65535) return fail(-1, "ny, nz must be <= 65535");
  if ((long long)nx * ny * nz > 0x7fffffffLL) return fail(-1, "volume exceeds 2^31 voxels");
  return 0;
}

static int gauss_in(GaussIn& G, int nx, int ny, int nz, const double* h_w, int r, int mode,
                    double cval) {
  if (!h_w || r < 0 || r > VDVC_GAUSS_MAX_R) return fail(-1, "radius must be 0..%d", VDVC_GAUSS_MAX_R);
  if (mode != 0 && mode != 1) return fail(-1, "mode must be 0 (nearest) or 1 (constant)");
  std::memset(&G, 0, sizeof(G));
  G.nx = nx; G.ny = ny; G.nz = nz; G.mode = mode; G.r = r; G.cval = cval;
  for (int k = 0; k <= r; ++k) G.w[k] = h_w[k];
  return 0;
}

// three axis passes in -> out -> tmp -> out
static int gauss3(vdvc_ctx* ctx, GaussIn G, const float* in, float* out, float* tmp) {
  G.axis = 0;
  VDVC_LAUNCH(vdvc::launch_gauss_axis(G, in, out, ctx->stream));
  G.axis = 1;
  VDVC_LAUNCH(vdvc::launch_gauss_axis(G, out, tmp, ctx->stream));
  G.axis = 2;
  VDVC_LAUNCH(vdvc::launch_gauss_axis(G, tmp, out, ctx->stream));
  return 0;
}

extern "C" {

int vdvc_gaussian_filter(vdvc_ctx* ctx, int nx, int ny, int nz, const float* d_in, float* d_out,
                         const double* h_w, int r, int mode, double cval) {
  if (!ctx || !d_in || !d_out) return fail(-1, "NULL argument");
  if (d_in == d_out) return fail(-1, "in-place filtering is not supported");
  int rc = dims_ok(nx, ny, nz);
  if (rc) return rc;
  GaussIn G;
  if ((rc = gauss_in(G, nx, ny, nz, h_w, r, mode, cval))) return rc;
  const size_t n = (size_t)nx * ny * nz;
  if ((rc = ctx->vtmp[0].ensure(n * sizeof(float)))) return rc;
  return gauss3(ctx, G, d_in, d_out, ctx->vtmp[0].as<float>());
}

int vdvc_smooth_masked(vdvc_ctx* ctx, int nx, int ny, int nz, const float* d_g, double cval,
                       const double* h_w, int r, float* d_out, uint8_t* d_valid) {
  if (!ctx || !d_g || !d_out) return fail(-1, "NULL argument");
  int rc = dims_ok(nx, ny, nz);
  if (rc) return rc;
  const long long n = (long long)nx * ny * nz;
  if (r < 0) {  // no smoothing: the validity mask only
    if (!d_valid) return fail(-1, "NULL valid mask");
    VDVC_LAUNCH(vdvc::launch_mask_prep(n, d_g, (float)cval, nullptr, nullptr, d_valid,
                                       ctx->stream));
    return 0;
  }
  GaussIn G;
  if ((rc = gauss_in(G, nx, ny, nz, h_w, r, 1, 0.0))) return rc;
  for (int k = 0; k < 4; ++k)
    if ((rc = ctx->vtmp[k].ensure((size_t)n * sizeof(float)))) return rc;
  float* gv = ctx->vtmp[0].as<float>();
  float* vf = ctx->vtmp[1].as<float>();
  float* num = ctx->vtmp[2].as<float>();
  float* tmp = ctx->vtmp[3].as<float>();
  VDVC_LAUNCH(vdvc::launch_mask_prep(n, d_g, (float)cval, gv, vf, d_valid, ctx->stream));
  if ((rc = gauss3(ctx, G, gv, num, tmp))) return rc;
  if ((rc = gauss3(ctx, G, vf, gv, tmp))) return rc;   // den into gv
  VDVC_LAUNCH(vdvc::launch_masked_ratio(n, num, gv, d_g, d_out, ctx->stream));
  return 0;
}

int vdvc_gradient3(vdvc_ctx* ctx, int nx, int ny, int nz, const float* d_f, float* d_gx,
                   float* d_gy, float* d_gz) {
  if (!ctx || !d_f || !d_gx || !d_gy || !d_gz) return fail(-1, "NULL argument");
  int rc = dims_ok(nx, ny, nz);
  if (rc) return rc;
  VDVC_LAUNCH(vdvc::launch_gradient(nx, ny, nz, d_f, d_gx, d_gy, d_gz, ctx->stream));
  return 0;
}

static void aff_in(AffIn& A, int nx, int ny, int nz, const double* M, const double* tr,
                   const double* c, double cval) {
  std::memset(&A, 0, sizeof(A));
  A.nx = nx; A.ny = ny; A.nz = nz; A.cval = cval;
  for (int k = 0; k < 9; ++k) A.M[k] = (float)M[k];
  for (int k = 0; k < 3; ++k) {
    A.tr[k] = (float)tr[k];
    A.c[k] = (float)c[k];
  }
}

int vdvc_affine_normal(vdvc_ctx* ctx, int nx, int ny, int nz, const float* d_f, const float* d_g,
                       const float* d_gx, const float* d_gy, const float* d_gz,
                       const uint8_t* d_valid, const double* h_M, const double* h_tr,
                       const double* h_c, double cval, double* h_H, double* h_b) {
  if (!ctx || !d_f || !d_g || !d_gx || !d_gy || !d_gz || !d_valid || !h_M || !h_tr || !h_c ||
      !h_H || !h_b)
    return fail(-1, "NULL argument");
  int rc = dims_ok(nx, ny, nz);
  if (rc) return rc;
  AffIn A;
  aff_in(A, nx, ny, nz, h_M, h_tr, h_c, cval);
  A.f = d_f; A.g = d_g; A.gx = d_gx; A.gy = d_gy; A.gz = d_gz; A.valid = d_valid;
  // two 4-wavefront blocks per CU at the kernel's occupancy of 2 per SIMD
  const long long nvox = (long long)nx * ny * nz;
  const int nb = (int)std::min<long long>(512, (nvox + VDVC_BLOCK - 1) / VDVC_BLOCK);
  const int nw = nb * VDVC_WAVES_PER_BLOCK;
  if ((rc = ctx->stats.ensure((size_t)nw * 90 * sizeof(double)))) return rc;
  if ((rc = ctx->red.ensure(96 * sizeof(double)))) return rc;
  if ((rc = timed_event(ctx))) return rc;
  VDVC_LAUNCH(vdvc::launch_affine_normal(A, ctx->stats.as<double>(), nb, ctx->stream));
  if ((rc = timed_event(ctx))) return rc;
  VDVC_LAUNCH(vdvc::launch_reduce(ctx->stats.as<double>(), nw, 90, ctx->red.as<double>(),
                                  ctx->stream));
  double s[90];
  if ((rc = fetch(ctx, ctx->red.p, 90 * sizeof(double), s))) return rc;
  int e = 0;
  for (int a = 0; a < 12; ++a)
    for (int b = a; b < 12; ++b, ++e) h_H[12 * a + b] = h_H[12 * b + a] = s[e];
  for (int a = 0; a < 12; ++a) h_b[a] = s[78 + a];
  return 0;
}

int vdvc_affine_rms(vdvc_ctx* ctx, int nx, int ny, int nz, const float* d_f, const float* d_g,
                    const uint8_t* d_valid, const double* h_M, const double* h_tr,
                    const double* h_c, double cval, double* h_out2) {
  if (!ctx || !d_f || !d_g || !d_valid || !h_M || !h_tr || !h_c || !h_out2)
    return fail(-1, "NULL argument");
  int rc = dims_ok(nx, ny, nz);
  if (rc) return rc;
  AffIn A;
  aff_in(A, nx, ny, nz, h_M, h_tr, h_c, cval);
  A.f = d_f; A.g = d_g; A.valid = d_valid;
  const int nb = 2048;
  if ((rc = ctx->stats.ensure((size_t)nb * 2 * sizeof(double)))) return rc;
  if ((rc = ctx->red.ensure(96 * sizeof(double)))) return rc;
  VDVC_LAUNCH(vdvc::launch_affine_rms(A, ctx->stats.as<double>(), nb, ctx->stream));
  VDVC_LAUNCH(vdvc::launch_reduce(ctx->stats.as<double>(), nb, 2, ctx->red.as<double>(),
                                  ctx->stream));
  return fetch(ctx, ctx->red.p, 2 * sizeof(double), h_out2);
}

int vdvc_warp_affine(vdvc_ctx* ctx, int nx, int ny, int nz, const float* d_g, int gnx, int gny,
                     int gnz, const double* h_A, const double* h_t, const double* h_c,
                     float* d_out) {
  if (!ctx || !d_g || !d_out || !h_A || !h_t || !h_c) return fail(-1, "NULL argument");
  if (nx < 1 || ny < 1 || nz < 1 || ny > 65535 || nz > 65535)
    return fail(-1, "invalid output grid %d %d %d", nx, ny, nz);
  if (gnx < 1 || gny < 1 || gnz < 1) return fail(-1, "invalid image dims");
  WarpIn W;
  std::memset(&W, 0, sizeof(W));
  W.nx = nx; W.ny = ny; W.nz = nz; W.gnx = gnx; W.gny = gny; W.gnz = gnz;
  for (int k = 0; k < 9; ++k) W.A[k] = (float)h_A[k];
  for (int k = 0; k < 3; ++k) {
    W.t[k] = (float)h_t[k];
    W.c[k] = (float)h_c[k];
  }
  VDVC_LAUNCH(vdvc::launch_warp_affine(W, d_g, d_out, ctx->stream));
  return 0;
}

int vdvc_overlap_bounds(vdvc_ctx* ctx, int nx, int ny, int nz, const void* d_f, int ftype,
                        const float* d_gw, double glt, int* h_lo3, int* h_hi3) {
  if (!ctx || !d_f || !d_gw || !h_lo3 || !h_hi3) return fail(-1, "NULL argument");
  if (nx < 1 || ny < 1 || nz < 1 || ny > 65535 || nz > 65535) return fail(-1, "invalid dims");
  if (!itype_ok(ftype)) return fail(-1, "unsupported image type %d", ftype);
  int rc;
  const size_t nf = (size_t)nx + ny + nz;
  if ((rc = ctx->vtmp[0].ensure(nf))) return rc;
  uint8_t* flags = ctx->vtmp[0].as<uint8_t>();
  VDVC_HIP(hipMemsetAsync(flags, 0, nf, ctx->stream));
  VDVC_LAUNCH(vdvc::launch_overlap_proj(nx, ny, nz, d_f, ftype, d_gw, glt, flags, ctx->stream));
  std::vector<uint8_t> h(nf);
  VDVC_HIP(hipMemcpyAsync(h.data(), flags, nf, hipMemcpyDeviceToHost, ctx->stream));
  VDVC_HIP(hipStreamSynchronize(ctx->stream));
  const int n3[3] = {nx, ny, nz};
  size_t off = 0;
  for (int a = 0; a < 3; ++a) {
    h_lo3[a] = h_hi3[a] = -1;
    for (int i = 0; i < n3[a]; ++i)
      if (h[off + i]) {
        if (h_lo3[a] < 0) h_lo3[a] = i;
        h_hi3[a] = i;
      }
    off += n3[a];
  }
  return 0;
}

}  // extern "C"

// ---- data movement at the solve's boundary (no reference arithmetic) ------

namespace {
int any_type_ok(int t) { return t >= VDVC_F32 && t <= VDVC_BOOL; }
}  // namespace

extern "C" {

int vdvc_convert(vdvc_ctx* ctx, int64_t n, const void* d_src, int stype, void* d_dst,
                 int dtype) {
  if (!ctx || (n > 0 && (!d_src || !d_dst))) return fail(-1, "NULL argument");
  if (n < 0) return fail(-1, "n < 0");
  if (!any_type_ok(stype) || !any_type_ok(dtype)) return fail(-1, "unsupported element type");
  VDVC_LAUNCH(vdvc::launch_convert(n, d_src, stype, d_dst, dtype, ctx->stream));
  return 0;
}

int vdvc_copy(vdvc_ctx* ctx, int64_t nbytes, const void* d_src, void* d_dst) {
  if (!ctx || (nbytes > 0 && (!d_src || !d_dst))) return fail(-1, "NULL argument");
  if (nbytes < 0) return fail(-1, "nbytes < 0");
  if (nbytes == 0) return 0;
  int rc = timed_event(ctx);
  if (rc) return rc;
  VDVC_LAUNCH(vdvc::launch_copy(nbytes, d_src, d_dst, ctx->stream));
  return timed_event(ctx);
}

int vdvc_stream_read(vdvc_ctx* ctx, int64_t nbytes, const void* d_src, int grid, uint32_t key,
                     uint32_t* d_out) {
  if (!ctx || !d_src || !d_out) return fail(-1, "NULL argument");
  if (nbytes < 16 || (nbytes & 15) || (reinterpret_cast<uintptr_t>(d_src) & 15))
    return fail(-1, "the probe reads a positive multiple of 16 bytes from a 16-byte aligned buffer");
  if (grid < 1) return fail(-1, "grid < 1");
  int rc = timed_event(ctx);
  if (rc) return rc;
  VDVC_LAUNCH(vdvc::launch_stream_read(nbytes, d_src, grid, key, d_out, ctx->stream));
  return timed_event(ctx);
}

int vdvc_fill(vdvc_ctx* ctx, int64_t n, int dtype, double value, void* d_dst) {
  if (!ctx || (n > 0 && !d_dst)) return fail(-1, "NULL argument");
  if (n < 0) return fail(-1, "n < 0");
  if (!any_type_ok(dtype)) return fail(-1, "unsupported element type");
  VDVC_LAUNCH(vdvc::launch_fill(n, d_dst, dtype, value, ctx->stream));
  return 0;
}

int vdvc_copy_volume(vdvc_ctx* ctx, int nx, int ny, int nz, const void* d_src, int stype,
                     int64_t sx, int64_t sy, int64_t sz, void* d_dst, int dtype, int dpitch) {
  if (!ctx || !d_src || !d_dst) return fail(-1, "NULL argument");
  if (nx < 1 || ny < 1 || nz < 1) return fail(-1, "invalid volume %d x %d x %d", nx, ny, nz);
  if (dpitch < nx) return fail(-1, "dpitch %d < nx %d", dpitch, nx);
  if (!any_type_ok(stype) || !any_type_ok(dtype)) return fail(-1, "unsupported element type");
  if (ny > 65535 || nz > 65535) return fail(-1, "ny and nz must be <= 65535");
  if (sx < 0 || sy < 0 || sz < 0) return fail(-1, "negative source stride");
  VolCopy C;
  C.nx = nx; C.ny = ny; C.nz = nz;
  C.sx = sx; C.sy = sy; C.sz = sz;
  C.src = d_src; C.st = stype;
  C.dst = d_dst; C.dt = dtype;
  C.dp = dpitch;
  VDVC_LAUNCH(vdvc::launch_copy_volume(C, ctx->stream));
  return 0;
}

int vdvc_resample_nodes(vdvc_ctx* ctx, int Mx, int My, int Mz, int hs, const void* d_src,
                        int Nx, int Ny, int Nz, int hd, void* d_dst, int rtype) {
  if (!ctx || !d_src || !d_dst) return fail(-1, "NULL argument");
  if (Mx < 1 || My < 1 || Mz < 1 || hs < 1 || Nx < 1 || Ny < 1 || Nz < 1 || hd < 1)
    return fail(-1, "invalid meshes");
  if ((long long)Mx * hs != (long long)Nx * hd || (long long)My * hs != (long long)Ny * hd ||
      (long long)Mz * hs != (long long)Nz * hd)
    return fail(-1, "the meshes must cover the same voxels (M*hs == N*hd per axis)");
  if (!rtype_ok(rtype)) return fail(-1, "rtype must be VDVC_F32 or VDVC_F64");
  if (3LL * (Nx + 1) * (Ny + 1) * (Nz + 1) > 0x7fffffffLL) return fail(-1, "mesh too large");
  NodeResample R;
  R.Mx = Mx; R.My = My; R.Mz = Mz; R.hs = hs;
  R.Nx = Nx; R.Ny = Ny; R.Nz = Nz; R.hd = hd;
  R.src = d_src; R.dst = d_dst; R.rt = rtype;
  VDVC_LAUNCH(vdvc::launch_resample_nodes(R, ctx->stream));
  return 0;
}

int vdvc_axpby_x(vdvc_ctx* ctx, int64_t n, double a, const void* d_x, int xtype, double b,
                 void* d_y, int ytype) {
  if (!ctx || (n > 0 && (!d_x || !d_y))) return fail(-1, "NULL argument");
  if (n < 0) return fail(-1, "n < 0");
  if (!ftype_ok(xtype) && xtype != VDVC_BF16) return fail(-1, "unsupported x type");
  if (!rtype_ok(ytype)) return fail(-1, "y must be VDVC_F32 or VDVC_F64");
  VDVC_LAUNCH(vdvc::launch_axpby_x(n, a, d_x, xtype, b, d_y, ytype, ctx->stream));
  return 0;
}

}  // extern "C"
