// Streaming-bandwidth probe: the practical HBM ceiling the voxel-streaming
// kernels are compared against.  Times a read-only 16-byte-per-lane
// reduction over 2 GiB (default and nontemporal cache policy, several grid
// sizes) and a 2 GiB copy with HIP events; prints one line per variant.
//
//   python -m voxeldvc_amd.build --probe && ./voxeldvc_amd/probe_hbm_stream
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

#define CK(x)                                                                  \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      std::exit(1);                                                            \
    }                                                                          \
  } while (0)

// grid-stride read of n 16-byte words, UNR independent loads in flight per
// lane; the xor of everything read is stored per thread so no load is dead
template <bool NT, int UNR>
__global__ __launch_bounds__(256) void read_kernel(const u32x4* __restrict__ a, long long n,
                                                   unsigned* __restrict__ out) {
  unsigned acc = 0;
  const long long stride = (long long)gridDim.x * 256;
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  for (; i + (UNR - 1) * stride < n; i += UNR * stride) {
    u32x4 v[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u)
      v[u] = NT ? __builtin_nontemporal_load(a + i + u * stride) : a[i + u * stride];
#pragma unroll
    for (int u = 0; u < UNR; ++u) acc ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
  }
  for (; i < n; i += stride) {
    const u32x4 v = a[i];
    acc ^= v.x ^ v.y ^ v.z ^ v.w;
  }
  out[(long long)blockIdx.x * 256 + threadIdx.x] = acc;
}

template <bool NT>
__global__ __launch_bounds__(256) void copy_kernel(const u32x4* __restrict__ a,
                                                   u32x4* __restrict__ b, long long n) {
  const long long stride = (long long)gridDim.x * 256;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
    const u32x4 v = NT ? __builtin_nontemporal_load(a + i) : a[i];
    if (NT) __builtin_nontemporal_store(v, b + i);
    else b[i] = v;
  }
}

int main() {
  const long long bytes = 1LL << 31;
  const long long n = bytes / 16;
  const int max_grid = 16384;
  u32x4 *a, *b;
  unsigned* out;
  CK(hipMalloc(&a, bytes));
  CK(hipMalloc(&b, bytes));
  CK(hipMalloc(&out, (size_t)max_grid * 256 * sizeof(unsigned)));
  CK(hipMemset(a, 1, bytes));
  CK(hipMemset(b, 0, bytes));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  auto time_it = [&](const char* name, int grid, auto launch, double moved) {
    launch();
    CK(hipGetLastError());
    CK(hipDeviceSynchronize());
    const int reps = 10;
    CK(hipEventRecord(e0));
    for (int k = 0; k < reps; ++k) launch();
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    float ms = 0.0f;
    CK(hipEventElapsedTime(&ms, e0, e1));
    const double us = ms / reps * 1e3;
    std::printf("%-16s grid %6d  %8.1f us  %7.1f GB/s\n", name, grid, us, moved / (us * 1e-6) / 1e9);
    std::fflush(stdout);
  };
  for (int g : {1024, 2048, 4096, 8192, 16384}) {
    time_it("read", g, [&] { hipLaunchKernelGGL((read_kernel<false, 4>), dim3(g), dim3(256), 0, 0, a, n, out); },
            (double)bytes);
    time_it("read nt", g, [&] { hipLaunchKernelGGL((read_kernel<true, 4>), dim3(g), dim3(256), 0, 0, a, n, out); },
            (double)bytes);
  }
  for (int g : {2048, 8192}) {
    time_it("copy", g, [&] { hipLaunchKernelGGL((copy_kernel<false>), dim3(g), dim3(256), 0, 0, a, b, n); },
            2.0 * bytes);
    time_it("copy nt", g, [&] { hipLaunchKernelGGL((copy_kernel<true>), dim3(g), dim3(256), 0, 0, a, b, n); },
            2.0 * bytes);
  }
  CK(hipFree(a));
  CK(hipFree(b));
  CK(hipFree(out));
  return 0;
}
