// ogs_noalloc.hip -- ogs_setup_global_to_local_noalloc: the global-to-local
// mapping of an ogs handle (ogs_t::SetupGlobalToLocalMapping,
// ogsSetup.cpp:867-886) without any device allocation, for callers that
// cannot allocate (code under graph capture, tight serving loops, harnesses
// that forbid hipMalloc).
//
// ogs_setup_global_to_local stages the node -> row map in device buffers it
// allocates per call.  This variant stages it in `scratch`, a fixed
// file-scope device array of kNoAllocMax entries: each GPU the library runs
// on holds its own copy until the process ends, and every call on that GPU
// shares it (calls that overlap on one GPU race on it; each call rewrites
// all the entries it uses before reading them).  In host code the array's
// name is not a device address: its address on each device comes from
// hipGetSymbolAddress, asked once per device with that device current, and
// only that address is handed to kernels and copies.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "ogs.hpp"
#include "ogs_hip.h"

namespace {

// Largest N the variant serves: 2^22 nodes (16 MiB per GPU).
constexpr int32_t kNoAllocMax = 1 << 22;
constexpr int kMaxDevices = 64;
constexpr int kThreads = 256;

__device__ __align__(256) int32_t scratch[kNoAllocMax];

// Per device: the address of `scratch` there (null until first asked).
int32_t *g_scratch[kMaxDevices] = {};
std::mutex g_lock;

// The device address of `scratch` on the current device.
int scratch_on_current(int32_t **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= kMaxDevices) return OGS_E_ARG;
  std::lock_guard<std::mutex> g(g_lock);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    g_scratch[dev] = static_cast<int32_t *>(p);
  }
  *out = g_scratch[dev];
  return OGS_SUCCESS;
}

__global__ __launch_bounds__(kThreads) void k_fill(int32_t *__restrict__ a, int32_t n, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) a[i] = v;
}

// The NoTrans Scatter of the row indices over the T lists: map[col] = row
// for every member of rows [0, nrows) (ogsOperator.cpp:206-238 with gv[n] =
// n, as ogsSetup.cpp:876-883 uses it).
__global__ __launch_bounds__(kThreads) void k_rows_to_cols(const int32_t *__restrict__ rs,
                                                           const int32_t *__restrict__ ci, int32_t nrows,
                                                           int32_t *__restrict__ map) {
  for (int64_t r = (int64_t)blockIdx.x * kThreads + threadIdx.x; r < nrows; r += (int64_t)gridDim.x * kThreads)
    for (int32_t j = rs[r]; j < rs[r + 1]; ++j) map[ci[j]] = (int32_t)r;
}

int grid_of(int64_t n) {
  const int64_t g = (n + kThreads - 1) / kThreads;
  return (int)(g < 1 ? 1 : g > 4096 ? 4096 : g);
}

}  // namespace

extern "C" int ogs_setup_global_to_local_noalloc(ogs_handle_t h, int32_t *GlobalToLocal) {
  if (!h) return OGS_E_ARG;
  const libp::ogs::ogs_t &o = h->ogs;
  const int32_t N = o.N;
  if (N > 0 && !GlobalToLocal) return OGS_E_ARG;
  if (N > kNoAllocMax) return OGS_E_TOO_LARGE;
  if (N == 0) return OGS_SUCCESS;
  int32_t *map = nullptr;
  int rc = scratch_on_current(&map);
  if (rc) return rc;
  const hipStream_t st = o.platform.stream;
  const libp::ogs::ogsOperator_t &L = o.local();
  hipLaunchKernelGGL(k_fill, dim3(grid_of(N)), dim3(kThreads), 0, st, map, N, -1);
  if (L.NrowsT > 0 && L.nnzT > 0)
    hipLaunchKernelGGL(k_rows_to_cols, dim3(grid_of(L.NrowsT)), dim3(kThreads), 0, st, L.o_rowStartsT.ptr(),
                       L.o_colIdsT.ptr(), L.NrowsT, map);
  rc = (int)hipGetLastError();
  // at one rank the halo operator has no rows (ogsSetup.cpp:884-885 scatters none)
  if (!rc) rc = (int)hipMemcpyAsync(GlobalToLocal, map, sizeof(int32_t) * (size_t)N, hipMemcpyDeviceToHost, st);
  if (!rc) rc = (int)hipStreamSynchronize(st);
  return rc;
}
