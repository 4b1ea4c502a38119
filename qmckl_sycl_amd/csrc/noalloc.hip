// Allocation-free form of the low-level radial launcher
// qmckl_compute_ao_basis_shell_gaussian_vgl_device (src/qmckl_ao_sycl.cpp:12-102).
//
// The allocating launcher copies the index arrays to the host and uploads a
// shell table it builds there.  This form builds the same table on the
// device, from the caller's device arrays, into a file-scope device array
// `scratch` of fixed size, so it calls no hipMalloc/hipFree and allocates
// nothing on the host: for callers that cannot allocate (graph capture,
// serving loops).  Each GPU holds its own copy of `scratch` until the
// process ends; every call on one GPU uses it, so calls that overlap on one
// GPU (two threads, two streams) race and must be serialised by the caller.
// A call leaves the table of its own inputs behind; the next call rebuilds
// it (its coverage counters are cleared by a fill kernel first).
//
// Scratch layout (bytes): [0, 2 MiB) shell table, kNoAllocMaxShells
// ShellDesc entries; [2 MiB, 2.25 MiB) coverage counters, one int32 per
// shell; then one int32 error word.
#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.hpp"

namespace qmckl_amd {

namespace {

constexpr int64_t kCoverOff = kNoAllocMaxShells * (int64_t)sizeof(ShellDesc);
constexpr int64_t kErrOff = kCoverOff + kNoAllocMaxShells * (int64_t)sizeof(int32_t);
constexpr int64_t kScratchBytes = kErrOff + 256;
static_assert(sizeof(ShellDesc) == 32, "the scratch layout assumes 32-byte shell descriptors");

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void *g_ptr[kMaxDevices] = {nullptr};

// error bits of the table build
constexpr int kErrNucleusRange = 1;  // nucleus_index / nucleus_shell_num out of range
constexpr int kErrCoverage = 2;      // a shell of no nucleus, or of two
constexpr int kErrPrimRange = 4;     // shell_prim_index / shell_prim_num out of range

__global__ void clear_kernel(int S, int32_t *cover, int32_t *err) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < S) cover[s] = 0;
  if (s == 0) *err = 0;
}

// one thread per nucleus: claims its shells
__global__ void nucleus_kernel(int N, int S, const int64_t *__restrict__ nucleus_index,
                               const int64_t *__restrict__ nucleus_shell_num, ShellDesc *shells,
                               int32_t *cover, int32_t *err) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int64_t s0 = nucleus_index[n], ns = nucleus_shell_num[n];
  if (ns < 0 || s0 < 0 || s0 + ns > S) {
    atomicOr(err, kErrNucleusRange);
    return;
  }
  for (int64_t s = s0; s < s0 + ns; ++s) {
    atomicAdd(&cover[s], 1);
    shells[s].nucl = n;
  }
}

// one thread per shell: checks that exactly one nucleus claimed it and
// fills the rest of its descriptor
__global__ void shell_kernel(int S, int Np, const int64_t *__restrict__ shell_prim_index,
                             const int64_t *__restrict__ shell_prim_num, ShellDesc *shells,
                             const int32_t *cover, int32_t *err) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  if (cover[s] != 1) atomicOr(err, kErrCoverage);
  const int64_t p0 = shell_prim_index[s], np = shell_prim_num[s];
  const bool ok = np >= 0 && p0 >= 0 && p0 + np <= Np;
  if (!ok) atomicOr(err, kErrPrimRange);
  ShellDesc d = shells[s];
  d.prim_start = ok ? (int32_t)p0 : 0;
  d.prim_num = ok ? (int32_t)np : 0;
  d.ao_start = s;
  d.l = 0;
  d.shell = s;
  d.a_min = 0.0;
  shells[s] = d;
}

}  // namespace

hipError_t noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(g_mutex);
  if (g_ptr[device] != nullptr) return hipSuccess;
  int cur = -1;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess) return e;
  if (cur != device) return hipErrorInvalidDevice;  // the caller makes `device` current
  void *p = nullptr;
  e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (p == nullptr) return hipErrorInvalidSymbol;
  g_ptr[device] = p;
  return hipSuccess;
}

hipError_t launch_shell_vgl_noalloc(hipStream_t s, int device, int64_t P, int64_t N, int64_t S,
                                    int64_t Np, const int64_t *nucleus_shell_num,
                                    const int64_t *nucleus_index, const double *nucleus_range,
                                    const int64_t *shell_prim_index, const int64_t *shell_prim_num,
                                    const double *coord, const double *nucl, const double *expo,
                                    const double *coef_normalized, double *shell_vgl, int *table_err) {
  *table_err = 0;
  if (S <= 0 || S > kNoAllocMaxShells || N <= 0 || Np <= 0 || N > INT32_MAX || Np > INT32_MAX)
    return hipErrorInvalidValue;
  void *base = nullptr;
  {
    std::lock_guard<std::mutex> g(g_mutex);
    if (device >= 0 && device < kMaxDevices) base = g_ptr[device];
  }
  if (base == nullptr) {
    const hipError_t e = noalloc_scratch_init(device);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> g(g_mutex);
    base = g_ptr[device];
  }
  ShellDesc *shells = static_cast<ShellDesc *>(base);
  int32_t *cover = reinterpret_cast<int32_t *>(static_cast<unsigned char *>(base) + kCoverOff);
  int32_t *err = reinterpret_cast<int32_t *>(static_cast<unsigned char *>(base) + kErrOff);
  const int bs = 256;
  hipLaunchKernelGGL(clear_kernel, dim3((unsigned)((S + bs - 1) / bs)), dim3(bs), 0, s, (int)S, cover, err);
  hipLaunchKernelGGL(nucleus_kernel, dim3((unsigned)((N + bs - 1) / bs)), dim3(bs), 0, s, (int)N, (int)S,
                     nucleus_index, nucleus_shell_num, shells, cover, err);
  hipLaunchKernelGGL(shell_kernel, dim3((unsigned)((S + bs - 1) / bs)), dim3(bs), 0, s, (int)S, (int)Np,
                     shell_prim_index, shell_prim_num, shells, cover, err);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // the radial kernel runs only on a table that passed every check
  int32_t h_err = 0;
  e = hipMemcpyAsync(&h_err, err, sizeof(h_err), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  if (h_err != 0) {
    *table_err = h_err;
    return hipSuccess;
  }
  DeviceBasis view;
  view.nucl_num = N;
  view.shell_num = S;
  view.prim_num = Np;
  view.shells = shells;
  view.expo = const_cast<double *>(expo);
  view.coef_norm = const_cast<double *>(coef_normalized);
  view.nucl_range = const_cast<double *>(nucleus_range);
  return launch_shell_vgl(s, view, P, coord, nucl, shell_vgl);
}

}  // namespace qmckl_amd
