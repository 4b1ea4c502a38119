// Layout kernels and the electron-nucleus / electron-electron distance
// kernels.
//
// en_distance[p][n] = |r_p - R_n| (src/qmckl_electron.cpp:302-331 calls
// qmckl_distance_device('T','T') whose kernel is at 536-549; here the plain
// Euclidean distance).  Coordinates are structure-of-arrays, the nucleus
// table (3N doubles) is staged in LDS once per workgroup of 64 points, and
// each workgroup's stores cover one contiguous run of the output.  The
// kernel is HBM-bound: 24 B read + 8N B written per point.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "internal.hpp"

namespace qmckl_amd {

namespace {

constexpr int kThreads = 256;

__global__ void __launch_bounds__(kThreads) soa_from_aos_kernel(int64_t n, const double *__restrict__ aos,
                                                                double *__restrict__ soa) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < 3 * n;
       i += (int64_t)gridDim.x * blockDim.x) {
    // soa[d][p] = aos[p][d]; read coalesced along the AoS array.
    const int64_t p = i / 3, d = i - 3 * p;
    soa[d * n + p] = aos[i];
  }
}

__global__ void __launch_bounds__(kThreads) aos_from_soa_kernel(int64_t n, const double *__restrict__ soa,
                                                                double *__restrict__ aos) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < 3 * n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t p = i / 3, d = i - 3 * p;
    aos[i] = soa[d * n + p];
  }
}

// out[c][r] = in[r][c], 32x32 tiles through LDS.
__global__ void __launch_bounds__(kThreads) transpose_kernel(int64_t rows, int64_t cols,
                                                             const double *__restrict__ in,
                                                             double *__restrict__ out) {
  __shared__ double tile[32][33];
  const int64_t r0 = (int64_t)blockIdx.y * 32, c0 = (int64_t)blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int j = ty; j < 32; j += 8) {
    const int64_t r = r0 + j, c = c0 + tx;
    tile[j][tx] = (r < rows && c < cols) ? in[r * cols + c] : 0.0;
  }
  __syncthreads();
  for (int j = ty; j < 32; j += 8) {
    const int64_t c = c0 + j, r = r0 + tx;
    if (r < rows && c < cols) out[c * rows + r] = tile[tx][j];
  }
}

__global__ void __launch_bounds__(kThreads) strided_copy_kernel(int64_t rows, int64_t cols,
                                                                const double *__restrict__ in,
                                                                int64_t ld_in, double *__restrict__ out,
                                                                int64_t ld_out) {
  const int64_t total = rows * cols;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = i / cols, c = i - r * cols;
    out[r * ld_out + c] = in[r * ld_in + c];
  }
}

constexpr int kMaxNuclLds = 2048;  // nuclei staged in LDS (at most 48 KiB)

// One workgroup per kEnPoints consecutive points: it stages its points and
// the nuclei in LDS and writes its points' rows, one contiguous run of the
// output.  With an even nucleus count and a 16-byte aligned output each
// thread writes pairs of neighbouring elements (same point, nuclei n and
// n + 1) as one nontemporal 16-byte store, thread t taking pairs t,
// t + 256, ...; the (point, nucleus) position advances by a constant step
// with one wrap test, no division in the loop.  Otherwise single elements.
constexpr int kEnPoints = 128;
template <int V>  // elements per store: 2 or 1
__global__ void __launch_bounds__(kThreads) en_distance_kernel(int64_t P, int64_t N,
                                                               const double *__restrict__ coord,
                                                               const double *__restrict__ nucl,
                                                               double *__restrict__ out) {
  extern __shared__ double sm[];  // [3][kEnPoints] points, then [3][N] nuclei when N <= kMaxNuclLds
  double *sp = sm;
  const bool staged = N <= kMaxNuclLds;
  const int64_t pb = (int64_t)blockIdx.x * kEnPoints;
  const int np = (int)(P - pb < kEnPoints ? P - pb : kEnPoints);
  for (int i = threadIdx.x; i < 3 * kEnPoints; i += kThreads) {
    const int d = i / kEnPoints, j = i - d * kEnPoints;
    sp[i] = j < np ? coord[d * P + pb + j] : 0.0;
  }
  if (staged)
    for (int i = threadIdx.x; i < 3 * N; i += kThreads) sm[3 * kEnPoints + i] = nucl[i];
  __syncthreads();
  const double *R = staged ? sm + 3 * kEnPoints : nucl;
  const int64_t units = (int64_t)np * N / V;  // stores of this workgroup
  constexpr int kStep = V * kThreads;         // elements between a thread's stores
  const int step_p = (int)(kStep / N), step_n = (int)(kStep - (int64_t)step_p * N);
  int pl = (int)((int64_t)V * threadIdx.x / N);
  int n = (int)((int64_t)V * threadIdx.x - (int64_t)pl * N);
  double *o = out + pb * N;
  for (int64_t u = threadIdx.x; u < units; u += kThreads) {
    const double px = sp[pl], py = sp[kEnPoints + pl], pz = sp[2 * kEnPoints + pl];
    if constexpr (V == 2) {
      typedef double d2 __attribute__((ext_vector_type(2)));
      const double x0 = px - R[n], y0 = py - R[N + n], z0 = pz - R[2 * N + n];
      const double x1 = px - R[n + 1], y1 = py - R[N + n + 1], z1 = pz - R[2 * N + n + 1];
      const d2 v = {sqrt(fma(x0, x0, fma(y0, y0, z0 * z0))), sqrt(fma(x1, x1, fma(y1, y1, z1 * z1)))};
      __builtin_nontemporal_store(v, reinterpret_cast<d2 *>(o + 2 * u));
    } else {
      const double x = px - R[n], y = py - R[N + n], z = pz - R[2 * N + n];
      __builtin_nontemporal_store(sqrt(fma(x, x, fma(y, y, z * z))), o + u);
    }
    pl += step_p;
    n += step_n;
    if (n >= N) {
      n -= (int)N;
      ++pl;
    }
  }
}

// C[i + j ldc] = |a_i - b_j| for i < m, j < n (src/qmckl_electron.cpp:373-553,
// qmckl_distance_device), where a_i = (A[3i], A[3i+1], A[3i+2]) for 'N' and
// (A[i], A[i+lda], A[i+2lda]) for 'T', b_j likewise from B and ldb.
// Threads run along i (the contiguous index of C); each workgroup covers 256
// values of i and walks the columns j from blockIdx.y in steps of gridDim.y.
template <bool TA, bool TB>
__global__ void __launch_bounds__(kThreads) distance_kernel(int64_t m, int64_t n, const double *__restrict__ A,
                                                            int64_t lda, const double *__restrict__ B, int64_t ldb,
                                                            double *__restrict__ C, int64_t ldc) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= m) return;
  const double ax = TA ? A[i] : A[i * lda], ay = TA ? A[i + lda] : A[i * lda + 1],
               az = TA ? A[i + 2 * lda] : A[i * lda + 2];
  for (int64_t j = blockIdx.y; j < n; j += gridDim.y) {
    const double x = ax - (TB ? B[j] : B[j * ldb]);
    const double y = ay - (TB ? B[j + ldb] : B[j * ldb + 1]);
    const double z = az - (TB ? B[j + 2 * ldb] : B[j * ldb + 2]);
    C[i + j * ldc] = sqrt(fma(x, x, fma(y, y, z * z)));
  }
}

unsigned grid_for(int64_t work) {
  int64_t g = (work + kThreads - 1) / kThreads;
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (unsigned)g;
}

// ee_distance[w][j][i] = |r_(w,i) - r_(w,j)| (src/qmckl_electron.cpp:333-371,
// qmckl_distance('T','T') of each walker's electrons against themselves).
// Workgroup = one walker's 64 x 64 tile of (i, j): the 64 j coordinates are
// staged in LDS, thread (i, j0) computes j = j0, j0+4, ..., and every store
// instruction of a wavefront writes 64 consecutive i of one row.  HBM-bound:
// 8 B written per pair.
constexpr int kEeTile = 64;
__global__ void __launch_bounds__(kThreads) ee_distance_kernel(int64_t E, int64_t W, int64_t tiles,
                                                               const double *__restrict__ coord,
                                                               double *__restrict__ out) {
  __shared__ double sj[3][kEeTile];
  const int64_t P = E * W;
  const int64_t b = blockIdx.x;
  const int64_t w = b / (tiles * tiles);
  const int64_t rem = b - w * tiles * tiles;
  const int64_t i0 = (rem % tiles) * kEeTile, j0 = (rem / tiles) * kEeTile;
  const int tid = threadIdx.x;
  if (tid < kEeTile) {
    const int64_t j = j0 + tid;
    const int64_t pj = w * E + (j < E ? j : 0);
    sj[0][tid] = coord[pj];
    sj[1][tid] = coord[P + pj];
    sj[2][tid] = coord[2 * P + pj];
  }
  __syncthreads();
  const int64_t i = i0 + (tid & 63);
  if (i >= E) return;
  const int64_t pi = w * E + i;
  const double xi = coord[pi], yi = coord[P + pi], zi = coord[2 * P + pi];
  for (int jj = tid >> 6; jj < kEeTile; jj += kThreads / 64) {
    const int64_t j = j0 + jj;
    if (j >= E) break;
    const double x = xi - sj[0][jj], y = yi - sj[1][jj], z = zi - sj[2][jj];
    out[(w * E + j) * E + i] = sqrt(fma(x, x, fma(y, y, z * z)));
  }
}

// Spatial order of the points for the tiled AO/MO kernels: points are
// binned into a 16 x 16 x 16 grid of cells over the bounding box of the
// nuclei (points outside fall into the border cells), the cells are taken
// in Z (Morton) order, and perm[i] is the i-th point in that order.  64
// consecutive entries of perm then lie close together, so whole 64-AO
// stages of a tile are screened out far more often than for 64 points
// spread over a walker.  Within a cell the order is that of atomic slot
// allocation; it does not matter, since every point's values are computed
// independently of the other points of its tile.
// Workspace: bbox[6] doubles (64 bytes), counts[kCells] int32, cell[P] int32.
constexpr int kCellBits = 4;
constexpr int kCells = 1 << (3 * kCellBits);
constexpr int kScanThreads = 1024;

__global__ void __launch_bounds__(kThreads) order_setup_kernel(int64_t N, const double *__restrict__ nucl,
                                                               double *__restrict__ bbox,
                                                               int32_t *__restrict__ counts) {
  __shared__ double red[6][kThreads];
  const int t = threadIdx.x;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t n = t; n < N; n += kThreads)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lo[d] = fmin(lo[d], nucl[d * N + n]);
      hi[d] = fmax(hi[d], nucl[d * N + n]);
    }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    red[d][t] = lo[d];
    red[3 + d][t] = hi[d];
  }
  __syncthreads();
  for (int h = kThreads / 2; h > 0; h >>= 1) {
    if (t < h)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        red[d][t] = fmin(red[d][t], red[d][t + h]);
        red[3 + d][t] = fmax(red[3 + d][t], red[3 + d][t + h]);
      }
    __syncthreads();
  }
  if (t < 6) bbox[t] = red[t][0];
  for (int i = t; i < kCells; i += kThreads) counts[i] = 0;
}

__device__ inline int spread_bits(int v) {  // bit b of v -> bit 3b
  int r = 0;
#pragma unroll
  for (int b = 0; b < kCellBits; ++b) r |= ((v >> b) & 1) << (3 * b);
  return r;
}

__global__ void __launch_bounds__(kThreads) order_cells_kernel(int64_t P, const double *__restrict__ coord,
                                                               const double *__restrict__ bbox,
                                                               int32_t *__restrict__ counts,
                                                               int32_t *__restrict__ cell) {
  const int64_t p = blockIdx.x * (int64_t)kThreads + threadIdx.x;
  if (p >= P) return;
  int code = 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double lo = bbox[d], w = bbox[3 + d] - lo;
    const double g = (w > 0.0 ? (coord[d * P + p] - lo) / w : 0.5) * (1 << kCellBits);
    // out-of-box values go to the border cells, NaN to cell 0
    const int q = g >= (double)((1 << kCellBits) - 1) ? (1 << kCellBits) - 1 : (g > 0.0 ? (int)g : 0);
    code |= spread_bits(q) << d;
  }
  cell[p] = code;
  atomicAdd(&counts[code], 1);
}

// counts -> exclusive offsets, one workgroup of kScanThreads threads, each
// owning kCells / kScanThreads consecutive cells
__global__ void __launch_bounds__(kScanThreads) order_scan_kernel(int32_t *__restrict__ counts) {
  constexpr int per = kCells / kScanThreads;
  __shared__ int32_t part[kScanThreads];
  const int t = threadIdx.x;
  int32_t v[per], sum = 0;
#pragma unroll
  for (int i = 0; i < per; ++i) {
    v[i] = counts[t * per + i];
    sum += v[i];
  }
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < kScanThreads; off <<= 1) {  // inclusive scan of the partial sums
    const int32_t add = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += add;
    __syncthreads();
  }
  int32_t run = part[t] - sum;
#pragma unroll
  for (int i = 0; i < per; ++i) {
    counts[t * per + i] = run;
    run += v[i];
  }
}

__global__ void __launch_bounds__(kThreads) order_scatter_kernel(int64_t P, const int32_t *__restrict__ cell,
                                                                 int32_t *__restrict__ offsets,
                                                                 int32_t *__restrict__ perm) {
  const int64_t p = blockIdx.x * (int64_t)kThreads + threadIdx.x;
  if (p >= P) return;
  perm[atomicAdd(&offsets[cell[p]], 1)] = (int32_t)p;
}

}  // namespace

hipError_t launch_soa_from_aos(hipStream_t s, int64_t n, const double *aos, double *soa) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(soa_from_aos_kernel, dim3(grid_for(3 * n)), dim3(kThreads), 0, s, n, aos, soa);
  return hipGetLastError();
}

hipError_t launch_aos_from_soa(hipStream_t s, int64_t n, const double *soa, double *aos) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(aos_from_soa_kernel, dim3(grid_for(3 * n)), dim3(kThreads), 0, s, n, soa, aos);
  return hipGetLastError();
}

hipError_t launch_transpose(hipStream_t s, int64_t rows, int64_t cols, const double *in, double *out) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  if ((rows + 31) / 32 > 65535) return hipErrorInvalidValue;
  dim3 grid((unsigned)((cols + 31) / 32), (unsigned)((rows + 31) / 32));
  hipLaunchKernelGGL(transpose_kernel, grid, dim3(kThreads), 0, s, rows, cols, in, out);
  return hipGetLastError();
}

hipError_t launch_strided_copy(hipStream_t s, int64_t rows, int64_t cols, const double *in,
                               int64_t ld_in, double *out, int64_t ld_out) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  hipLaunchKernelGGL(strided_copy_kernel, dim3(grid_for(rows * cols)), dim3(kThreads), 0, s, rows,
                     cols, in, ld_in, out, ld_out);
  return hipGetLastError();
}

hipError_t launch_en_distance(hipStream_t s, int64_t P, int64_t N, const double *coord,
                              const double *nucl, double *out) {
  if (P <= 0 || N <= 0) return hipSuccess;
  if (N > 0x7fffffff / kEnPoints) return hipErrorInvalidValue;  // per-workgroup element counts in int
  const size_t lds = sizeof(double) * (3 * kEnPoints + (N <= kMaxNuclLds ? 3 * (size_t)N : 0));
  const int64_t blocks = (P + kEnPoints - 1) / kEnPoints;
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  if (N % 2 == 0 && reinterpret_cast<uintptr_t>(out) % 16 == 0)
    hipLaunchKernelGGL(en_distance_kernel<2>, dim3((unsigned)blocks), dim3(kThreads), lds, s, P, N, coord, nucl,
                       out);
  else
    hipLaunchKernelGGL(en_distance_kernel<1>, dim3((unsigned)blocks), dim3(kThreads), lds, s, P, N, coord, nucl,
                       out);
  return hipGetLastError();
}

hipError_t launch_distance(hipStream_t s, bool ta, bool tb, int64_t m, int64_t n, const double *A, int64_t lda,
                           const double *B, int64_t ldb, double *C, int64_t ldc) {
  if (m <= 0 || n <= 0) return hipSuccess;
  const int64_t bx = (m + kThreads - 1) / kThreads;
  if (bx > 0x7fffffff) return hipErrorInvalidValue;
  // enough workgroups to fill the chip, each walking several columns
  const int64_t want = (4096 + bx - 1) / bx;
  const unsigned by = (unsigned)(n < want ? n : (want > 65535 ? 65535 : want));
  const dim3 grid((unsigned)bx, by);
  if (ta && tb)
    hipLaunchKernelGGL((distance_kernel<true, true>), grid, dim3(kThreads), 0, s, m, n, A, lda, B, ldb, C, ldc);
  else if (ta)
    hipLaunchKernelGGL((distance_kernel<true, false>), grid, dim3(kThreads), 0, s, m, n, A, lda, B, ldb, C, ldc);
  else if (tb)
    hipLaunchKernelGGL((distance_kernel<false, true>), grid, dim3(kThreads), 0, s, m, n, A, lda, B, ldb, C, ldc);
  else
    hipLaunchKernelGGL((distance_kernel<false, false>), grid, dim3(kThreads), 0, s, m, n, A, lda, B, ldb, C, ldc);
  return hipGetLastError();
}

hipError_t launch_ee_distance(hipStream_t s, int64_t E, int64_t W, const double *coord, double *out) {
  if (E <= 0 || W <= 0) return hipSuccess;
  const int64_t tiles = (E + kEeTile - 1) / kEeTile;
  const int64_t blocks = W * tiles * tiles;
  if (blocks > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL(ee_distance_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s, E, W, tiles, coord,
                     out);
  return hipGetLastError();
}

int64_t point_order_workspace_bytes(int64_t P) { return 64 + 4 * (int64_t)kCells + 4 * P; }

// The same cells, points ordered by a stable radix sort of their cell codes:
// within a cell in ascending point index, so the order is a function of the
// coordinates alone (the atomic scatter above leaves it open).
namespace {
size_t stable_sort_temp_bytes(int64_t P) {
  size_t bytes = 0;
  if (rocprim::radix_sort_pairs(nullptr, bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                rocprim::counting_iterator<int32_t>(0), (int32_t *)nullptr, (size_t)P, 0,
                                3 * kCellBits) != hipSuccess)
    return 0;
  return bytes;
}
}  // namespace

int64_t point_order_stable_workspace_bytes(int64_t P) {
  return 64 + 4 * (int64_t)kCells + 8 * P + 256 + (int64_t)stable_sort_temp_bytes(P);
}

hipError_t launch_point_order_stable(hipStream_t s, int64_t P, const double *coord, int64_t N, const double *nucl,
                                     void *workspace, int32_t *perm) {
  if (P <= 0) return hipSuccess;
  if (P > 0x7fffffff || N <= 0) return hipErrorInvalidValue;
  char *ws = static_cast<char *>(workspace);
  double *bbox = reinterpret_cast<double *>(ws);
  int32_t *counts = reinterpret_cast<int32_t *>(ws + 64);
  int32_t *cell = counts + kCells;
  uint32_t *sorted = reinterpret_cast<uint32_t *>(cell + P);
  const int64_t used = 64 + 4 * (int64_t)kCells + 8 * P;
  void *temp = ws + (used + 255) / 256 * 256;
  size_t temp_bytes = stable_sort_temp_bytes(P);
  const unsigned blocks = (unsigned)((P + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(order_setup_kernel, dim3(1), dim3(kThreads), 0, s, N, nucl, bbox, counts);
  hipLaunchKernelGGL(order_cells_kernel, dim3(blocks), dim3(kThreads), 0, s, P, coord, bbox, counts, cell);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return rocprim::radix_sort_pairs(temp, temp_bytes, reinterpret_cast<const uint32_t *>(cell), sorted,
                                   rocprim::counting_iterator<int32_t>(0), perm, (size_t)P, 0, 3 * kCellBits, s);
}

hipError_t launch_point_order(hipStream_t s, int64_t P, const double *coord, int64_t N, const double *nucl,
                              void *workspace, int32_t *perm) {
  if (P <= 0) return hipSuccess;
  if (P > 0x7fffffff || N <= 0) return hipErrorInvalidValue;
  double *bbox = static_cast<double *>(workspace);
  int32_t *counts = reinterpret_cast<int32_t *>(static_cast<char *>(workspace) + 64);
  int32_t *cell = counts + kCells;
  const unsigned blocks = (unsigned)((P + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(order_setup_kernel, dim3(1), dim3(kThreads), 0, s, N, nucl, bbox, counts);
  hipLaunchKernelGGL(order_cells_kernel, dim3(blocks), dim3(kThreads), 0, s, P, coord, bbox, counts, cell);
  hipLaunchKernelGGL(order_scan_kernel, dim3(1), dim3(kScanThreads), 0, s, counts);
  hipLaunchKernelGGL(order_scatter_kernel, dim3(blocks), dim3(kThreads), 0, s, P, cell, counts, perm);
  return hipGetLastError();
}

}  // namespace qmckl_amd
