// Internal state of a qmckl context and the launcher interface between the
// host-side provider layer (C++) and the HIP kernels (.hip files).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/qmckl_gpu.h"

#define QMCKL_AMD_SPLIT_FP32 QMCKL_PRECISION_FP32_DEVICE
#define QMCKL_AMD_SPLIT_FP16X2 QMCKL_PRECISION_FP16X2_DEVICE
#define QMCKL_AMD_SPLIT_BF16X2 QMCKL_PRECISION_BF16X2_DEVICE
#define QMCKL_AMD_SPLIT_BF16X3 QMCKL_PRECISION_BF16X3_DEVICE
#define QMCKL_AMD_INT8X5 QMCKL_PRECISION_INT8X5_DEVICE

namespace qmckl_amd {

constexpr uint64_t kValidTag = 0xBEEFFACEull;    // include/qmckl_types.hpp:12
constexpr uint64_t kInvalidTag = 0xDEADBEEFull;  // include/qmckl_types.hpp:13
constexpr double kCutoff = 27.631021115928547;   // -ln(1e-12), src/qmckl_ao_sycl.cpp:22
constexpr int kMaxL = 6;                          // highest shell angular momentum on the device
constexpr int kMaxCart = (kMaxL + 1) * (kMaxL + 2) / 2;
constexpr int kChunkShells = 16;  // shells per AO-kernel chunk (LDS radial tile)
constexpr int kChunkAos = 64;     // AOs per AO-kernel chunk (one per lane)

// Device-side description of one shell, in AO order (nucleus-major, shells
// of a nucleus in order), built at finalize.
struct ShellDesc {
  int32_t nucl;        // nucleus index
  int32_t prim_start;  // first primitive
  int32_t prim_num;    // number of primitives
  int32_t ao_start;    // first AO
  int32_t l;           // angular momentum
  int32_t shell;       // index in the user's shell arrays
  double a_min;        // smallest exponent: the shell is zero where a_min * r^2 > kCutoff
};

// Output of the AO kernel: values at out[(p*ncomp + c)*ld + k] and optional
// nonzero flags per (64-row tile of the (p,c) rows, 32-AO block).
template <typename OutT>
struct AoOut {
  OutT *out;
  int64_t ld;
  uint8_t *flags;  // [ceil(ncomp*P/64)][nkb] or nullptr
  int nkb;
};

// Basis tables on the device (owned by the context, rebuilt by finalize).
struct DeviceBasis {
  int64_t nucl_num = 0, shell_num = 0, prim_num = 0, ao_num = 0;
  ShellDesc *shells = nullptr;     // [S] in AO order
  int32_t *ao_shell = nullptr;     // [A] position of the AO's shell in `shells`
  int32_t *ao_abc = nullptr;       // [A] a | b<<8 | c<<16
  double *expo = nullptr;          // [Np]
  double *coef_norm = nullptr;     // [Np]
  double *nucl_range = nullptr;    // [N]
  int32_t *nucl_max_l = nullptr;   // [N]
  int64_t *nucl_prim_index = nullptr;  // [N+1]
  // AO-kernel chunks: positions in `shells` where chunk c starts, [nchunks+1];
  // a chunk holds at most kChunkShells shells and kChunkAos AOs
  int32_t *chunks = nullptr;
  int nchunks = 0;
  // [ceil(A/64)+1] position in `shells` of the shell holding AO 64*st (S
  // past the last AO)
  int32_t *stage_shell = nullptr;
  int max_shell_prims = 0;         // host copy: largest prim count of a shell
  int max_l = 0;                   // host copy
};

struct PointState {
  int64_t num = 0;
  uint64_t date = 0;
  double *coord = nullptr;  // [3][num]
  int64_t capacity = 0;     // points the allocation holds
};

struct Walker {
  int64_t num = 0;
  PointState point;
};

struct Nucleus {
  int64_t num = 0;
  double *coord = nullptr;   // [3][num]
  double *charge = nullptr;  // [num]
  int32_t uninitialized = (1 << 3) - 1;
  bool provided = false;
};

struct Electron {
  int64_t num = 0, up_num = 0, down_num = 0;
  Walker walker, walker_old;
  double *en_distance = nullptr;  // [P][N]
  int64_t en_distance_cap = 0;
  uint64_t en_distance_date = 0;
  double *ee_distance = nullptr;  // [walk][E][E]
  int64_t ee_distance_cap = 0;
  uint64_t ee_distance_date = 0;
  int32_t uninitialized = (1 << 1) - 1;
  bool provided = false;
};

struct AoBasis {
  char type = 0;
  int64_t shell_num = 0, prim_num = 0, ao_num = 0;
  int64_t *nucleus_index = nullptr, *nucleus_shell_num = nullptr;
  int32_t *shell_ang_mom = nullptr;
  int64_t *shell_prim_num = nullptr, *shell_prim_index = nullptr;
  double *shell_factor = nullptr, *exponent = nullptr, *coefficient = nullptr,
         *prim_factor = nullptr, *ao_factor = nullptr;
  DeviceBasis dev;
  double *shell_vgl = nullptr;  // [P][5][S]
  int64_t shell_vgl_cap = 0;
  uint64_t shell_vgl_date = 0;
  double *ao_vgl = nullptr;     // [P][5][A]
  int64_t ao_vgl_cap = 0;
  uint64_t ao_vgl_date = 0;
  double *ao_value = nullptr;   // [P][A]
  int64_t ao_value_cap = 0;
  uint64_t ao_value_date = 0;
  // nonzero flags of ao_vgl per (64 rows of [5P], 32 AOs), written with it
  uint8_t *ao_vgl_flags = nullptr;
  int64_t ao_vgl_flags_cap = 0;
  // FP32 AO operand of the mixed contraction (blocked layout, AOs padded to
  // ld = A rounded up to 64), its row maxima and nonzero flags
  float *ao_f32 = nullptr;
  int64_t ao_f32_cap = 0;
  // the INT8X5 form's residuals (float)(v - (double)(float)v), same layout
  float *ao_f32_lo = nullptr;
  int64_t ao_f32_lo_cap = 0;
  float *ao_f32_rowmax = nullptr;
  int64_t ao_f32_rowmax_cap = 0;
  uint8_t *ao_f32_flags = nullptr;
  int64_t ao_f32_flags_cap = 0;
  // spatial order of the points for the operand's tiles, and its workspace
  int32_t *point_perm = nullptr;
  int64_t point_perm_cap = 0;
  void *order_ws = nullptr;
  int64_t order_ws_cap = 0;
  int32_t uninitialized = (1 << 14) - 1;
  bool provided = false;
};

// Column plan of the MO contraction.  Column j of the prepared coefficient
// copies holds MO col_mo[j] (null: MO j), so MOs whose nonzero coefficients
// sit on the same AOs share 16-column blocks.  The order only permutes the
// MOs within each group of `group` consecutive MOs (the columns of one
// workgroup of the kernel that uses it), so a workgroup still owns one MO
// range and writes it as contiguous runs.  Bit s of blk[b * nw + s / 64] is
// set when column block b (columns [16b, 16b + 16)) has a nonzero
// coefficient among the AOs [16s, 16s + 16); blocks >= nblk hold no MO.
// blk null: every block is live.
struct ColPlan {
  const int32_t *col_mo = nullptr;
  const uint64_t *blk = nullptr;
  int64_t nw = 0;
  int64_t nblk = 0;
};

// A column plan kept by the MO basis for one group width (device arrays;
// col_mo null when the planned order leaves too many blocks live).
struct ColPlanSlot {
  int64_t group = 0;  // 0: unused
  int32_t *col_mo = nullptr;
  uint64_t *blk = nullptr;
  double live = 1.0;  // live (16-MO, 16-AO) blocks of the planned order over all of them
  std::vector<uint64_t> host_blk;  // host copy of blk (when kept)
};

// One operand of launch_ao_rows_f32 (no residual array) as an item of a
// batched launch: the arguments and the job's wavefronts [wave0, wave0 +
// waves) of the launch (waves = point tiles x groups).
struct AoRowsJob {
  int64_t P, S, A, ld, npt;
  int N, groups, nkb, skip_halves;
  const double *coord;
  const int32_t *perm;
  const double *nucl;
  const ShellDesc *shells;
  const int32_t *stage_shell;
  const double *ao_factor, *expo, *coef, *nucl_range;
  float *buf, *rowmax;
  uint8_t *flags;
  int64_t wave0, waves;
};

// One molecule's split contraction in a persistent batch launch (the
// arguments of launch_mo_gemm_mixed, the tile width and its work items).
struct MixedGemmJob {
  int64_t P, npt, K, Ncol;
  const float *A;
  int64_t lda;
  const float *rowmax;
  const uint16_t *Bs;
  int64_t ldb;
  const float *colinv;
  const uint8_t *flags;
  int nkb;
  const int32_t *perm;
  double *C;
  int64_t ldc;
  ColPlan plan;
  int cbw;                     // tile width: 64 * cbw MOs
  int dense;                   // 1: the DENSE contraction kernel, 0: LIVE
  int64_t item0, gx, items;    // first item, row-tile halves, items (gx x column groups)
};
struct MoBasis {
  int64_t mo_num = 0;
  double *coefficient = nullptr;    // [M][A]
  double *coefficient_t = nullptr;  // [A][M]
  double *coef_frag = nullptr;      // coefficient_t in FP64 MFMA fragment order (planned column order)
  int64_t coef_frag_nk4 = 0;
  // nonzero map of the coefficients made at finalize (host, [M][nz_nw]
  // words of 16-AO blocks) and the column plans made from it: slot 0 for
  // the FP64 contraction's workgroup width, slot 1 for another width of the
  // split forms
  std::vector<uint64_t> nz_words;
  int64_t nz_nw = 0;
  ColPlanSlot col_plans[2];
  double *mo_vgl = nullptr;         // [P][5][M]
  int64_t mo_vgl_cap = 0;
  uint64_t mo_vgl_date = 0;
  double *mo_value = nullptr;       // [P][M]
  int64_t mo_value_cap = 0;
  uint64_t mo_value_date = 0;
  qmckl_precision_device precision = QMCKL_PRECISION_FP64_DEVICE;
  // MO coefficients prepared for the mixed forms: [terms][M][ld] 16-bit
  // split terms (or [M][ld] FP32), inverse column scales [M]
  uint16_t *coef_split = nullptr;
  float *coef_colinv = nullptr;
  // the INT8X5 form's coefficient slices (k_gemm_i8.hip fragment order);
  // coef_colinv then holds the inverse MO scales
  int8_t *coef_i8 = nullptr;
  qmckl_precision_device coef_split_form = QMCKL_PRECISION_FP64_DEVICE;
  // tile choices of the mixed forms (results do not depend on them):
  // spatial point order -1 auto / 0 off / 1 on; AO-operand wavefronts per
  // CU 0 auto / else that many (the AOs of a tile are split into groups to
  // reach it); tuned: set by the autotuner
  int32_t tile_point_order = -1;
  int32_t tile_group_waves = 0;
  int32_t tile_gemm = 0;  // contraction kernel of the split forms (QMCKL_GEMM_TILE_*)
  bool tile_tuned = false;
  // adjoint scratch: G = w C [5P][A], per-block partials, split-K partials
  double *adj_G = nullptr;
  int64_t adj_G_cap = 0;
  double *adj_partR = nullptr;
  int64_t adj_partR_cap = 0;
  double *adj_partA = nullptr;
  int64_t adj_partA_cap = 0;
  double *adj_partC = nullptr;
  int64_t adj_partC_cap = 0;
  // the adjoint's B operands in FP64 fragment order: the coefficients [M][A]
  // (built on first use after finalize) and the weights [5P][M] (per call)
  double *adj_cfrag = nullptr;
  int64_t adj_cfrag_cap = 0, adj_cfrag_nk4 = 0;
  bool adj_cfrag_ok = false;
  double *adj_wfrag = nullptr;
  int64_t adj_wfrag_cap = 0;
  int32_t *adj_kmap = nullptr;   // (point, component) rows in point-tile order
  int64_t adj_kmap_cap = 0;
  int32_t *adj_perm = nullptr;   // the points in that order (stable, so the sums are reproducible)
  int64_t adj_perm_cap = 0;
  void *adj_order_ws = nullptr;
  int64_t adj_order_ws_cap = 0;
  uint8_t *adj_tflags = nullptr;  // screen flags per (64-point tile, 16-AO block)
  int64_t adj_tflags_cap = 0;
  // with a column plan, G = w C runs with K over the MOs in the plan's
  // order: the weights' columns gathered into that order (per call), the
  // coefficient fragments built with their rows in it, and the live bits of
  // (16-AO column block, 16-MO K block) pairs [ceil(A/16)][adj_blk_nw]
  double *adj_wperm = nullptr;
  int64_t adj_wperm_cap = 0;
  uint64_t *adj_blk = nullptr;
  int64_t adj_blk_nw = 0;
  bool adj_planned = false;
  int32_t uninitialized = (1 << 2) - 1;
  bool provided = false;
};

struct Context {
  uint64_t tag = kValidTag;
  int device_id = QMCKL_NO_DEVICE;
  hipStream_t stream = nullptr;
  std::recursive_mutex mutex;
  int lock_count = 0;
  qmckl_exit_code_device err_code = QMCKL_SUCCESS_DEVICE;
  char err_function[QMCKL_MAX_FUN_LEN_DEVICE] = {0};
  char err_message[QMCKL_MAX_MSG_LEN_DEVICE] = {0};
  std::unordered_map<void *, size_t> device_allocs;
  std::unordered_map<void *, size_t> host_allocs;
  uint64_t date = 0;
  // bumped by every AO or MO finalize: a recorded batch graph compares it
  // (the device tables a finalize rebuilds may land at the old addresses)
  uint64_t basis_generation = 0;
  bool timing = false;
  // set while a batch enqueues this context's work: the compute chain's
  // stream synchronisations are skipped and the batch waits once at the end
  bool defer_sync = false;
  // set while a batch graph is captured: kernel timers record nothing
  bool capturing = false;
  // set while an eager batch enqueues this context: a split-form
  // contraction of tile width <= 3 is appended here instead of launched (the
  // batch runs all of them in persistent launches)
  std::vector<MixedGemmJob> *collect_jobs = nullptr;
  // ... and, with it, its AO operand here (the batch writes all of them in
  // one launch)
  std::vector<AoRowsJob> *collect_ao = nullptr;
  // the batch's job table and work counters (device), and an event marking
  // the end of this context's part of a batch
  MixedGemmJob *batch_jobs = nullptr;
  int64_t batch_jobs_cap = 0;
  std::vector<MixedGemmJob> batch_jobs_host;  // the last eager batch's table, as uploaded
  AoRowsJob *batch_ao_jobs = nullptr;         // the batch's AO job table (device)
  int64_t batch_ao_jobs_cap = 0;
  std::vector<AoRowsJob> batch_ao_host;       // the last eager batch's AO table, as uploaded
  void *batch_jobs_pinned = nullptr;  // pinned host staging of the table's upload
  size_t batch_jobs_pinned_cap = 0;
  hipEvent_t batch_event = nullptr;
  struct TimerAcc {
    int64_t launches = 0;
    double total_ms = 0.0;
  };
  std::unordered_map<std::string, TimerAcc> timers;
  PointState point;
  Nucleus nucleus;
  Electron electron;
  AoBasis ao;
  MoBasis mo;
};

// Brackets one kernel launch with HIP events on the context's stream when
// timing is enabled; the elapsed time is added to the context's timers when
// the object goes out of scope (after the caller's synchronisation).
class KernelTimer {
 public:
  KernelTimer(Context *c, const char *name);
  void stop();
  ~KernelTimer();
  KernelTimer(const KernelTimer &) = delete;
  KernelTimer &operator=(const KernelTimer &) = delete;

 private:
  Context *c_;
  const char *name_;
  hipEvent_t start_ = nullptr, stop_ = nullptr;
  bool stopped_ = false;
};

// ---- host helpers (context.cpp) ----
Context *ctx_of(qmckl_context_device context);  // nullptr when invalid
qmckl_exit_code_device fail(qmckl_context_device context, qmckl_exit_code_device code,
                            const char *function, const char *message);
// Tracked device allocation on the context's device (nullptr on failure or
// when no device is bound).
void *dmalloc(Context *c, size_t bytes);
qmckl_exit_code_device dfree(Context *c, void *ptr);
// Replace *slot by a fresh allocation of `bytes` (freeing the old one).
bool drealloc(Context *c, void **slot, size_t bytes);
// Grow-only buffer: reallocates *slot when cap < need (bytes).
bool grow_bytes(Context *c, void **slot, int64_t *cap, int64_t need);
constexpr int kMaxFlagBlocks = 2048;  // 32-AO blocks the AO kernel can flag (65,536 AOs)
bool has_device(const Context *c);
hipError_t sync(Context *c);
qmckl_exit_code_device hip_check(qmckl_context_device context, hipError_t e, const char *function);

// Device tables from host copies of the basis arrays (ao.cpp).  Returns an
// error message, empty on success.  expo_dev is stored as b.expo.
std::string build_basis_tables(Context *c, DeviceBasis &b, int64_t N, int64_t S, int64_t Np,
                               int64_t A, const std::vector<int64_t> &nucleus_index,
                               const std::vector<int64_t> &nucleus_shell_num,
                               const std::vector<int32_t> &shell_ang_mom,
                               const std::vector<int64_t> &shell_prim_num,
                               const std::vector<int64_t> &shell_prim_index,
                               const std::vector<double> &shell_factor,
                               const std::vector<double> &exponent,
                               const std::vector<double> &coefficient,
                               const std::vector<double> &prim_factor, double *expo_dev);
void release_basis_tables(Context *c, DeviceBasis &b);

// ---- kernel launchers (.hip files); all enqueue on `stream` ----
hipError_t launch_soa_from_aos(hipStream_t s, int64_t n, const double *aos, double *soa);
hipError_t launch_aos_from_soa(hipStream_t s, int64_t n, const double *soa, double *aos);
hipError_t launch_transpose(hipStream_t s, int64_t rows, int64_t cols, const double *in,
                            double *out);
// ee_distance [W][E][E] of the W walkers' E electrons, coord [3][W*E]
// Spatial order of the points (cells of a grid over the nuclei's bounding
// box, in Z order): perm[P] receives the point indices; workspace holds
// point_order_workspace_bytes(P) bytes.
int64_t point_order_workspace_bytes(int64_t P);
hipError_t launch_point_order(hipStream_t s, int64_t P, const double *coord, int64_t N, const double *nucl,
                              void *workspace, int32_t *perm);
// The same cells with the points of a cell in ascending index (a stable
// radix sort of the cell codes): a function of the coordinates alone.
int64_t point_order_stable_workspace_bytes(int64_t P);
hipError_t launch_point_order_stable(hipStream_t s, int64_t P, const double *coord, int64_t N, const double *nucl,
                                     void *workspace, int32_t *perm);
hipError_t launch_distance(hipStream_t s, bool ta, bool tb, int64_t m, int64_t n, const double *A, int64_t lda,
                           const double *B, int64_t ldb, double *C, int64_t ldc);
hipError_t launch_ee_distance(hipStream_t s, int64_t E, int64_t W, const double *coord, double *out);
hipError_t launch_en_distance(hipStream_t s, int64_t P, int64_t N, const double *coord,
                              const double *nucl, double *out);
hipError_t launch_shell_vgl(hipStream_t s, const DeviceBasis &b, int64_t P, const double *coord,
                            const double *nucl, double *shell_vgl);
// ao_vgl (ncomp 5) or ao_value (ncomp 1); perm (with ao_vgl_lane_kernel
// only, or null) is an order of the points into spatially compact 64-point
// tiles
hipError_t launch_ao_vgl(hipStream_t s, const DeviceBasis &b, int64_t P, const double *coord,
                         const double *nucl, const double *ao_factor, double *ao_vgl,
                         int ncomp, uint8_t *flags = nullptr, int nkb = 0, const int32_t *perm = nullptr);
// Which kernel writes FP64 ao_vgl and ao_value: the lane-per-point one (k_ao_rows.hip)
// for bases of 1024 AOs or more whose rows are whole 64-byte runs, the
// lane-per-AO one (k_ao.hip) otherwise.  Measured on MI355X: water64 (4,160
// AOs, 81,920 points) 4.00 against 5.26 ms; the 48-molecule ragged batch
// (30-620 AOs, rows mostly not 64-byte aligned) 22.3 against 10.0 ms in all;
// on its molecules with 64-byte rows alone 1.40 against 1.31 ms (the lane
// kernel ahead only at 560 AOs).
inline bool ao_vgl_lane_kernel(int64_t A) { return A >= 1024 && A % 8 == 0; }
// FP64 ao_vgl [P][5][A] (ncomp 5) or ao_value [P][A] (ncomp 1) with one
// lane per point (k_ao_rows.hip), the points of lane-tile t being
// perm[64 t ..] (null: 64 t ..), and, for ao_vgl if flags is not null,
// nonzero flags [(5P + 63) / 64][nkb], nkb = ceil(A / 32), per (64 rows of
// [point * 5 + component], 32 AOs).
hipError_t launch_ao_vgl_f64(hipStream_t s, const DeviceBasis &b, int64_t P, const double *coord,
                             const int32_t *perm, const double *nucl, const double *ao_factor, double *out,
                             int ncomp, uint8_t *flags, int nkb);
// The point order for spatially compact 64-point tiles of the lane-per-point
// AO kernels: *perm = ao.point_perm, computed here on the context's stream,
// or null for the caller's order.  The tile configuration's point_order
// forces it on (1) or off (0); by default (-1) it is on for 1024 AOs or more
// and 4096 points or more.  hipErrorOutOfMemory if its buffers cannot grow.
hipError_t tile_point_order(Context *c, const int32_t **perm);
// AO operand of the mixed contraction, FP32 in the blocked layout
// buf[c][64-point tile][AO/4][point in tile][4] (AOs zero-padded to ld, a
// multiple of 64), tile row i holding point perm[i] (null: point i), row
// maxima rowmax[c][tile row] (the AOs are split over `groups` wavefronts
// per tile, ao_rows_groups) and, if flags is not null, nonzero flags
// [5 * tiles][nkb] with nkb = ld / 32.  With flags, a 64-AO stage that the
// screen zeroes for all 64 points of a tile is not written (its two flags
// are zero, so the contraction never reads it); with skip_halves, neither
// is a 32-AO half that the screen zeroes (its flag is zero): only for
// contractions that read no half whose flag is zero (not the dense split
// kernel, which reads whole stages).  With buf_lo (same size as buf) each
// value's FP32 residual v - (double)(float)v goes there as well.
int ao_rows_groups(int64_t P, int64_t ld, int group_waves = 0);
// Fills *job with the checks of launch_ao_rows_f32; wave0 is left to the
// caller.
hipError_t make_ao_rows_job(const DeviceBasis &b, int64_t P, const double *coord, const int32_t *perm,
                            const double *nucl, const double *ao_factor, int64_t ld, int groups, float *buf,
                            float *rowmax, uint8_t *flags, int nkb, bool skip_halves, AoRowsJob *job);
// Every job's row maxima zeroed, then all jobs' operands in one launch of
// `total` wavefronts; jobs (device memory) in order of wave0.
hipError_t launch_ao_rows_batch(hipStream_t s, const AoRowsJob *jobs, int njobs, int64_t total);
hipError_t launch_ao_rows_f32(hipStream_t s, const DeviceBasis &b, int64_t P, const double *coord,
                              const int32_t *perm, const double *nucl, const double *ao_factor, int64_t ld,
                              int groups, float *buf, float *rowmax, uint8_t *flags, int nkb,
                              float *buf_lo = nullptr, bool skip_halves = false);
hipError_t launch_ao_from_shell_table(hipStream_t s, const DeviceBasis &b, int64_t P,
                                      const double *coord, const double *nucl,
                                      const double *ao_factor, const double *shell_vgl,
                                      int64_t S_table, double *out, int ncomp);
hipError_t launch_mo_gemm_f64(hipStream_t s, int64_t rows, int64_t K, int64_t Ncol,
                              const double *A, int64_t lda, const double *B, int64_t ldb,
                              double *C, int64_t ldc, const uint8_t *flags = nullptr, int nkb = 0);
// Coefficients in the MFMA fragment order of the FP64 forward contraction:
// Bf holds (Ncol rounded up to 64 * f64_frag_cbw(Ncol)) x (K rounded up to
// 16) doubles; *nk4 receives the K steps of 4 per column block.
int f64_frag_cbw(int64_t Ncol);
// Column j of Bf takes column col_mo[j] of B (null: column j), K row k
// takes row row_map[k] of B (null: row k).
hipError_t launch_frag_b_f64(hipStream_t s, int64_t K, int64_t Ncol, const double *B, int64_t ldb,
                             double *Bf, int64_t *nk4, const int32_t *col_mo = nullptr,
                             const int32_t *row_map = nullptr);
// out[r][j] = in[rmap[r]][cmap[j]] for r < rows, j < cols (a null map: the identity).
hipError_t launch_gather(hipStream_t s, int64_t rows, int64_t cols, const double *in, int64_t ld_in,
                         const int32_t *rmap, const int32_t *cmap, double *out, int64_t ld_out);
// With a column plan, Bf is in the plan's column order (launch_frag_b_f64
// with plan.col_mo), column j of the product is stored at C column
// plan.col_mo[j], and the MFMAs and coefficient loads of (column block,
// 16-AO block) pairs whose coefficients are all zero are skipped.  ocols
// (onkb bytes per tile of 320 output rows, one per 16 output columns; not
// with plan.col_mo): a 16-column block of a tile whose flag is zero is
// neither computed nor stored.
hipError_t launch_mo_gemm_f64_frag(hipStream_t s, int64_t R, int64_t K, int64_t Ncol, const double *A,
                                   int64_t lda, const double *Bf, int64_t nk4, double *C, int64_t ldc,
                                   const uint8_t *flags = nullptr, int nkb = 0, ColPlan plan = ColPlan(),
                                   const uint8_t *ocols = nullptr, int onkb = 0);
// Nonzero map of coef [M][K] at 16-AO granularity into words [M][nw],
// nw = ceil(ceil(K / 16) / 64): bit s of words[m * nw + s / 64] is set when
// MO m has a nonzero (or NaN) coefficient among the AOs [16s, 16s + 16).
hipError_t launch_coef_blocks(hipStream_t s, int64_t M, int64_t K, const double *coef, int64_t nw,
                              uint64_t *words);
// The column plan of the FP64 contraction (mo.cpp; built on first use after
// finalize): *plan is empty when the MO order is kept; *host_blk points to
// the host copy of its live bits.
hipError_t mo_f64_column_plan(Context *c, ColPlan *plan, const std::vector<uint64_t> **host_blk);
// Column order from that map (mo_plan.cpp) for the M MOs whose words start
// at `words`: the MOs in groups of 16 that each touch few AO blocks, the
// groups chained so that neighbours touch similar blocks.  Fills col_mo [M]
// (indices into those M MOs) and blk [ceil(M/16)][nw] and returns the
// number of live (column block, AO block) pairs over nkb16 = ceil(K/16) AO
// blocks per column block.
int64_t plan_mo_columns(const uint64_t *words, int64_t M, int64_t nw, int64_t nkb16,
                        std::vector<int32_t> &col_mo, std::vector<uint64_t> &blk);
// out_t[Ncol][R] = (A^T Bf)^T with A stored K-major [K][lda] (element (r, k)
// at A[k*lda + r]) and B in the fragment order of launch_frag_b_f64, split
// over K into `splits` partials [splits][R][Ncol] reduced in a fixed order.
// With kmap (K-row map, [K]), row k of the product is row kmap[k] of A (B
// already in the mapped order); with tflags ([K/320 rounded up][tnkb]), the
// K rows of 320-row tile t contribute nothing to the 16-row output blocks b
// whose flag (t, b) is zero, and are skipped.
int gemm_tn_frag_splits(int64_t R, int64_t K);
hipError_t launch_gemm_tn_f64_frag(hipStream_t s, int64_t R, int64_t K, int64_t Ncol, const double *A,
                                   int64_t lda, const double *Bf, int64_t nk4, double *out_t,
                                   double *partial, int splits, const int32_t *kmap = nullptr,
                                   const uint8_t *tflags = nullptr, int tnkb = 0);
// kmap[5 i + c] = 5 perm[i] + c for the (point, component) rows of P points.
hipError_t launch_point_rows_map(hipStream_t s, int64_t P, const int32_t *perm, int32_t *kmap);
// Screen flags of P points in 64-point tiles (tile t = points perm[64 t ..],
// perm null: in order): tflags[t][b] = 1 when some shell overlapping 16-AO
// block b passes the AO screen for some point of the tile, else 0 (every
// AO of the block is then zero at every point of the tile); nkb = A / 16
// rounded up.
hipError_t launch_tile_screen(hipStream_t s, const DeviceBasis &b, int64_t P, const double *coord,
                              const int32_t *perm, const double *nucl, uint8_t *tflags, int nkb);
// Per-block partials of dL/dR and dL/dalpha from G: block j holds the
// points perm[64 j ..] (perm null: 64 j ..), whose G rows are 5 i + c for
// i = 64 j .. (G's rows in the order of perm).
hipError_t launch_ao_adjoint(hipStream_t s, const DeviceBasis &b, int64_t P, const double *coord,
                             const int32_t *perm, const double *nucl, const double *ao_factor, const double *G,
                             double *partR, double *partA);
int64_t ao_adjoint_blocks(int64_t P);
hipError_t launch_adjoint_reduce(hipStream_t s, const DeviceBasis &b, int64_t nblk,
                                 const int64_t *nucleus_index, const int64_t *nucleus_shell_num,
                                 const double *partR, const double *partA, double *dR, double *dA);
inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }
int split_terms(int form);
// MOs per workgroup of the split contraction: the smallest of 64, 128,
// 192, 256, 320 that covers mo_num, at most 320 (192 for the three-term
// form: register budget); more MOs take several workgroup columns.
inline int split_wg_cols(int form, int64_t M) {
  const int cap = form == QMCKL_AMD_SPLIT_BF16X3 ? 192 : 320;
  const int64_t w = (M + 63) / 64 * 64;
  return (int)(w < 64 ? 64 : (w > cap ? cap : w));
}
// MOs the prepared coefficient buffer holds: M rounded up to the workgroup
// width of the split contraction (zero columns), M for the FP32 form.
inline int64_t split_cols_padded(int form, int64_t M) {
  if (form == QMCKL_AMD_SPLIT_FP32) return M;
  return round_up(M, split_wg_cols(form, M));
}
// Prepared column j (and colinv[j]) holds MO col_mo[j] (null: MO j); the
// FP32 form takes no column map.
hipError_t launch_split_coef(hipStream_t s, int form, int64_t M, int64_t K, const double *coef,
                             int64_t ldb, uint16_t *B, float *colinv, const int32_t *col_mo = nullptr);
// Mixed MO contraction of the blocked AO operand (launch_ao_rows_f32) of P
// points with the prepared coefficients into mo_vgl C [P][5][Ncol] (ldc);
// tile row i holds point perm[i] (null: point i).  The split forms take a
// column plan as launch_mo_gemm_f64_frag does (the coefficients prepared
// with its col_mo); the FP32 form takes none.
hipError_t launch_mo_gemm_mixed(hipStream_t s, int form, int64_t P, int64_t K, int64_t Ncol,
                                const float *A, int64_t lda, const float *rowmax, const void *B, int64_t ldb,
                                const float *colinv, const uint8_t *flags, int nkb, const int32_t *perm,
                                double *C, int64_t ldc, ColPlan plan = ColPlan(), int gemm_tile = 0);
// Fills *job for a split form (the checks of launch_mo_gemm_mixed); item0 is
// left to the caller.  The FP32 form has no persistent launch.
hipError_t make_mixed_gemm_job(int form, int64_t P, int64_t K, int64_t Ncol, const float *A, int64_t lda,
                               const float *rowmax, const void *B, int64_t ldb, const float *colinv,
                               const uint8_t *flags, int nkb, const int32_t *perm, double *C, int64_t ldc,
                               ColPlan plan, MixedGemmJob *job, int gemm_tile = 0);
// One launch for every job of the same split form and tile width cbw: jobs
// [njobs] in device memory with consecutive item ranges covering [0, total),
// counter a device int the launch zeroes first (a kernel, so it can be
// captured).
hipError_t launch_mo_gemm_mixed_persistent(hipStream_t s, int form, int cbw, int dense, const MixedGemmJob *jobs,
                                           int njobs, int64_t total, int *counter);
// Fused AO evaluation and split contraction (k_fused.hip) of P points
// (coordinates [3][P], tile row i = point perm[i], null: point i) with the
// coefficients prepared without a column plan (launch_split_coef, col_mo
// null, ld = ldb) into mo_vgl C [P][5][M] (ldc): one workgroup per 64
// points and 64 MOs, no AO operand in memory.
constexpr int64_t kFusedMaxLd = 8192;  // AOs (rounded up to 64) the fused kernel's stage mask covers
hipError_t launch_fused_ao_mo(hipStream_t s, int form, const DeviceBasis &b, int64_t P, const double *coord,
                              const int32_t *perm, const double *nucl, const double *ao_factor, int64_t ld,
                              const uint16_t *Bs, const float *colinv, int64_t M, double *C, int64_t ldc);
hipError_t launch_strided_copy(hipStream_t s, int64_t rows, int64_t cols, const double *in,
                               int64_t ld_in, double *out, int64_t ld_out);
// INT8X5 form (k_gemm_i8.hip): the coefficient slices of M MOs (prepared
// column j = MO col_mo[j]) with ld = ldb, slice_coef_i8_bytes(M, ldb) bytes,
// and colinv [M] their inverse MO scales; the contraction of the blocked AO
// operand A and its residuals Alo (launch_ao_rows_f32 with buf_lo) into
// mo_vgl C, with the flags, point order and column plan of the split forms.
int64_t slice_coef_i8_bytes(int64_t M, int64_t ldb);
hipError_t launch_slice_coef_i8(hipStream_t s, int64_t M, int64_t K, const double *coef, int64_t ldb, int8_t *Bq,
                                float *colinv, const int32_t *col_mo);
hipError_t launch_mo_gemm_i8(hipStream_t s, int64_t P, int64_t K, int64_t Ncol, const float *A, const float *Alo,
                             int64_t lda, const int8_t *Bq, int64_t ldb, const float *colinv, const uint8_t *flags,
                             int nkb, const int32_t *perm, double *C, int64_t ldc, ColPlan plan);

// Allocation-free radial launcher (noalloc.hip): the shell table is built on
// the device into a fixed file-scope device array of kNoAllocMaxShells
// entries.  noalloc_scratch_init(device) looks up that array's address on
// `device` (which must be current) once per device; context creation calls
// it.  launch_shell_vgl_noalloc validates the index arrays on the device,
// waits for that check on `s` and sets *table_err (bit 1: nucleus ranges,
// 2: shell coverage, 4: primitive ranges) instead of running the radial
// kernel when it fails.
constexpr int64_t kNoAllocMaxShells = 65536;
hipError_t noalloc_scratch_init(int device);
hipError_t launch_shell_vgl_noalloc(hipStream_t s, int device, int64_t P, int64_t N, int64_t S,
                                    int64_t Np, const int64_t *nucleus_shell_num,
                                    const int64_t *nucleus_index, const double *nucleus_range,
                                    const int64_t *shell_prim_index, const int64_t *shell_prim_num,
                                    const double *coord, const double *nucl, const double *expo,
                                    const double *coef_normalized, double *shell_vgl, int *table_err);

}  // namespace qmckl_amd
