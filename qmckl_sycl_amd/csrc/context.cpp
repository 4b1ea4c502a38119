// Context, error record and tracked memory.
//
// Follows src/qmckl_context.cpp (create 187-288, destroy 294-345, touch
// 47-62, check 64-79, lock/unlock 81-110), src/qmckl_basic_functions.cpp
// (failwith 50-79) and src/qmckl_memory.cpp (tracked device allocations
// 70-125 and 173-220, blocking memcpys 226-362).  The context owns one HIP
// stream on its device; the memory table is a hash map keyed by pointer.
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <new>
#include <unordered_set>

#include "internal.hpp"

namespace qmckl_amd {

// Live contexts: a handle is dereferenced only if create returned it and
// destroy has not retired it, so a stale or arbitrary handle is reported as
// an invalid context instead of being read.
namespace {
std::mutex g_live_mutex;
std::unordered_set<const Context *> &live_contexts() {
  static std::unordered_set<const Context *> live;
  return live;
}
}  // namespace

Context *ctx_of(qmckl_context_device context) {
  if (context == QMCKL_NULL_CONTEXT_DEVICE) return nullptr;
  Context *c = reinterpret_cast<Context *>(context);
  {
    std::lock_guard<std::mutex> g(g_live_mutex);
    if (live_contexts().count(c) == 0) return nullptr;
  }
  if (c->tag != kValidTag) return nullptr;
  return c;
}

static void copy_str(char *dst, size_t cap, const char *src) {
  if (!src) src = "";
  std::strncpy(dst, src, cap - 1);
  dst[cap - 1] = '\0';
}

qmckl_exit_code_device fail(qmckl_context_device context, qmckl_exit_code_device code,
                            const char *function, const char *message) {
  return qmckl_failwith_device(context, code, function, message);
}

bool has_device(const Context *c) { return c->device_id >= 0 && c->stream != nullptr; }

bool grow_bytes(Context *c, void **slot, int64_t *cap, int64_t need) {
  if (*slot != nullptr && *cap >= need) return true;
  if (!drealloc(c, slot, (size_t)(need > 0 ? need : 1))) {
    *cap = 0;
    return false;
  }
  *cap = need;
  return true;
}

void *dmalloc(Context *c, size_t bytes) {
  if (!has_device(c)) return nullptr;
  if (bytes == 0) bytes = 1;
  void *p = nullptr;
  if (hipSetDevice(c->device_id) != hipSuccess) return nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  c->device_allocs[p] = bytes;
  return p;
}

qmckl_exit_code_device dfree(Context *c, void *ptr) {
  if (ptr == nullptr) return QMCKL_SUCCESS_DEVICE;
  {
    std::lock_guard<std::recursive_mutex> g(c->mutex);
    auto it = c->device_allocs.find(ptr);
    if (it == c->device_allocs.end()) return QMCKL_DEALLOCATION_FAILED_DEVICE;
    c->device_allocs.erase(it);
  }
  // Work queued on the context's stream may still use the buffer.
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (hipFree(ptr) != hipSuccess) return QMCKL_DEALLOCATION_FAILED_DEVICE;
  return QMCKL_SUCCESS_DEVICE;
}

bool drealloc(Context *c, void **slot, size_t bytes) {
  if (*slot) {
    dfree(c, *slot);
    *slot = nullptr;
  }
  *slot = dmalloc(c, bytes);
  return *slot != nullptr;
}

hipError_t sync(Context *c) {
  if (!has_device(c)) return hipErrorNoDevice;
  if (c->defer_sync) return hipSuccess;
  return hipStreamSynchronize(c->stream);
}

qmckl_exit_code_device hip_check(qmckl_context_device context, hipError_t e, const char *function) {
  if (e == hipSuccess) return QMCKL_SUCCESS_DEVICE;
  return fail(context, QMCKL_FAILURE_DEVICE, function, hipGetErrorString(e));
}

KernelTimer::KernelTimer(Context *c, const char *name) : c_(c), name_(name) {
  if (!c_->timing || c_->capturing || !has_device(c_)) return;
  if (hipEventCreate(&start_) != hipSuccess || hipEventCreate(&stop_) != hipSuccess ||
      hipEventRecord(start_, c_->stream) != hipSuccess) {
    (void)hipGetLastError();
    if (start_) (void)hipEventDestroy(start_);
    if (stop_) (void)hipEventDestroy(stop_);
    start_ = stop_ = nullptr;
  }
}

void KernelTimer::stop() {
  if (start_ && !stopped_) stopped_ = hipEventRecord(stop_, c_->stream) == hipSuccess;
}

KernelTimer::~KernelTimer() {
  if (!start_) return;
  stop();
  float ms = 0.0f;
  if (stopped_ && hipEventSynchronize(stop_) == hipSuccess &&
      hipEventElapsedTime(&ms, start_, stop_) == hipSuccess) {
    std::lock_guard<std::recursive_mutex> g(c_->mutex);
    auto &t = c_->timers[name_];
    t.launches += 1;
    t.total_ms += ms;
  }
  (void)hipEventDestroy(start_);
  (void)hipEventDestroy(stop_);
}

}  // namespace qmckl_amd

using namespace qmckl_amd;

extern "C" {

qmckl_exit_code_device qmckl_failwith_device(qmckl_context_device context,
                                             qmckl_exit_code_device exit_code,
                                             const char *function, const char *message) {
  Context *c = ctx_of(context);
  if (c == nullptr) return QMCKL_INVALID_CONTEXT_DEVICE;
  if (exit_code == QMCKL_SUCCESS_DEVICE) return QMCKL_SUCCESS_DEVICE;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  c->err_code = exit_code;
  copy_str(c->err_function, sizeof(c->err_function), function);
  copy_str(c->err_message, sizeof(c->err_message), message);
  return exit_code;
}

const char *qmckl_string_of_error_device(qmckl_exit_code_device error) {
  switch (error) {
    case QMCKL_SUCCESS_DEVICE: return "Success";
    case QMCKL_FAILURE_DEVICE: return "Failure";
    case QMCKL_ERRNO_DEVICE: return "errno set";
    case QMCKL_INVALID_CONTEXT_DEVICE: return "Invalid context";
    case QMCKL_ALLOCATION_FAILED_DEVICE: return "Allocation failed";
    case QMCKL_DEALLOCATION_FAILED_DEVICE: return "De-allocation failed";
    case QMCKL_NOT_PROVIDED_DEVICE: return "Not provided";
    case QMCKL_OUT_OF_BOUNDS_DEVICE: return "Index out of bounds";
    case QMCKL_ALREADY_SET_DEVICE: return "Attempt to modify an immutable value";
    case QMCKL_INVALID_EXIT_CODE_DEVICE: return "Invalid exit code";
    default: break;
  }
  if (error >= 1 && error <= 20) {
    static const char *names[21] = {
        "", "Invalid argument 1", "Invalid argument 2", "Invalid argument 3",
        "Invalid argument 4", "Invalid argument 5", "Invalid argument 6",
        "Invalid argument 7", "Invalid argument 8", "Invalid argument 9",
        "Invalid argument 10", "Invalid argument 11", "Invalid argument 12",
        "Invalid argument 13", "Invalid argument 14", "Invalid argument 15",
        "Invalid argument 16", "Invalid argument 17", "Invalid argument 18",
        "Invalid argument 19", "Invalid argument 20"};
    return names[error];
  }
  return "Unknown error";
}

qmckl_exit_code_device qmckl_get_error_device(qmckl_context_device context,
                                              qmckl_exit_code_device *exit_code,
                                              char *function_name, char *message) {
  Context *c = ctx_of(context);
  if (c == nullptr) return QMCKL_INVALID_CONTEXT_DEVICE;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  if (exit_code) *exit_code = c->err_code;
  if (function_name) copy_str(function_name, QMCKL_MAX_FUN_LEN_DEVICE, c->err_function);
  if (message) copy_str(message, QMCKL_MAX_MSG_LEN_DEVICE, c->err_message);
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_init_point_device(qmckl_context_device context) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  c->point = PointState();
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_init_ao_basis_device(qmckl_context_device context) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  c->ao.uninitialized = (1 << 14) - 1;  // src/qmckl_ao.cpp:18
  c->ao.provided = false;
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_init_mo_basis_device(qmckl_context_device context) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  c->mo.uninitialized = (1 << 2) - 1;  // src/qmckl_context.cpp:141
  c->mo.provided = false;
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_context_device qmckl_context_create_device(int device_id) {
  Context *c = new (std::nothrow) Context();
  if (c == nullptr) return QMCKL_NULL_CONTEXT_DEVICE;
  if (device_id != QMCKL_NO_DEVICE) {
    int count = 0;
    if (device_id < 0 || hipGetDeviceCount(&count) != hipSuccess || device_id >= count ||
        hipSetDevice(device_id) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
      (void)hipGetLastError();
      delete c;
      return QMCKL_NULL_CONTEXT_DEVICE;
    }
    c->device_id = device_id;
    // address of the allocation-free launcher's device scratch on this
    // device (once per device); on failure that launcher retries and
    // reports the error itself, every other entry point is unaffected
    if (noalloc_scratch_init(device_id) != hipSuccess) (void)hipGetLastError();
  }
  {
    std::lock_guard<std::mutex> g(g_live_mutex);
    live_contexts().insert(c);
  }
  return reinterpret_cast<qmckl_context_device>(c);
}

qmckl_exit_code_device qmckl_context_destroy_device(qmckl_context_device context) {
  Context *c = ctx_of(context);
  if (c == nullptr) return QMCKL_INVALID_CONTEXT_DEVICE;
  {
    std::lock_guard<std::mutex> g(g_live_mutex);
    live_contexts().erase(c);
  }
  if (c->stream) {
    (void)hipSetDevice(c->device_id);
    (void)hipStreamSynchronize(c->stream);
  }
  for (auto &kv : c->device_allocs) (void)hipFree(kv.first);
  c->device_allocs.clear();
  for (auto &kv : c->host_allocs) std::free(kv.first);
  c->host_allocs.clear();
  if (c->batch_event) (void)hipEventDestroy(c->batch_event);
  c->batch_event = nullptr;
  if (c->batch_jobs_pinned) (void)hipHostFree(c->batch_jobs_pinned);
  c->batch_jobs_pinned = nullptr;
  if (c->stream) (void)hipStreamDestroy(c->stream);
  c->stream = nullptr;
  c->tag = kInvalidTag;
  delete c;
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_context_touch_device(qmckl_context_device context) {
  Context *c = ctx_of(context);
  if (c == nullptr)
    return QMCKL_INVALID_CONTEXT_DEVICE;
  c->date += 1;
  c->point.date = c->date;
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_context_device qmckl_context_check_device(qmckl_context_device context) {
  return ctx_of(context) ? context : QMCKL_NULL_CONTEXT_DEVICE;
}

void qmckl_lock_device(qmckl_context_device context) {
  Context *c = ctx_of(context);
  if (!c) return;
  c->mutex.lock();
  c->lock_count += 1;
}

void qmckl_unlock_device(qmckl_context_device context) {
  Context *c = ctx_of(context);
  if (!c) return;
  c->lock_count -= 1;
  c->mutex.unlock();
}

int64_t qmckl_get_device_id(qmckl_context_device context) {
  Context *c = ctx_of(context);
  return c ? c->device_id : (int64_t)QMCKL_NO_DEVICE;
}

void *qmckl_malloc_host(qmckl_context_device context, qmckl_memory_info_struct_device info) {
  Context *c = ctx_of(context);
  if (!c) return nullptr;
  void *p = std::calloc(1, info.size ? info.size : 1);
  if (!p) return nullptr;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  c->host_allocs[p] = info.size;
  return p;
}

qmckl_exit_code_device qmckl_free_host(qmckl_context_device context, void *ptr) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  if (ptr == nullptr)
    return fail(context, QMCKL_INVALID_ARG_2_DEVICE, "qmckl_free_host", "NULL pointer");
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  auto it = c->host_allocs.find(ptr);
  if (it == c->host_allocs.end())
    return fail(context, QMCKL_INVALID_ARG_2_DEVICE, "qmckl_free_host",
                "Pointer not found in context");
  c->host_allocs.erase(it);
  std::free(ptr);
  return QMCKL_SUCCESS_DEVICE;
}

void *qmckl_malloc_device(qmckl_context_device context, size_t size) {
  Context *c = ctx_of(context);
  if (!c) return nullptr;
  void *p = dmalloc(c, size);
  if (p == nullptr)
    fail(context, QMCKL_ALLOCATION_FAILED_DEVICE, "qmckl_malloc_device",
         has_device(c) ? "hipMalloc failed" : "no device bound to this context");
  return p;
}

qmckl_exit_code_device qmckl_free_device(qmckl_context_device context, void *ptr) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  if (ptr == nullptr)
    return fail(context, QMCKL_INVALID_ARG_2_DEVICE, "qmckl_free_device", "NULL pointer");
  qmckl_exit_code_device rc = dfree(c, ptr);
  if (rc != QMCKL_SUCCESS_DEVICE)
    return fail(context, QMCKL_INVALID_ARG_2_DEVICE, "qmckl_free_device",
                "Pointer not found in context");
  return QMCKL_SUCCESS_DEVICE;
}

static qmckl_exit_code_device memcpy_kind(qmckl_context_device context, void *dest, void *src,
                                          size_t size, hipMemcpyKind kind, const char *fn) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  if (dest == nullptr) return fail(context, QMCKL_INVALID_ARG_2_DEVICE, fn, "NULL pointer");
  if (src == nullptr) return fail(context, QMCKL_INVALID_ARG_3_DEVICE, fn, "NULL pointer");
  if (!has_device(c)) return fail(context, QMCKL_FAILURE_DEVICE, fn, "no device bound");
  if (size == 0) return QMCKL_SUCCESS_DEVICE;
  hipError_t e = hipMemcpyAsync(dest, src, size, kind, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(context, QMCKL_FAILURE_DEVICE, fn, hipGetErrorString(e));
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_memcpy_H2D(qmckl_context_device context, void *dest, void *src,
                                        size_t size) {
  return memcpy_kind(context, dest, src, size, hipMemcpyHostToDevice, "qmckl_memcpy_H2D");
}

qmckl_exit_code_device qmckl_memcpy_D2H(qmckl_context_device context, void *dest, void *src,
                                        size_t size) {
  return memcpy_kind(context, dest, src, size, hipMemcpyDeviceToHost, "qmckl_memcpy_D2H");
}

qmckl_exit_code_device qmckl_memcpy_D2D(qmckl_context_device context, void *dest, void *src,
                                        size_t size) {
  return memcpy_kind(context, dest, src, size, hipMemcpyDeviceToDevice, "qmckl_memcpy_D2D");
}

qmckl_exit_code_device qmckl_set_kernel_timing_device(qmckl_context_device context, bool enable) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  c->timing = enable;
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_reset_kernel_timing_device(qmckl_context_device context) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  c->timers.clear();
  return QMCKL_SUCCESS_DEVICE;
}

qmckl_exit_code_device qmckl_get_kernel_timing_device(qmckl_context_device context,
                                                      const char *kernel, int64_t *launches,
                                                      double *total_ms) {
  Context *c = ctx_of(context);
  if (!c) return QMCKL_INVALID_CONTEXT_DEVICE;
  if (kernel == nullptr)
    return fail(context, QMCKL_INVALID_ARG_2_DEVICE, "qmckl_get_kernel_timing_device", "NULL name");
  std::lock_guard<std::recursive_mutex> g(c->mutex);
  auto it = c->timers.find(kernel);
  if (launches) *launches = it == c->timers.end() ? 0 : it->second.launches;
  if (total_ms) *total_ms = it == c->timers.end() ? 0.0 : it->second.total_ms;
  return QMCKL_SUCCESS_DEVICE;
}

}  // extern "C"
