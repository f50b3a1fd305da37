// lbm_common.hip -- error state, device memory entry points and the copy
// kernel used to calibrate bandwidth measurements.
#include "lbm_internal.hpp"

namespace lbm {

static thread_local std::string g_last_error;

// Records the message of a failed call.  It also resets HIP's per-thread
// last-error state, which a failed runtime call such as an out-of-memory
// hipMalloc leaves set: otherwise the next launch check of any handle would
// report that old failure.
void set_error(const char* fmt, ...) {
  (void)hipGetLastError();
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

void clear_error() { g_last_error.clear(); }

// 16-byte vector copy, grid-stride.  Reads `n16` uint4 and writes them: 32 B of
// HBM traffic per element.  Non-temporal, like the streaming time-step
// kernels, and the faster of the two forms on the MI355X (4 GiB each way:
// about 6.5 vs 6.2-6.3 TB/s cached).
__global__ void __launch_bounds__(256) k_copy16(const uint4* __restrict__ src,
                                                uint4* __restrict__ dst, size_t n16) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  using v4 = unsigned int __attribute__((ext_vector_type(4)));
  const v4* s = reinterpret_cast<const v4*>(src);
  v4* d = reinterpret_cast<v4*>(dst);
  for (; i < n16; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(s + i), d + i);
}

}  // namespace lbm

extern "C" {

const char* lbm_last_error(void) { return lbm::g_last_error.c_str(); }
const char* lbm_version(void) { return "lbmminiapp_amd 0.1 (gfx950)"; }

lbm_status lbm_device_count(int* count) {
  LBM_REQUIRE(count, LBM_ERR_ARG, "lbm_device_count: null pointer");
  LBM_HIP_CHECK(hipGetDeviceCount(count));
  return LBM_OK;
}

lbm_status lbm_set_device(int device) {
  LBM_HIP_CHECK(hipSetDevice(device));
  return LBM_OK;
}

lbm_status lbm_device_alloc(size_t bytes, void** ptr) {
  LBM_REQUIRE(ptr, LBM_ERR_ARG, "lbm_device_alloc: null pointer");
  *ptr = nullptr;
  if (bytes == 0) return LBM_OK;
  LBM_HIP_CHECK(hipMalloc(ptr, bytes));
  return LBM_OK;
}

lbm_status lbm_device_free(void* ptr) {
  if (ptr) LBM_HIP_CHECK(hipFree(ptr));
  return LBM_OK;
}

lbm_status lbm_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  LBM_REQUIRE(dst && src, LBM_ERR_ARG, "lbm_memcpy_h2d: null pointer");
  LBM_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return LBM_OK;
}

lbm_status lbm_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  LBM_REQUIRE(dst && src, LBM_ERR_ARG, "lbm_memcpy_d2h: null pointer");
  LBM_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return LBM_OK;
}

lbm_status lbm_device_synchronize(void) {
  LBM_HIP_CHECK(hipDeviceSynchronize());
  return LBM_OK;
}

namespace {
// The timed copy of lbm_copy_bandwidth on resources the caller owns and frees.
lbm_status copy_timed(const void* a, void* b, size_t bytes, int32_t reps, hipStream_t st,
                      hipEvent_t e0, hipEvent_t e1, double* elapsed_ms) {
  LBM_HIP_CHECK(hipMemsetAsync(const_cast<void*>(a), 0, bytes, st));
  const size_t n16 = bytes / 16;
  const int block = 256;
  size_t grid = (n16 + block - 1) / block;
  if (grid > 1048576) grid = 1048576;
  // one untimed launch to page in both buffers
  lbm::k_copy16<<<dim3((unsigned)grid), dim3(block), 0, st>>>((const uint4*)a, (uint4*)b, n16);
  LBM_LAUNCH_CHECK();
  LBM_HIP_CHECK(hipEventRecord(e0, st));
  for (int r = 0; r < reps; r++)
    lbm::k_copy16<<<dim3((unsigned)grid), dim3(block), 0, st>>>((const uint4*)a, (uint4*)b, n16);
  LBM_LAUNCH_CHECK();
  LBM_HIP_CHECK(hipEventRecord(e1, st));
  LBM_HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  LBM_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  *elapsed_ms = ms;
  return LBM_OK;
}
}  // namespace

lbm_status lbm_copy_bandwidth(size_t bytes, int32_t reps, double* elapsed_ms) {
  LBM_REQUIRE(elapsed_ms && reps > 0 && bytes >= 16 && bytes % 16 == 0, LBM_ERR_ARG,
              "lbm_copy_bandwidth: bytes must be a positive multiple of 16, reps > 0");
  void *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  lbm_status status = LBM_OK;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    lbm::set_error("lbm_copy_bandwidth: hipMalloc of 2 x %zu bytes failed", bytes);
    status = LBM_ERR_OOM;
  } else if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
             hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    lbm::set_error("lbm_copy_bandwidth: stream/event creation failed");
    status = LBM_ERR_HIP;
  } else {
    status = copy_timed(a, b, bytes, reps, st, e0, e1, elapsed_ms);
  }
  if (st) (void)hipStreamSynchronize(st);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return status;
}

}  // extern "C"
