// Host ingest: a pageable host array into device memory through pinned
// staging (no reference counterpart: the reference uploads pageable memory
// with one synchronous copy, src/rcabeam/bindings/python.cpp:527-531).
//
// The array is cut into 64 MiB chunks.  Host threads copy chunk k into pinned
// staging buffer k % 2 while the DMA engine moves chunk k - 1 out of the
// other buffer to the device, so the upload runs at the slower of the
// threaded host copy and the host-to-device link instead of their sum.  The
// two staging buffers (128 MiB of pinned memory) and a stream are created
// per device on first use and kept until the process ends; calls on one
// device are serialized.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "rcabeam_mi355x.h"

namespace rcb {
namespace {

constexpr size_t kChunkBytes = size_t{64} << 20;
constexpr int kStagingBuffers = 2;
constexpr int kMaxDevices = 64;
constexpr int kMaxThreads = 64;

struct Staging {
  std::mutex lock;
  bool ready = false;
  void* buffer[kStagingBuffers] = {};
  hipEvent_t drained[kStagingBuffers] = {};  // the copy out of buffer b has finished
  hipEvent_t ordered = nullptr;               // null-stream work enqueued before the call
  hipStream_t stream = nullptr;
};
Staging g_staging[kMaxDevices];

hipError_t prepare(Staging& s) {
  if (s.ready) return hipSuccess;
  hipError_t e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&s.ordered, hipEventDisableTiming);
  for (int b = 0; b < kStagingBuffers && e == hipSuccess; ++b) {
    e = hipHostMalloc(&s.buffer[b], kChunkBytes, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s.drained[b], hipEventDisableTiming);
  }
  s.ready = e == hipSuccess;
  return e;
}

// dst[0, n) = src[0, n) with `threads` host threads (the calling thread takes the first slice).
void parallel_copy(unsigned char* dst, const unsigned char* src, size_t n, int threads) {
  const size_t slice = ((n + threads - 1) / threads + 4095) & ~size_t{4095};  // ceil(n / threads), whole pages
  std::vector<std::thread> pool;
  for (int i = 1; i < threads; ++i) {
    const size_t begin = std::min(n, slice * i), end = std::min(n, slice * (i + 1));
    if (begin < end) pool.emplace_back([=] { std::memcpy(dst + begin, src + begin, end - begin); });
  }
  std::memcpy(dst, src, std::min(n, slice));
  for (auto& t : pool) t.join();
}

}  // namespace
}  // namespace rcb

using namespace rcb;

static rcabeam_status upload(const void* host, void* device, size_t bytes, int threads, bool ordered);

extern "C" rcabeam_status rcabeam_upload_host(const void* host, void* device, size_t bytes, int threads) {
  return upload(host, device, bytes, threads, true);
}

extern "C" rcabeam_status rcabeam_upload_host_concurrent(const void* host, void* device, size_t bytes, int threads) {
  return upload(host, device, bytes, threads, false);
}

static rcabeam_status upload(const void* host, void* device, size_t bytes, int threads, bool ordered) {
  if (!bytes) return RCABEAM_SUCCESS;
  if (!host || !device || threads < 1) return RCABEAM_ERROR_ARGUMENT;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return RCABEAM_ERROR_CUDA;
  Staging& s = g_staging[dev];
  std::lock_guard<std::mutex> guard(s.lock);
  if (prepare(s) != hipSuccess) return RCABEAM_ERROR_CUDA;
  const int t = std::min(threads, kMaxThreads);
  const auto* src = static_cast<const unsigned char*>(host);
  auto* dst = static_cast<unsigned char*>(device);
  // The copies run on a non-blocking stream: order them after everything already enqueued on the
  // null stream (where every other entry point launches), which may still read or write `device`.
  hipError_t e = hipSuccess;
  if (ordered) {
    e = hipEventRecord(s.ordered, nullptr);
    if (e == hipSuccess) e = hipStreamWaitEvent(s.stream, s.ordered, 0);
  }
  for (size_t off = 0, k = 0; off < bytes && e == hipSuccess; off += kChunkBytes, ++k) {
    const size_t n = std::min(kChunkBytes, bytes - off);
    const int b = static_cast<int>(k % kStagingBuffers);
    e = hipEventSynchronize(s.drained[b]);
    if (e != hipSuccess) break;
    parallel_copy(static_cast<unsigned char*>(s.buffer[b]), src + off, n, t);
    e = hipMemcpyAsync(dst + off, s.buffer[b], n, hipMemcpyHostToDevice, s.stream);
    if (e == hipSuccess) e = hipEventRecord(s.drained[b], s.stream);
  }
  const hipError_t sync = hipStreamSynchronize(s.stream);
  return e == hipSuccess && sync == hipSuccess ? RCABEAM_SUCCESS : RCABEAM_ERROR_CUDA;
}
