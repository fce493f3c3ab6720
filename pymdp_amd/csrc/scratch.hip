// Stream-scoped device scratch shared by the plain entry points (host code).
// The bookkeeping is ScratchBook (scratch_book.hpp); this file adds the HIP
// allocation, the capture test and the locking.
#include "scratch.hpp"

#include <mutex>
#include <string>

#include "../../include/pymdp_hip.h"
#include "common.hpp"
#include "scratch_book.hpp"

namespace pmh {
namespace {

// Never destroyed (like the plan registries that retire tables into it), so
// no destructor runs at process exit, when the order of static destruction
// across translation units is unspecified.
std::mutex& g_mu = *new std::mutex;
ScratchBook& g_book = *new ScratchBook;

}  // namespace

bool stream_capturing(hipStream_t stream) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream, &st) != hipSuccess) return false;
  return st != hipStreamCaptureStatusNone;
}

void scratch_retire(void* p) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_book.retire(p);
}

int scratch_get(int pool, hipStream_t stream, size_t bytes, void** out, const char* pfx) {
  int dev = 0;
  PMH_HIP_CHECK(hipGetDevice(&dev), pfx);
  const bool capturing = stream_capturing(stream);
  hipError_t err = hipSuccess;
  auto alloc = [&](size_t n) -> void* {
    void* p = nullptr;
    err = hipMalloc(&p, n);
    return err == hipSuccess ? p : nullptr;
  };
  std::lock_guard<std::mutex> lk(g_mu);
  switch (g_book.get(ScratchBook::Key(dev, reinterpret_cast<uintptr_t>(stream), pool), bytes, capturing, alloc,
                     out)) {
    case ScratchBook::kOk:
      return 0;
    case ScratchBook::kCapturing:
      set_error(std::string(pfx) + "scratch for this call is not allocated on the capturing stream; make one call "
                                   "of this shape on that stream outside stream capture first");
      return PYMDP_HIP_EINVAL;
    default:
      set_error(std::string(pfx) + "hipMalloc of " + std::to_string(bytes) + " scratch bytes: " +
                hipGetErrorString(err));
      return PYMDP_HIP_EDEVICE;
  }
}

}  // namespace pmh

// Frees retired scratch and plan tables and every per-stream scratch buffer
// of the current device (they are allocated again on next use).
// Synchronises the device.  Graphs captured before this call must not be
// replayed after it.
extern "C" int pymdp_hip_release_scratch(void) {
  pmh::set_error("");
  int dev = 0;
  PMH_HIP_CHECK(hipGetDevice(&dev), "scratch: ");
  PMH_HIP_CHECK(hipDeviceSynchronize(), "scratch: ");
  std::lock_guard<std::mutex> lk(pmh::g_mu);
  int rc = 0;
  for (void* p : pmh::g_book.release(dev))
    if (hipFree(p) != hipSuccess) rc = PYMDP_HIP_EDEVICE;
  if (rc) pmh::set_error("scratch: hipFree failed");
  return rc;
}

// Bytes of per-stream scratch held (all devices); *retired_count receives
// the number of retired allocations awaiting release.
extern "C" int64_t pymdp_hip_scratch_bytes(int64_t* retired_count) {
  std::lock_guard<std::mutex> lk(pmh::g_mu);
  if (retired_count) *retired_count = (int64_t)pmh::g_book.retired_count();
  return (int64_t)pmh::g_book.live_bytes();
}
