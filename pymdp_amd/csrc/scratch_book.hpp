// Bookkeeping of the stream-scoped scratch (host only, no HIP calls, so it
// is testable on a machine without a GPU).
//
// One grow-only buffer per key (device, stream, pool).  A buffer that has to
// grow is moved to the retired list, never freed here; retired buffers and
// the live buffers of one device are handed back only by release(), which
// the owner calls after synchronising the device.  Under capture nothing is
// allocated.
#pragma once

#include <cstddef>
#include <cstdint>
#include <map>
#include <tuple>
#include <vector>

namespace pmh {

class ScratchBook {
 public:
  using Key = std::tuple<int, uintptr_t, int>;  // device, stream, pool
  enum Result { kOk = 0, kCapturing = 1, kNoMemory = 2 };

  static constexpr size_t kGranule = size_t(2) << 20;  // allocations are rounded to 2 MiB

  // Scratch of at least `bytes` for `key`; `alloc(n)` returns n bytes or
  // nullptr.  A grown buffer's predecessor is retired.
  template <class Alloc>
  Result get(const Key& key, size_t bytes, bool capturing, Alloc&& alloc, void** out) {
    if (bytes == 0) bytes = 1;
    Buf& b = bufs_[key];
    if (b.bytes >= bytes) {
      *out = b.ptr;
      return kOk;
    }
    if (capturing) return kCapturing;
    const size_t want = (bytes + kGranule - 1) / kGranule * kGranule;
    void* p = alloc(want);
    if (!p) return kNoMemory;
    if (b.ptr) retired_.push_back(b.ptr);
    b.ptr = p;
    b.bytes = want;
    *out = p;
    return kOk;
  }

  void retire(void* p) {
    if (p) retired_.push_back(p);
  }

  // Every retired buffer and the live buffers of `device`; they leave the book.
  std::vector<void*> release(int device) {
    std::vector<void*> out;
    out.swap(retired_);
    for (auto it = bufs_.begin(); it != bufs_.end();) {
      if (std::get<0>(it->first) == device) {
        if (it->second.ptr) out.push_back(it->second.ptr);
        it = bufs_.erase(it);
      } else {
        ++it;
      }
    }
    return out;
  }

  size_t live_bytes() const {
    size_t t = 0;
    for (const auto& kv : bufs_) t += kv.second.bytes;
    return t;
  }
  size_t retired_count() const { return retired_.size(); }

 private:
  struct Buf {
    void* ptr = nullptr;
    size_t bytes = 0;
  };
  std::map<Key, Buf> bufs_;
  std::vector<void*> retired_;
};

}  // namespace pmh
