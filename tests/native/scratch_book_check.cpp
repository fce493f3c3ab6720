// Host check of the scratch bookkeeping (pymdp_amd/csrc/scratch_book.hpp)
// with a fake allocator: a pointer handed to a caller is never given back
// for freeing before release(); buffers of different streams are distinct;
// nothing is allocated under capture.  Prints "ok" or the first failure.
#include <cstdio>
#include <cstdlib>
#include <set>
#include <vector>

#include "scratch_book.hpp"

using pmh::ScratchBook;

static int fail(const char* what) {
  std::printf("FAIL: %s\n", what);
  return 1;
}

int main() {
  ScratchBook book;
  std::set<void*> live;     // every allocation the fake allocator made and nobody freed
  std::set<void*> handed;   // every pointer a caller has received
  int allocs = 0;
  auto alloc = [&](size_t n) -> void* {
    void* p = std::malloc(n > 64 ? 64 : n);
    live.insert(p);
    ++allocs;
    return p;
  };
  const int dev = 0;
  const uintptr_t s1 = 0x1000, s2 = 0x2000;
  void* p = nullptr;

  // first call on stream 1
  if (book.get({dev, s1, 0}, 100, false, alloc, &p) != ScratchBook::kOk) return fail("first get");
  handed.insert(p);
  void* small = p;
  // same size again: same buffer, no allocation
  if (book.get({dev, s1, 0}, 100, false, alloc, &p) != ScratchBook::kOk || p != small || allocs != 1)
    return fail("reuse of the stream's buffer");
  // a capture of the same shape reuses it
  if (book.get({dev, s1, 0}, 100, true, alloc, &p) != ScratchBook::kOk || p != small)
    return fail("capture of a warm shape");
  // a larger call (outside capture) grows: the old buffer is retired, not freed
  if (book.get({dev, s1, 0}, ScratchBook::kGranule + 1, false, alloc, &p) != ScratchBook::kOk || p == small)
    return fail("growth");
  handed.insert(p);
  if (book.retired_count() != 1) return fail("grown buffer not retired");
  // a larger call under capture allocates nothing
  const int before = allocs;
  if (book.get({dev, s1, 0}, 10 * ScratchBook::kGranule, true, alloc, &p) != ScratchBook::kCapturing ||
      allocs != before)
    return fail("allocation under capture");
  // another stream gets its own buffer
  void* q = nullptr;
  if (book.get({dev, s2, 0}, 100, false, alloc, &q) != ScratchBook::kOk) return fail("second stream");
  handed.insert(q);
  if (q == small || q == p) return fail("streams share a buffer");
  // another pool on the same stream is distinct too
  void* r = nullptr;
  if (book.get({dev, s1, 1}, 100, false, alloc, &r) != ScratchBook::kOk || r == small) return fail("pools");
  handed.insert(r);
  // a plan table retired by its owner joins the list
  void* plan_tab = alloc(32);
  book.retire(plan_tab);
  handed.insert(plan_tab);
  // another device's buffer survives a release of device 0
  void* other = nullptr;
  if (book.get({1, s1, 0}, 100, false, alloc, &other) != ScratchBook::kOk) return fail("device 1");
  // up to here nothing has been handed back for freeing: every pointer ever
  // handed out is still live
  for (void* h : handed)
    if (!live.count(h)) return fail("a handed-out pointer was freed before release");
  std::vector<void*> freed = book.release(dev);
  std::set<void*> fs(freed.begin(), freed.end());
  if (fs.size() != freed.size()) return fail("release returns a pointer twice");
  for (void* h : handed)
    if (!fs.count(h)) return fail("release misses a buffer");
  if (fs.count(other)) return fail("release frees another device's buffer");
  for (void* f : freed) {
    live.erase(f);
    std::free(f);
  }
  if (book.retired_count() != 0) return fail("retired list not emptied");
  // after release the next call allocates again
  if (book.get({dev, s1, 0}, 100, false, alloc, &p) != ScratchBook::kOk || !live.count(p))
    return fail("allocation after release");
  std::printf("ok\n");
  return 0;
}
