// Stream-scoped device scratch for the plain entry points.
//
// One grow-only buffer per (device, stream, pool).  Calls on one stream are
// ordered by the stream and share its buffer; calls on different streams
// never share one, so entry points may be called concurrently on different
// streams.  A buffer that has to grow is retired, not freed: a HIP graph
// captured earlier may still hold its address.  Retired buffers (and plan
// tables retired by their owners) are freed only by
// pymdp_hip_release_scratch(), which the caller must not invoke while such a
// graph may still be replayed.  Under stream capture nothing is allocated: a
// buffer too small for the call is an error that says to make one call of
// the shape on that stream outside capture first.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace pmh {

enum ScratchPool { kPoolFpi = 0, kPoolEfe = 1, kPoolAdjoint = 2, kPoolCount = 3 };

// Device scratch of at least `bytes` for `pool` on the current device and
// `stream`.  Returns 0, or an error code with the message set under `pfx`.
int scratch_get(int pool, hipStream_t stream, size_t bytes, void** out, const char* pfx);

// True if `stream` is capturing into a graph.
bool stream_capturing(hipStream_t stream);

// Hands a device allocation to the retired list; freed by
// pymdp_hip_release_scratch().
void scratch_retire(void* p);

}  // namespace pmh
