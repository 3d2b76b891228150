// Device workspace kept by the library per (device, stream), for launchers
// whose callers pass none (mega_zf's three-kernel form, the MLP's split-K
// partial sums). Work queued on one stream runs in order, so every user of
// that stream's buffer is done with it before the next one starts; streams
// never share a buffer. It grows with hipFree (which synchronizes the device,
// so earlier work has finished with the old buffer) and hipMalloc, and is
// never handed out while the stream is being captured into a graph: a
// captured pointer would dangle once the buffer grew.
#include <map>
#include <mutex>

#include "internal.h"

namespace mega {
namespace detail {

hipError_t stream_workspace(hipStream_t st, size_t need, void** out) {
  *out = nullptr;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  hipError_t e = hipStreamIsCapturing(st, &cap);
  if (e != hipSuccess) return e;
  if (cap != hipStreamCaptureStatusNone) return hipErrorStreamCaptureUnsupported;
  struct Ws {
    void* p = nullptr;
    size_t bytes = 0;
  };
  static std::mutex mu;
  static std::map<std::pair<int, hipStream_t>, Ws> ws;
  int dev = 0;
  e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(mu);
  Ws& w = ws[{dev, st}];
  if (w.bytes < need) {
    if (w.p) {
      e = hipFree(w.p);
      if (e != hipSuccess) return e;
      w = Ws{};
    }
    e = hipMalloc(&w.p, need);
    if (e != hipSuccess) return e;
    w.bytes = need;
  }
  *out = w.p;
  return hipSuccess;
}

}  // namespace detail
}  // namespace mega
