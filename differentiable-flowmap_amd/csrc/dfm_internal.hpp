// Host-side error plumbing and stream waits shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

namespace dfm {
// Records a thread-local message for dfm_last_error() and returns `code`.
int set_error(int code, const char* what, const char* detail = nullptr);

// Waits for the work queued on `st` so far by spinning on an event recorded
// behind it.  The calling thread stays runnable: a blocking stream wait
// sleeps, and its wake-up on an idle host core measured ~2 ms per solver
// poll in some processes (the GPU idling meanwhile), against none when the
// process ran on one busy core.  `ev` is the caller's own event
// (hipEventDisableTiming); an error of the stream ends the wait.
inline hipError_t spin_sync(hipStream_t st, hipEvent_t ev) {
  hipError_t e = hipEventRecord(ev, st);
  if (e != hipSuccess) return e;
  while ((e = hipEventQuery(ev)) == hipErrorNotReady) {
  }
  return e;
}
}  // namespace dfm
