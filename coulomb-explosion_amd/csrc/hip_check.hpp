// HIP keeps a thread's last error until hipGetLastError reads it, so a failed
// call stays visible to every later launch check of the thread.  The library
// reads (clears) each error it reports, and clears before each launch it checks
// with hipGetLastError, so that check sees that launch only (not an earlier
// failure of its own or of the caller's).
#pragma once

#include <hip/hip_runtime.h>

namespace ce {
inline void clear_stale_error() { (void)hipGetLastError(); }
}  // namespace ce
