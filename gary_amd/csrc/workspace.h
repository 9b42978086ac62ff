// workspace.h — grow-only device scratch buffers owned by the library.
#pragma once

#include <cstddef>

namespace gary {

// Returns (in *ptr) a device buffer of at least `bytes` bytes for (device, stream). The buffer
// is reused by every later call with the same key and is held until the process ends. Calls that
// share a key must be ordered by that stream; the caller holds no lock after return.
int workspace_get(int device, void *stream, size_t bytes, void **ptr);

// Kernel timing hooks around the dominant counting kernel (no-ops unless enabled through
// gary_kernel_timing).
int timing_begin(void *stream);
int timing_end(void *stream);

// Current device (hipGetDevice) with error reporting.
int current_device(int *device);

}  // namespace gary
