// workspace.h — grow-only device scratch buffers owned by the library.
#pragma once

#include <cstddef>

namespace gary {

// Returns (in *ptr) a device buffer of at least `bytes` bytes for the key (device, stream), where
// `stream` is a stream handle or any other distinct tag. The buffer is reused by every later call
// with the same key and is held until the process ends; growing it waits for the device. Calls
// that share a key must be ordered (same stream); the caller holds no lock after return.
int workspace_get(int device, void *stream, size_t bytes, void **ptr);

// Kernel timing hooks around the dominant counting kernel (no-ops unless enabled through
// gary_kernel_timing).
int timing_begin(void *stream);
int timing_end(void *stream);

// Current device (hipGetDevice) with error reporting.
int current_device(int *device);

}  // namespace gary
