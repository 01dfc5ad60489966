// Launch tables to the device without a host wait.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace mba_host {

// Copies `bytes` from host memory `src` to device memory `dst` on stream
// `st`, asynchronously: the bytes go first into a pinned staging slot (a ring
// of slots per device, each guarded by an event recorded after its copy), so
// `src` may be released on return and the stream is not synchronised.  A slot
// is reused only after its previous copy has completed.  Returns 0, or -1
// after a failed runtime call.
int upload_tables(void* dst, const void* src, size_t bytes, hipStream_t st);

}  // namespace mba_host
