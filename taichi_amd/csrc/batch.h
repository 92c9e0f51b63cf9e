// Internal interface between the SpMV kernels (spmv.hip) and the serving
// runtime (serve.hip): descriptor-level enqueue of the batched kernels, so a
// caller can keep descriptors in its own pinned buffers.
#pragma once

#include <hip/hip_runtime.h>

#include "taichi_amd.h"

namespace tc {

// bytes of descriptor memory (host and device) for up to max_requests
// requests of k <= 64 columns
size_t batch_desc_bytes(int max_requests);
// Fills the descriptors of n requests in host_desc (which must stay valid
// until the copy on stream s has run) and enqueues their copy to dev_desc
// plus one launch per (column chunk width, x type) group of the batch. Both
// buffers hold cap bytes; a batch whose descriptors need more is rejected
// (TC_ERR_CAPACITY) before anything is written or enqueued.
int batch_enqueue(const tc_request *reqs, int n, void *host_desc, void *dev_desc, hipStream_t s, size_t cap);

}  // namespace tc
