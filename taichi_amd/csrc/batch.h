// Internal interface between the SpMV kernels (spmv.hip) and the serving
// runtime (serve.hip): descriptor-level enqueue of the persistent batched
// kernel, so a caller can keep descriptors in its own pinned buffers.
#pragma once

#include <hip/hip_runtime.h>

#include "taichi_amd.h"

namespace tc {

// bytes of one work descriptor
size_t batch_desc_size();
// Fills n descriptors in host_desc (which must stay valid until the copy on
// stream s has run) and enqueues their copy to dev_desc plus one persistent
// launch covering all n requests.
int batch_enqueue(const tc_request *reqs, int n, void *host_desc, void *dev_desc, hipStream_t s);

}  // namespace tc
