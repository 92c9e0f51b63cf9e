// Streaming calibration kernels for the benchmark's roofline figures: a
// device copy and a read-only pass at the highest rate plain HIP code reaches
// on this part. They set the practical ceilings that the SpMV kernels'
// measured traffic is compared with; they have no counterpart in TaiChi and
// nothing in the SpMV path calls them.
//
// Both are grid-stride loops over 16-byte vectors, 256 threads a workgroup;
// the copy writes through non-temporal stores (the written lines do not
// displace the source's lines in L2 and the MALL), which is what lifts a
// MALL-resident copy well above an ordinary store's rate. The caller picks the
// grid (workgroups); the benchmark sweeps a few per CU and keeps the fastest.
#include <hip/hip_runtime.h>

#include <stddef.h>
#include <stdint.h>

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void stream_copy_kernel(const f4 *__restrict__ src, f4 *__restrict__ dst,
                                                          size_t n4)
{
    const size_t step = (size_t)gridDim.x * 256;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += step)
        __builtin_nontemporal_store(src[i], dst + i);
}

// every thread leaves its partial sum in out[thread] (out holds grid x 256
// floats), so the loads cannot be dropped
__global__ __launch_bounds__(256) void stream_read_kernel(const f4 *__restrict__ src, float *__restrict__ out,
                                                          size_t n4)
{
    const size_t step = (size_t)gridDim.x * 256;
    const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
    f4 s = {0.f, 0.f, 0.f, 0.f};
    for (size_t i = t; i < n4; i += step) s += src[i];
    out[t] = s.x + s.y + s.z + s.w;
}

bool usable(const void *p, size_t bytes) { return p && bytes % 16 == 0 && (uintptr_t)p % 16 == 0; }

}  // namespace

extern "C" {

// Copies `bytes` (a multiple of 16, both pointers 16-byte aligned) from src
// to dst on `stream` with `grid` workgroups. Returns 0, or -1 for bad
// arguments, or the HIP error of the launch.
int tcs_copy(const void *src, void *dst, size_t bytes, int grid, hipStream_t stream)
{
    if (!usable(src, bytes) || !usable(dst, bytes) || grid < 1) return -1;
    stream_copy_kernel<<<grid, 256, 0, stream>>>((const f4 *)src, (f4 *)dst, bytes / 16);
    return (int)hipGetLastError();
}

// Reads `bytes` of src on `stream` with `grid` workgroups; out must hold
// out_len >= grid x 256 floats (the per-thread sums). Returns as tcs_copy.
int tcs_read(const void *src, size_t bytes, float *out, size_t out_len, int grid, hipStream_t stream)
{
    if (!usable(src, bytes) || !out || grid < 1 || out_len < (size_t)grid * 256) return -1;
    stream_read_kernel<<<grid, 256, 0, stream>>>((const f4 *)src, out, bytes / 16);
    return (int)hipGetLastError();
}

}  // extern "C"
