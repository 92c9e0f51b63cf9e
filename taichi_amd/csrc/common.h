// Internal helpers shared by the host (g++) and device (hipcc) translation
// units of libtaichi_amd.so.
#pragma once

#include "taichi_amd.h"

#include <cstdarg>
#include <cstdint>
#include <functional>

namespace tc {

// Records a printf-style message as the calling thread's last error and
// returns `status` unchanged, so call sites read `return fail(...)`.
int fail(int status, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// Status for an exception caught at a C entry point (called from a catch
// handler): TC_ERR_NOMEM for std::bad_alloc, TC_ERR_INVALID otherwise, with
// the function name in the thread's last error. No exception crosses the C ABI.
int exception_status(const char *fn);

// Runs fn(begin, end) over [0, n) split into contiguous chunks on up to
// `max_threads` std::threads (0 = hardware concurrency, capped at 16); below
// `grain` items it runs on the calling thread.
void parallel_for(long long n, const std::function<void(long long, long long)> &fn,
                  int max_threads = 0, long long grain = 4096);

// Counter-based random numbers: value i of stream `stream` under `seed`.
// Independent of evaluation order, so generators give identical output on
// any thread count and any machine.
static inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline uint64_t rand_key(uint64_t seed, uint64_t stream)
{
    return mix64(mix64(seed) ^ (stream * 0xD1342543DE82EF95ull + 0x2545F4914F6CDD1Dull));
}
static inline uint64_t rand_at(uint64_t key, uint64_t i)
{
    return mix64(key + i * 0x9E3779B97F4A7C15ull);
}
// uniform integer in [0, n) from a 64-bit random value
static inline uint64_t bounded(uint64_t r, uint64_t n)
{
    return (uint64_t)(((unsigned __int128)r * n) >> 64);
}
// uniform float in [0, 1) with 24 random mantissa bits
static inline float unit_float(uint64_t r) { return (float)(r >> 40) * (1.0f / 16777216.0f); }

// Device layout constants shared by the upload code and the kernels.
constexpr int kGroupRows = 64;        // rows per offset group
constexpr int kDefaultLongRow = 64;   // rows longer than this get a wavefront

}  // namespace tc
