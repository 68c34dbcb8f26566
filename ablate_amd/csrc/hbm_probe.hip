// Achievable-bandwidth reference for the roofline: a streaming copy and a streaming read of device buffers with
// one 16-byte load per lane per step (the access shape that reaches the highest HBM rate on gfx950 in our
// measurements), timed with HIP events.  It has no counterpart in the reference; bench.py reports it beside the
// RHS so that the RHS's measured traffic can be compared with what the memory system delivers.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "fv_internal.hpp"
#include "fv_layout.hpp"

namespace abl {
namespace {

typedef double d2 __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(256) void k_probe_copy(const d2* __restrict__ a, d2* __restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// Sums its slice; the sum is stored only when it equals `never` (a value the zero-filled input cannot produce),
// which keeps the loads without a write stream.
__global__ __launch_bounds__(256) void k_probe_read(const d2* __restrict__ a, int64_t n, double never,
                                                    double* __restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const d2 v = a[i];
    const double s = v.x + v.y;
    if (s == never) out[0] = s;
}

}  // namespace
}  // namespace abl

extern "C" int abl_hbm_probe(int32_t kind, int64_t bytes, int32_t reps, double* gb_per_s, double* ms) {
    if (!gb_per_s || !ms || (kind != 0 && kind != 1) || bytes < (1 << 20) || reps < 1) {
        abl::set_global_error("abl_hbm_probe: kind must be 0 (copy) or 1 (read), bytes >= 1 MiB, reps >= 1");
        return 1;
    }
    const int64_t n = bytes / 16;
    void *a = nullptr, *b = nullptr, *o = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto done = [&](int rc, const char* what) {
        if (a) (void)hipFree(a);
        if (b) (void)hipFree(b);
        if (o) (void)hipFree(o);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (rc) abl::set_global_error(std::string("abl_hbm_probe: ") + what);
        return rc;
    };
    if (hipMalloc(&a, (size_t)n * 16) != hipSuccess) return done(1, "out of device memory");
    if (kind == 0 && hipMalloc(&b, (size_t)n * 16) != hipSuccess) return done(1, "out of device memory");
    if (hipMalloc(&o, 64) != hipSuccess) return done(1, "out of device memory");
    if (hipMemset(a, 0, (size_t)n * 16) != hipSuccess) return done(1, "hipMemset failed");
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return done(1, "hipEventCreate failed");
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    auto launch = [&] {
        if (kind == 0)
            hipLaunchKernelGGL(abl::k_probe_copy, grid, block, 0, 0, (const abl::d2*)a, (abl::d2*)b, n);
        else
            hipLaunchKernelGGL(abl::k_probe_read, grid, block, 0, 0, (const abl::d2*)a, n, 1.2345, (double*)o);
    };
    for (int w = 0; w < 2; w++) launch();
    if (hipEventRecord(e0, 0) != hipSuccess) return done(1, "hipEventRecord failed");
    for (int r = 0; r < reps; r++) launch();
    if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
        return done(1, "the probe kernels failed");
    if (hipGetLastError() != hipSuccess) return done(1, "launch failed");
    float t = 0.f;
    if (hipEventElapsedTime(&t, e0, e1) != hipSuccess) return done(1, "hipEventElapsedTime failed");
    *ms = (double)t / reps;
    const double moved = (double)n * 16 * (kind == 0 ? 2.0 : 1.0);
    *gb_per_s = moved / (*ms * 1e-3) / 1e9;
    return done(0, "");
}
