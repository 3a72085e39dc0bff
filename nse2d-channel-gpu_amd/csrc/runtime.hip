// Runtime plumbing of the library: last-error string, per-device scratch, memory helpers.
#include "nse2d_common.hpp"
#include "../../include/nse2d_hip.h"

#include <mutex>
#include <string>
#include <vector>

namespace nse2d {

namespace {
thread_local std::string g_last_error;
std::mutex g_scratch_mu;
std::vector<std::pair<double*, size_t>> g_scratch;  // per device: (pointer, capacity in doubles)
}  // namespace

void set_last_error(const std::string& msg) { g_last_error = msg; }

double* scratch_doubles(size_t n) {
    int dev = 0;
    NSE2D_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((size_t)dev >= g_scratch.size()) g_scratch.resize((size_t)dev + 1, {nullptr, 0});
    auto& e = g_scratch[(size_t)dev];
    if (e.second < n) {
        // hipFree waits for the device, so a buffer still read by queued kernels is not freed early.
        if (e.first) NSE2D_HIP_CHECK(hipFree(e.first));
        size_t cap = 4096;
        while (cap < n) cap *= 2;
        NSE2D_HIP_CHECK(hipMalloc(&e.first, cap * sizeof(double)));
        e.second = cap;
    }
    return e.first;
}

}  // namespace nse2d

extern "C" {

void* nse2d_dev_alloc(size_t bytes) {
    void* p = nullptr;
    NSE2D_HIP_CHECK(hipMalloc(&p, bytes ? bytes : 8));
    return p;
}
void nse2d_dev_free(void* p) {
    if (p) NSE2D_HIP_CHECK(hipFree(p));
}
void nse2d_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    NSE2D_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}
void nse2d_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    NSE2D_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}
void nse2d_memcpy_d2d(void* dst, const void* src, size_t bytes) {
    NSE2D_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
}
void nse2d_device_sync(void) { NSE2D_HIP_CHECK(hipDeviceSynchronize()); }
hipStream_t nse2d_stream_create(void) {
    hipStream_t s = nullptr;
    NSE2D_HIP_CHECK(hipStreamCreate(&s));  // default priority
    return s;
}
void nse2d_stream_destroy(hipStream_t s) {
    if (s) NSE2D_HIP_CHECK(hipStreamDestroy(s));
}
void nse2d_stream_sync(hipStream_t s) { NSE2D_HIP_CHECK(hipStreamSynchronize(s)); }
int nse2d_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}
const char* nse2d_last_error(void) { return nse2d::g_last_error.c_str(); }
const char* nse2d_version(void) { return "nse2d-hip 0.1 gfx950"; }

}  // extern "C"
