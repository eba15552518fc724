// Shared pieces of libgtpeec: errors (thrown inside, turned into codes at the C ABI), device buffers, complex
// arithmetic for device code, launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gtpeec.h"

namespace gt {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error(code, msg); }

inline void hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    int code = (e == hipErrorOutOfMemory) ? GT_ERR_OOM : GT_ERR_DEVICE;
    fail(code, std::string(what) + ": " + hipGetErrorString(e));
}

#define GT_HIP(call) ::gt::hip_check((call), #call)
#define GT_LAUNCH(what) ::gt::hip_check(hipGetLastError(), what)

// A device allocation owned by its holder.
template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) GT_HIP(hipMalloc(&p, count * sizeof(T)));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // Set-up copies wait for completion: their host sources are often temporaries.
    void upload(const T* host, size_t count, hipStream_t s) {
        if (count > n) fail(GT_ERR_ARG, "upload larger than the buffer");
        if (count) {
            GT_HIP(hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, s));
            GT_HIP(hipStreamSynchronize(s));
        }
    }
    void upload(const std::vector<T>& v, hipStream_t s) {
        if (v.size() != n) alloc(v.size());
        upload(v.data(), v.size(), s);
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Page-locked host memory (hipHostMalloc): the small per-iteration copies of the solver go straight to and from it.
template <class T>
struct HostBuf {
    T* p = nullptr;
    size_t n = 0;
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    ~HostBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) GT_HIP(hipHostMalloc((void**)&p, count * sizeof(T), hipHostMallocDefault));
        for (size_t i = 0; i < count; ++i) p[i] = T();
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
    }
    T& operator[](size_t i) { return p[i]; }
    const T& operator[](size_t i) const { return p[i]; }
};

// Waits for the stream by polling: the solver's host steps sit between short device phases, where a sleeping wait
// adds its wake-up time to every step.
inline void spin_sync(hipStream_t s) {
    for (;;) {
        const hipError_t e = hipStreamQuery(s);
        if (e == hipSuccess) return;
        if (e != hipErrorNotReady) GT_HIP(e);
    }
}

// Interleaved complex number for device and host code.
template <class T>
struct cpx {
    T re, im;
    __host__ __device__ cpx() : re(0), im(0) {}
    __host__ __device__ cpx(T r, T i = 0) : re(r), im(i) {}
};

template <class T>
__host__ __device__ inline cpx<T> operator+(cpx<T> a, cpx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <class T>
__host__ __device__ inline cpx<T> operator-(cpx<T> a, cpx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <class T>
__host__ __device__ inline cpx<T> operator*(cpx<T> a, cpx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class T>
__host__ __device__ inline cpx<T> operator*(T s, cpx<T> a) { return {s * a.re, s * a.im}; }
template <class T>
__host__ __device__ inline cpx<T> conj_mul(cpx<T> a, cpx<T> b) {   // conj(a) * b
    return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
template <class T>
__host__ __device__ inline T norm2(cpx<T> a) { return a.re * a.re + a.im * a.im; }

typedef cpx<double> zd;
typedef cpx<float> zf;

inline unsigned blocks_for(size_t n, unsigned threads) { return (unsigned)((n + threads - 1) / threads); }

}  // namespace gt
