// common.hpp -- shared host/device plumbing of liboagpu (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace oagpu {

// Thrown inside the library, caught at the C ABI and turned into a status
// code plus oagpu_last_error() text (the reference throws std::runtime_error
// from errorMsg(), src/basic/macros.cpp:21-24; nothing crosses the ABI).
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define OAGPU_HIP(call)                                                         \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess)                                                   \
            throw ::oagpu::Error(3, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define OAGPU_REQUIRE(cond, msg)                                                \
    do {                                                                        \
        if (!(cond)) throw ::oagpu::Error(2, msg);                              \
    } while (0)

// Device buffer owning a hipMalloc'd allocation.
template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) OAGPU_HIP(hipMalloc(&p, count * sizeof(T)));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T* h, size_t count, hipStream_t s = 0) {
        OAGPU_REQUIRE(count <= n, "upload larger than buffer");
        if (count) OAGPU_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& v, hipStream_t s = 0) {
        if (v.size() != n) alloc(v.size());
        upload(v.data(), v.size(), s);
    }
    void download(T* h, size_t count, hipStream_t s = 0) const {
        OAGPU_REQUIRE(count <= n, "download larger than buffer");
        if (count) OAGPU_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    void zero(hipStream_t s = 0) {
        if (n) OAGPU_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s));
    }
};

constexpr int kNumXcd = 8;  // MI355X: 8 XCDs, workgroups dispatched round-robin

// Grid rounded up to a multiple of the XCD count so that xcd_block() is a
// bijection; blocks beyond the work simply exit.
inline unsigned xcd_grid(size_t work, unsigned block) {
    size_t g = (work + block - 1) / block;
    g = (g + kNumXcd - 1) / kNumXcd * kNumXcd;
    return static_cast<unsigned>(g ? g : kNumXcd);
}

// Maps the hardware workgroup id (dispatched round-robin over the XCDs) to a
// logical id such that each XCD works on one contiguous range of rows: the
// element/node data a row gathers then stays in that XCD's L2.
__device__ inline unsigned xcd_block() {
    const unsigned nb = gridDim.x;
    const unsigned per = nb / kNumXcd;
    const unsigned b = blockIdx.x;
    return (b % kNumXcd) * per + b / kNumXcd;
}

}  // namespace oagpu

#include <utility>

namespace oagpu {

// Compile-time loop: f(std::integral_constant<int, I>) for I = 0..N-1, so
// that indices into per-lane arrays are constants (registers, not scratch).
template <typename F, int... Is>
__host__ __device__ __forceinline__ void sfor_impl(F&& f, std::integer_sequence<int, Is...>) {
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F>
__host__ __device__ __forceinline__ void sfor(F&& f) {
    sfor_impl(f, std::make_integer_sequence<int, N>{});
}

}  // namespace oagpu
