// common.hpp - HIP error handling, device buffers and wave64 helpers shared
// by the stepping kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace tfd {

struct HipError : std::runtime_error {
    int code;
    HipError(const std::string& s, int c) : std::runtime_error(s), code(c) {}
};

#define TF_HIP(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw ::tfd::HipError(std::string(#call) + ": " + hipGetErrorString(e_),        \
                                  e_ == hipErrorOutOfMemory ? 3 : 2);                      \
    } while (0)

// Owning device allocation.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t n) {
        release();
        if (n == 0) return;
        TF_HIP(hipMalloc(&p, n));
        bytes = n;
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kXcd = 8;  // MI355X: 8 XCDs, each with its own L2

// The hardware hands workgroup b to XCD b % 8. This maps the workgroup to a
// logical block index such that XCD x holds one contiguous range of logical
// blocks: XCD x < r gets [x(q+1), (x+1)(q+1)), XCD x >= r gets
// [xq + r, (x+1)q + r), with q = grid / 8, r = grid % 8 (a bijection on
// [0, grid)). Kernels that gather neighbours index rows by the logical block,
// so each L2 caches the part of the gathered vector near its own rows.
__device__ __forceinline__ int xcd_block_of(int b, int g) {
    const int x = b % kXcd, i = b / kXcd;
    const int q = g / kXcd, r = g % kXcd;
    return x * q + (x < r ? x : r) + i;
}
__device__ __forceinline__ int xcd_block() { return xcd_block_of((int)blockIdx.x, (int)gridDim.x); }

inline int blocks_for(int64_t n, int per_block = kBlock) {
    int64_t b = (n + per_block - 1) / per_block;
    return (int)(b < 1 ? 1 : b);
}

// Sum over `width` consecutive lanes (width a power of two <= 64).
template <int W>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
    for (int o = W / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, W);
    return v;
}

// Block-wide fp64 sum of one value per thread; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* smem) {
    v = group_sum<kWave>(v);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    if (lane == 0) smem[w] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + kWave - 1) / kWave;
        for (int i = 0; i < nw; ++i) r += smem[i];
    }
    __syncthreads();
    return r;
}

}  // namespace tfd
