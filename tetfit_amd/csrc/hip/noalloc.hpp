// noalloc.hpp - tf_mlp_grad_noalloc: the preconditioner MLP's forward and
// backward pass (the kernels of tf_mlp_grad) on caller-owned device buffers,
// with every work buffer in one fixed device array, so that a call allocates
// no device memory (for callers under HIP stream capture and loops that must
// not allocate). Included once, at the end of tetfit_hip.hip.
#pragma once

#include <mutex>

namespace {

// Work buffers of one call, in `scratch`: the backward pass's partial
// gradient vectors (one per block: at most kNoAllocParts blocks over the full
// 16-row tiles, plus one for the last, partial tile), then that partial
// tile's features (16 x 16 bf16), gy and outputs (16 floats each), padded
// with zero rows. The size does not depend on the number of rows.
constexpr int kNoAllocParts = 128;
constexpr int64_t kNoAllocPartFloats = (int64_t)(kNoAllocParts + 1) * kMlpParams;
constexpr int64_t kNoAllocFloats = kNoAllocPartFloats + 16 * kMlpIn / 2 + 16 + 16;
// Largest number of feature rows one call takes (2^31 rows: 64 GiB of bf16
// features), so that tile counts stay far from any integer limit.
constexpr int64_t kNoAllocMaxRows = (int64_t)1 << 31;

// One copy per GPU, held until the process ends (about 2.7 MB). In host code
// the name `scratch` is not a device address: the host reaches each GPU's
// copy only through the address hipGetSymbolAddress returns with that GPU
// current (scratch_address).
static __device__ __align__(256) float scratch[kNoAllocFloats];

constexpr int kNoAllocDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kNoAllocDevices] = {};

// Device address of the current GPU's `scratch`, looked up on the first call
// per device and kept.
float* scratch_address() {
    int dev = 0;
    TF_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocDevices) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        TF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw std::runtime_error("hipGetSymbolAddress returned a null address");
        g_scratch_ptr[dev] = p;
    }
    return static_cast<float*>(g_scratch_ptr[dev]);
}

// Copies the last `rows` (1..15) feature rows and gy entries into the tile
// buffers and zero-fills the rest of the tile. One block of 256 threads:
// thread k stores feature k % 16 of tile row k / 16.
__global__ void __launch_bounds__(256) noalloc_tail_in_kernel(const short* __restrict__ x,
                                                              const float* __restrict__ gy, int rows,
                                                              short* __restrict__ xt, float* __restrict__ gyt) {
    const int k = threadIdx.x;
    const int r = k / kMlpIn;
    xt[k] = r < rows ? x[k] : (short)0;
    if (k < 16) gyt[k] = k < rows ? gy[k] : 0.f;
}

// The tile's first `rows` outputs to the caller's y.
__global__ void __launch_bounds__(64) noalloc_tail_out_kernel(const float* __restrict__ yt, int rows,
                                                              float* __restrict__ y) {
    const int k = threadIdx.x;
    if (k < rows) y[k] = yt[k];
}

static_assert(16 * kMlpIn == 256, "the tail kernel maps one thread per feature of a 16-row tile");

}  // namespace

extern "C" tf_status tf_mlp_grad_noalloc(const float* params, const uint16_t* x, const float* gy, int64_t n,
                                         float* y, float* grad, void* stream) {
    if (!params || !grad || n < 0 || (n && (!x || !gy || !y))) { g_err = "bad argument"; return TF_ERR_ARG; }
    if (n > kNoAllocMaxRows) {
        g_err = "tf_mlp_grad_noalloc: n = " + std::to_string(n) + " feature rows exceeds the maximum of " +
                std::to_string(kNoAllocMaxRows);
        return TF_ERR_ARG;
    }
    return guarded([&] {
        hipStream_t s = static_cast<hipStream_t>(stream);
        float* base = scratch_address();
        float* part = base;
        short* xt = reinterpret_cast<short*>(base + kNoAllocPartFloats);
        float* gyt = base + kNoAllocPartFloats + 16 * kMlpIn / 2;
        float* yt = gyt + 16;
        const short* xs = reinterpret_cast<const short*>(x);
        const int64_t full = n / 16;
        const int tail = (int)(n % 16);
        int nparts = 0;
        if (full > 0) {
            mlp_forward_kernel<<<(int)std::min<int64_t>(4096, (full + 3) / 4), kBlock, 0, s>>>(full, params, xs, y);
            nparts = (int)std::min<int64_t>(kNoAllocParts, (full + 3) / 4);
            mlp_backward_kernel<<<nparts, kBlock, 0, s>>>(full, params, xs, gy, part);
        }
        if (tail) {
            // every slot of the tile buffers is rewritten here, so nothing of
            // an earlier call's contents is read
            noalloc_tail_in_kernel<<<1, 16 * kMlpIn, 0, s>>>(xs + full * 16 * kMlpIn, gy + full * 16, tail, xt, gyt);
            mlp_forward_kernel<<<1, kBlock, 0, s>>>(1, params, xt, yt);
            mlp_backward_kernel<<<1, kBlock, 0, s>>>(1, params, xt, gyt, part + (int64_t)nparts * kMlpParams);
            noalloc_tail_out_kernel<<<1, 64, 0, s>>>(yt, tail, y + full * 16);
            ++nparts;
        }
        // grad = sum of the nparts partials written above (lr = 0: no update)
        mlp_adam_kernel<<<blocks_for(kMlpParams), kBlock, 0, s>>>(nparts, part, nullptr, nullptr, nullptr, grad, 0.f,
                                                                  0.9f, 0.999f, 1e-8f, 1, 1.f);
        TF_HIP(hipGetLastError());
    });
}
