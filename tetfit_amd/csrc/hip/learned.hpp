// learned.hpp - device side of the learned CG preconditioner: expansion of
// the MLP output onto A's SELL pattern, the training products over a bank of
// 64 (residual, correction) samples and the sample capture from CG.
//
// Sample bank layout: [row][64] fp32 (sample s of row e at e*64 + s), so that
// with one wave per matrix row, lane = sample, every gather of a column j is
// one 256-byte coalesced read. Quantities are in the Jacobi-scaled system
// Â = D^-1/2 A D^-1/2:  r~ = D^-1/2 r,  c~ = D^1/2 c,  Ĝ = D^1/2 G D^1/2.
//
// Training loss over the bank (weights w_s, zero for empty slots):
//   L = sum_s w_s e_s^T Â e_s,   e_s = Ĝ r~_s - c~_s = -(I - ĜÂ) c~_s,
// the Â-norm of the error left by one preconditioned Richardson step.
//   dL/dĜ_ej = 2 sum_s w_s (Â e_s)_e r~_s,j
#pragma once

#include "common.hpp"
#include "sell.hpp"

namespace tfd {

constexpr int kBank = 64;  // samples in the bank = lanes of a wave

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Training window: bank row e is matrix row w0 + e (e < nw); a column j
// enters as bank row j - w0 and only when that lies in [0, nw) (the
// principal submatrix of the window's rows).
//
// out[e][s] = sum_j M_ej in[j][s] (- sub[e][s]); M on A's SELL pattern with
// fp32 values mval (0 on padding). One wave per row, lane = sample.
__global__ void __launch_bounds__(kBlock) bank_spmm_kernel(int32_t nw, int32_t w0, const int64_t* __restrict__ sptr,
                                                           const int32_t* __restrict__ col,
                                                           const float* __restrict__ mval,
                                                           const float* __restrict__ in,
                                                           const float* __restrict__ sub, float* __restrict__ out) {
    const int lane = lane_id();
    for (int e = uniform(wave_id()); e < nw; e += wave_count()) {
        const int row = w0 + e;
        const int s = row / kSlice, r = row % kSlice;
        const int64_t bs = sptr[s];
        const int w = (int)((sptr[s + 1] - bs) / kSlice);
        float acc = 0.f;
        int k = 0;
        for (; k + 4 <= w; k += 4) {
            int32_t j[4];
            float v[4], x[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int64_t pos = bs + (int64_t)(k + u) * kSlice + r;
                j[u] = col[pos] - w0;
                v[u] = mval[pos];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) x[u] = (uint32_t)j[u] < (uint32_t)nw ? in[(int64_t)j[u] * kBank + lane] : 0.f;
#pragma unroll
            for (int u = 0; u < 4; ++u) acc += v[u] * x[u];
        }
        for (; k < w; ++k) {
            const int64_t pos = bs + (int64_t)k * kSlice + r;
            const int32_t j = col[pos] - w0;
            if ((uint32_t)j < (uint32_t)nw) acc += mval[pos] * in[(int64_t)j * kBank + lane];
        }
        const int64_t o = (int64_t)e * kBank + lane;
        if (sub) acc -= sub[o];
        out[o] = acc;
    }
}

// Per SELL position (e, j) of the window's rows: gpos = coef * sum_s w_s
// AE[e][s] R[j][s] (0 for a column outside the window); loss partial per
// block: sum_e sum_s w_s E[e][s] AE[e][s].
__global__ void __launch_bounds__(kBlock) bank_grad_kernel(int32_t nw, int32_t w0, const int64_t* __restrict__ sptr,
                                                           const int32_t* __restrict__ col,
                                                           const float* __restrict__ R, const float* __restrict__ E,
                                                           const float* __restrict__ AE,
                                                           const float* __restrict__ wts, float coef,
                                                           float* __restrict__ gpos, double* __restrict__ lpart) {
    __shared__ double red[kBlock / kWave];
    const int lane = lane_id();
    const float ws = wts[lane];
    double loss = 0;
    for (int e = uniform(wave_id()); e < nw; e += wave_count()) {
        const int row = w0 + e;
        const int s = row / kSlice, r = row % kSlice;
        const int64_t bs = sptr[s];
        const int w = (int)((sptr[s + 1] - bs) / kSlice);
        const int64_t o = (int64_t)e * kBank + lane;
        const float ae = AE[o] * ws;
        loss += (double)ae * (double)E[o];
        for (int k = 0; k < w; ++k) {
            const int64_t pos = bs + (int64_t)k * kSlice + r;
            const int32_t j = col[pos] - w0;
            float g = (uint32_t)j < (uint32_t)nw ? ae * R[(int64_t)j * kBank + lane] : 0.f;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) g += __shfl_xor(g, off, 64);
            if (lane == 0) gpos[pos] = coef * g;
        }
    }
    loss = block_sum(loss, red);
    if (threadIdx.x == 0) lpart[blockIdx.x] = loss;
}

// Pair gradient: the (e, j) and (j, e) positions of an off-diagonal pair
// share one MLP output; the pad tail of gy stays zero.
__global__ void pair_grad_kernel(int32_t n_pairs, const int64_t* __restrict__ pa, const int64_t* __restrict__ pb,
                                 const float* __restrict__ gpos, float* __restrict__ gy) {
    const int32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pairs) return;
    float g = gpos[pa[p]];
    if (pb[p] >= 0) g += gpos[pb[p]];
    gy[p] = g;
}

// fp32 -> bf16 bits, round to nearest even (finite inputs).
__device__ __forceinline__ uint16_t float_to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Ĝ (training, scaled, fp32) and G (CG, unscaled, bf16) values at every
// SELL position. The two positions of a pair get the same value, so G stays
// exactly symmetric after rounding.
// With c16 (A's 16-bit column offsets) the bf16 value goes to gpk packed
// as (c16 << 16 | bf16), otherwise to gval.
__global__ void expand_kernel(int64_t entries, const int32_t* __restrict__ pos2pair, const float* __restrict__ base,
                              const float* __restrict__ scale, const float* __restrict__ y, float* __restrict__ ghat,
                              uint16_t* __restrict__ gval, const uint16_t* __restrict__ c16,
                              uint32_t* __restrict__ gpk) {
    const int64_t pos = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= entries) return;
    const int32_t p = pos2pair[pos];
    const float g = p >= 0 ? base[p] + y[p] : 0.f;
    if (ghat) ghat[pos] = g;
    const uint16_t b = float_to_bf16(p >= 0 ? scale[p] * g : 0.f);
    if (gval) gval[pos] = b;
    if (gpk) gpk[pos] = ((uint32_t)c16[pos] << 16) | b;
}

// Bank column `slot` <- x (the CG iterate x_k being captured), the
// training window's rows (x points at its first row).
template <typename T>
__global__ void bank_capture_kernel(int32_t nw, int slot, const T* __restrict__ x, float* __restrict__ Cb) {
    const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nw) return;
    Cb[(int64_t)e * kBank + slot] = (float)x[e];
}

// After the solve: c~ = D^1/2 (x_final - x_k) (xf and isd point at the
// window's first row). The residual r~ = Â c~ is formed at training time.
template <typename T>
__global__ void bank_finalize_kernel(int32_t nw, int slot, const T* __restrict__ xf, const float* __restrict__ isd,
                                     float* __restrict__ Cb) {
    const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nw) return;
    const int64_t o = (int64_t)e * kBank + slot;
    Cb[o] = (float)(((double)xf[e] - (double)Cb[o]) / (double)isd[e]);
}

// Uniform(-sqrt3, sqrt3) probe entries (unit variance) from a counter hash,
// for the bank columns flagged in mask (bit s).
__device__ __forceinline__ float hash_unit(uint64_t k) {
    k ^= k >> 33; k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ULL;
    k ^= k >> 33;
    return (float)((double)(k >> 11) * (1.0 / 9007199254740992.0));
}
__global__ void bank_random_kernel(int32_t n, uint64_t mask, uint64_t seed, float* __restrict__ Cb) {
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= (int64_t)n * kBank) return;
    const int s = (int)(o % kBank);
    if (!((mask >> s) & 1)) return;
    Cb[o] = 1.7320508f * (2.f * hash_unit(seed * 0x9E3779B97F4A7C15ULL + (uint64_t)o) - 1.f);
}

// Rb[:, s] <- T[:, s] for the columns flagged in mask.
__global__ void bank_copy_cols_kernel(int32_t n, uint64_t mask, const float* __restrict__ src, float* __restrict__ Rb) {
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= (int64_t)n * kBank) return;
    if ((mask >> (o % kBank)) & 1) Rb[o] = src[o];
}

// Column dots: part[b][s] = sum over this block's rows of X[e][s] * Y[e][s].
__global__ void __launch_bounds__(kBlock) bank_coldot_kernel(int32_t n, const float* __restrict__ X,
                                                             const float* __restrict__ Y, double* __restrict__ part) {
    __shared__ double red[kBlock / kWave][kBank];
    const int lane = lane_id(), w = threadIdx.x / kWave;
    double acc = 0;
    for (int row = wave_id(); row < n; row += wave_count()) {
        const int64_t o = (int64_t)row * kBank + lane;
        acc += (double)X[o] * (double)Y[o];
    }
    red[w][lane] = acc;
    __syncthreads();
    if (w == 0) {
        double t = 0;
        for (int k = 0; k < kBlock / kWave; ++k) t += red[k][lane];
        part[(int64_t)blockIdx.x * kBank + lane] = t;
    }
}

// x[k] = U[free[k]]: the warm start of a solve, restored for a Jacobi redo.
template <typename T>
__global__ void gather_free_kernel(int32_t n, const int32_t* __restrict__ free_edge, const T* __restrict__ U,
                                   T* __restrict__ x) {
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) x[k] = U[free_edge[k]];
}

}  // namespace tfd
