// mlp.hpp - the learned CG preconditioner: a 16 -> 64 -> 64 -> 1 MLP
// evaluated once per stored entry pair (e, j) of the CG system matrix A,
// with forward and backward passes on bf16 MFMA (fp32 accumulate).
//
// The network predicts the Jacobi-scaled entries of a symmetric sparse
// approximate inverse G with the pattern of A:
//   G_ej = (base_ej + y_ej) / sqrt(A_ee A_jj),
//   base = 1 on the diagonal, -A_ej/sqrt(A_ee A_jj) off it
// so that with y = 0 G is the first-order Neumann inverse D^-1/2 (2I - Â) D^-1/2.
// Features are symmetric in (e, j), hence G is symmetric.
//
// MFMA used: v_mfma_f32_16x16x16_bf16 (wave64). Operand layouts, lane l:
//   A (16x16, [i][k]):  i = l%16, k = 4*(l/16) + 0..3   (4 bf16)
//   B (16x16, [k][j]):  j = l%16, k = 4*(l/16) + 0..3   (4 bf16)
//   D (16x16, [i][j]):  i = 4*(l/16) + 0..3, j = l%16   (4 fp32)
// Activations are kept transposed (hidden x pairs) so that the D tile of one
// layer is directly the B operand of the next.
#pragma once

#include "common.hpp"

namespace tfd {

constexpr int kMlpIn = 16, kMlpHid = 64;
// parameter vector layout (fp32 master copy)
constexpr int kW1 = 0, kB1 = kW1 + kMlpHid * kMlpIn, kW2 = kB1 + kMlpHid, kB2 = kW2 + kMlpHid * kMlpHid,
              kW3 = kB2 + kMlpHid, kB3 = kW3 + kMlpHid, kMlpParams = kB3 + 1;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (short)(u >> 16);
}
__device__ __forceinline__ float bf2f(short b) { return __uint_as_float(((uint32_t)(uint16_t)b) << 16); }

__device__ __forceinline__ floatx4 mfma16(bf16x4 a, bf16x4 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Per-lane weight operands, loaded once per wave.
struct MlpRegs {
    bf16x4 w1[4];       // A of layer 1, block hb: W1[16hb + l%16][4(l/16)..]
    bf16x4 w2[4][4];    // A of layer 2, block (hb2, kb): W2[16hb2 + l%16][16kb + 4(l/16)..]
    float b1[4][4], b2[4][4], w3[4][4];  // [block][r] at hidden 16b + 4(l/16) + r
    float b3;
};

__device__ __forceinline__ void mlp_load(const float* __restrict__ P, MlpRegs& R) {
    const int l = threadIdx.x & 63, i = l & 15, q = l >> 4;
#pragma unroll
    for (int hb = 0; hb < 4; ++hb) {
        bf16x4 a;
        for (int e = 0; e < 4; ++e) a[e] = f2bf(P[kW1 + (16 * hb + i) * kMlpIn + 4 * q + e]);
        R.w1[hb] = a;
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) {
            bf16x4 w;
            for (int e = 0; e < 4; ++e) w[e] = f2bf(P[kW2 + (16 * hb + i) * kMlpHid + 16 * kb + 4 * q + e]);
            R.w2[hb][kb] = w;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            R.b1[hb][r] = P[kB1 + 16 * hb + 4 * q + r];
            R.b2[hb][r] = P[kB2 + 16 * hb + 4 * q + r];
            R.w3[hb][r] = P[kW3 + 16 * hb + 4 * q + r];
        }
    }
    R.b3 = P[kB3];
}

// Forward on a tile of 16 pairs. x: this lane's 4 features of pair l%16
// (features 4(l/16)..+3). Returns y for pair l%16 (same in the 4 lanes that
// share it) and keeps z1, z2 (pre-activations, D layout).
__device__ __forceinline__ float mlp_tile_forward(const MlpRegs& R, bf16x4 x, floatx4 z1[4], floatx4 z2[4]) {
    bf16x4 h1b[4];
#pragma unroll
    for (int hb = 0; hb < 4; ++hb) {
        floatx4 c = {R.b1[hb][0], R.b1[hb][1], R.b1[hb][2], R.b1[hb][3]};
        z1[hb] = mfma16(R.w1[hb], x, c);
        bf16x4 h;
        for (int r = 0; r < 4; ++r) h[r] = f2bf(fmaxf(z1[hb][r], 0.f));
        h1b[hb] = h;
    }
    float y = 0;
#pragma unroll
    for (int hb = 0; hb < 4; ++hb) {
        floatx4 c = {R.b2[hb][0], R.b2[hb][1], R.b2[hb][2], R.b2[hb][3]};
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) c = mfma16(R.w2[hb][kb], h1b[kb], c);
        z2[hb] = c;
        for (int r = 0; r < 4; ++r) y += R.w3[hb][r] * fmaxf(c[r], 0.f);
    }
    // sum over the 4 lanes that hold the same pair (l%16 equal)
    y += __shfl_xor(y, 16, 64);
    y += __shfl_xor(y, 32, 64);
    return y + R.b3;
}

// Features: 16 bf16 per pair, row-major [pair][16].
__device__ __forceinline__ bf16x4 load_x(const short* __restrict__ X, int64_t pair) {
    const int q = (threadIdx.x & 63) >> 4;
    const short* p = X + pair * kMlpIn + 4 * q;
    bf16x4 v = {p[0], p[1], p[2], p[3]};
    return v;
}

// y for every pair (n_pairs padded to a multiple of 16 by the caller).
__global__ void __launch_bounds__(kBlock) mlp_forward_kernel(int64_t n_tiles, const float* __restrict__ P,
                                                             const short* __restrict__ X, float* __restrict__ y) {
    MlpRegs R;
    mlp_load(P, R);
    const int l = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t t = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / 64; t < n_tiles; t += waves) {
        floatx4 z1[4], z2[4];
        const int64_t pair = t * 16 + (l & 15);
        const float v = mlp_tile_forward(R, load_x(X, pair), z1, z2);
        if (l < 16) y[pair] = v;
    }
}

// Backward: given gy = dL/dy per pair, accumulate dL/dparams. Each block
// writes one partial gradient vector (kMlpParams floats) to part[blockIdx].
// The tile loop is uniform over the block (tile = base + wave, idle waves run
// with x = 0 and gy = 0, which adds nothing), so the LDS staging can use block
// barriers; the block reduction is a fixed-order sum, so the result does not
// depend on scheduling.
__global__ void __launch_bounds__(kBlock) mlp_backward_kernel(int64_t n_tiles, const float* __restrict__ P,
                                                              const short* __restrict__ X, const float* __restrict__ gy,
                                                              float* __restrict__ part) {
    constexpr int kWaves = kBlock / 64;
    MlpRegs R;
    mlp_load(P, R);
    const int l = threadIdx.x & 63, i = l & 15, q = l >> 4, w = threadIdx.x >> 6;
    // W2^T operands: A[i][k] = W2^T[16ib + i][16kb + 4q + e] = W2[16kb + 4q + e][16ib + i]
    bf16x4 w2t[4][4];
    for (int ib = 0; ib < 4; ++ib)
        for (int kb = 0; kb < 4; ++kb) {
            bf16x4 a;
            for (int e = 0; e < 4; ++e) a[e] = f2bf(P[kW2 + (16 * kb + 4 * q + e) * kMlpHid + 16 * ib + i]);
            w2t[ib][kb] = a;
        }
    floatx4 dW2[4][4], dW1[4];
    float db1[4][4], db2[4][4], dw3[4][4], db3 = 0;
    for (int a = 0; a < 4; ++a) {
        dW1[a] = floatx4{0, 0, 0, 0};
        for (int b = 0; b < 4; ++b) { dW2[a][b] = floatx4{0, 0, 0, 0}; db1[a][b] = db2[a][b] = dw3[a][b] = 0; }
    }
    // per-wave LDS tiles: G2, H1 and G1 as [hidden][16 pairs] bf16, X as [16 pairs][16]
    __shared__ short sG2[kWaves][kMlpHid][16], sH1[kWaves][kMlpHid][16], sG1[kWaves][kMlpHid][16];
    __shared__ short sX[kWaves][16][16];
    for (int64_t base = (int64_t)blockIdx.x * kWaves; base < n_tiles; base += (int64_t)gridDim.x * kWaves) {
        const int64_t t = base + w;
        const bool live = t < n_tiles;
        const int64_t pair = t * 16 + i;
        bf16x4 x = {0, 0, 0, 0};
        if (live) x = load_x(X, pair);
        floatx4 z1[4], z2[4];
        (void)mlp_tile_forward(R, x, z1, z2);
        const float g = live ? gy[pair] : 0.f;
        db3 += (q == 0) ? g : 0.f;
        // layer 3 and g2 = dL/dz2 = g * w3 * (z2 > 0)
        bf16x4 g2b[4];
        for (int hb = 0; hb < 4; ++hb) {
            bf16x4 gb;
            for (int r = 0; r < 4; ++r) {
                const float h2 = fmaxf(z2[hb][r], 0.f);
                dw3[hb][r] += g * h2;
                const float g2 = z2[hb][r] > 0.f ? g * R.w3[hb][r] : 0.f;
                db2[hb][r] += g2;
                gb[r] = f2bf(g2);
                sG2[w][16 * hb + 4 * q + r][i] = gb[r];
                sH1[w][16 * hb + 4 * q + r][i] = f2bf(fmaxf(z1[hb][r], 0.f));
            }
            g2b[hb] = gb;
        }
        for (int e = 0; e < 4; ++e) sX[w][i][4 * q + e] = x[e];
        // g1 = (W2^T g2) * (z1 > 0)
        for (int ib = 0; ib < 4; ++ib) {
            floatx4 c = {0, 0, 0, 0};
            for (int kb = 0; kb < 4; ++kb) c = mfma16(w2t[ib][kb], g2b[kb], c);
            for (int r = 0; r < 4; ++r) {
                const float g1 = z1[ib][r] > 0.f ? c[r] : 0.f;
                db1[ib][r] += g1;
                sG1[w][16 * ib + 4 * q + r][i] = f2bf(g1);
            }
        }
        __syncthreads();
        // dW2 += G2 H1^T : A = G2[16ib + i][pairs 4q..], B[k=pair][j] = H1[16kb + i][pairs 4q..]
        for (int ib = 0; ib < 4; ++ib) {
            bf16x4 a = {sG2[w][16 * ib + i][4 * q], sG2[w][16 * ib + i][4 * q + 1], sG2[w][16 * ib + i][4 * q + 2],
                        sG2[w][16 * ib + i][4 * q + 3]};
            for (int kb = 0; kb < 4; ++kb) {
                bf16x4 b = {sH1[w][16 * kb + i][4 * q], sH1[w][16 * kb + i][4 * q + 1], sH1[w][16 * kb + i][4 * q + 2],
                            sH1[w][16 * kb + i][4 * q + 3]};
                dW2[ib][kb] = mfma16(a, b, dW2[ib][kb]);
            }
            // dW1 += G1 X : A = G1[16ib + i][pairs 4q..], B[k=pair][j=feature] = X[pair 4q + e][i]
            bf16x4 a1 = {sG1[w][16 * ib + i][4 * q], sG1[w][16 * ib + i][4 * q + 1], sG1[w][16 * ib + i][4 * q + 2],
                         sG1[w][16 * ib + i][4 * q + 3]};
            bf16x4 bx = {sX[w][4 * q][i], sX[w][4 * q + 1][i], sX[w][4 * q + 2][i], sX[w][4 * q + 3][i]};
            dW1[ib] = mfma16(a1, bx, dW1[ib]);
        }
        __syncthreads();
    }
    // bias / w3 partials: sum over the 16 lanes that share q (different pairs)
    for (int a = 0; a < 4; ++a)
        for (int r = 0; r < 4; ++r)
            for (int o = 1; o < 16; o <<= 1) {
                db1[a][r] += __shfl_xor(db1[a][r], o, 64);
                db2[a][r] += __shfl_xor(db2[a][r], o, 64);
                dw3[a][r] += __shfl_xor(dw3[a][r], o, 64);
            }
    for (int o = 1; o < 64; o <<= 1) db3 += __shfl_xor(db3, o, 64);
    // fixed-order block reduction: wave 0 stores, waves 1..3 add in turn
    __shared__ float acc[kMlpParams];
    for (int ww = 0; ww < kWaves; ++ww) {
        if (w == ww) {
            auto put = [&](int k, float v) { acc[k] = ww == 0 ? v : acc[k] + v; };
            for (int ib = 0; ib < 4; ++ib)
                for (int r = 0; r < 4; ++r) {
                    const int hid = 16 * ib + 4 * q + r;
                    // D tile [i = hid][j = l%16]
                    for (int kb = 0; kb < 4; ++kb) put(kW2 + hid * kMlpHid + 16 * kb + i, dW2[ib][kb][r]);
                    put(kW1 + hid * kMlpIn + i, dW1[ib][r]);
                    if (i == 0) {
                        put(kB1 + hid, db1[ib][r]);
                        put(kB2 + hid, db2[ib][r]);
                        put(kW3 + hid, dw3[ib][r]);
                    }
                }
            if (l == 0) put(kB3, db3);
        }
        __syncthreads();
    }
    for (int k = threadIdx.x; k < kMlpParams; k += blockDim.x) part[(int64_t)blockIdx.x * kMlpParams + k] = acc[k];
}

// grad[k] = sum over blocks of part; then Adam on the fp32 master parameters.
__global__ void mlp_adam_kernel(int nblocks, const float* __restrict__ part, float* __restrict__ P,
                                float* __restrict__ m, float* __restrict__ v, float* __restrict__ grad_out, float lr,
                                float b1, float b2, float eps, int step, float scale) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kMlpParams) return;
    float g = 0;
    for (int b = 0; b < nblocks; ++b) g += part[(int64_t)b * kMlpParams + k];
    g *= scale;
    if (grad_out) grad_out[k] = g;
    if (lr == 0.f) return;
    const float mk = b1 * m[k] + (1 - b1) * g;
    const float vk = b2 * v[k] + (1 - b2) * g * g;
    m[k] = mk;
    v[k] = vk;
    const float mh = mk / (1 - powf(b1, (float)step)), vh = vk / (1 - powf(b2, (float)step));
    P[k] -= lr * mh / (sqrtf(vh) + eps);
}

}  // namespace tfd
