// sell.hpp - SELL-64 sparse layout and the kernels that stream it.
//
// Rows are grouped in slices of 64 (one wave64 lane per row); a slice of
// width w (its longest row) stores entry k of its rows at
// sptr[s] + 64*k + lane, so every load of values, column indices and signs
// is a fully coalesced 64-lane access. Short rows are padded with value 0
// and their own last column (a valid index). Within a row the entries keep
// the CSR column order, so the sums run in the same order as a sorted-CSR
// sequential product.
#pragma once

#include "common.hpp"
#include "kernels.hpp"

namespace tfd {

constexpr int kSlice = 64;

// Sum over the slice row of val*gather(col). Loads are issued in batches of
// 8 entries (16 independent loads, then 8 independent gathers) so that each
// lane keeps enough memory traffic in flight; the tail batch is predicated.
// Accumulation stays in column order.
template <typename T, class G>
__device__ __forceinline__ double sell_dot(int64_t base, int w, const int32_t* __restrict__ col,
                                           const T* __restrict__ val, int lane, G gather) {
    constexpr int U = 8;
    const int32_t* c = col + base + lane;
    const T* v = val + base + lane;
    double acc = 0;
    for (int k0 = 0; k0 < w; k0 += U) {
        int32_t ci[U];
        double vi[U], gi[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kk = min(k0 + u, w - 1);  // clamped: every load stays inside the slice
            ci[u] = c[kk * 64];
            vi[u] = (double)v[kk * 64];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) gi[u] = gather(ci[u]);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (k0 + u < w) acc += vi[u] * gi[u];
    }
    return acc;
}

// Same with int8 signs instead of values (incidence matrices).
template <typename T>
__device__ __forceinline__ double sell_sgn_dot(int64_t base, int w, const int32_t* __restrict__ col,
                                               const int8_t* __restrict__ sg, int lane, const T* __restrict__ x) {
    constexpr int U = 8;
    const int32_t* c = col + base + lane;
    const int8_t* s = sg + base + lane;
    double acc = 0;
    for (int k0 = 0; k0 < w; k0 += U) {
        int32_t ci[U];
        double si[U], gi[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kk = min(k0 + u, w - 1);
            ci[u] = c[kk * 64];
            si[u] = (double)s[kk * 64];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) gi[u] = (double)x[ci[u]];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (k0 + u < w) acc += si[u] * gi[u];
    }
    return acc;
}

__device__ __forceinline__ int wave_id() { return (blockIdx.x * blockDim.x + threadIdx.x) / kWave; }
__device__ __forceinline__ int wave_count() { return gridDim.x * blockDim.x / kWave; }
__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// curl_f = Ct F over edges (dga, D:878-879).
template <typename T>
__global__ void __launch_bounds__(kBlock) sell_ct_kernel(int32_t nrows, int32_t nslices, const int64_t* __restrict__ sptr,
                                                         const int32_t* __restrict__ col, const int8_t* __restrict__ sg,
                                                         const T* __restrict__ x, T* __restrict__ y) {
    const int lane = lane_id();
    for (int s = wave_id(); s < nslices; s += wave_count()) {
        const int64_t b = sptr[s];
        const int w = (int)((sptr[s + 1] - b) / kSlice);
        const double v = sell_sgn_dot<T>(b, w, col, sg, lane, x);
        const int32_t r = s * kSlice + lane;
        if (r < nrows) y[r] = (T)v;
    }
}

// U[edge(r)] += dt * (H curl_f)[r] (dga, D:889).
template <typename T>
__global__ void __launch_bounds__(kBlock) sell_h_kernel(int32_t nrows, int32_t nslices, const int64_t* __restrict__ sptr,
                                                        const int32_t* __restrict__ col, const T* __restrict__ val,
                                                        const int32_t* __restrict__ row_edge, const T* __restrict__ x,
                                                        T* __restrict__ U, double dt) {
    const int lane = lane_id();
    for (int s = wave_id(); s < nslices; s += wave_count()) {
        const int64_t b = sptr[s];
        const int w = (int)((sptr[s + 1] - b) / kSlice);
        const double v = sell_dot<T>(b, w, col, val, lane, [&](int32_t j) { return (double)x[j]; });
        const int32_t r = s * kSlice + lane;
        if (r < nrows) {
            const int32_t e = row_edge[r];
            U[e] = (T)((double)U[e] + dt * v);
        }
    }
}

// mode 0: F -= dt * N x (dga, D:917); mode 1: F = N x (gfdtd, D:1468).
template <typename T>
__global__ void __launch_bounds__(kBlock) sell_n_kernel(int32_t nrows, int32_t nslices, const int64_t* __restrict__ sptr,
                                                        const int32_t* __restrict__ col, const T* __restrict__ val,
                                                        const T* __restrict__ x, T* __restrict__ Fv, double dt, int mode) {
    const int lane = lane_id();
    for (int s = wave_id(); s < nslices; s += wave_count()) {
        const int64_t b = sptr[s];
        const int w = (int)((sptr[s + 1] - b) / kSlice);
        const double v = sell_dot<T>(b, w, col, val, lane, [&](int32_t j) { return (double)x[j]; });
        const int32_t r = s * kSlice + lane;
        if (r < nrows) Fv[r] = mode == 0 ? (T)((double)Fv[r] - dt * v) : (T)v;
    }
}

// rhs[k] = E U_old - RHS1 SrcFld + dt (Ct F + fb) at free edge e = free[k]
// (D:1516-1523; sigma and I are zero on the supported inputs). The three
// operators are stored as SELL-64 over the free rows; SrcFld is U on
// Dirichlet edges at the new time, held in `dir_new` (edge-indexed).
template <typename T>
__global__ void __launch_bounds__(kBlock) gf_rhs_kernel(
    int32_t n_free, int32_t nslices, const int32_t* __restrict__ free_edge, const int64_t* __restrict__ e_ptr,
    const int32_t* __restrict__ e_col, const T* __restrict__ e_val, const int64_t* __restrict__ r_ptr,
    const int32_t* __restrict__ r_col, const T* __restrict__ r_val, const int64_t* __restrict__ c_ptr,
    const int32_t* __restrict__ c_col, const int8_t* __restrict__ c_sg, const T* __restrict__ U,
    const T* __restrict__ dir_new, const T* __restrict__ Fv, const T* __restrict__ fb, T* __restrict__ rhs,
    double dt) {
    const int lane = lane_id();
    for (int s = wave_id(); s < nslices; s += wave_count()) {
        const int64_t eb = e_ptr[s], rb = r_ptr[s], cb = c_ptr[s];
        const int ew = (int)((e_ptr[s + 1] - eb) / kSlice), rw = (int)((r_ptr[s + 1] - rb) / kSlice);
        const int cw = (int)((c_ptr[s + 1] - cb) / kSlice);
        const double eu = sell_dot<T>(eb, ew, e_col, e_val, lane, [&](int32_t j) { return (double)U[j]; });
        const double rs = rw ? sell_dot<T>(rb, rw, r_col, r_val, lane, [&](int32_t j) { return (double)dir_new[j]; }) : 0.0;
        const double cf = sell_sgn_dot<T>(cb, cw, c_col, c_sg, lane, Fv);
        const int32_t k = s * kSlice + lane;
        if (k < n_free) rhs[k] = (T)((eu - rs) + dt * cf + dt * (double)fb[free_edge[k]]);
    }
}

// ---- CG (ConjugateGradientSolver.hpp:52-103) ------------------------------
// r = b - A x, z = pre r, partials of r.r and r.z into part_b[0..nb), [nb..2nb)
// With jacobi = false z is left to cg_c_kernel (z = G r) and the r.z partials are 0.
template <typename T>
__global__ void __launch_bounds__(kBlock) cg_init_kernel(int32_t n, int32_t nslices, const int64_t* __restrict__ sptr,
                                                         const int32_t* __restrict__ col, const T* __restrict__ val,
                                                         const T* __restrict__ pre, const T* __restrict__ b,
                                                         const T* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
                                                         double* __restrict__ part_b, CgState* st, int jacobi) {
    __shared__ double red[kBlock / kWave];
    const int lane = lane_id();
    double rr = 0, rz = 0;
    for (int s = wave_id(); s < nslices; s += wave_count()) {
        const int64_t bs = sptr[s];
        const int w = (int)((sptr[s + 1] - bs) / kSlice);
        const double ax = sell_dot<T>(bs, w, col, val, lane, [&](int32_t j) { return (double)x[j]; });
        const int32_t i = s * kSlice + lane;
        if (i < n) {
            const T ri = (T)((double)b[i] - ax);
            r[i] = ri;
            rr += (double)ri * (double)ri;
            if (jacobi) {
                const T zi = (T)((double)pre[i] * (double)ri);
                z[i] = zi;
                rz += (double)ri * (double)zi;
            }
        }
    }
    rr = block_sum(rr, red);
    rz = block_sum(rz, red);
    if (threadIdx.x == 0) {
        part_b[blockIdx.x] = rr;
        part_b[gridDim.x + blockIdx.x] = rz;
        if (blockIdx.x == 0) { st->done = 0; st->iter = 0; }
    }
}

// Iteration `it`, first half: stop test on the residual of the previous
// half (ratio = ||r||/||r0||, while (ratio > tol && it < maxiter)), then
// p_new = z + beta p_old (p = z at it 0) fused into the gather of
// q = A p_new, partial p.q into part_a.
template <typename T>
__global__ void __launch_bounds__(kBlock) cg_a_kernel(int32_t it, int32_t n, int32_t nslices,
                                                      const int64_t* __restrict__ sptr, const int32_t* __restrict__ col,
                                                      const T* __restrict__ val, const T* __restrict__ z,
                                                      T* __restrict__ p0, T* __restrict__ p1, T* __restrict__ q,
                                                      const double* __restrict__ part_rr, int32_t nb_rr,
                                                      const double* __restrict__ part_rz, int32_t nb_rz,
                                                      double* __restrict__ part_a, CgState* st) {
    if (st->done) return;
    __shared__ double red[kBlock / kWave];
    const double RR = sum_partials(part_rr, nb_rr, red);
    const double RZ = sum_partials(part_rz, nb_rz, red);
    const double rr0 = it == 0 ? RR : st->rr0;
    const double ratio = std::sqrt(RR) / std::sqrt(rr0);
    const bool stop = !(ratio > st->tol && it < st->maxiter);
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (it == 0) st->rr0 = RR;
        st->rho[it & 1] = RZ;
        st->relres = ratio;
        if (stop) { st->done = 1; st->iter = it; }
    }
    if (stop) return;
    const double beta = it == 0 ? 0.0 : RZ / st->rho[(it - 1) & 1];
    const T* pold = (it & 1) ? p1 : p0;
    T* pnew = (it & 1) ? p0 : p1;
    const int lane = lane_id();
    double pq = 0;
    for (int s = wave_id(); s < nslices; s += wave_count()) {
        const int64_t bs = sptr[s];
        const int w = (int)((sptr[s + 1] - bs) / kSlice);
        double qv;
        if (it == 0) qv = sell_dot<T>(bs, w, col, val, lane, [&](int32_t j) { return (double)z[j]; });
        else
            qv = sell_dot<T>(bs, w, col, val, lane,
                             [&](int32_t j) { return (double)(T)((double)z[j] + beta * (double)pold[j]); });
        const int32_t i = s * kSlice + lane;
        if (i < n) {
            const T pi = it == 0 ? z[i] : (T)((double)z[i] + beta * (double)pold[i]);
            pnew[i] = pi;
            const T qi = (T)qv;
            q[i] = qi;
            pq += (double)pi * (double)qi;
        }
    }
    pq = block_sum(pq, red);
    if (threadIdx.x == 0) part_a[blockIdx.x] = pq;
}

// Iteration `it`, second half: alpha = rho/(p.q); x += alpha p; r -= alpha q;
// z = pre r (the next iteration's z); partials of r.r and r.z into part_b.
template <typename T>
__global__ void __launch_bounds__(kBlock) cg_b_kernel(int32_t it, int32_t n, const T* __restrict__ pre,
                                                      T* __restrict__ x, T* __restrict__ r, T* __restrict__ z,
                                                      const T* __restrict__ p0, const T* __restrict__ p1,
                                                      const T* __restrict__ q, const double* __restrict__ part_a,
                                                      int32_t nb_a, double* __restrict__ part_b, CgState* st,
                                                      int jacobi) {
    if (st->done) return;
    __shared__ double red[kBlock / kWave];
    const double PQ = sum_partials(part_a, nb_a, red);
    const double alpha = st->rho[it & 1] / PQ;
    const T* pnew = (it & 1) ? p0 : p1;
    double rr = 0, rz = 0;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        x[i] = (T)((double)x[i] + alpha * (double)pnew[i]);
        const T ri = (T)((double)r[i] - alpha * (double)q[i]);
        r[i] = ri;
        rr += (double)ri * (double)ri;
        if (jacobi) {
            const T zi = (T)((double)pre[i] * (double)ri);
            z[i] = zi;
            rz += (double)ri * (double)zi;
        }
    }
    rr = block_sum(rr, red);
    rz = block_sum(rz, red);
    if (threadIdx.x == 0) {
        part_b[blockIdx.x] = rr;
        part_b[gridDim.x + blockIdx.x] = rz;
    }
}

// z = G r with G the learned approximate inverse stored on A's SELL pattern
// (same column array, its own values in fp32), partial r.z into part_c.
template <typename T>
__global__ void __launch_bounds__(kBlock) cg_c_kernel(int32_t n, int32_t nslices, const int64_t* __restrict__ sptr,
                                                      const int32_t* __restrict__ col, const float* __restrict__ gval,
                                                      const T* __restrict__ r, T* __restrict__ z,
                                                      double* __restrict__ part_c, const CgState* st) {
    if (st->done) return;
    __shared__ double red[kBlock / kWave];
    const int lane = lane_id();
    double rz = 0;
    for (int s = wave_id(); s < nslices; s += wave_count()) {
        const int64_t bs = sptr[s];
        const int w = (int)((sptr[s + 1] - bs) / kSlice);
        const double zv = sell_dot<float>(bs, w, col, gval, lane, [&](int32_t j) { return (double)r[j]; });
        const int32_t i = s * kSlice + lane;
        if (i < n) {
            const T zi = (T)zv;
            z[i] = zi;
            rz += (double)r[i] * (double)zi;
        }
    }
    rz = block_sum(rz, red);
    if (threadIdx.x == 0) part_c[blockIdx.x] = rz;
}

}  // namespace tfd
