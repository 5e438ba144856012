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

// Sum over the slice row of val*gather(col), accumulated in A (the field
// type: fp64 on the plain path, fp32 on the fast path). Loads are issued in
// batches of U entries (2U independent loads, then U independent gathers)
// so that each lane keeps enough memory traffic in flight; the tail batch is
// clamped to the row's last entry and predicated. Accumulation stays in
// column order, so the result does not depend on U or on the launch shape.
template <typename A, int U, typename T, class G>
__device__ __forceinline__ double sell_dot(int64_t base, int w, const int32_t* __restrict__ col,
                                           const T* __restrict__ val, int lane, G gather) {
    const int32_t* c = col + base + lane;
    const T* v = val + base + lane;
    A acc = 0;
    for (int k0 = 0; k0 < w; k0 += U) {
        int32_t ci[U];
        A vi[U], gi[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kk = min(k0 + u, w - 1);  // clamped: every load stays inside the slice
            ci[u] = c[kk * 64];
            vi[u] = (A)v[kk * 64];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) gi[u] = (A)gather(ci[u]);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (k0 + u < w) acc += vi[u] * gi[u];
    }
    return (double)acc;
}

// Same with int8 signs instead of values (incidence matrices).
template <typename T, int U>
__device__ __forceinline__ double sell_sgn_dot(int64_t base, int w, const int32_t* __restrict__ col,
                                               const int8_t* __restrict__ sg, int lane, const T* __restrict__ x) {
    const int32_t* c = col + base + lane;
    const int8_t* s = sg + base + lane;
    T acc = 0;
    for (int k0 = 0; k0 < w; k0 += U) {
        int32_t ci[U];
        T si[U], gi[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kk = min(k0 + u, w - 1);
            ci[u] = c[kk * 64];
            si[u] = (T)s[kk * 64];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) gi[u] = x[ci[u]];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (k0 + u < w) acc += si[u] * gi[u];
    }
    return (double)acc;
}

__device__ __forceinline__ int wave_id() { return (blockIdx.x * blockDim.x + threadIdx.x) / kWave; }
__device__ __forceinline__ int wave_count() { return gridDim.x * blockDim.x / kWave; }
__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// Slices [s0, s1) of this workgroup: the logical blocks (xcd_block) split
// the slices into contiguous chunks, so each XCD streams one contiguous
// part of the matrix and its L2 holds the part of the gathered vector that
// those rows touch. The waves of a block take the chunk's slices in turn.
struct SliceRange {
    int s, end, step;
};
__device__ __forceinline__ SliceRange block_slices(int nslices) {
    const int lb = xcd_block();
    const int s0 = (int)((int64_t)nslices * lb / gridDim.x);
    const int s1 = (int)((int64_t)nslices * (lb + 1) / gridDim.x);
    return SliceRange{s0 + (int)(threadIdx.x / kWave), s1, (int)(blockDim.x / kWave)};
}

// curl_f = Ct F over edges (dga, D:878-879).
template <typename T, int NB>
__global__ void __launch_bounds__(kBlock) sell_ct_kernel(int32_t nrows, int32_t nslices, const int64_t* __restrict__ sptr,
                                                         const int32_t* __restrict__ col, const int8_t* __restrict__ sg,
                                                         const T* __restrict__ x, T* __restrict__ y) {
    const int lane = lane_id();
    for (SliceRange sr = block_slices(nslices); sr.s < sr.end; sr.s += sr.step) {
        const int s = sr.s;
        const int64_t b = sptr[s];
        const int w = (int)((sptr[s + 1] - b) / kSlice);
        const double v = sell_sgn_dot<T, NB>(b, w, col, sg, lane, x);
        const int32_t r = s * kSlice + lane;
        if (r < nrows) y[r] = (T)v;
    }
}

// U[edge(r)] += dt * (H curl_f)[r] (dga, D:889); row_dt (batches) gives
// each row its own simulation's dt.
template <typename T, int NB>
__global__ void __launch_bounds__(kBlock) sell_h_kernel(int32_t nrows, int32_t nslices, const int64_t* __restrict__ sptr,
                                                        const int32_t* __restrict__ col, const T* __restrict__ val,
                                                        const int32_t* __restrict__ row_edge, const T* __restrict__ x,
                                                        T* __restrict__ U, double dt,
                                                        const double* __restrict__ row_dt) {
    const int lane = lane_id();
    for (SliceRange sr = block_slices(nslices); sr.s < sr.end; sr.s += sr.step) {
        const int s = sr.s;
        const int64_t b = sptr[s];
        const int w = (int)((sptr[s + 1] - b) / kSlice);
        const double v = sell_dot<T, NB>(b, w, col, val, lane, [&](int32_t j) { return x[j]; });
        const int32_t r = s * kSlice + lane;
        if (r < nrows) {
            const int32_t e = row_edge[r];
            U[e] = (T)((double)U[e] + (row_dt ? row_dt[r] : dt) * v);
        }
    }
}

// mode 0: F -= dt * N x (dga, D:917); mode 1: F = N x (gfdtd, D:1468);
// row_dt as in sell_h_kernel.
template <typename T, int NB>
__global__ void __launch_bounds__(kBlock) sell_n_kernel(int32_t nrows, int32_t nslices, const int64_t* __restrict__ sptr,
                                                        const int32_t* __restrict__ col, const T* __restrict__ val,
                                                        const T* __restrict__ x, T* __restrict__ Fv, double dt, int mode,
                                                        const double* __restrict__ row_dt) {
    const int lane = lane_id();
    for (SliceRange sr = block_slices(nslices); sr.s < sr.end; sr.s += sr.step) {
        const int s = sr.s;
        const int64_t b = sptr[s];
        const int w = (int)((sptr[s + 1] - b) / kSlice);
        const double v = sell_dot<T, NB>(b, w, col, val, lane, [&](int32_t j) { return x[j]; });
        const int32_t r = s * kSlice + lane;
        if (r < nrows) Fv[r] = mode == 0 ? (T)((double)Fv[r] - (row_dt ? row_dt[r] : dt) * v) : (T)v;
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
    for (SliceRange sr = block_slices(nslices); sr.s < sr.end; sr.s += sr.step) {
        const int s = sr.s;
        const int64_t eb = e_ptr[s], rb = r_ptr[s], cb = c_ptr[s];
        const int ew = (int)((e_ptr[s + 1] - eb) / kSlice), rw = (int)((r_ptr[s + 1] - rb) / kSlice);
        const int cw = (int)((c_ptr[s + 1] - cb) / kSlice);
        const double eu = sell_dot<T, 8>(eb, ew, e_col, e_val, lane, [&](int32_t j) { return U[j]; });
        const double rs = rw ? sell_dot<T, 8>(rb, rw, r_col, r_val, lane, [&](int32_t j) { return dir_new[j]; }) : 0.0;
        const double cf = sell_sgn_dot<T, 8>(cb, cw, c_col, c_sg, lane, Fv);
        const int32_t k = s * kSlice + lane;
        if (k < n_free) rhs[k] = (T)((eu - rs) + dt * cf + dt * (double)fb[free_edge[k]]);
    }
}

// ---- CG (ConjugateGradientSolver.hpp:52-103) ------------------------------
// Preconditioned CG in the single-reduction form of Chronopoulos and Gear
// (1989): with u = M r, w = A u, gamma = r.u, delta = w.u,
//   beta_i  = gamma_i / gamma_{i-1},  alpha_i = gamma_i / (delta_i - beta_i gamma_i / alpha_{i-1}),
//   p = u + beta p,  s = w + beta s  (s = A p),  x += alpha p,  r -= alpha s.
// In exact arithmetic the iterates are those of the reference's PCG, and the
// stop test is the reference's: ||r|| / ||r0|| > tol && it < maxiter. Its
// point on this device: every SpMV gathers one vector (A u, G r), and the
// vector updates are one streaming kernel, so an iteration is
//   Jacobi (M = 1/diag A):  cg_w (stop test, w = A u), cg_upd (updates, u = M r)
//   SpMV M = G (learned):   cg_g (stop test, u = G r), cg_w, cg_upd.
// Reductions go through per-block fp64 partials that the next kernel sums in
// a fixed order (no atomics, no grid-wide fences).

// Stop test at the start of pass `it` on the residual left by the previous
// update (the partials part_rr). Every block computes the same sums.
__device__ __forceinline__ bool cg_stop_test(int32_t it, const double* __restrict__ part_rr, int32_t nb_rr,
                                             CgState* st, double* red) {
    const double RR = sum_partials(part_rr, nb_rr, red);
    const double rr0 = it == 0 ? RR : st->rr0;
    const double ratio = std::sqrt(RR) / std::sqrt(rr0);
    const bool stop = !(ratio > st->tol && it < st->maxiter);
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (it == 0) st->rr0 = RR;
        st->relres = ratio;
        if (stop) { st->done = 1; st->iter = it; }
    }
    return stop;
}

// r = b - A x, partial r.r; with jacobi also u = pre r and partial r.u.
template <typename T>
__global__ void __launch_bounds__(kBlock) cg_init_kernel(int32_t n, int32_t nslices, const int64_t* __restrict__ sptr,
                                                         const int32_t* __restrict__ col, const T* __restrict__ val,
                                                         const T* __restrict__ pre, const T* __restrict__ b,
                                                         const T* __restrict__ x, T* __restrict__ r, T* __restrict__ u,
                                                         double* __restrict__ part_rr, double* __restrict__ part_gam,
                                                         CgState* st, int jacobi) {
    __shared__ double red[kBlock / kWave];
    const int lane = lane_id();
    double rr = 0, ru = 0;
    for (SliceRange sr = block_slices(nslices); sr.s < sr.end; sr.s += sr.step) {
        const int s = sr.s;
        const int64_t bs = sptr[s];
        const int w = (int)((sptr[s + 1] - bs) / kSlice);
        const double ax = sell_dot<T, 8>(bs, w, col, val, lane, [&](int32_t j) { return x[j]; });
        const int32_t i = s * kSlice + lane;
        if (i < n) {
            const T ri = (T)((double)b[i] - ax);
            r[i] = ri;
            rr += (double)ri * (double)ri;
            if (jacobi) {
                const T ui = (T)((double)pre[i] * (double)ri);
                u[i] = ui;
                ru += (double)ri * (double)ui;
            }
        }
    }
    rr = block_sum(rr, red);
    ru = block_sum(ru, red);
    if (threadIdx.x == 0) {
        part_rr[blockIdx.x] = rr;
        if (jacobi) part_gam[blockIdx.x] = ru;
        if (blockIdx.x == 0) { st->done = 0; st->iter = 0; }
    }
}

// Learned preconditioner pass: stop test (test_stop = 1 in solves), then
// u = G r with G on A's SELL pattern (same column array, fp32 values),
// partial r.u.
template <typename T, int NB>
__global__ void __launch_bounds__(kBlock) cg_g_kernel(int32_t it, int32_t n, int32_t nslices,
                                                      const int64_t* __restrict__ sptr, const int32_t* __restrict__ col,
                                                      const float* __restrict__ gval, const T* __restrict__ r,
                                                      T* __restrict__ u, int test_stop,
                                                      const double* __restrict__ part_rr, int32_t nb_rr,
                                                      double* __restrict__ part_gam, CgState* st) {
    if (st->done) return;
    __shared__ double red[kBlock / kWave];
    if (test_stop && cg_stop_test(it, part_rr, nb_rr, st, red)) return;
    const int lane = lane_id();
    double ru = 0;
    for (SliceRange sr = block_slices(nslices); sr.s < sr.end; sr.s += sr.step) {
        const int s = sr.s;
        const int64_t bs = sptr[s];
        const int w = (int)((sptr[s + 1] - bs) / kSlice);
        const double uv = sell_dot<T, NB>(bs, w, col, gval, lane, [&](int32_t j) { return r[j]; });
        const int32_t i = s * kSlice + lane;
        if (i < n) {
            const T ui = (T)uv;
            u[i] = ui;
            ru += (double)r[i] * (double)ui;
        }
    }
    ru = block_sum(ru, red);
    if (threadIdx.x == 0) part_gam[blockIdx.x] = ru;
}

// w = A u, partial w.u; with test_stop the pass starts with the stop test.
template <typename T, int NB>
__global__ void __launch_bounds__(kBlock) cg_w_kernel(int32_t it, int32_t n, int32_t nslices,
                                                      const int64_t* __restrict__ sptr, const int32_t* __restrict__ col,
                                                      const T* __restrict__ val, const T* __restrict__ u,
                                                      T* __restrict__ w, int test_stop,
                                                      const double* __restrict__ part_rr, int32_t nb_rr,
                                                      double* __restrict__ part_del, CgState* st) {
    if (st->done) return;
    __shared__ double red[kBlock / kWave];
    if (test_stop && cg_stop_test(it, part_rr, nb_rr, st, red)) return;
    const int lane = lane_id();
    double wu = 0;
    for (SliceRange sr = block_slices(nslices); sr.s < sr.end; sr.s += sr.step) {
        const int s = sr.s;
        const int64_t bs = sptr[s];
        const int wd = (int)((sptr[s + 1] - bs) / kSlice);
        const double wv = sell_dot<T, NB>(bs, wd, col, val, lane, [&](int32_t j) { return u[j]; });
        const int32_t i = s * kSlice + lane;
        if (i < n) {
            const T wi = (T)wv;
            w[i] = wi;
            wu += (double)wi * (double)u[i];
        }
    }
    wu = block_sum(wu, red);
    if (threadIdx.x == 0) part_del[blockIdx.x] = wu;
}

// Vector updates of pass `it`; gamma and delta from their partials, the
// previous gamma and alpha from slot (it-1)&1 of the state. With jacobi the
// next u = pre r and its partial r.u go to part_gam_next.
template <typename T>
__global__ void __launch_bounds__(kBlock) cg_upd_kernel(int32_t it, int32_t n, const T* __restrict__ pre,
                                                        T* __restrict__ x, T* __restrict__ r, T* __restrict__ u,
                                                        const T* __restrict__ w, T* __restrict__ p, T* __restrict__ s,
                                                        const double* __restrict__ part_gam, int32_t nb_gam,
                                                        const double* __restrict__ part_del, int32_t nb_del,
                                                        double* __restrict__ part_rr, double* __restrict__ part_gam_next,
                                                        CgState* st, int jacobi) {
    if (st->done) return;
    __shared__ double red[kBlock / kWave];
    const double gam = sum_partials(part_gam, nb_gam, red);
    const double del = sum_partials(part_del, nb_del, red);
    double alpha, beta;
    if (it == 0) {
        beta = 0.0;
        alpha = gam / del;
    } else {
        beta = gam / st->gam[(it - 1) & 1];
        alpha = gam / (del - beta * gam / st->alpha[(it - 1) & 1]);
    }
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        st->gam[it & 1] = gam;
        st->alpha[it & 1] = alpha;
    }
    double rr = 0, ru = 0;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const T pi = it == 0 ? u[i] : (T)((double)u[i] + beta * (double)p[i]);
        const T si = it == 0 ? w[i] : (T)((double)w[i] + beta * (double)s[i]);
        p[i] = pi;
        s[i] = si;
        x[i] = (T)((double)x[i] + alpha * (double)pi);
        const T ri = (T)((double)r[i] - alpha * (double)si);
        r[i] = ri;
        rr += (double)ri * (double)ri;
        if (jacobi) {
            const T ui = (T)((double)pre[i] * (double)ri);
            u[i] = ui;
            ru += (double)ri * (double)ui;
        }
    }
    rr = block_sum(rr, red);
    ru = block_sum(ru, red);
    if (threadIdx.x == 0) {
        part_rr[blockIdx.x] = rr;
        if (jacobi) part_gam_next[blockIdx.x] = ru;
    }
}

}  // namespace tfd
