// kernels.hpp - gfx950 kernels of the three leapfrog loops.
//
// Field and operator values are of type T (double on the plain path, float on
// the fast path); every reduction accumulates in double. Incidence is stored
// as int32 indices plus signs: 2-bit codes packed four to a byte
// (0 = absent, 1 = +1, 3 = -1) for the fixed-width FIT stencils, int8 per
// entry for the CSR incidence of the tetrahedral meshes.
#pragma once

#include "common.hpp"

#include <algorithm>

namespace tfd {

__device__ __forceinline__ double sgn2(uint32_t code) {
    // 0 -> 0, 1 -> +1, 3 -> -1
    return code == 1 ? 1.0 : (code == 3 ? -1.0 : 0.0);
}

// Value injected at a source edge for the current step: sum of its terms of
// carrier(source, step) * weight.
__device__ __forceinline__ double source_value(int tb, int te, const int32_t* __restrict__ term_src,
                                               const double* __restrict__ term_w,
                                               const double* __restrict__ carr_step) {
    double v = 0;
    for (int t = tb; t < te; ++t) v += carr_step[term_src[t]] * term_w[t];
    return v;
}

// Device state of one CG solve (kernels in sell.hpp). Each kernel writes one
// fp64 partial per block; the next kernel sums them (every block, in index
// order, so all blocks take the same decision) - no cross-block fences.
struct CgState {
    double rr0;        // ||r0||^2 of this solve
    double gam[2];     // r.u of pass it, slot it&1
    double alpha[2];   // step length of pass it, slot it&1
    double beta[2];    // beta of pass it, slot it&1 (learned passes)
    double relres;     // ||r||/||r0|| at the last stop test
    double tol;
    int32_t maxiter;
    int32_t iter;      // iterations done when the solve stopped
    int32_t done;
    int32_t need_conv;  // the solve must meet tol (learned solves; Jacobi ones may stop at maxiter)
    int32_t rr0_given;  // rr0 holds the warm-start residual of an extrapolated start (cg_set_rr0_kernel)
};

// Guard of the kernels a gfdtd step enqueues ahead of the host's check of
// the previous CG solve (gf_scatter of that step, the next step's
// dga_curl, sell_n, probe_kernel, gf_inject and gf_rhs_init): they run only
// if that solve has stopped, and, for a solve that must meet the tolerance,
// met it. Otherwise they return at once and the host enqueues them again
// once the solve is complete. A null guard always passes.
__device__ __forceinline__ bool cg_guard_ok(const CgState* g) {
    return !g || (g->done && (!g->need_conv || g->relres <= g->tol));
}

// Extrapolated start of a gfdtd solve (an option of the fast path; the
// stop test stays relative to the warm-start residual b - A U_old). On the
// free rows x holds U_n (the warm start), h1 = U_{n-1}, h2 = U_{n-2}:
// order 1: x0 = 2 U_n - U_{n-1}; order 2: x0 = 3 U_n - 3 U_{n-1} + U_{n-2};
// order 0: x unchanged. d = x0 - U_n as the rounded x0 gives it (exact when
// x0 and U_n lie within a factor of two, Sterbenz), for r0 - A d; the
// history then shifts by one step. Guard as in cg_guard_ok.
template <typename T>
__global__ void __launch_bounds__(kBlock) cg_extrap_kernel(int32_t n, T* __restrict__ x, T* __restrict__ h1,
                                                           T* __restrict__ h2, T* __restrict__ d, int order,
                                                           const CgState* guard) {
    if (!cg_guard_ok(guard)) return;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const T xi = x[i], a = h1[i], b = h2[i];
        const double x0 = order == 2 ? 3.0 * (double)xi - 3.0 * (double)a + (double)b
                          : order == 1 ? 2.0 * (double)xi - (double)a
                                       : (double)xi;
        const T x0t = (T)x0;
        d[i] = (T)((double)x0t - (double)xi);
        x[i] = x0t;
        h2[i] = a;
        h1[i] = xi;
    }
}

// ------------------------------------------------------------------ FIT
//
// Grid order (tf_fit_ops edge_sid / face_sid): with the fields stored at
// the grid positions of their edges and faces, the four faces around an
// interior edge and the four edges of a face follow from the position.
// Slots: x edge (i,j,k): y faces (i,j,k-1), (i,j,k), z faces (i,j-1,k),
// (i,j,k); y edge: x faces (i,j,k-1), (i,j,k), z faces (i-1,j,k), (i,j,k);
// z edge: x faces (i,j-1,k), (i,j,k), y faces (i-1,j,k), (i,j,k). x face
// (i,j,k): y edges (i,j,k), (i,j,k+1), z edges (i,j,k), (i,j+1,k); y face:
// x edges (i,j,k), (i,j,k+1), z edges (i,j,k), (i+1,j,k); z face: x edges
// (i,j,k), (i,j+1,k), y edges (i,j,k), (i+1,j,k). Slots of boundary
// entities may lie outside the grid; those are never read (code 0).
// Division of n < 2^31 by a fixed divisor d >= 1 with a multiply-high and a
// shift (Granlund and Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994): l = ceil(log2 d), m = 2^32 (2^l - d) / d + 1,
// n / d = (mulhi(m, n) + n) >> l; mulhi(m, n) < n, so the sum stays below
// 2^32.
struct FastDiv {
    uint32_t d, m, l;
};
inline FastDiv make_fastdiv(uint32_t d) {
    uint32_t l = 0;
    while (((uint64_t)1 << l) < d) ++l;
    const uint64_t m = ((((uint64_t)1 << l) - d) << 32) / d + 1;
    return FastDiv{d, (uint32_t)m, l};
}
__host__ __device__ inline uint32_t fastdiv(const FastDiv& f, uint32_t n) {
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t t = __umulhi(f.m, n);
#else
    const uint32_t t = (uint32_t)(((uint64_t)f.m * n) >> 32);
#endif
    return (t + n) >> f.l;
}

struct FitGrid {
    int32_t nx, ny, nz;
    FastDiv dx, dx1, dy, dy1;  // by nx, nx + 1, ny, ny + 1
};
inline FitGrid make_fit_grid(int32_t nx, int32_t ny, int32_t nz) {
    return FitGrid{nx, ny, nz, make_fastdiv(nx), make_fastdiv(nx + 1), make_fastdiv(ny), make_fastdiv(ny + 1)};
}
// n = q d + r with the fixed divisor of f
__host__ __device__ inline void fit_divmod(const FastDiv& f, uint32_t n, uint32_t& q, uint32_t& r) {
    q = fastdiv(f, n);
    r = n - q * f.d;
}
// Items of each direction in grid order: x, y, z edges (faces = false) or
// x, y, z faces (faces = true).
__host__ __device__ inline void fit_counts(const FitGrid g, bool faces, uint32_t n[3]) {
    const uint32_t nx = g.nx, ny = g.ny, nz = g.nz;
    if (faces) {
        n[0] = (nx + 1) * ny * nz; n[1] = nx * (ny + 1) * nz; n[2] = nx * ny * (nz + 1);
    } else {
        n[0] = nx * (ny + 1) * (nz + 1); n[1] = (nx + 1) * ny * (nz + 1); n[2] = (nx + 1) * (ny + 1) * nz;
    }
}
// Block mapping for grid order. The items of one direction are stored k-plane
// by k-plane, so the XCD that takes the x-th eighth of the blocks of each of
// the three directions works on one k-slab of the grid, and the neighbours its
// items gather (faces of an edge, edges of a face) stay in its own L2; with a
// single contiguous range per XCD, each gathered plane would be read through
// the L2s of two XCDs. Workgroup b runs on XCD b % 8 and takes the (b / 8)-th
// block of that XCD's list (its x, then y, then z blocks). The grid has
// 8 * (longest list) workgroups (slab_grid); returns the item, or -1 for a
// workgroup past its XCD's list or a lane past its direction's end.
__host__ __device__ inline uint32_t slab_len(uint32_t n, uint32_t x) {
    const uint32_t nb = (n + kBlock - 1) / kBlock;
    return (x + 1) * nb / kXcd - x * nb / kXcd;
}
inline int slab_grid(const uint32_t n[3]) {
    uint32_t m = 1;
    for (uint32_t x = 0; x < (uint32_t)kXcd; ++x)
        m = std::max(m, slab_len(n[0], x) + slab_len(n[1], x) + slab_len(n[2], x));
    return (int)(m * kXcd);
}
// The XCD's list interleaves the three directions block by block (x, y, z,
// x, y, z, ...; the longer directions' last few blocks follow), so that all
// three advance through the slab's planes together and a plane of gathered
// neighbours, read by two directions, is fetched once while it is in L2.
__device__ __forceinline__ int32_t slab_item(int b, const uint32_t n[3]) {
    const uint32_t x = (uint32_t)b % kXcd, i = (uint32_t)b / kXcd;
    uint32_t lo[3], len[3];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
        const uint32_t nb = (n[t] + kBlock - 1) / kBlock;
        lo[t] = x * nb / kXcd;
        len[t] = (x + 1) * nb / kXcd - lo[t];
    }
    const uint32_t m = min(len[0], min(len[1], len[2]));
    int t = -1;
    uint32_t j = 0;
    if (i < 3 * m) {
        t = (int)(i % 3);
        j = i / 3;
    } else {
        uint32_t k = i - 3 * m;
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            if (t < 0 && k < len[u] - m) {
                t = u;
                j = m + k;
            } else if (t < 0) {
                k -= len[u] - m;
            }
        }
        if (t < 0) return -1;
    }
    const uint32_t base = t == 0 ? 0u : (t == 1 ? n[0] : n[0] + n[1]);
    const uint32_t r = (lo[t] + j) * kBlock + threadIdx.x;
    return r < n[t] ? (int32_t)(base + r) : -1;
}

// 32-bit index arithmetic: grid order is used only when edge and face
// counts are below 2^31 (fit_grid_codes).
__host__ __device__ inline void fit_edge_slots(const FitGrid g, uint32_t s, int32_t f[4]) {
    const uint32_t nx = g.nx, ny = g.ny, nz = g.nz;
    const uint32_t nEx = nx * (ny + 1) * (nz + 1), nEy = (nx + 1) * ny * (nz + 1);
    const uint32_t nFx = (nx + 1) * ny * nz, nFy = nx * (ny + 1) * nz;
    auto xf = [&](uint32_t i, uint32_t j, uint32_t k) { return (int32_t)((k * ny + j) * (nx + 1) + i); };
    auto yf = [&](uint32_t i, uint32_t j, uint32_t k) { return (int32_t)(nFx + (k * (ny + 1) + j) * nx + i); };
    auto zf = [&](uint32_t i, uint32_t j, uint32_t k) { return (int32_t)(nFx + nFy + (k * ny + j) * nx + i); };
    // out-of-grid slots (boundary edges) wrap around in unsigned arithmetic;
    // they are computed but never read
    uint32_t i, j, k, t;
    if (s < nEx) {
        fit_divmod(g.dx, s, t, i);
        fit_divmod(g.dy1, t, k, j);
        f[0] = yf(i, j, k - 1); f[1] = yf(i, j, k); f[2] = zf(i, j - 1, k); f[3] = zf(i, j, k);
    } else if (s < nEx + nEy) {
        fit_divmod(g.dx1, s - nEx, t, i);
        fit_divmod(g.dy, t, k, j);
        f[0] = xf(i, j, k - 1); f[1] = xf(i, j, k); f[2] = zf(i - 1, j, k); f[3] = zf(i, j, k);
    } else {
        fit_divmod(g.dx1, s - nEx - nEy, t, i);
        fit_divmod(g.dy1, t, k, j);
        f[0] = xf(i, j - 1, k); f[1] = xf(i, j, k); f[2] = yf(i - 1, j, k); f[3] = yf(i, j, k);
    }
}
__host__ __device__ inline void fit_face_slots(const FitGrid g, uint32_t s, int32_t e[4]) {
    const uint32_t nx = g.nx, ny = g.ny, nz = g.nz;
    const uint32_t nEx = nx * (ny + 1) * (nz + 1), nEy = (nx + 1) * ny * (nz + 1);
    const uint32_t nFx = (nx + 1) * ny * nz, nFy = nx * (ny + 1) * nz;
    auto xe = [&](uint32_t i, uint32_t j, uint32_t k) { return (int32_t)((k * (ny + 1) + j) * nx + i); };
    auto ye = [&](uint32_t i, uint32_t j, uint32_t k) { return (int32_t)(nEx + (k * ny + j) * (nx + 1) + i); };
    auto ze = [&](uint32_t i, uint32_t j, uint32_t k) {
        return (int32_t)(nEx + nEy + (k * (ny + 1) + j) * (nx + 1) + i);
    };
    uint32_t i, j, k, t;
    if (s < nFx) {
        fit_divmod(g.dx1, s, t, i);
        fit_divmod(g.dy, t, k, j);
        e[0] = ye(i, j, k); e[1] = ye(i, j, k + 1); e[2] = ze(i, j, k); e[3] = ze(i, j + 1, k);
    } else if (s < nFx + nFy) {
        fit_divmod(g.dx, s - nFx, t, i);
        fit_divmod(g.dy1, t, k, j);
        e[0] = xe(i, j, k); e[1] = xe(i, j, k + 1); e[2] = ze(i, j, k); e[3] = ze(i + 1, j, k);
    } else {
        fit_divmod(g.dx, s - nFx - nFy, t, i);
        fit_divmod(g.dy, t, k, j);
        e[0] = xe(i, j, k); e[1] = xe(i, j + 1, k); e[2] = ye(i, j, k); e[3] = ye(i + 1, j, k);
    }
}

// Sum over the four positions p of sign_p * v[slot_p] in position order,
// with code = slot map (2 bits per position) << 8 | sign code (2 bits per
// position: 1 -> +1, 2 -> -1, 0 -> 0): the reference's F0 - F1 + F2 - F3
// order, whatever slot each of its faces sits in.
__device__ __forceinline__ double slot_sum(uint32_t code, const double v[4]) {
    const uint32_t m = code >> 8;
    double c = sgn2(code & 3) * v[m & 3];
    c += sgn2((code >> 2) & 3) * v[(m >> 2) & 3];
    c += sgn2((code >> 4) & 3) * v[(m >> 4) & 3];
    c += sgn2((code >> 6) & 3) * v[(m >> 6) & 3];
    return c;
}

// E phase (D:1850-1895). The first sp_blocks workgroups take the special
// items, so that their longer dependent chains (source value, Poynting sum)
// start with the launch instead of trailing it; the rest cover every edge in
// index order (XCD-mapped among themselves). Edges whose sign byte is 0
// (pec / source / non-common) are left to the special items. U = ca*U_old + cb*(sum_k s_k F[f_k])
//   common:   ca = upml^2 M_h M_q, cb = upml^3 M_h dt, s = dual_curl*(+,-,+,-)
// Special items (uncommon h-source edges, e/j source edges):
//   kind 0 (h):  U = ca*U_old + cb*(sum s_k F + src)     (ca = M_h M_q, cb = M_h dt)
//   kind 1 (e):  U = src
//   kind 2 (j):  U = ca*U_old + cb*(sum s_k F - src)     (common coefficients)
// With code16 (grid order) the main blocks take their edges by slab_item,
// each edge's faces from its position and its sign and order code from code16
// (slot_sum); otherwise edges by index order over XCD ranges, faces from the
// packed face quadruple f4 and the sign byte sg.
template <typename T>
__global__ void __launch_bounds__(kBlock) fit_e_kernel(
    int32_t n_edges, int32_t sp_blocks, const uint8_t* __restrict__ sg, const T* __restrict__ ca,
    const T* __restrict__ cb, const int4* __restrict__ f4, const uint16_t* __restrict__ code16, FitGrid grid,
    const T* __restrict__ Fv, T* __restrict__ U,
    int32_t n_sp, const int32_t* __restrict__ sp_edge, const int4* __restrict__ sp_f4,
    const uint8_t* __restrict__ sp_sg, const double* __restrict__ sp_ca, const double* __restrict__ sp_cb,
    const int8_t* __restrict__ sp_kind, const int32_t* __restrict__ sp_tb, const int32_t* __restrict__ sp_te,
    const int32_t* __restrict__ term_src, const double* __restrict__ term_w,
    const double* __restrict__ carr_step, double* __restrict__ poy_part) {
    const int hb = (int)blockIdx.x;
    if (hb >= sp_blocks) {
        if (code16) {
            uint32_t cnt[3];
            fit_counts(grid, false, cnt);
            const int32_t e = slab_item(hb - sp_blocks, cnt);
            if (e < 0) return;
            const uint32_t code = code16[e];
            if (code == 0) return;
            int32_t f[4];
            fit_edge_slots(grid, (uint32_t)e, f);
            const double v[4] = {(double)Fv[f[0]], (double)Fv[f[1]], (double)Fv[f[2]], (double)Fv[f[3]]};
            U[e] = (T)((double)ca[e] * (double)U[e] + (double)cb[e] * slot_sum(code, v));
            return;
        }
        const int32_t e = xcd_block_of(hb - sp_blocks, (int)gridDim.x - sp_blocks) * kBlock + threadIdx.x;
        if (e >= n_edges) return;
        const uint32_t s = sg[e];
        if (s == 0) return;
        const int4 f = f4[e];
        double c = sgn2(s & 3) * (double)Fv[f.x] + sgn2((s >> 2) & 3) * (double)Fv[f.y] +
                   sgn2((s >> 4) & 3) * (double)Fv[f.z] + sgn2((s >> 6) & 3) * (double)Fv[f.w];
        U[e] = (T)((double)ca[e] * (double)U[e] + (double)cb[e] * c);
        return;
    }
    __shared__ double red[kBlock / kWave];
    const int32_t i = hb * kBlock + threadIdx.x;
    double poy = 0;
    if (i < n_sp) {
        const int32_t e = sp_edge[i];
        const double src = source_value(sp_tb[i], sp_te[i], term_src, term_w, carr_step);
        const int k = sp_kind[i];
        const double uold = (double)U[e];
        double unew;
        if (k == 1) {
            unew = src;
        } else {
            const int4 f = sp_f4[i];
            const uint32_t s = sp_sg[i];
            double c = 0;
            if (s & 3) c += sgn2(s & 3) * (double)Fv[f.x];
            if ((s >> 2) & 3) c += sgn2((s >> 2) & 3) * (double)Fv[f.y];
            if ((s >> 4) & 3) c += sgn2((s >> 4) & 3) * (double)Fv[f.z];
            if ((s >> 6) & 3) c += sgn2((s >> 6) & 3) * (double)Fv[f.w];
            c += k == 0 ? src : -src;
            unew = sp_ca[i] * uold + sp_cb[i] * c;
        }
        U[e] = (T)unew;
        // Poynting term 0.5*(U_old+U)*curl_fb, nonzero on h-source edges only (D:1901)
        if (k == 0) poy = 0.5 * (uold + (double)(T)unew) * src;
    }
    const double s = block_sum(poy, red);
    if (threadIdx.x == 0) poy_part[hb] = s;
}

// H phase over tbc faces (D:1908-1915): B = cm*F_old - dt*sum s_k U[e_k];
// F = cn*B with cm = M_mu*fpml, cn = fpml*M_nu, s = curl*(+,-,+,-).
// code16 / grid as in fit_e_kernel.
template <typename T>
__global__ void __launch_bounds__(kBlock) fit_h_kernel(int32_t n_faces, const uint8_t* __restrict__ sg,
                                                       const T* __restrict__ cm, const T* __restrict__ cn,
                                                       const int4* __restrict__ e4, const uint16_t* __restrict__ code16,
                                                       FitGrid grid, const T* __restrict__ U,
                                                       T* __restrict__ Fv, T* __restrict__ B, double dt) {
    if (code16) {
        uint32_t cnt[3];
        fit_counts(grid, true, cnt);
        const int32_t f = slab_item((int)blockIdx.x, cnt);
        if (f < 0) return;
        const uint32_t code = code16[f];
        if (code == 0) return;
        int32_t e[4];
        fit_face_slots(grid, (uint32_t)f, e);
        const double v[4] = {(double)U[e[0]], (double)U[e[1]], (double)U[e[2]], (double)U[e[3]]};
        const double b = (double)cm[f] * (double)Fv[f] - dt * slot_sum(code, v);
        B[f] = (T)b;
        Fv[f] = (T)((double)cn[f] * b);
        return;
    }
    const int32_t f = xcd_block() * kBlock + threadIdx.x;
    if (f >= n_faces) return;
    const uint32_t s = sg[f];
    if (s == 0) return;
    const int4 e = e4[f];
    double c = sgn2(s & 3) * (double)U[e.x] + sgn2((s >> 2) & 3) * (double)U[e.y] +
               sgn2((s >> 4) & 3) * (double)U[e.z] + sgn2((s >> 6) & 3) * (double)U[e.w];
    const double b = (double)cm[f] * (double)Fv[f] - dt * c;
    B[f] = (T)b;
    Fv[f] = (T)((double)cn[f] * b);
}

// ------------------------------------------------------------------ probes
// One thread per (probe, component): E from U, H from B (D:2400-2470,
// D:3093-3158). Output sample layout [probe][6].
struct ProbeDev {
    int32_t ne, nh;
    int32_t e_idx[12];
    double e_w[12][3];
    int32_t h_idx[6];
    double h_w[6][3];
};

template <typename T>
__global__ void probe_kernel(int32_t n_probes, const ProbeDev* __restrict__ pr, const T* __restrict__ U,
                             const T* __restrict__ B, double* __restrict__ out, const CgState* guard) {
    if (!cg_guard_ok(guard)) return;
    const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_probes * 6) return;
    const int p = t / 6, c = t % 6;
    const ProbeDev& q = pr[p];
    double v = 0;
    if (c < 3) {
        for (int k = 0; k < q.ne; ++k) v += (double)U[q.e_idx[k]] * q.e_w[k][c];
    } else {
        for (int k = 0; k < q.nh; ++k) v += (double)B[q.h_idx[k]] * q.h_w[k][c - 3];
    }
    out[t] = v;
}

// ------------------------------------------------------------------ DGA
// Source injection for the tetrahedral methods: h sources add Fb to curl_f
// (Ctb*Fb, one +1 per boundary edge), e sources set U (class-4 edges).
// The Poynting partial uses U before the E update: 0.5*U_old*fb.
template <typename T>
__global__ void __launch_bounds__(kBlock) dga_inject_kernel(
    int32_t n_inj, const int32_t* __restrict__ inj_edge, const int8_t* __restrict__ inj_kind,
    const int32_t* __restrict__ inj_tb, const int32_t* __restrict__ inj_te, const int32_t* __restrict__ term_src,
    const double* __restrict__ term_w, const double* __restrict__ carr_step, T* __restrict__ curl_f,
    T* __restrict__ U, double* __restrict__ poy_part, int init) {
    __shared__ double red[kBlock / kWave];
    const int32_t i = blockIdx.x * kBlock + threadIdx.x;
    double poy = 0;
    if (i < n_inj) {
        const int32_t e = inj_edge[i];
        const double v = source_value(inj_tb[i], inj_te[i], term_src, term_w, carr_step);
        if (init && inj_kind[i] == 0) {
            // before the loop every source edge takes its e-source sum at
            // t = 0 (D:746-765), which is 0 on an h edge
            U[e] = (T)0;
        } else if (inj_kind[i] == 0) {
            curl_f[e] = (T)((double)curl_f[e] + v);
            poy = 0.5 * (double)U[e] * v;
        } else {
            U[e] = (T)v;
        }
    }
    const double s = block_sum(poy, red);
    if (threadIdx.x == 0) poy_part[blockIdx.x] = s;
}

// Second Poynting partial after the E update: 0.5*U_new*fb on h-source edges.
template <typename T>
__global__ void __launch_bounds__(kBlock) dga_poynting_kernel(
    int32_t n_inj, const int32_t* __restrict__ inj_edge, const int8_t* __restrict__ inj_kind,
    const int32_t* __restrict__ inj_tb, const int32_t* __restrict__ inj_te, const int32_t* __restrict__ term_src,
    const double* __restrict__ term_w, const double* __restrict__ carr_step, const T* __restrict__ U,
    double* __restrict__ poy_part) {
    __shared__ double red[kBlock / kWave];
    const int32_t i = blockIdx.x * kBlock + threadIdx.x;
    double poy = 0;
    if (i < n_inj && inj_kind[i] == 0) {
        const double v = source_value(inj_tb[i], inj_te[i], term_src, term_w, carr_step);
        poy = 0.5 * (double)U[inj_edge[i]] * v;
    }
    const double s = block_sum(poy, red);
    if (threadIdx.x == 0) poy_part[blockIdx.x] = s;
}

// ---- lossy dga: fractional unknowns (D:883-920, tf_dga_frac)
// Sources of a lossy dga step. h edges as in dga_inject_kernel (curl_f and
// the first Poynting partial); the other kinds write their value into the
// edge's U_frac entry slot[e] (U_d of a class-4 edge, U_a of a class-1
// edge; none for classes 2/3, whose entries the update rebuilds from their
// old values, D:888-890). With init (the pass before the loop, D:726-790)
// every source edge writes its e-source value at t = 0 (0 unless an e
// source) into all of its fractional entries, half of it into U_c entries
// ([q0, q1)); mu_a / mu_b are the entries of each edge's Mu row.
template <typename T>
__global__ void __launch_bounds__(kBlock) frac_inject_kernel(
    int32_t n_inj, const int32_t* __restrict__ inj_edge, const int8_t* __restrict__ inj_kind,
    const int32_t* __restrict__ inj_tb, const int32_t* __restrict__ inj_te, const int32_t* __restrict__ term_src,
    const double* __restrict__ term_w, const double* __restrict__ carr_step, T* __restrict__ curl_f,
    const T* __restrict__ U, T* __restrict__ Uf, const int32_t* __restrict__ slot, const int32_t* __restrict__ mu_a,
    const int32_t* __restrict__ mu_b, int32_t q0, int32_t q1, double* __restrict__ poy_part, int init) {
    __shared__ double red[kBlock / kWave];
    const int32_t i = blockIdx.x * kBlock + threadIdx.x;
    double poy = 0;
    if (i < n_inj) {
        const int32_t e = inj_edge[i];
        const double v = source_value(inj_tb[i], inj_te[i], term_src, term_w, carr_step);
        if (init) {
            const double u = inj_kind[i] == 0 ? 0.0 : v;
            for (int32_t c : {mu_a[e], mu_b[e]})
                if (c >= 0) Uf[c] = (T)(c >= q0 && c < q1 ? 0.5 * u : u);
        } else if (inj_kind[i] == 0) {
            curl_f[e] = (T)((double)curl_f[e] + v);
            poy = 0.5 * (double)U[e] * v;
        } else if (slot[e] >= 0) {
            Uf[slot[e]] = (T)v;
        }
    }
    const double s = block_sum(poy, red);
    if (threadIdx.x == 0) poy_part[blockIdx.x] = s;
}

// out[i] = X[a[i]] (+ X[b[i]] when b[i] >= 0): U = Mu U_frac, F = Tf F_frac
// (rows of one or two unit entries).
template <typename T>
__global__ void __launch_bounds__(kBlock) frac_gather_kernel(int32_t n, const int32_t* __restrict__ a,
                                                             const int32_t* __restrict__ b, const T* __restrict__ X,
                                                             T* __restrict__ out) {
    const int32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int32_t j = b[i];
    out[i] = j >= 0 ? (T)((double)X[a[i]] + (double)X[j]) : X[a[i]];
}

// Source work taken by workgroups appended to dga_curl_kernel's grid (after
// the face workgroups in logical order), one injection item per thread:
//   mode 1 (dga): the second Poynting half of the step, 0.5 U_new v on the
//     h-source edges;
//   mode 2 (gfdtd, fem): the step's sources as gf_inject_kernel forms them
//     (fb, dir_new and the Poynting term from U_old).
// One fp64 partial per appended workgroup into poy_part; mode 0: none.
template <typename T>
struct CurlSources {
    int32_t n_inj = 0;
    int32_t mode = 0;
    const int32_t* edge = nullptr;
    const int8_t* kind = nullptr;
    const int32_t* tb = nullptr;
    const int32_t* te = nullptr;
    const int32_t* term_src = nullptr;
    const double* term_w = nullptr;
    const double* carr_step = nullptr;
    const double* carr_prev = nullptr;  // mode 2: the previous step's carriers (null at step 1)
    T* fb = nullptr;
    T* dir_new = nullptr;
    int32_t delta = 0;                  // mode 2, fem: an h edge's forcing is Fb - Fb_old
    double* poy_part = nullptr;
};

// gfdtd sources of one item (D:1478-1502): returns its Poynting term.
template <typename T>
__device__ __forceinline__ double gf_source_item(int32_t i, const int32_t* __restrict__ inj_edge,
                                                 const int8_t* __restrict__ inj_kind, const int32_t* __restrict__ inj_tb,
                                                 const int32_t* __restrict__ inj_te,
                                                 const int32_t* __restrict__ term_src,
                                                 const double* __restrict__ term_w, const double* __restrict__ carr_step,
                                                 const double* __restrict__ carr_prev, T* __restrict__ fb,
                                                 T* __restrict__ dir_new, const T* __restrict__ U, int delta) {
    const int32_t e = inj_edge[i];
    const double v = source_value(inj_tb[i], inj_te[i], term_src, term_w, carr_step);
    if (inj_kind[i] == 0) {
        const double vo = carr_prev ? source_value(inj_tb[i], inj_te[i], term_src, term_w, carr_prev) : 0.0;
        fb[e] = (T)(delta ? v - vo : v);
        return (double)U[e] * (0.5 * (v + vo));
    }
    if (inj_kind[i] == 2) {
        fb[e] = (T)(-v);  // current I at t (D:1491-1492): -dt I joins dt Ctb Fb in Psi
    } else {
        dir_new[e] = (T)v;
    }
    return 0.0;
}

// curl_u = C U (3 signed edges per face), B -= dt*curl_u (D:902-910).
template <typename T>
__global__ void __launch_bounds__(kBlock) dga_curl_kernel(int32_t n_faces, const int32_t* __restrict__ c_col,
                                                          const int8_t* __restrict__ c_sg, const T* __restrict__ U,
                                                          T* __restrict__ curl_u, T* __restrict__ B, double dt,
                                                          const double* __restrict__ face_dt, const CgState* guard,
                                                          CurlSources<T> py) {
    if (!cg_guard_ok(guard)) return;
    // the appended workgroups are the last ones dispatched, spread over the
    // XCDs; the face workgroups keep the XCD-aware map over their own count
    const int face_blocks = (n_faces + kBlock - 1) / kBlock;
    if (py.mode && (int)blockIdx.x >= face_blocks) {
        __shared__ double red[kBlock / kWave];
        const int lb = (int)blockIdx.x;
        const int32_t i = (lb - face_blocks) * kBlock + threadIdx.x;
        double poy = 0;
        if (i < py.n_inj) {
            if (py.mode == 1) {
                if (py.kind[i] == 0)
                    poy = 0.5 * (double)U[py.edge[i]] *
                          source_value(py.tb[i], py.te[i], py.term_src, py.term_w, py.carr_step);
            } else {
                poy = gf_source_item<T>(i, py.edge, py.kind, py.tb, py.te, py.term_src, py.term_w, py.carr_step,
                                        py.carr_prev, py.fb, py.dir_new, U, py.delta);
            }
        }
        const double s = block_sum(poy, red);
        if (threadIdx.x == 0) py.poy_part[lb - face_blocks] = s;
        return;
    }
    const int32_t f = xcd_block_of((int)blockIdx.x, py.mode ? face_blocks : (int)gridDim.x) * kBlock + threadIdx.x;
    if (f >= n_faces) return;
    const int32_t* c = c_col + 3 * (int64_t)f;
    double cu;
    if (c_sg) {
        const int8_t* s = c_sg + 3 * (int64_t)f;
        cu = (double)s[0] * (double)U[c[0]] + (double)s[1] * (double)U[c[1]] + (double)s[2] * (double)U[c[2]];
    } else {
        // every tet face has the fixed pattern +1,-1,+1 on (p0p1, p0p2, p1p2) (D:6345-6352)
        cu = (double)U[c[0]] - (double)U[c[1]] + (double)U[c[2]];
    }
    curl_u[f] = (T)cu;
    B[f] = (T)((double)B[f] - (face_dt ? face_dt[f] : dt) * cu);
}

// ------------------------------------------------------------------ gfdtd
__device__ __forceinline__ double sum_partials(const double* __restrict__ part, int n, double* red) {
    double v = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v += part[i];
    v = block_sum(v, red);
    __shared__ double bcast;
    if (threadIdx.x == 0) bcast = v;
    __syncthreads();
    v = bcast;
    __syncthreads();
    return v;
}

// U[free[k]] = x[k]; Dirichlet edges take their new values.
// U_older (fem, or null): the old U of every written edge is kept first
// (U_older = U_old, D:1764).
template <typename T>
__global__ void gf_scatter_kernel(int32_t n_free, const int32_t* __restrict__ free_edge, const T* __restrict__ x,
                                  T* __restrict__ U, int32_t n_dir, const int32_t* __restrict__ dir_edge,
                                  const T* __restrict__ dir_new, const CgState* guard, T* __restrict__ U_older) {
    if (!cg_guard_ok(guard)) return;
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_free) {
        const int32_t e = free_edge[i];
        if (U_older) U_older[e] = U[e];
        U[e] = x[i];
    } else if (i - n_free < n_dir) {
        const int32_t e = dir_edge[i - n_free];
        if (U_older) U_older[e] = U[e];
        U[e] = dir_new[e];
    }
}

// The same in edge order: edge e takes x[edge_row[e]], or its new
// Dirichlet value when edge_row[e] < 0. The free rows are ordered for the
// gathers of the CG, not by edge, so one thread per edge gathers from x and
// writes U (and U_older) in whole cache lines.
template <typename T>
__global__ void gf_scatter_edges_kernel(int32_t n_edges, const int32_t* __restrict__ edge_row,
                                        const T* __restrict__ x, T* __restrict__ U, const T* __restrict__ dir_new,
                                        const CgState* guard, T* __restrict__ U_older) {
    if (!cg_guard_ok(guard)) return;
    const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_edges) return;
    const int32_t r = edge_row[e];
    if (U_older) U_older[e] = U[e];
    U[e] = r >= 0 ? x[r] : dir_new[e];
}

// fem: the vectors the step's initial residual gathers (U = U_old here, Src
// = the Dirichlet values at the new time): free edges W = W2 = U - U_older;
// Dirichlet edges W = 2 U - U_older - Src, W2 = Src - U_older, so that with
// x0 = U_old on the free edges
//   Psi_f - A x0 = 1/dt^2 (E W)_f - 0.5/dt (SigMat W2)_f - (Ct N C U_old)_f
//                  + 1/dt (Ctb (Fb - Fb_old) - I)_f
// without the cancelling 2/dt^2 E U_old - A x0 of the reference's Psi
// (D:1731-1733). guard as in cg_guard_ok.
// One thread per edge; edge_row[e] < 0 marks a Dirichlet edge.
template <typename T>
__global__ void fem_w_kernel(int32_t n_edges, const int32_t* __restrict__ edge_row, const T* __restrict__ U,
                             const T* __restrict__ U_older, const T* __restrict__ src, T* __restrict__ W,
                             T* __restrict__ W2, const CgState* guard) {
    if (!cg_guard_ok(guard)) return;
    const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_edges) return;
    if (edge_row[e] >= 0) {
        const T v = (T)((double)U[e] - (double)U_older[e]);
        W[e] = v;
        W2[e] = v;
    } else {
        const double u = U[e], uo = U_older[e], sv = src[e];
        W[e] = (T)((u - uo) + (u - sv));
        W2[e] = (T)(sv - uo);
    }
}

// gfdtd sources at step i (D:1478-1502): fb, the edge-indexed forcing
// Ctb Fb - I of Psi (h sources at t - dt/2: Fb; j dipoles at t: -I), and
// dir_new (e sources); the Poynting term U_old . 0.5 Ctb (Fb + Fb_old)
// (D:1518) with fb_old from the previous step's carriers. With delta (fem,
// D:1701-1733) an h edge's forcing is Fb - Fb_old.
template <typename T>
__global__ void __launch_bounds__(kBlock) gf_inject_kernel(
    int32_t n_inj, const int32_t* __restrict__ inj_edge, const int8_t* __restrict__ inj_kind,
    const int32_t* __restrict__ inj_tb, const int32_t* __restrict__ inj_te, const int32_t* __restrict__ term_src,
    const double* __restrict__ term_w, const double* __restrict__ carr_step, const double* __restrict__ carr_prev,
    T* __restrict__ fb, T* __restrict__ dir_new, const T* __restrict__ U, double* __restrict__ poy_part,
    const CgState* guard, int delta) {
    if (!cg_guard_ok(guard)) return;
    __shared__ double red[kBlock / kWave];
    const int32_t i = blockIdx.x * kBlock + threadIdx.x;
    double poy = 0;
    if (i < n_inj)
        poy = gf_source_item<T>(i, inj_edge, inj_kind, inj_tb, inj_te, term_src, term_w, carr_step, carr_prev, fb,
                                dir_new, U, delta);
    const double s = block_sum(poy, red);
    if (threadIdx.x == 0) poy_part[blockIdx.x] = s;
}

// ------------------------------------------------------------------ energy
// part[block] = sum over this block's entries of w_i x_i^2 (fp64).
template <typename T>
__global__ void __launch_bounds__(kBlock) weighted_sq_kernel(int64_t n, const T* __restrict__ x,
                                                             const double* __restrict__ w,
                                                             double* __restrict__ part) {
    __shared__ double red[kBlock / kWave];
    double acc = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double v = (double)x[i];
        acc += w[i] * v * v;
    }
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

}  // namespace tfd
