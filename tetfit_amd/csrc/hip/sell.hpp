// sell.hpp - SELL-64 sparse layout and the kernels that stream it.
//
// Rows are grouped in slices of 64 (one wave64 lane per row); a slice of
// width w (its longest row) stores entry k of its rows at
// sptr[s] + 64*k + lane, so every load of values, column indices and signs
// is a fully coalesced 64-lane access. Short rows are padded with value 0
// and their own last column (a valid index). Within a row the entries keep
// the order of the CSR row they come from, so a row sum runs in the same
// order as a sequential CSR product.
//
// Column indices are stored either as int32, or, when every slice's columns
// lie within 65536 of the slice's smallest column, as a per-slice int32 base
// plus a uint16 offset per entry (2 bytes instead of 4 on every entry).
#pragma once

#include "common.hpp"
#include "kernels.hpp"

namespace tfd {

constexpr int kSlice = 64;

// Device view of a SELL-64 matrix with values of type V (int8 for signed
// incidence, uint16 for bf16 bits). Exactly one of col / col16 is set.
// 16-bit column codes come in two forms: with cwin = 0 the code is the
// offset from the slice's one base cbase[slice] (slices spanning at most
// 65,536 columns); with cwin = 1 the slice has four bases cbase[4 slice + w]
// and the code is w << 14 | offset (each of up to four windows spans at most
// 16,384 columns). Lattice-like meshes numbered plane by plane need the
// second form once a row's neighbours in the planes below and above lie more
// than 65,536 apart (Kuhn meshes from about 70^3): their columns still fall
// into three narrow clusters.
template <typename V>
struct SellView {
    int32_t nrows, nslices;
    const int64_t* sptr;
    const int32_t* col;     // int32 columns, or null
    const uint16_t* col16;  // 16-bit column codes, or null
    const int32_t* cbase;
    const V* val;
    int32_t cwin;
    int32_t band;  // waves take each XCD's slices round-robin (block_slices)
};

// Column of a 16-bit code in the windowed form, with the slice's four bases
// in scalar registers. The base is picked with masks rather than a select
// over the four, which the compiler may lower to a dynamically indexed array
// (placed in LDS, enlarging every workgroup's allocation).
struct ColWin {
    int32_t b0, x01, b2, x23;  // b0, b0 ^ b1, b2, b2 ^ b3
};
template <typename V>
__device__ __forceinline__ ColWin col_windows(const SellView<V>& m, int s) {
    const int32_t* b = m.cbase + 4 * s;
    const int32_t b0 = __builtin_amdgcn_readfirstlane(b[0]), b1 = __builtin_amdgcn_readfirstlane(b[1]);
    const int32_t b2 = __builtin_amdgcn_readfirstlane(b[2]), b3 = __builtin_amdgcn_readfirstlane(b[3]);
    return ColWin{b0, b0 ^ b1, b2, b2 ^ b3};
}
__device__ __forceinline__ int32_t win_col(const ColWin& w, uint32_t code) {
    const int32_t m1 = -(int32_t)((code >> 14) & 1u), m2 = -(int32_t)(code >> 15);
    const int32_t lo = w.b0 ^ (w.x01 & m1), hi = w.b2 ^ (w.x23 & m1);
    return (lo ^ ((lo ^ hi) & m2)) + (int32_t)(code & 0x3FFFu);
}

__device__ __forceinline__ float bf16_to_float(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// Loads of the operator's values and column codes, which each SpMV reads
// exactly once. With NT (the wide-mesh instantiation) they are non-temporal,
// so that the caches evict them first and keep the gathered vector, which
// every row touches again about two mesh planes of rows later: on a mesh
// whose operators exceed the 256 MB Infinity Cache (96^3) the matrix bytes
// streamed in between also exceed an XCD's 4 MB L2, and this cuts the
// gathers' re-fetches. On a smaller mesh the operators stay resident in the
// Infinity Cache from one pass to the next, which non-temporal loads give up.
template <bool NT, typename X>
__device__ __forceinline__ X stream_load(const X* p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

template <typename A, typename V>
__device__ __forceinline__ A sell_value(V v) {
    return (A)v;
}
template <>
__device__ __forceinline__ float sell_value<float, uint16_t>(uint16_t v) {
    return bf16_to_float(v);
}
template <>
__device__ __forceinline__ double sell_value<double, uint16_t>(uint16_t v) {
    return (double)bf16_to_float(v);
}

// Sum over row (slice s, lane) of val * gather(col), accumulated in A (the
// field type: fp64 on the plain path, fp32 on the fast path). The slice is
// uniform across the wave, so its base, width and column base live in
// scalar registers and each entry's loads use one lane offset plus a
// constant offset. Loads are issued in batches of U entries (2U independent
// loads, then U independent gathers) so that each lane keeps enough memory
// traffic in flight; the tail batch is clamped to the row's last entry and
// predicated. Accumulation stays in entry order, so the result does not
// depend on U or on the launch shape. W selects the wide-mesh instantiation:
// the windowed 16-bit column decode compiled in and non-temporal operator
// loads (stream_load); the kernels below take W from their launch.
template <typename A, int U, bool W, typename V, class G>
__device__ __forceinline__ double sell_dot(const SellView<V>& m, int s, int lane, G gather) {
    s = __builtin_amdgcn_readfirstlane(s);
    const int64_t base = m.sptr[s];
    const int w = __builtin_amdgcn_readfirstlane((int)((m.sptr[s + 1] - base) / kSlice));
    const V* v = m.val + base;
    A acc = 0;
    auto batch = [&](int k0, auto col_at, bool full) {
        int32_t ci[U];
        A vi[U], gi[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kk = full ? k0 + u : min(k0 + u, w - 1);  // clamped: loads stay inside the slice
            ci[u] = col_at(kk * kSlice + lane);
            vi[u] = sell_value<A, V>(stream_load<W>(v + kk * kSlice + lane));
        }
#pragma unroll
        for (int u = 0; u < U; ++u) gi[u] = (A)gather(ci[u]);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (full || k0 + u < w) acc += vi[u] * gi[u];
    };
    auto rows = [&](auto col_at) {
        int k0 = 0;
        for (; k0 + U <= w; k0 += U) batch(k0, col_at, true);
        if (k0 < w) batch(k0, col_at, false);
    };
    if (W && m.col16 && m.cwin) {
        const uint16_t* c = m.col16 + base;
        const ColWin cw = col_windows(m, s);
        rows([&](int o) { return win_col(cw, stream_load<W>(c + o)); });
    } else if (m.col16) {
        const uint16_t* c = m.col16 + base;
        const int32_t cb = __builtin_amdgcn_readfirstlane(m.cbase[s]);
        rows([&](int o) { return cb + (int32_t)stream_load<W>(c + o); });
    } else {
        const int32_t* c = m.col + base;
        rows([&](int o) { return stream_load<W>(c + o); });
    }
    return (double)acc;
}

__device__ __forceinline__ int wave_id() { return (blockIdx.x * blockDim.x + threadIdx.x) / kWave; }
__device__ __forceinline__ int wave_count() { return gridDim.x * blockDim.x / kWave; }
__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// Slices of this wave. Each XCD takes one contiguous eighth of the slices,
// so its L2 holds the part of the gathered vector that those rows touch.
// Default: the XCD's part is split into one contiguous chunk per workgroup
// (logical blocks, xcd_block) and the waves of a workgroup take the chunk's
// slices in turn. With band set (batched dga: many small block-diagonal
// instances), the waves of all the XCD's workgroups take its slices
// round-robin, so at any moment they work on one band of consecutive
// slices; with fewer than 8 workgroups the parts are as many as the
// workgroups. Either way every slice goes to exactly one wave.
struct SliceRange {
    int s, end, step;
};
__device__ __forceinline__ SliceRange block_slices(int nslices, int band) {
    const int wpb = (int)(blockDim.x / kWave), w = (int)(threadIdx.x / kWave);
    if (band) {
        const int g = (int)gridDim.x, b = (int)blockIdx.x;
        const int nx = g < kXcd ? g : kXcd, x = b % nx, i = b / nx;
        const int bx = g / nx + (x < g % nx ? 1 : 0);  // workgroups of part x
        const int x0 = (int)((int64_t)nslices * x / nx), x1 = (int)((int64_t)nslices * (x + 1) / nx);
        return SliceRange{x0 + i * wpb + w, x1, bx * wpb};
    }
    const int lb = xcd_block();
    const int s0 = (int)((int64_t)nslices * lb / gridDim.x);
    const int s1 = (int)((int64_t)nslices * (lb + 1) / gridDim.x);
    return SliceRange{s0 + w, s1, wpb};
}

#define TF_FOR_SLICES(m)                                                               \
    for (SliceRange sr_ = block_slices((m).nslices, (m).band); sr_.s < sr_.end; sr_.s += sr_.step)

// ---- software-pipelined slice stream (windowed 16-bit columns) ------------
// A wave's slices as a stream of batches of U entries: (s, 0), (s, U), ...
// up to the slice width, then the next slice. While the gathers of batch b
// are in flight, the entries of batch b+1 are loaded, so each lane keeps two
// batches of loads outstanding instead of waiting out the entry load and
// then the gather of every slice in turn. Vector memory counters retire in
// issue order, so the order of issue is: gathers of b, the row loads of b's
// slice when b is its last batch, entries of b+1; the wait before b's sums
// then leaves b+1's entries in flight. Entries are summed in entry order, so
// the results equal those of the unpipelined slice loop bit for bit.
struct SliceMeta {
    int64_t base;  // first entry of the slice
    int w;         // width (entries per row)
    ColWin cw;     // column windows
};
// A wave-uniform load through the constant address space, which the
// compiler issues as a scalar memory load: slice pointers and column bases
// are never written by the kernels, and a vector load of them would retire
// in order with (and so wait out) the batch loads in flight.
template <typename X>
__device__ __forceinline__ X scalar_load(const X* p) {
    return *(const __attribute__((address_space(4))) X*)p;
}
template <typename V>
__device__ __forceinline__ SliceMeta slice_meta(const SellView<V>& m, int s) {
    s = __builtin_amdgcn_readfirstlane(s);
    const int64_t base = scalar_load(m.sptr + s);
    const int w = (int)((scalar_load(m.sptr + s + 1) - base) / kSlice);
    const int32_t* b = m.cbase + 4 * s;
    const int32_t b0 = scalar_load(b), b1 = scalar_load(b + 1), b2 = scalar_load(b + 2), b3 = scalar_load(b + 3);
    return SliceMeta{base, w, ColWin{b0, b0 ^ b1, b2, b2 ^ b3}};
}

// E: one entry as loaded (column code and value); load(pos) loads the entry
// at SELL position pos; gather(e, cw) fetches the gathered operand of e;
// add(acc, e, g) accumulates; row_pre(s) issues the loads of slice s's row
// vectors and returns them; row_post(s, acc, rin) finishes slice s;
// next(sn), when given, runs once the next slice sn is known, before its
// entries are loaded (loads of a row indirection one slice ahead).
struct NoNext {
    __device__ __forceinline__ void operator()(int) const {}
};
template <int U, typename E, typename Acc, typename V>
struct SellStream {
    const SellView<V>& m;
    SliceRange sr;
    int lane;
    SliceMeta cur;
    int k0 = 0;
    E e[U];

    __device__ __forceinline__ SellStream(const SellView<V>& mv, SliceRange r, int l) : m(mv), sr(r), lane(l) {}

    template <class Load>
    __device__ __forceinline__ void fetch(const SliceMeta& mm, int kk0, E (&ee)[U], Load load) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kk = min(kk0 + u, mm.w - 1);  // clamped: loads stay inside the slice
            ee[u] = load(mm.base + (int64_t)kk * kSlice + lane);
        }
    }

    // Issues the first batch's loads (before a reduction the caller waits on).
    template <class Load>
    __device__ __forceinline__ void start(Load load) {
        if (sr.s >= sr.end) return;
        cur = slice_meta(m, sr.s);
        k0 = 0;
        fetch(cur, 0, e, load);
    }

    // Every slice must have width >= 1 (the host checks it for the matrices
    // streamed this way). The loop body issues the same loads on every pass,
    // whatever the batch's position in its slice (the row loads of a
    // slice's earlier batches and the last pass's re-fetch of its own batch
    // are spare loads of cached lines), so the compiler's wait before the
    // sums can leave exactly the next batch's U loads in flight.
    template <class Load, class Gather, class Add, class RowPre, class RowPost, class Next = NoNext>
    __device__ __forceinline__ void run(Load load, Gather gather, Add add, RowPre row_pre, RowPost row_post,
                                        Next next = Next{}) {
        int s = sr.s;
        if (s >= sr.end) return;
        Acc acc{};
        for (;;) {
            decltype(gather(e[0], cur.cw)) g[U];
#pragma unroll
            for (int u = 0; u < U; ++u) g[u] = gather(e[u], cur.cw);
            const auto rin = row_pre(s);
            const bool last = k0 + U >= cur.w;
            const int sn = last ? s + sr.step : s;
            const bool more = sn < sr.end;
            SliceMeta mn = cur;
            int kn = k0;
            if (more) {
                kn = last ? 0 : k0 + U;
                if (last) {
                    mn = slice_meta(m, sn);
                    next(sn);
                }
            }
            E en[U];
            fetch(mn, kn, en, load);
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (k0 + u < cur.w) add(acc, e[u], g[u]);
            if (last) {
                row_post(s, acc, rin);
                acc = Acc{};
            }
            if (!more) break;
            s = sn;
            k0 = kn;
            cur = mn;
#pragma unroll
            for (int u = 0; u < U; ++u) e[u] = en[u];
        }
    }
};

// curl_f = Ct F over edges (dga, D:878-879).
template <typename T, int NB, bool W = false>
__global__ void __launch_bounds__(kBlock) sell_ct_kernel(SellView<int8_t> m, const T* __restrict__ x,
                                                         T* __restrict__ y) {
    const int lane = lane_id();
    if constexpr (W) {
        if (m.col16 && m.cwin) {
            struct Ent { uint16_t c; int8_t v; };
            SellStream<NB, Ent, T, int8_t> ss(m, block_slices(m.nslices, m.band), lane);
            auto load = [&](int64_t pos) { return Ent{stream_load<W>(m.col16 + pos), stream_load<W>(m.val + pos)}; };
            ss.start(load);
            ss.run(
                load, [&](const Ent& e, const ColWin& cw) { return x[win_col(cw, e.c)]; },
                [&](T& a, const Ent& e, T gv) { a += (T)e.v * gv; }, [&](int) { return 0; },
                [&](int sl, T v, int) {
                    const int32_t r = sl * kSlice + lane;
                    if (r < m.nrows) y[r] = v;
                });
            return;
        }
    }
    TF_FOR_SLICES(m) {
        const int s = sr_.s;
        const double v = sell_dot<T, NB, W>(m, s, lane, [&](int32_t j) { return x[j]; });
        const int32_t r = s * kSlice + lane;
        if (r < m.nrows) y[r] = (T)v;
    }
}

// U[edge(r)] += dt * (H curl_f)[r] (dga, D:889); row_dt (batches) gives
// each row its own simulation's dt.
template <typename T, int NB, bool W = false>
__global__ void __launch_bounds__(kBlock) sell_h_kernel(SellView<T> m, const int32_t* __restrict__ row_edge,
                                                        const T* __restrict__ x, T* __restrict__ U, double dt,
                                                        const double* __restrict__ row_dt) {
    const int lane = lane_id();
    if constexpr (W) {
        if (m.col16 && m.cwin) {
            struct Ent { uint16_t c; T v; };
            const SliceRange sr = block_slices(m.nslices, m.band);
            // edge of the lane's row, one slice ahead (clamped row past the end)
            auto edge_of = [&](int sl) { return row_edge[min(sl * kSlice + lane, m.nrows - 1)]; };
            int32_t e_cur = sr.s < sr.end ? edge_of(sr.s) : 0, e_nxt = 0;
            SellStream<NB, Ent, T, T> ss(m, sr, lane);
            auto load = [&](int64_t pos) { return Ent{stream_load<W>(m.col16 + pos), stream_load<W>(m.val + pos)}; };
            ss.start(load);
            ss.run(
                load, [&](const Ent& e, const ColWin& cw) { return x[win_col(cw, e.c)]; },
                [&](T& a, const Ent& e, T gv) { a += e.v * gv; }, [&](int) { return U[e_cur]; },
                [&](int sl, T a, T uold) {
                    const int32_t r = sl * kSlice + lane;
                    if (r < m.nrows) U[e_cur] = (T)((double)uold + (row_dt ? row_dt[r] : dt) * (double)a);
                    e_cur = e_nxt;
                },
                [&](int sn) { e_nxt = edge_of(sn); });
            return;
        }
    }
    TF_FOR_SLICES(m) {
        const int s = sr_.s;
        const double v = sell_dot<T, NB, W>(m, s, lane, [&](int32_t j) { return x[j]; });
        const int32_t r = s * kSlice + lane;
        if (r < m.nrows) {
            const int32_t e = row_edge[r];
            U[e] = (T)((double)U[e] + (row_dt ? row_dt[r] : dt) * v);
        }
    }
}

// Lossy dga, one fractional row r per lane (D:888-890, D:916-919):
// Y[r] = (R X)[r] + sdt (K v)[r], K the row of H / Mp / Mq (N / Tr / Ts)
// applied to curl_f (curl_u), R the row's old entries (the identity, P_p
// or R_r, a node-star Q or tet S block) applied to the old U_frac (F_frac);
// sdt = dt for U, -dt for F. K and R have the same rows.
template <typename T, int NB, bool W = false>
__global__ void __launch_bounds__(kBlock) sell_frac_kernel(SellView<T> K, SellView<T> R, const T* __restrict__ v,
                                                           const T* __restrict__ X, T* __restrict__ Y, double sdt) {
    const int lane = lane_id();
    TF_FOR_SLICES(K) {
        const int s = sr_.s;
        const double kv = sell_dot<T, NB, W>(K, s, lane, [&](int32_t j) { return v[j]; });
        const double rx = sell_dot<T, NB, W>(R, s, lane, [&](int32_t j) { return X[j]; });
        const int32_t r = s * kSlice + lane;
        if (r < K.nrows) Y[r] = (T)(rx + sdt * kv);
    }
}

// mode 0: F -= dt * N x (dga, D:917); mode 1: F = N x (gfdtd, D:1468);
// row_dt as in sell_h_kernel; guard as in cg_guard_ok.
template <typename T, int NB, bool W = false>
__global__ void __launch_bounds__(kBlock) sell_n_kernel(SellView<T> m, const T* __restrict__ x, T* __restrict__ Fv,
                                                        double dt, int mode, const double* __restrict__ row_dt,
                                                        const CgState* guard) {
    if (!cg_guard_ok(guard)) return;
    const int lane = lane_id();
    // (the slice stream measured slower here: N's rows hold about 7 entries,
    // one batch, and its extra registers cost more occupancy than the
    // overlap gains)
    TF_FOR_SLICES(m) {
        const int s = sr_.s;
        const double v = sell_dot<T, NB, W>(m, s, lane, [&](int32_t j) { return x[j]; });
        const int32_t r = s * kSlice + lane;
        if (r < m.nrows) Fv[r] = mode == 0 ? (T)((double)Fv[r] - (row_dt ? row_dt[r] : dt) * v) : (T)v;
    }
}

// part[block] = sum over the block's rows r of x[r] (M x)[r], accumulated
// in fp64: twice the discrete energy x.(M x) / 2 of a symmetric mass matrix
// (We = 1/2 U.E U, Wh = 1/2 B.N B, D:2773-2775), or the Joule term
// U.SigMat U of gfdtd's losses (D:1519, D:1544); guard as in cg_guard_ok.
template <typename T, int NB, bool W = false>
__global__ void __launch_bounds__(kBlock) quad_form_kernel(SellView<T> m, const T* __restrict__ x,
                                                           double* __restrict__ part, const CgState* guard) {
    if (!cg_guard_ok(guard)) return;
    __shared__ double red[kBlock / kWave];
    const int lane = lane_id();
    double acc = 0;
    TF_FOR_SLICES(m) {
        const int s = sr_.s;
        const double mx = sell_dot<double, NB, W>(m, s, lane, [&](int32_t j) { return x[j]; });
        const int32_t r = s * kSlice + lane;
        if (r < m.nrows) acc += (double)x[r] * mx;
    }
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// Energies of one output sample from the partials of the two reductions:
// out = (1/2 sum pe[0..ge), 1/2 sum pn[0..gn)), each sum taken in partial
// order on one thread (the order in which tf_energy sums them on the host).
__global__ void __launch_bounds__(64) energy_sum_kernel(const double* __restrict__ pe, int ge,
                                                        const double* __restrict__ pn, int gn,
                                                        double* __restrict__ out, const CgState* guard) {
    if (!cg_guard_ok(guard) || threadIdx.x != 0) return;
    double a = 0, b = 0;
    for (int k = 0; k < ge; ++k) a += pe[k];
    for (int k = 0; k < gn; ++k) b += pn[k];
    out[0] = 0.5 * a;
    out[1] = 0.5 * b;
}

// ---- CG (ConjugateGradientSolver.hpp:52-103) ------------------------------
// Preconditioned CG in the single-reduction form of Chronopoulos and Gear
// (1989): with u = M r, w = A u, gamma = r.u, delta = w.u,
//   beta_i  = gamma_i / gamma_{i-1},  alpha_i = gamma_i / (delta_i - beta_i gamma_i / alpha_{i-1}),
//   p = u + beta p,  s = w + beta s  (s = A p),  x += alpha p,  r -= alpha s.
// In exact arithmetic the iterates are those of the reference's PCG, and the
// stop test is the reference's: ||r|| / ||r0|| > tol && it < maxiter.
// Kernels per iteration:
//   Jacobi (M = 1/diag A):  cg_w (stop test, w = A u), cg_upd (updates, u = M r)
//   learned (M = G, SpMV):  cg_pg (x, r updates, u = G r), cg_wu (stop test,
//                           w = A u, p and s updates)
// Reductions go through per-block fp64 partials that the next kernel sums in
// a fixed order (no atomics, no grid-wide fences).

// ---- scalar prologue of the CG kernels
// A kernel needs global scalars (sums of the previous kernel's per-block
// partials, and values in the CgState) before its first store. Every load
// of the prologue is issued before any of them is used: each thread holds
// its share of up to kMaxPart partials per array in registers (the same
// per-thread order as a strided loop, so the sums are unchanged), and both
// arrays are reduced in one block reduction. The prologue then costs one
// memory round trip instead of one per loop trip and per scalar.
constexpr int kMaxPart = 2048;
constexpr int kPartPer = kMaxPart / kBlock;

struct Partials {
    double v[kPartPer];
};
// Loads without branches: an index past n reads p[0] (every partial array
// is allocated with at least one entry) and is masked to zero.
__device__ __forceinline__ Partials load_partials(const double* __restrict__ p, int n) {
    Partials r;
    double t[kPartPer];
#pragma unroll
    for (int k = 0; k < kPartPer; ++k) {
        const int i = threadIdx.x + k * kBlock;
        t[k] = p[i < n ? i : 0];
    }
#pragma unroll
    for (int k = 0; k < kPartPer; ++k) r.v[k] = threadIdx.x + k * kBlock < n ? t[k] : 0.0;
    return r;
}
__device__ __forceinline__ double thread_sum(const Partials& r) {
    double s = 0;
#pragma unroll
    for (int k = 0; k < kPartPer; ++k) s += r.v[k];
    return s;
}

// Block-wide sums of two per-thread values, returned to every thread.
// red holds 2 * (waves per block) + 2 doubles.
__device__ __forceinline__ void block_sum2_all(double& a, double& b, double* red) {
    constexpr int nw = kBlock / kWave;
    a = group_sum<kWave>(a);
    b = group_sum<kWave>(b);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    if (lane == 0) {
        red[w] = a;
        red[nw + w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0, sb = 0;
        for (int i = 0; i < nw; ++i) {
            sa += red[i];
            sb += red[nw + i];
        }
        red[2 * nw] = sa;
        red[2 * nw + 1] = sb;
    }
    __syncthreads();
    a = red[2 * nw];
    b = red[2 * nw + 1];
    __syncthreads();
}
constexpr int kRed2 = 2 * (kBlock / kWave) + 2;

// Stop decision of pass `it` from the residual norm RR = ||r_it||^2 (the
// reference's test, ConjugateGradientSolver.hpp:68-74): stop unless
// ||r|| / ||r0|| > tol and it < maxiter. Every block decides the same;
// block 0 records it. A zero initial residual (a quiet step: the warm start
// already solves the system) stops at once as in the reference, whose 0/0
// fails the `> tol` test; its ratio is recorded as 0 so that a learned solve
// counts as converged. After an extrapolated start (rr0_given) r0 is the
// warm-start residual b - A U_old set before pass 0, so the test is the
// reference's for the warm start while the iterates start elsewhere; a zero
// warm-start residual there falls back to the start's own ||r_0||.
__device__ __forceinline__ bool cg_decide_stop(int32_t it, double RR, double rr0_prev, int32_t rr0_given,
                                               double tol, int32_t maxiter, CgState* st, bool leader) {
    const bool own = it == 0 && !(rr0_given && rr0_prev > 0.0);
    const double rr0 = own ? RR : rr0_prev;
    const double ratio = rr0 == 0.0 ? 0.0 : std::sqrt(RR) / std::sqrt(rr0);
    const bool stop = !(ratio > tol && it < maxiter);
    if (threadIdx.x == 0 && leader) {
        if (own) st->rr0 = RR;
        st->relres = ratio;
        if (stop) { st->done = 1; st->iter = it; }
    }
    return stop;
}

// Residual and search-direction image of one row, stored side by side for
// the learned passes: cg_pg gathers both with one 8- (fp32) or 16-byte
// (fp64) load per matrix entry.
template <typename T>
struct alignas(2 * sizeof(T)) RS {
    T r, s;
};

// The residual update r - alpha s in one rounding, used identically by the
// lane that owns a row and by every lane that gathers it.
template <typename T>
__device__ __forceinline__ T cg_r_next(T r, T s, double alpha) {
    return (T)__builtin_fma(-alpha, (double)s, (double)r);
}

// Batched solves (tf_create_batch of gfdtd instances): independent systems
// stacked block-diagonally, each instance's rows a whole number of 64-row
// slices. Workgroup b works on slices [s0[b], s1[b]) of instance inst[b];
// the per-workgroup partials of instance i are those of workgroups
// [w0[i], w1[i]) (the same table for every kernel of a solve), so each
// instance sums only its own partials in a fixed order, and its state is
// st + 2 i (two slots per instance). inst == null: one solve over the grid.
struct CgBatch {
    const int32_t* inst = nullptr;
    const int32_t* s0 = nullptr;
    const int32_t* s1 = nullptr;
    const int32_t* w0 = nullptr;
    const int32_t* w1 = nullptr;
};
struct CgSeg {
    int p0, n;       // partials of this workgroup's solve
    CgState* st;
    bool leader;     // the workgroup that records the solve's scalars
};
__device__ __forceinline__ CgSeg cg_seg(const CgBatch& bt, int nb, CgState* st) {
    if (!bt.inst) return CgSeg{0, nb, st, blockIdx.x == 0};
    const int i = bt.inst[blockIdx.x], w0 = bt.w0[i];
    return CgSeg{w0, bt.w1[i] - w0, st + 2 * i, (int)blockIdx.x == w0};
}
__device__ __forceinline__ SliceRange cg_slices(int nslices, int band, const CgBatch& bt) {
    if (!bt.inst) return block_slices(nslices, band);
    const int wpb = (int)(blockDim.x / kWave), w = (int)(threadIdx.x / kWave);
    return SliceRange{bt.s0[blockIdx.x] + w, bt.s1[blockIdx.x], wpb};
}

// Restart of a solve from its saved initial residual: r = r0 (r[i * r_stride]),
// partial r.r; with jacobi also u = pre r and partial r.u. The solver's x
// must hold the same warm start as when r0 was formed.
template <typename T>
__global__ void __launch_bounds__(kBlock) cg_restart_kernel(int32_t n, const T* __restrict__ pre,
                                                            const T* __restrict__ r0, T* __restrict__ r,
                                                            int r_stride, T* __restrict__ u,
                                                            double* __restrict__ part_rr,
                                                            double* __restrict__ part_gam, CgState* st, int jacobi) {
    __shared__ double red[kRed2];
    double rr = 0, ru = 0;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const T ri = r0[i];
        r[(int64_t)i * r_stride] = ri;
        rr += (double)ri * (double)ri;
        if (jacobi) {
            const T ui = (T)((double)pre[i] * (double)ri);
            u[i] = ui;
            ru += (double)ri * (double)ui;
        }
    }
    block_sum2_all(rr, ru, red);
    if (threadIdx.x == 0) {
        part_rr[blockIdx.x] = rr;
        if (jacobi) part_gam[blockIdx.x] = ru;
        if (blockIdx.x == 0) { st->done = 0; st->iter = 0; }
    }
}

// w = A u, partial w.u; with test_stop the pass starts with the stop test on
// the partials part_rr of the previous update.
template <typename T, int NB, bool W = false>
__global__ void __launch_bounds__(kBlock) cg_w_kernel(int32_t it, SellView<T> m, const T* __restrict__ u,
                                                      T* __restrict__ w, int test_stop,
                                                      const double* __restrict__ part_rr, int32_t nb_rr,
                                                      double* __restrict__ part_del, CgState* st_in,
                                                      CgBatch bt) {
    __shared__ double red[kRed2];
    const CgSeg sg = cg_seg(bt, nb_rr, st_in);
    CgState* st = sg.st;
    const Partials prr = load_partials(part_rr + sg.p0, test_stop ? sg.n : 0);
    const int done = st->done;
    const double rr0 = st->rr0, tol = st->tol;
    const int32_t maxiter = st->maxiter, given = st->rr0_given;
    if (done) {
        // a batched instance that has stopped leaves zero partials for the
        // kernels of its later passes (they return at its done flag too)
        return;
    }
    if (test_stop) {
        double RR = thread_sum(prr), unused = 0;
        block_sum2_all(RR, unused, red);
        if (cg_decide_stop(it, RR, rr0, given, tol, maxiter, st, sg.leader)) return;
    }
    const int lane = lane_id();
    double wu = 0, none = 0;
    const SliceRange srng = cg_slices(m.nslices, m.band, bt);
    if constexpr (W) {
        if (m.col16 && m.cwin) {
            struct Ent { uint16_t c; T v; };
            SellStream<NB, Ent, T, T> ss(m, srng, lane);
            auto load = [&](int64_t pos) { return Ent{stream_load<W>(m.col16 + pos), stream_load<W>(m.val + pos)}; };
            ss.start(load);
            ss.run(
                load, [&](const Ent& e, const ColWin& cw) { return u[win_col(cw, e.c)]; },
                [&](T& a, const Ent& e, T gv) { a += e.v * gv; },
                [&](int sl) { return u[min(sl * kSlice + lane, m.nrows - 1)]; },
                [&](int sl, T wi, T ui) {
                    const int32_t i = sl * kSlice + lane;
                    if (i < m.nrows) {
                        w[i] = wi;
                        wu += (double)wi * (double)ui;
                    }
                });
            block_sum2_all(wu, none, red);
            if (threadIdx.x == 0) part_del[blockIdx.x] = wu;
            return;
        }
    }
    for (SliceRange sr_ = srng; sr_.s < sr_.end; sr_.s += sr_.step) {
        const int s = sr_.s;
        const double wv = sell_dot<T, NB, W>(m, s, lane, [&](int32_t j) { return u[j]; });
        const int32_t i = s * kSlice + lane;
        if (i < m.nrows) {
            const T wi = (T)wv;
            w[i] = wi;
            wu += (double)wi * (double)u[i];
        }
    }
    block_sum2_all(wu, none, red);
    if (threadIdx.x == 0) part_del[blockIdx.x] = wu;
}

// Jacobi vector updates of pass `it`; gamma and delta from their partials,
// the previous gamma and alpha from slot (it-1)&1 of the state; then the
// next u = pre r and its partial r.u go to part_gam_next.
template <typename T>
__global__ void __launch_bounds__(kBlock) cg_upd_kernel(int32_t it, int32_t n, const T* __restrict__ pre,
                                                        T* __restrict__ x, T* __restrict__ r, T* __restrict__ u,
                                                        const T* __restrict__ w, T* __restrict__ p, T* __restrict__ s,
                                                        const double* __restrict__ part_gam, int32_t nb_gam,
                                                        const double* __restrict__ part_del, int32_t nb_del,
                                                        double* __restrict__ part_rr, double* __restrict__ part_gam_next,
                                                        CgState* st_in, CgBatch bt) {
    __shared__ double red[kRed2];
    const CgSeg sg = cg_seg(bt, nb_gam, st_in);
    CgState* st = sg.st;
    // batched: the same workgroup table for both partial arrays
    const Partials pg = load_partials(part_gam + sg.p0, sg.n),
                   pd = load_partials(part_del + (bt.inst ? sg.p0 : 0), bt.inst ? sg.n : nb_del);
    const int done = st->done;
    const double gam_prev = st->gam[(it - 1) & 1], alpha_prev = st->alpha[(it - 1) & 1];
    if (done) return;
    // rows: a grid-stride loop, or (batched) the rows of the workgroup's slices
    const int32_t rstep = bt.inst ? (int32_t)blockDim.x : (int32_t)(gridDim.x * blockDim.x);
    const int32_t i0 = bt.inst ? bt.s0[blockIdx.x] * kSlice + (int32_t)threadIdx.x
                               : (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    const int32_t iend = bt.inst ? min(n, bt.s1[blockIdx.x] * kSlice) : n;
    // the thread's first element is loaded before the partial sums are reduced
    T u0 = 0, w0 = 0, p0 = 0, s0 = 0, x0 = 0, r0 = 0, q0 = 0;
    if (i0 < iend) {
        u0 = u[i0]; w0 = w[i0]; p0 = p[i0]; s0 = s[i0]; x0 = x[i0]; r0 = r[i0]; q0 = pre[i0];
    }
    double gam = thread_sum(pg), del = thread_sum(pd);
    block_sum2_all(gam, del, red);
    double alpha, beta;
    if (it == 0) {
        beta = 0.0;
        alpha = gam / del;
    } else {
        beta = gam / gam_prev;
        alpha = gam / (del - beta * gam / alpha_prev);
    }
    if (threadIdx.x == 0 && sg.leader) {
        st->gam[it & 1] = gam;
        st->alpha[it & 1] = alpha;
    }
    double rr = 0, ru = 0;
    for (int32_t i = i0; i < iend; i += rstep) {
        if (i != i0) {
            u0 = u[i]; w0 = w[i]; p0 = p[i]; s0 = s[i]; x0 = x[i]; r0 = r[i]; q0 = pre[i];
        }
        const T pi = it == 0 ? u0 : (T)((double)u0 + beta * (double)p0);
        const T si = it == 0 ? w0 : (T)((double)w0 + beta * (double)s0);
        p[i] = pi;
        s[i] = si;
        x[i] = (T)((double)x0 + alpha * (double)pi);
        const T ri = (T)((double)r0 - alpha * (double)si);
        r[i] = ri;
        rr += (double)ri * (double)ri;
        const T ui = (T)((double)q0 * (double)ri);
        u[i] = ui;
        ru += (double)ri * (double)ui;
    }
    block_sum2_all(rr, ru, red);
    if (threadIdx.x == 0) {
        part_rr[blockIdx.x] = rr;
        part_gam_next[blockIdx.x] = ru;
    }
}

// Both products of one row over the gathered (r, s) pairs: gr = (G r)_i and
// gs = (G s)_i, so that u = G (r - alpha s) = gr - alpha gs needs alpha only
// after the gathers. G is stored packed, 32 bits per entry (A's 16-bit
// column code << 16 | bf16 value), when A has 16-bit columns (pk), otherwise as
// bf16 values beside A's column indices (m.val).
template <typename T, int U, bool W>
__device__ __forceinline__ void sell_dot_rs(const SellView<uint16_t>& m, const uint32_t* __restrict__ pk, int s,
                                            int lane, const RS<T>* __restrict__ rs, T& gr, T& gs) {
    s = __builtin_amdgcn_readfirstlane(s);
    const int64_t base = m.sptr[s];
    const int w = __builtin_amdgcn_readfirstlane((int)((m.sptr[s + 1] - base) / kSlice));
    T ar = 0, as = 0;
    auto batch = [&](int k0, auto entry, bool full) {
        int32_t ci[U];
        T vi[U];
        RS<T> gi[U];
#pragma unroll
        for (int u = 0; u < U; ++u) entry((full ? k0 + u : min(k0 + u, w - 1)) * kSlice + lane, ci[u], vi[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) gi[u] = rs[ci[u]];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (full || k0 + u < w) {
                ar += vi[u] * gi[u].r;
                as += vi[u] * gi[u].s;
            }
    };
    auto rows = [&](auto entry) {
        int k0 = 0;
        for (; k0 + U <= w; k0 += U) batch(k0, entry, true);
        if (k0 < w) batch(k0, entry, false);
    };
    if (W && pk && m.cwin) {
        const uint32_t* q = pk + base;
        const ColWin cw = col_windows(m, s);
        rows([&](int o, int32_t& c, T& v) {
            const uint32_t e = stream_load<W>(q + o);
            c = win_col(cw, e >> 16);
            v = (T)bf16_to_float((uint16_t)(e & 0xFFFFu));
        });
    } else if (pk) {
        const uint32_t* q = pk + base;
        const int32_t cb = __builtin_amdgcn_readfirstlane(m.cbase[s]);
        rows([&](int o, int32_t& c, T& v) {
            const uint32_t e = stream_load<W>(q + o);
            c = cb + (int32_t)(e >> 16);
            v = (T)bf16_to_float((uint16_t)(e & 0xFFFFu));
        });
    } else if (W && m.col16 && m.cwin) {
        const uint16_t* c16 = m.col16 + base;
        const uint16_t* vv = m.val + base;
        const ColWin cw = col_windows(m, s);
        rows([&](int o, int32_t& c, T& v) {
            c = win_col(cw, stream_load<W>(c16 + o));
            v = (T)bf16_to_float(stream_load<W>(vv + o));
        });
    } else if (m.col16) {
        const uint16_t* c16 = m.col16 + base;
        const uint16_t* vv = m.val + base;
        const int32_t cb = __builtin_amdgcn_readfirstlane(m.cbase[s]);
        rows([&](int o, int32_t& c, T& v) {
            c = cb + (int32_t)stream_load<W>(c16 + o);
            v = (T)bf16_to_float(stream_load<W>(vv + o));
        });
    } else {
        const int32_t* c32 = m.col + base;
        const uint16_t* vv = m.val + base;
        rows([&](int o, int32_t& c, T& v) {
            c = stream_load<W>(c32 + o);
            v = (T)bf16_to_float(stream_load<W>(vv + o));
        });
    }
    gr = ar;
    gs = as;
}

// Learned pass `it`, first kernel. For it >= 1: alpha_{it-1} from gamma and
// beta of pass it-1 (state slot (it-1)&1), the partials of delta_{it-1}
// (part_del) and alpha_{it-2}; then on the owned rows x += alpha p and
// r_it = r_{it-1} - alpha s_{it-1} (rs_in holds (r_{it-1}, s_{it-1}); r_it
// goes to rs_out[i].r), and u = G r_it = G r_{it-1} - alpha G s_{it-1};
// partials r.u and r.r. For it = 0, u = G r with r = rs_in[].r. The
// products of the wave's first slice do not need alpha and are issued
// before the partial sums are reduced. G: bf16 values on A's pattern
// (sell_dot_rs).
template <typename T, int NB, bool W = false>
__global__ void __launch_bounds__(kBlock) cg_pg_kernel(int32_t it, SellView<uint16_t> g,
                                                       const uint32_t* __restrict__ gpk,
                                                       const RS<T>* __restrict__ rs_in, RS<T>* __restrict__ rs_out,
                                                       const T* __restrict__ p, T* __restrict__ x, T* __restrict__ u,
                                                       const double* __restrict__ part_del, int32_t nb_del,
                                                       double* __restrict__ part_gam, double* __restrict__ part_rr,
                                                       CgState* st) {
    __shared__ double red[kRed2];
    const int k = (it - 1) & 1;
    const Partials pd = load_partials(part_del, it > 0 ? nb_del : 0);
    const int done = st->done;
    const double gam = st->gam[k], beta = st->beta[k], alpha_pp = st->alpha[k ^ 1];
    if (done) return;
    const int lane = lane_id();
    SliceRange sr = block_slices(g.nslices, g.band);
    double alpha = 0;
    auto reduce_alpha = [&] {
        if (it > 0) {
            double del = thread_sum(pd), unused = 0;
            block_sum2_all(del, unused, red);
            alpha = it == 1 ? gam / del : gam / (del - beta * gam / alpha_pp);
            if (threadIdx.x == 0 && blockIdx.x == 0) st->alpha[k] = alpha;
        }
    };
    double ru = 0, rr = 0;
    if constexpr (W) {
        if (gpk && g.cwin) {
            // the slice stream: packed entries (A's column code << 16 | bf16 G value)
            struct Acc { T r, s; };
            struct RowIn { RS<T> o; T pv, xv; };
            SellStream<NB, uint32_t, Acc, uint16_t> ss(g, sr, lane);
            auto load = [&](int64_t pos) { return stream_load<W>(gpk + pos); };
            ss.start(load);  // in flight during the reduction
            reduce_alpha();
            ss.run(
                load, [&](uint32_t e, const ColWin& cw) { return rs_in[win_col(cw, e >> 16)]; },
                [&](Acc& a, uint32_t e, const RS<T>& gv) {
                    const T v = (T)bf16_to_float((uint16_t)(e & 0xFFFFu));
                    a.r += v * gv.r;
                    a.s += v * gv.s;
                },
                [&](int sl) {
                    // unconditional loads (a clamped row past the end): no
                    // branch whose join would wait on them
                    const int32_t i = min(sl * kSlice + lane, g.nrows - 1);
                    return RowIn{rs_in[i], p[i], x[i]};
                },
                [&](int sl, const Acc& a, const RowIn& q) {
                    const int32_t i = sl * kSlice + lane;
                    if (i < g.nrows) {
                        T ri = q.o.r, ui = a.r;
                        if (it > 0) {
                            ri = cg_r_next(q.o.r, q.o.s, alpha);
                            rs_out[i] = RS<T>{ri, q.o.s};
                            x[i] = (T)((double)q.xv + alpha * (double)q.pv);
                            ui = (T)((double)a.r - alpha * (double)a.s);
                        }
                        u[i] = ui;
                        ru += (double)ri * (double)ui;
                        rr += (double)ri * (double)ri;
                    }
                });
            block_sum2_all(ru, rr, red);
            if (threadIdx.x == 0) {
                part_gam[blockIdx.x] = ru;
                part_rr[blockIdx.x] = rr;
            }
            return;
        }
    }
    const bool first = sr.s < sr.end;
    T gr0 = 0, gs0 = 0;
    if (first) sell_dot_rs<T, NB, W>(g, gpk, sr.s, lane, rs_in, gr0, gs0);
    reduce_alpha();
    auto row = [&](int sl, T gr, T gs) {
        const int32_t i = sl * kSlice + lane;
        if (i < g.nrows) {
            const RS<T> o = rs_in[i];
            T ri = o.r, ui = gr;
            if (it > 0) {
                ri = cg_r_next(o.r, o.s, alpha);
                rs_out[i] = RS<T>{ri, o.s};  // whole pair: s is replaced by cg_wu of this pass
                x[i] = (T)((double)x[i] + alpha * (double)p[i]);
                ui = (T)((double)gr - alpha * (double)gs);
            }
            u[i] = ui;
            ru += (double)ri * (double)ui;
            rr += (double)ri * (double)ri;
        }
    };
    if (first) {
        row(sr.s, gr0, gs0);
        sr.s += sr.step;
    }
    for (; sr.s < sr.end; sr.s += sr.step) {
        T gr, gs;
        sell_dot_rs<T, NB, W>(g, gpk, sr.s, lane, rs_in, gr, gs);
        row(sr.s, gr, gs);
    }
    block_sum2_all(ru, rr, red);
    if (threadIdx.x == 0) {
        part_gam[blockIdx.x] = ru;
        part_rr[blockIdx.x] = rr;
    }
}

// Learned pass `it`, second kernel: stop test on r_it (part_rr), gamma_it
// and beta_it (stored in slot it&1), w = A u, p = u + beta p and
// s_it = w + beta s_{it-1} on the owned rows (rs_cur[i] holds (r_it,
// s_{it-1}) as cg_pg wrote it; s_it replaces its s), partial w.u (delta_it).
template <typename T, int NB, bool W = false>
__global__ void __launch_bounds__(kBlock) cg_wu_kernel(int32_t it, SellView<T> m, const T* __restrict__ u,
                                                       T* __restrict__ p, RS<T>* __restrict__ rs_cur,
                                                       const double* __restrict__ part_rr,
                                                       const double* __restrict__ part_gam, int32_t nb_pg,
                                                       double* __restrict__ part_del, CgState* st) {
    __shared__ double red[kRed2];
    const Partials prr = load_partials(part_rr, nb_pg), pga = load_partials(part_gam, nb_pg);
    const int done = st->done;
    const double rr0 = st->rr0, tol = st->tol, gam_prev = st->gam[(it - 1) & 1];
    const int32_t maxiter = st->maxiter, given = st->rr0_given;
    if (done) return;
    double RR = thread_sum(prr), gam = thread_sum(pga);
    block_sum2_all(RR, gam, red);
    if (cg_decide_stop(it, RR, rr0, given, tol, maxiter, st, blockIdx.x == 0)) return;
    const double beta = it == 0 ? 0.0 : gam / gam_prev;
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        st->gam[it & 1] = gam;
        st->beta[it & 1] = beta;
    }
    const int lane = lane_id();
    double wu = 0, none = 0;
    if constexpr (W) {
        if (m.col16 && m.cwin) {
            struct Ent { uint16_t c; T v; };
            struct RowIn { T ui, pv; RS<T> o; };
            SellStream<NB, Ent, T, T> ss(m, block_slices(m.nslices, m.band), lane);
            auto load = [&](int64_t pos) { return Ent{stream_load<W>(m.col16 + pos), stream_load<W>(m.val + pos)}; };
            ss.start(load);
            ss.run(
                load, [&](const Ent& e, const ColWin& cw) { return u[win_col(cw, e.c)]; },
                [&](T& a, const Ent& e, T gv) { a += e.v * gv; },
                [&](int sl) {
                    const int32_t i = min(sl * kSlice + lane, m.nrows - 1);
                    return RowIn{u[i], p[i], rs_cur[i]};
                },
                [&](int sl, T wi, const RowIn& q) {
                    const int32_t i = sl * kSlice + lane;
                    if (i < m.nrows) {
                        if (it == 0) {
                            p[i] = q.ui;
                            rs_cur[i] = RS<T>{q.o.r, wi};
                        } else {
                            p[i] = (T)((double)q.ui + beta * (double)q.pv);
                            rs_cur[i] = RS<T>{q.o.r, (T)((double)wi + beta * (double)q.o.s)};
                        }
                        wu += (double)wi * (double)q.ui;
                    }
                });
            block_sum2_all(wu, none, red);
            if (threadIdx.x == 0) part_del[blockIdx.x] = wu;
            return;
        }
    }
    TF_FOR_SLICES(m) {
        const int sl = sr_.s;
        const double wv = sell_dot<T, NB, W>(m, sl, lane, [&](int32_t j) { return u[j]; });
        const int32_t i = sl * kSlice + lane;
        if (i < m.nrows) {
            const T wi = (T)wv;
            const T ui = u[i];
            const RS<T> o = rs_cur[i];  // (r', s of the previous pass), as cg_pg wrote it
            if (it == 0) {
                p[i] = ui;
                rs_cur[i] = RS<T>{o.r, wi};
            } else {
                p[i] = (T)((double)ui + beta * (double)p[i]);
                rs_cur[i] = RS<T>{o.r, (T)((double)wi + beta * (double)o.s)};
            }
            wu += (double)wi * (double)ui;
        }
    }
    block_sum2_all(wu, none, red);
    if (threadIdx.x == 0) part_del[blockIdx.x] = wu;
}

// Initial residual of the step's solve at free edge e = free[k] (D:1516-1536:
// Psi = E U_old - RHS1 SrcFld + dt (Ct F + Ctb Fb - I), rhs = Psi on free rows,
// r0 = rhs - A x with the warm start x = U_old on free edges). Since
// E_f U_old - A x = (E_f - [A 0]) U_old =: K U_old exactly, the residual is
// formed without the cancelling products E_f U_old and A x:
//   r0 = K U_old - RHS1 SrcFld + dt (Ct F + fb),  fb = Ctb Fb - I (edge-indexed)
// K (free rows, full edge columns) holds E's couplings to Dirichlet edges and
// whatever E_ff - A leaves (zero for the supported lossless media), so many
// of its slices are empty and skipped, as are RHS1's; with dir_terms = 0 (the
// host knows U_old and SrcFld to be zero on every Dirichlet edge) both
// products are skipped. r0 is kept in `r0save`
// for a restart; with jacobi also u = pre r and partial r.u. guard as in
// cg_guard_ok. SrcFld is U on Dirichlet edges at the new time (`dir_new`).
//
// The same kernel forms fem's residual (fem_w_kernel) with other operators,
// vectors and coefficients: r0 = (ck (kop xk) + cr (r1 xr)) + cc (ct F) + cf
// fb; gfdtd: xk = U, xr = dir_new, ck = 1, cr = -1, cc = cf = dt; fem: kop =
// E_f, r1 = SigMat_f, xk = W, xr = W2, ck = 1/dt^2, cr = -0.5/dt, cc = -1,
// cf = 1/dt, F = N C U_old.
struct RhsCoef {
    double ck, cr, cc, cf;
};
template <typename T, int NB, bool W = false>
__global__ void __launch_bounds__(kBlock) gf_rhs_init_kernel(
    SellView<T> kop, SellView<T> r1, SellView<int8_t> ct, int dir_terms, const int32_t* __restrict__ free_edge,
    const T* __restrict__ U, const T* __restrict__ dir_new, const T* __restrict__ Fv, const T* __restrict__ fb,
    T* __restrict__ r0save, RhsCoef co, const T* __restrict__ pre, T* __restrict__ r, int r_stride,
    T* __restrict__ u, double* __restrict__ part_rr, double* __restrict__ part_gam, CgState* st_in, int jacobi,
    const CgState* guard, SellView<T> am, const T* __restrict__ d, double* __restrict__ part_rr0,
    CgBatch bt, const double* __restrict__ row_dt) {
    if (!cg_guard_ok(guard)) return;
    __shared__ double red[kRed2];
    const int lane = lane_id();
    const CgSeg sg = cg_seg(bt, 0, st_in);
    double rr = 0, ru = 0, rr0 = 0;
    for (SliceRange sr_ = cg_slices(ct.nslices, ct.band, bt); sr_.s < sr_.end; sr_.s += sr_.step) {
        const int s = sr_.s;
        const double ku = dir_terms && kop.sptr[s + 1] > kop.sptr[s]
                              ? sell_dot<T, 8, W>(kop, s, lane, [&](int32_t j) { return U[j]; })
                              : 0.0;
        const double rs = dir_terms && r1.sptr[s + 1] > r1.sptr[s]
                              ? sell_dot<T, 8, W>(r1, s, lane, [&](int32_t j) { return dir_new[j]; })
                              : 0.0;
        const double cf = sell_dot<T, NB, W>(ct, s, lane, [&](int32_t j) { return Fv[j]; });
        // extrapolated start x0 = U_old + d: r = r0 - A d (r0 = b - A U_old)
        const double ad = d ? sell_dot<T, NB, W>(am, s, lane, [&](int32_t j) { return d[j]; }) : 0.0;
        const int32_t k = s * kSlice + lane;
        if (k < ct.nrows) {
            // batched instances: each row with its instance's dt (gfdtd: cc = cf = dt)
            const double cc = row_dt ? row_dt[k] : co.cc, cfb = row_dt ? row_dt[k] : co.cf;
            const T rk0 = (T)((co.ck * ku + co.cr * rs) + cc * cf + cfb * (double)fb[free_edge[k]]);
            r0save[k] = rk0;
            T rk = rk0;
            if (d) {
                rr0 += (double)rk0 * (double)rk0;
                rk = (T)((double)rk0 - ad);
            }
            r[(int64_t)k * r_stride] = rk;
            rr += (double)rk * (double)rk;
            if (jacobi) {
                const T uk = (T)((double)pre[k] * (double)rk);
                u[k] = uk;
                ru += (double)rk * (double)uk;
            }
        }
    }
    block_sum2_all(rr, ru, red);
    if (d) {
        double none = 0;
        block_sum2_all(rr0, none, red);
    }
    if (threadIdx.x == 0) {
        part_rr[blockIdx.x] = rr;
        if (jacobi) part_gam[blockIdx.x] = ru;
        if (d) part_rr0[blockIdx.x] = rr0;
        if (sg.leader) { sg.st->done = 0; sg.st->iter = 0; }
    }
}

// rr0 of an extrapolated start: ||b - A U_old||^2 from gf_rhs_init's
// partials (a fixed assignment of partials to threads and a fixed reduction
// tree: the same sum every time), marked as given for the stop test of
// pass 0 (cg_decide_stop); guard as in cg_guard_ok.
__global__ void __launch_bounds__(kBlock) cg_set_rr0_kernel(const double* __restrict__ part, int nb, CgState* st,
                                                            const CgState* guard) {
    if (!cg_guard_ok(guard)) return;
    __shared__ double red[kRed2];
    double s = 0, none = 0;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) s += part[k];
    block_sum2_all(s, none, red);
    if (threadIdx.x == 0) {
        st->rr0 = s;
        st->rr0_given = 1;
    }
}

}  // namespace tfd
