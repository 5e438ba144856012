// tetfit_hip.hip - libtetfit_hip.so: device context, per-method step
// sequences and the extern "C" tf_* entry points of include/tetfit_hip.h.
//
// Reference loops replaced: fdtd D:1840-1933, dga D:810-939,
// gfdtd D:1459-1565 (D: = src/DGA/Discretization.hpp), with the CG of
// src/DGA/ConjugateGradientSolver.hpp:52-103 run on the device.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <future>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "learned.hpp"
#include "mlp.hpp"
#include "sell.hpp"
#include "tetfit_hip.h"

using namespace tfd;

namespace {

thread_local std::string g_err;

// Grid caps of the grid-stride kernels: the CG partial buffers hold one
// fp64 per block of the producing kernel.
constexpr int kCgPartA = 2048;   // blocks of cg_w (SpMV with A)
constexpr int kCgPartB = 1024;   // blocks of gf_rhs_init / cg_restart / cg_upd
constexpr int kCgPartC = 2048;   // blocks of cg_pg (SpMV with the learned G)
static_assert(kCgPartA <= kMaxPart && kCgPartB <= kMaxPart && kCgPartC <= kMaxPart,
              "CG kernels read at most kMaxPart partials per array");
// SELL-C-sigma: rows of A (gfdtd) and H (dga) are sorted by decreasing length
// within windows of this many rows, so the 64 rows of a slice pad little.
constexpr int kSigma = 1024;
constexpr int kSpmvCap = 8192;   // blocks of the other SELL kernels

// Carrier of Source::Compute in long double (Source.hpp:316-327); 0 for t<0.
double carrier_value(const tf_source& s, double t) {
    if (t < 0) return 0;
    const long double PI_L = 3.141592653589793238L;
    double ret = s.amp;
    if (s.field == TF_FIELD_J) ret *= s.impedance;
    const long double arg = t - s.delay;
    switch (s.carrier) {
        case TF_CARRIER_SIN: ret *= std::sin(2 * PI_L * s.freq * arg); break;
        case TF_CARRIER_COS: ret *= s.amp * std::cos(2 * PI_L * s.freq * arg); break;
        case TF_CARRIER_GAUSSIAN: ret *= std::exp(-std::pow(t - s.delay, 2) / std::pow(s.width, 2)); break;
        case TF_CARRIER_RICKER: {
            long double a = PI_L * s.freq * arg;
            ret *= (1 - 2 * a * a) * std::exp(-(a * a));
            break;
        }
        case TF_CARRIER_GAUSSMOD:
            ret *= std::exp(-std::pow(t - s.delay, 2) / std::pow(s.width, 2));
            ret *= std::sin(2 * PI_L * s.freq * arg);
            break;
        default: break;
    }
    return ret;
}

template <class T>
void upload(DevBuf& b, const std::vector<T>& v) {
    b.alloc(v.size() * sizeof(T));
    if (!v.empty()) TF_HIP(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
}

struct Csr32 {
    DevBuf ptr, col, val, sg;
    int32_t nrows = 0;
    int64_t nnz = 0;
};

struct Sell {
    DevBuf sptr, col, val, sg, col16, cbase;
    int32_t nrows = 0, nslices = 0;
    int64_t nnz = 0, entries = 0;
    bool c16 = false;   // 16-bit column codes (col16 + cbase) in the kernels' view
    bool cwin = false;  // ... in the four-window form (SellView::cwin)
    bool band = false;  // banded slice order (SellView::band), batched dga
    int64_t min_width = 0;  // narrowest slice (the CG kernels' slice streams need >= 1)
    template <class V>
    SellView<V> view(const DevBuf& v) const {
        return SellView<V>{nrows, nslices, sptr.as<int64_t>(), c16 ? nullptr : col.as<int32_t>(),
                           c16 ? col16.as<uint16_t>() : nullptr, cbase.as<int32_t>(), v.as<V>(), cwin ? 1 : 0,
                           band ? 1 : 0};
    }
    template <class T>
    SellView<T> vals() const { return view<T>(val); }
    SellView<int8_t> signs() const { return view<int8_t>(sg); }
};

// Launch shape of the tuned SELL kernels: grid size (0 = one slice per
// wave), entries per gather batch (4, 8 or 16) and the slice-to-wave mapping
// (band 0: one contiguous chunk per workgroup; 1: the waves of an XCD take
// its slices round-robin, block_slices). Results do not depend on any of
// them (each row is summed by one lane in column order).
enum TuneKernel { TK_CT = 0, TK_H, TK_N, TK_W, TK_G, TK_WU, TK_COUNT };
const char* const kTuneName[TK_COUNT] = {"sell_ct_kernel", "sell_h_kernel", "sell_n_kernel", "cg_w_kernel",
                                         "cg_pg_kernel", "cg_wu_kernel"};
struct KTune {
    int blocks = 0;
    int unroll = 8;
    int band = 0;
};

// SELL kernels take the gather batch NB and W, the wide-mesh instantiation
// (windowed 16-bit column decode, non-temporal operator loads), launched for
// contexts with tf_ctx::wide set.
#define TF_LAUNCH_NB(w, nb, kern, T, grid, stream, ...)                                      \
    do {                                                                                    \
        if (w) {                                                                            \
            switch (nb) {                                                                   \
                case 4: kern<T, 4, true><<<(grid), kBlock, 0, (stream)>>>(__VA_ARGS__); break;   \
                case 16: kern<T, 16, true><<<(grid), kBlock, 0, (stream)>>>(__VA_ARGS__); break; \
                default: kern<T, 8, true><<<(grid), kBlock, 0, (stream)>>>(__VA_ARGS__); break;  \
            }                                                                               \
        } else {                                                                            \
            switch (nb) {                                                                   \
                case 4: kern<T, 4><<<(grid), kBlock, 0, (stream)>>>(__VA_ARGS__); break;         \
                case 16: kern<T, 16><<<(grid), kBlock, 0, (stream)>>>(__VA_ARGS__); break;       \
                default: kern<T, 8><<<(grid), kBlock, 0, (stream)>>>(__VA_ARGS__); break;        \
            }                                                                               \
        }                                                                                   \
    } while (0)

// Runs f(lo, hi) over [0, n) in chunks of `chunk` entries on up to 16 host
// threads (chunk boundaries do not depend on the thread count). f must not
// throw.
template <class Fn>
void par_chunks(int64_t n, int64_t chunk, const Fn& f) {
    const int64_t nc = (n + chunk - 1) / chunk;
    const int64_t hw = std::max(1u, std::thread::hardware_concurrency());
    const int nt = (int)std::min<int64_t>(nc, std::min<int64_t>(hw, 16));
    if (nt <= 1) {
        for (int64_t k = 0; k < nc; ++k) f(k * chunk, std::min(n, (k + 1) * chunk));
        return;
    }
    std::atomic<int64_t> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&] {
            for (int64_t k; (k = next++) < nc;) f(k * chunk, std::min(n, (k + 1) * chunk));
        });
    for (auto& t : th) t.join();
}

// Rows of the free-row order the learned preconditioner is trained on: all
// of them up to this many, else a window of this many in the middle of the
// order (a slab of the mesh); the loss is over the principal submatrix of
// those rows, with every sample's residual formed from its correction.
constexpr int32_t kTrainRows = 1 << 20;

// Host tables of the learned preconditioner (learned_tables()).
struct LearnedTables {
    int32_t n_pairs = 0, n_pad = 0;
    int64_t pair_w0 = 0, pair_w1 = 0;  // pairs whose first row lies in the training window
    std::vector<uint16_t> X;           // standardised bf16 features [pair][16]
    std::vector<float> base, scale, ahat;
    std::vector<int64_t> pa, pb;
    std::vector<int32_t> pos2pair;
};

// State of the learned preconditioner of a gfdtd context.
struct LearnedPc {
    enum { kOff = 0, kCollecting = 1, kActive = 2, kDisabled = 3 };
    int state = kOff;
    int64_t steps = 0;              // solves seen since set-up
    int learn_steps = 200, train_iters = 150;
    double lr = 3e-3;
    int32_t n_pairs = 0, n_pad = 0;
    int32_t w0 = 0, nw = 0;            // training window: free rows [w0, w0 + nw)
    int64_t pair_w0 = 0, pair_w1 = 0;  // pairs whose first row lies in the window
    std::future<std::unique_ptr<LearnedTables>> tables;  // host tables, built in the background
    DevBuf X, base, scale, pa, pb, pos2pair, y, ghat, gval, gpk, ahat, isd;
    DevBuf P, Pbest, m, v, part, gy, gpos, lpart, cpart;
    DevBuf Rb, Cb, T1, T2;
    std::vector<int> slot;          // 0 free, 1 captured correction pair, 2 random probe
    std::vector<int> jac_iters, pc_iters;
    double train_seconds = 0, loss0 = 0, loss = 0;
    int64_t fallbacks = 0, iters = 0, solves = 0;
    std::vector<const DevBuf*> bufs() const {
        return {&X, &base, &scale, &pa, &pb, &pos2pair, &y, &ghat, &gval, &gpk, &ahat, &isd, &P, &Pbest,
                &m, &v, &part, &gy, &gpos, &lpart, &cpart, &Rb, &Cb, &T1, &T2};
    }
};

}  // namespace

struct tf_ctx {
    int method = 0, prec = 0;
    int32_t E = 0, F = 0;
    double dt = 0;
    hipStream_t stream = nullptr;
    size_t tsize = 8;
    DevBuf U, Fv, B;

    // FIT
    DevBuf fe_sg, fe_ca, fe_cb, fe_f4, fh_sg, fh_cm, fh_cn, fh_e4;
    DevBuf fit_ep, fit_nu;  // fp64 energy weights
    std::vector<uint8_t> h_fe_sg;
    std::vector<int32_t> fit_uncommon, fit_src;
    std::vector<int4> fit_sp_f4_u;
    std::vector<uint8_t> fit_sp_sg_u;
    std::vector<double> fit_sp_ca_u, fit_sp_cb_u;
    std::vector<int4> h_fe_f4;
    // FIT grid order: fields, coefficients and codes stored at the grid
    // positions of edges and faces (fit_edge_slots / fit_face_slots)
    bool fit_grid = false;
    bool tet_relabel = false;  // tet fields and operators in the locality ranks' order (h_esid, h_fsid)
    // launch the SELL kernels' wide-mesh instantiation (W): set when some
    // operator uses windowed 16-bit columns, or when the operators stepped
    // each step exceed the 256 MB Infinity Cache
    bool wide = false;
    FitGrid grid{};
    std::vector<int32_t> h_esid, h_fsid;  // reference edge / face -> device position (FIT grid order, tet relabel)
    std::vector<uint16_t> h_fe_code;      // per grid edge: slot map << 8 | signs (0: not in the main pass)
    DevBuf fe_code, fh_code;
    std::vector<double> h_ca_common, h_cb_common;
    DevBuf sp_edge, sp_f4, sp_sg, sp_ca, sp_cb, sp_kind, sp_tb, sp_te;
    int32_t n_sp = 0;

    // DGA / gfdtd
    Csr32 Ct;
    Sell sCt, H, N, A, Kf, R1f, Ctf;  // Kf: E_f - [A 0] of the gfdtd residual (gf_rhs_init_kernel)
    Sell Efull;  // eps mass E of the energy diagnostic (all edges), tet methods
    Sell Sig;    // gfdtd SigMat (all edges) of the losses, lossy media only
    // dga with lossy media (tf_dga_frac): stacked fractional operators, the
    // Mu / Tf rows as (first, second or -1) entries, the source slot of each
    // edge, and U_frac / F_frac ping-pong buffers (fcur: the current one)
    bool frac = false;
    Sell fKu, fRu, fKf, fRf;
    // fem: U of the step before last (U_older) and the residual's gathered
    // vectors W, W2 (fem_w_kernel); Kf, R1f then hold E and SigMat on the free rows
    bool fem = false;
    DevBuf U_older, fw, fw2;
    DevBuf mu_a, mu_b, tf_a, tf_b, fslot, uf[2], ff[2];
    int fcur = 0;
    int32_t nU = 0, nF = 0, q0 = 0, q1 = 0;
    std::vector<int32_t> h_mu_a, h_mu_b, h_tf_a, h_tf_b;
    DevBuf loss_part;            // per step and block: U_old.SigMat U_old partials of a run
    int64_t loss_cap = 0;
    std::vector<double> loss_q;  // per step i of the simulation: U_old(i).SigMat U_old(i)
    bool c_fixed_signs = false;
    DevBuf c_col, c_sg, h_edge, curl_f, curl_u;
    DevBuf free_edge, dir_edge, dir_new, fb, pre, x, r, z, p0, p1, q, r0, cg_part, cg_state;
    DevBuf edge_row;  // per edge: its free row in the CG order, or -1 (Dirichlet)
    DevBuf rs0, rs1;  // (r, s) pairs of the learned passes, even / odd pass
    // extrapolated starts (tf_cg_opts.guess): U_{n-1}, U_{n-2} on the free
    // rows, the start's shift d, gf_rhs_init's partials of ||b - A U_old||^2;
    // hist = steps of history held, valid for the solve of step hist_next
    DevBuf hx1, hx2, dx, rr0part;
    int hist = 0;
    int64_t hist_next = -1;
    std::vector<int32_t> h_perm;  // SELL row r of H holds row h_perm[r] of the problem's H
    int32_t n_free = 0, n_dir = 0;
    // gfdtd: every Dirichlet value of every step is zero (pec values, no e
    // sources), and U is zero on the Dirichlet edges now; both together make
    // the Dirichlet terms of the initial residual vanish
    bool dir_zero_static = false, ud_zero = true;
    bool k_dir_only = true;  // K has entries in Dirichlet columns only
    std::vector<int32_t> h_dir_edge;
    std::vector<double> h_fixed_value;
    DevBuf inj_edge, inj_kind, inj_tb, inj_te;
    int32_t n_inj = 0;
    CgState* h_state = nullptr;  // pinned, two slots (as cg_state)
    CgState init_src[2] = {};     // host sources of the initial-state copies, per slot
    hipEvent_t solve_ev = nullptr;  // recorded after a solve's first batch of passes
    int cg_guess = 30;           // iterations of the previous solve
    int cg_guess_pc = 15;        // same, for solves with the learned preconditioner
    // host copy of A (free rows, CSR) and its SELL slice pointers, kept for
    // the learned preconditioner's pair list and features
    std::vector<int64_t> hA_ptr, hA_sptr;
    std::vector<int32_t> hA_col;
    std::vector<double> hA_val;
    LearnedPc lp;
    KTune tune[TK_COUNT];

    // batch of independent dga simulations (tf_create_batch): the instances
    // are concatenated (block-diagonal operators), each row/face carries its
    // own simulation's dt, each source its own carrier time base
    struct Batch {
        int32_t n = 0;
        std::vector<int32_t> eoff, foff, poff;   // n+1 prefix sums: edges, faces, probes
        std::vector<double> dt;
        std::vector<std::vector<tf_source>> src;
        std::vector<std::vector<tf_source_term>> terms;
        std::vector<std::vector<tf_probe>> probes;
        std::vector<int64_t> sample_steps;
        int64_t every = 0;                       // probe sample period in steps (0: off)
        std::vector<int32_t> ecnt;               // edges of each instance (gfdtd batches pad eoff with
                                                 // solver-only edges after them)
        bool gfdtd = false;                      // CG-solved instances (run_gfdtd_batch)
    };
    std::unique_ptr<Batch> batch;
    DevBuf row_dt_h, row_dt_f;                   // per H row / per face dt (batches)
    DevBuf row_dt_cg;                            // per CG row dt (gfdtd batches)
    DevBuf bt_inst, bt_s0, bt_s1, bt_w0, bt_w1;  // CG workgroup table of a gfdtd batch (CgBatch)
    int32_t bt_chunks = 0;
    std::vector<double> src_dt;                  // per source dt of the carrier table

    // sources
    std::vector<tf_source> src;
    DevBuf term_src, term_w, carr;
    int64_t carr_cap = 0;

    // probes and output schedule; with energy_out the samples also record
    // (We, Wh) in eser (2 per sample, epart: the reductions' partials)
    int32_t n_probes = 0;
    DevBuf probes, samples, eser, epart;
    bool energy_out = false;
    int64_t sample_cap = 0, n_samples = 0;
    std::vector<double> sample_times;
    double next_print = 0, period = 0;
    bool out_enabled = false;
    bool initial_done = false;

    // Poynting partials
    DevBuf poy;
    int64_t poy_cap = 0;
    std::vector<double> poynting;

    // Step loops of the explicit methods captured in a HIP graph: kGraphSteps
    // steps of launches whose per-step carrier rows and Poynting partial
    // slots point into the graph's own buffers (carr, poy), refilled and
    // drained around each replay. Rebuilt after anything that changes the
    // launches (sources, tuning).
    struct StepGraph {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        DevBuf carr, poy;
    };
    StepGraph sg;
    bool use_graphs = true;

    // profiling
    bool profiling = false;
    struct Rec { const char* name; hipEvent_t a, b; };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::map<std::string, std::pair<double, int64_t>> ktimes;

    size_t device_bytes() const;
    ~tf_ctx();
};

tf_ctx::~tf_ctx() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (sg.exec) (void)hipGraphExecDestroy(sg.exec);
    if (sg.graph) (void)hipGraphDestroy(sg.graph);
    if (solve_ev) (void)hipEventDestroy(solve_ev);
    for (auto e : ev_pool) (void)hipEventDestroy(e);
    if (h_state) (void)hipHostFree(h_state);
    if (stream) (void)hipStreamDestroy(stream);
}

size_t tf_ctx::device_bytes() const {
    size_t s = 0;
    const DevBuf* all[] = {&U, &Fv, &B, &fit_ep, &fit_nu, &fe_sg, &fe_ca, &fe_cb, &fe_f4, &fh_sg, &fh_cm, &fh_cn, &fh_e4, &fe_code, &fh_code,
                           &sp_edge, &sp_f4, &sp_sg, &sp_ca, &sp_cb, &sp_kind, &sp_tb, &sp_te, &c_col, &c_sg,
                           &h_edge, &curl_f, &curl_u, &free_edge, &edge_row, &dir_edge, &dir_new, &fb, &pre, &x, &r, &rs0, &rs1, &z,
                           &p0, &p1, &q, &r0, &hx1, &hx2, &dx, &rr0part, &cg_part, &cg_state, &inj_edge, &inj_kind, &inj_tb, &inj_te,
                           &term_src, &term_w, &carr, &probes, &samples, &eser, &epart, &poy, &row_dt_h, &row_dt_f,
                           &mu_a, &mu_b, &tf_a, &tf_b, &fslot, &row_dt_cg, &bt_inst, &bt_s0, &bt_s1, &bt_w0, &bt_w1, &uf[0], &uf[1], &ff[0], &ff[1], &U_older, &fw, &fw2};
    for (auto b : all) s += b->bytes;
    const Csr32* cs[] = {&Ct};
    for (auto c : cs) s += c->ptr.bytes + c->col.bytes + c->val.bytes + c->sg.bytes;
    const Sell* ss[] = {&sCt, &H, &N, &A, &Kf, &R1f, &Ctf, &Efull, &Sig, &fKu, &fRu, &fKf, &fRf};
    for (auto c : ss) s += c->sptr.bytes + c->col.bytes + c->val.bytes + c->sg.bytes + c->col16.bytes + c->cbase.bytes;
    for (auto b : lp.bufs()) s += b->bytes;
    return s;
}

namespace {

// ---------------------------------------------------------------- helpers

template <class T>
void upload_csr(Csr32& d, const tf_csr& c, bool values) {
    if (c.nnz >= (int64_t)INT32_MAX) throw std::runtime_error("matrix too large for int32 row pointers");
    d.nrows = c.nrows;
    d.nnz = c.nnz;
    std::vector<int32_t> ptr(c.nrows + 1);
    for (int32_t i = 0; i <= c.nrows; ++i) ptr[i] = (int32_t)c.rowptr[i];
    upload(d.ptr, ptr);
    d.col.alloc(c.nnz * 4);
    if (c.nnz) TF_HIP(hipMemcpy(d.col.p, c.col, c.nnz * 4, hipMemcpyHostToDevice));
    if (values) {
        std::vector<T> v(c.val, c.val + c.nnz);
        upload(d.val, v);
    } else {
        d.sg.alloc(c.nnz);
        if (c.nnz) TF_HIP(hipMemcpy(d.sg.p, c.sgn, c.nnz, hipMemcpyHostToDevice));
    }
}

// CSR -> SELL-64 (slices of 64 rows, column-major inside a slice, rows
// padded with value 0 on their last column).
std::vector<int64_t> sell_layout(const tf_csr& c) {
    const int32_t nslices = (c.nrows + kSlice - 1) / kSlice;
    std::vector<int64_t> sptr(nslices + 1, 0);
    for (int32_t s = 0; s < nslices; ++s) {
        int64_t w = 0;
        for (int32_t r = s * kSlice; r < std::min(c.nrows, (s + 1) * kSlice); ++r)
            w = std::max<int64_t>(w, c.rowptr[r + 1] - c.rowptr[r]);
        sptr[s + 1] = sptr[s] + w * kSlice;
    }
    return sptr;
}

template <class T>
void upload_sell(Sell& d, const tf_csr& c, bool values) {
    d.nrows = c.nrows;
    d.nnz = c.nnz;
    d.nslices = (c.nrows + kSlice - 1) / kSlice;
    const std::vector<int64_t> sptr = sell_layout(c);
    d.entries = sptr[d.nslices];
    d.min_width = d.nslices ? INT64_MAX : 0;
    for (int32_t s = 0; s < d.nslices; ++s) d.min_width = std::min(d.min_width, (sptr[s + 1] - sptr[s]) / kSlice);
    std::vector<int32_t> col(d.entries, 0);
    std::vector<T> val(values ? d.entries : 0, (T)0);
    std::vector<int8_t> sg(values ? 0 : d.entries, 0);
    for (int32_t s = 0; s < d.nslices; ++s) {
        const int64_t w = (sptr[s + 1] - sptr[s]) / kSlice;
        // rows without entries (and lanes past the last row) pad with a column
        // the slice already uses, so that padding does not widen its span
        int32_t slice_col = 0;
        for (int32_t r = s * kSlice; r < std::min(c.nrows, (s + 1) * kSlice); ++r)
            if (c.rowptr[r + 1] > c.rowptr[r]) { slice_col = c.col[c.rowptr[r]]; break; }
        for (int lane = 0; lane < kSlice; ++lane) {
            const int32_t r = s * kSlice + lane;
            int64_t b = 0, e = 0;
            if (r < c.nrows) { b = c.rowptr[r]; e = c.rowptr[r + 1]; }
            const int32_t padc = e > b ? c.col[e - 1] : slice_col;
            for (int64_t k = 0; k < w; ++k) {
                const int64_t o = sptr[s] + k * kSlice + lane;
                if (b + k < e) {
                    col[o] = c.col[b + k];
                    if (values) val[o] = (T)c.val[b + k];
                    else sg[o] = c.sgn[b + k];
                } else {
                    col[o] = padc;
                }
            }
        }
    }
    upload(d.sptr, sptr);
    upload(d.col, col);
    if (values) upload(d.val, val);
    else upload(d.sg, sg);
    // 16-bit column codes: one base per slice when every slice spans at most
    // 65,536 columns, else four windows of 16,384 columns per slice when
    // every slice's columns fit in four, else int32 columns.
    // $TETFIT_SELL_COL32 (non-empty, not 0) keeps int32 columns and
    // $TETFIT_SELL_WIN (likewise) skips the one-base form, to exercise those
    // layouts on small meshes.
    auto env_on = [](const char* name) {
        const char* v = std::getenv(name);
        return v && *v && std::strcmp(v, "0") != 0;
    };
    const bool force32 = env_on("TETFIT_SELL_COL32");
    std::vector<int32_t> cb(std::max(d.nslices, 1), 0);
    bool fits = !force32 && !env_on("TETFIT_SELL_WIN");
    for (int32_t s = 0; s < d.nslices && fits; ++s) {
        int32_t lo = INT32_MAX, hi = 0;
        for (int64_t o = sptr[s]; o < sptr[s + 1]; ++o) { lo = std::min(lo, col[o]); hi = std::max(hi, col[o]); }
        if (sptr[s + 1] == sptr[s]) lo = hi = 0;
        cb[s] = lo;
        fits = (int64_t)hi - lo <= 65535;
    }
    std::vector<uint16_t> c16;
    if (fits) {
        c16.assign(std::max<int64_t>(d.entries, 1), 0);
        for (int32_t s = 0; s < d.nslices; ++s)
            for (int64_t o = sptr[s]; o < sptr[s + 1]; ++o) c16[o] = (uint16_t)(col[o] - cb[s]);
    } else if (!force32) {
        // windows: greedy over the slice's sorted columns, each window opening
        // at the smallest column not yet covered
        cb.assign(4 * std::max(d.nslices, 1), 0);
        std::vector<uint8_t> nwin(std::max(d.nslices, 1), 0);
        fits = true;
        std::vector<int32_t> cs;
        for (int32_t s = 0; s < d.nslices && fits; ++s) {
            cs.assign(col.begin() + sptr[s], col.begin() + sptr[s + 1]);
            std::sort(cs.begin(), cs.end());
            int nw = 0;
            for (size_t k = 0; k < cs.size();) {
                if (nw == 4) { fits = false; break; }
                const int32_t b = cs[k];
                cb[4 * s + nw++] = b;
                while (k < cs.size() && (int64_t)cs[k] - b <= 16383) ++k;
            }
            nwin[s] = (uint8_t)nw;
        }
        if (fits) {
            c16.assign(std::max<int64_t>(d.entries, 1), 0);
            for (int32_t s = 0; s < d.nslices; ++s)
                for (int64_t o = sptr[s]; o < sptr[s + 1]; ++o) {
                    // the last opened window whose base is <= the column; the
                    // windows are disjoint and cover every column of the slice
                    int w = nwin[s] - 1;
                    while (w > 0 && cb[4 * s + w] > col[o]) --w;
                    const int64_t off = (int64_t)col[o] - cb[4 * s + w];
                    if (off < 0 || off > 16383) throw std::logic_error("SELL column window");
                    c16[o] = (uint16_t)(w << 14 | off);
                }
            d.cwin = true;
        }
    }
    d.c16 = fits;
    if (fits) upload(d.col16, c16);
    upload(d.cbase, cb);
}

// Row order for SELL-C-sigma: within each window of kSigma consecutive rows,
// rows sorted by decreasing length (stable). perm[new] = old.
std::vector<int32_t> sigma_order(const tf_csr& m) {
    std::vector<int32_t> perm(m.nrows);
    for (int32_t i = 0; i < m.nrows; ++i) perm[i] = i;
    for (int32_t w0 = 0; w0 < m.nrows; w0 += kSigma) {
        const int32_t w1 = std::min(m.nrows, w0 + kSigma);
        std::stable_sort(perm.begin() + w0, perm.begin() + w1, [&](int32_t a, int32_t b) {
            return m.rowptr[a + 1] - m.rowptr[a] > m.rowptr[b + 1] - m.rowptr[b];
        });
    }
    return perm;
}

// Row order of the CG's free rows: by the edges' locality rank when the
// problem carries one (tf_dga_ops.edge_rank: same-class edges along a grid
// line consecutive, so a slice's gathers touch few cache lines), else
// SELL-C-sigma. TETFIT_ROW_ORDER=sigma keeps SELL-C-sigma. perm[new] = old.
std::vector<int32_t> free_row_order(const tf_dga_ops& d) {
    const char* env = std::getenv("TETFIT_ROW_ORDER");
    if (!d.edge_rank || (env && std::strcmp(env, "sigma") == 0)) return sigma_order(d.A);
    std::vector<int32_t> perm(d.n_free);
    for (int32_t i = 0; i < d.n_free; ++i) perm[i] = i;
    std::stable_sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) {
        return d.edge_rank[d.free_edges[a]] < d.edge_rank[d.free_edges[b]];
    });
    return perm;
}

// Row order of a gfdtd batch's CG rows: instance k owns merged rows
// [row0[k], row0[k+1]) (multiples of 64, so no slice holds two instances);
// inside an instance, SELL-C-sigma windows that stop at its end.
std::vector<int32_t> batch_row_order(const tf_dga_ops& d, const std::vector<int32_t>& row0) {
    std::vector<int32_t> perm(d.n_free);
    for (int32_t i = 0; i < d.n_free; ++i) perm[i] = i;
    for (size_t k = 0; k + 1 < row0.size(); ++k)
        for (int32_t w0 = row0[k]; w0 < row0[k + 1]; w0 += kSigma) {
            const int32_t w1 = std::min(row0[k + 1], w0 + kSigma);
            std::stable_sort(perm.begin() + w0, perm.begin() + w1, [&](int32_t a, int32_t b) {
                return d.A.rowptr[a + 1] - d.A.rowptr[a] > d.A.rowptr[b + 1] - d.A.rowptr[b];
            });
        }
    return perm;
}

// Rows rows[0..n) of m in that order (entries keep their order), columns
// relabeled through cmap when given.
struct CsrBuf {
    std::vector<int64_t> rp;
    std::vector<int32_t> col;
    std::vector<double> val;
    std::vector<int8_t> sg;
    tf_csr v{};
};
void select_rows(const tf_csr& m, const int32_t* rows, int32_t n, const int32_t* cmap, CsrBuf& o) {
    o.rp.assign(1, 0);
    o.rp.reserve((size_t)n + 1);
    o.col.clear();
    o.val.clear();
    o.sg.clear();
    int64_t total = 0;
    for (int32_t k = 0; k < n; ++k) total += m.rowptr[rows[k] + 1] - m.rowptr[rows[k]];
    o.col.reserve(total);
    if (m.val) o.val.reserve(total);
    if (m.sgn) o.sg.reserve(total);
    for (int32_t k = 0; k < n; ++k) {
        const int32_t e = rows[k];
        for (int64_t q = m.rowptr[e]; q < m.rowptr[e + 1]; ++q) {
            o.col.push_back(cmap ? cmap[m.col[q]] : m.col[q]);
            if (m.val) o.val.push_back(m.val[q]);
            if (m.sgn) o.sg.push_back(m.sgn[q]);
        }
        o.rp.push_back((int64_t)o.col.size());
    }
    o.v = tf_csr{};
    o.v.nrows = n;
    o.v.ncols = m.ncols;
    o.v.nnz = o.rp.back();
    o.v.rowptr = o.rp.data();
    o.v.col = o.col.data();
    o.v.val = m.val ? o.val.data() : nullptr;
    o.v.sgn = m.sgn ? o.sg.data() : nullptr;
}

// A tet problem renumbered by its locality ranks (tf_dga_ops.edge_rank,
// face_rank): edge e of the problem is edge ep[e] here and face f is face
// fp[f]. Rows keep their entries in the problem's order (only the labels
// change), so every row sum runs in the same order and each field entry is
// bitwise the one of the problem's numbering. H's rows follow their edges'
// new labels (h_row relabeled to match); A keeps its rows, the free edges
// their order in A (with new labels). Not for lossy dga (its fractional
// unknowns keep the problem's order) or when TETFIT_TET_ORDER=numbering.
struct TetRelabel {
    CsrBuf C, Ct, H, N, E, RHS1, Sig;
    std::vector<int32_t> ep, fp, bnd_edge, h_row, free_edges, iota_e, iota_f;
    std::vector<int8_t> edge_class;
    std::vector<double> fixed_value;
    tf_problem p{};
};

bool relabel_tet(const tf_problem* src, TetRelabel& r) {
    const tf_dga_ops& d = src->dga;
    const int32_t E = src->n_edges, F = src->n_faces;
    const char* env = std::getenv("TETFIT_TET_ORDER");
    if (src->method == TF_METHOD_FDTD || !d.edge_rank || !d.face_rank || d.frac.enabled ||
        (env && std::strcmp(env, "numbering") == 0))
        return false;
    auto inverse = [](const int32_t* rank, int32_t n, std::vector<int32_t>& fwd, std::vector<int32_t>& inv) {
        fwd.assign(rank, rank + n);
        inv.assign(n, -1);
        for (int32_t i = 0; i < n; ++i) {
            if (rank[i] < 0 || rank[i] >= n || inv[rank[i]] >= 0) throw std::runtime_error("locality rank is not a permutation");
            inv[rank[i]] = i;
        }
    };
    std::vector<int32_t> einv, finv;
    inverse(d.edge_rank, E, r.ep, einv);
    inverse(d.face_rank, F, r.fp, finv);
    const int32_t* ep = r.ep.data();
    const int32_t* fp = r.fp.data();
    if (d.C.nrows != F || d.Ct.nrows != E || d.N.nrows != F) throw std::runtime_error("operator sizes");
    select_rows(d.C, finv.data(), F, ep, r.C);
    select_rows(d.Ct, einv.data(), E, fp, r.Ct);
    select_rows(d.N, finv.data(), F, fp, r.N);
    auto full = [&](const tf_csr& m, CsrBuf& o) -> tf_csr {
        if (m.nrows != E) return m;  // absent (0 rows)
        select_rows(m, einv.data(), E, ep, o);
        return o.v;
    };
    tf_problem& p = r.p;
    p = *src;
    tf_dga_ops& o = p.dga;
    o.C = r.C.v;
    o.Ct = r.Ct.v;
    o.N = r.N.v;
    o.E = full(d.E, r.E);
    o.RHS1 = full(d.RHS1, r.RHS1);
    o.Sig = full(d.Sig, r.Sig);
    r.bnd_edge.assign(E, -1);
    r.edge_class.assign(E, 0);
    r.fixed_value.assign(E, 0.0);
    for (int32_t e = 0; e < E; ++e) {
        if (d.bnd_edge) r.bnd_edge[ep[e]] = d.bnd_edge[e];
        r.edge_class[ep[e]] = d.edge_class[e];
        r.fixed_value[ep[e]] = d.fixed_value[e];
    }
    o.bnd_edge = d.bnd_edge ? r.bnd_edge.data() : nullptr;
    o.edge_class = r.edge_class.data();
    o.fixed_value = r.fixed_value.data();
    if (d.h_row && d.H.nrows > 0) {
        std::vector<int32_t> hrows;
        hrows.reserve(d.H.nrows);
        r.h_row.assign(E, -1);
        for (int32_t e2 = 0; e2 < E; ++e2) {
            const int32_t h = d.h_row[einv[e2]];
            if (h < 0) continue;
            if (h >= d.H.nrows) throw std::runtime_error("h_row out of range");
            r.h_row[e2] = (int32_t)hrows.size();
            hrows.push_back(h);
        }
        if ((int32_t)hrows.size() != d.H.nrows) throw std::runtime_error("H rows and class-1 edges differ");
        select_rows(d.H, hrows.data(), d.H.nrows, ep, r.H);
        o.H = r.H.v;
        o.h_row = r.h_row.data();
    }
    if (d.n_free > 0) {
        r.free_edges.resize(d.n_free);
        for (int32_t k = 0; k < d.n_free; ++k) r.free_edges[k] = ep[d.free_edges[k]];
        o.free_edges = r.free_edges.data();
    }
    r.iota_e.resize(E);
    r.iota_f.resize(F);
    for (int32_t e = 0; e < E; ++e) r.iota_e[e] = e;
    for (int32_t f = 0; f < F; ++f) r.iota_f[f] = f;
    o.edge_rank = r.iota_e.data();
    o.face_rank = r.iota_f.data();
    return true;
}

int sell_blocks(const Sell& m, int cap) {
    const int need = (m.nslices + kBlock / kWave - 1) / (kBlock / kWave);
    return std::max(1, std::min(need, cap));
}

// An operator's view with the slice mapping of launch shape t (a matrix
// banded by construction, batched dga, stays banded); tuned_view: with the
// context's tuned shape of kernel k.
template <class V>
SellView<V> with_band(SellView<V> v, const KTune& t) {
    v.band = v.band || t.band ? 1 : 0;
    return v;
}
template <class V>
SellView<V> tuned_view(const tf_ctx* c, int k, SellView<V> v) {
    return with_band(v, c->tune[k]);
}

int tuned_blocks(const tf_ctx* c, int k, const Sell& m, int cap) {
    const int b = c->tune[k].blocks;
    return b > 0 ? std::max(1, std::min(b, std::min(cap, sell_blocks(m, INT32_MAX)))) : sell_blocks(m, cap);
}

template <class T>
void zero_field(DevBuf& b, int64_t n, hipStream_t s) {
    b.alloc(std::max<int64_t>(n, 1) * sizeof(T));
    TF_HIP(hipMemsetAsync(b.p, 0, b.bytes, s));
}

// Device position of a reference edge / face (grid order or identity).
int32_t dev_edge(const tf_ctx* c, int32_t e) { return c->fit_grid || c->tet_relabel ? c->h_esid[e] : e; }
int32_t dev_face(const tf_ctx* c, int32_t f) { return c->fit_grid || c->tet_relabel ? c->h_fsid[f] : f; }

uint8_t code4(const int s[4]) {
    uint8_t c = 0;
    for (int k = 0; k < 4; ++k) c |= (uint8_t)((s[k] > 0 ? 1 : (s[k] < 0 ? 3 : 0)) << (2 * k));
    return c;
}

void ev_begin(tf_ctx* c, const char* name) {
    if (!c->profiling) return;
    if (c->ev_used + 2 > c->ev_pool.size()) {
        for (int i = 0; i < 256; ++i) {
            hipEvent_t e;
            TF_HIP(hipEventCreate(&e));
            c->ev_pool.push_back(e);
        }
    }
    tf_ctx::Rec r{name, c->ev_pool[c->ev_used], c->ev_pool[c->ev_used + 1]};
    c->ev_used += 2;
    TF_HIP(hipEventRecord(r.a, c->stream));
    c->recs.push_back(r);
}
void ev_end(tf_ctx* c) {
    if (!c->profiling) return;
    TF_HIP(hipEventRecord(c->recs.back().b, c->stream));
}
// Profiled launch whose event pair is bound to the kernel's own dispatch
// (hipExtLaunchKernelGGL records start and stop with the dispatch), so the
// time excludes the marker packets that a separate event record adds
// (about 2 us per launch here); plain launch when not profiling.
template <typename K, typename... Args>
void launch_timed(tf_ctx* c, const char* name, K kernel, int grid, Args... args) {
    if (!c->profiling) {
        kernel<<<grid, kBlock, 0, c->stream>>>(args...);
        return;
    }
    if (c->ev_used + 2 > c->ev_pool.size())
        for (int i = 0; i < 256; ++i) {
            hipEvent_t e;
            TF_HIP(hipEventCreate(&e));
            c->ev_pool.push_back(e);
        }
    tf_ctx::Rec r{name, c->ev_pool[c->ev_used], c->ev_pool[c->ev_used + 1]};
    c->ev_used += 2;
    c->recs.push_back(r);
    hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), 0, c->stream, r.a, r.b, 0, args...);
}
#define TF_LAUNCH_TIMED(c, name, nb, kern, T, grid, ...)                                       \
    do {                                                                                    \
        if ((c)->wide) {                                                                    \
            switch (nb) {                                                                   \
                case 4: launch_timed(c, name, kern<T, 4, true>, grid, __VA_ARGS__); break;     \
                case 16: launch_timed(c, name, kern<T, 16, true>, grid, __VA_ARGS__); break;   \
                default: launch_timed(c, name, kern<T, 8, true>, grid, __VA_ARGS__); break;    \
            }                                                                               \
        } else {                                                                            \
            switch (nb) {                                                                   \
                case 4: launch_timed(c, name, kern<T, 4>, grid, __VA_ARGS__); break;           \
                case 16: launch_timed(c, name, kern<T, 16>, grid, __VA_ARGS__); break;         \
                default: launch_timed(c, name, kern<T, 8>, grid, __VA_ARGS__); break;          \
            }                                                                               \
        }                                                                                   \
    } while (0)
void ev_flush(tf_ctx* c) {
    if (!c->profiling || c->recs.empty()) return;
    TF_HIP(hipStreamSynchronize(c->stream));
    for (auto& r : c->recs) {
        float ms = 0;
        TF_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        auto& k = c->ktimes[r.name];
        k.first += ms;
        k.second += 1;
    }
    c->recs.clear();
    c->ev_used = 0;
}

// ---------------------------------------------------------------- create

// Grid order for a FIT problem that carries grid positions: the slot map of
// every common edge and tbc face (which slot holds the reference's face /
// edge at each position, fit_edge_slots / fit_face_slots). False (the
// index-table layout is used) if any of them does not have exactly its
// slots' entities, or $TETFIT_FIT_INDEXED is set to a non-empty value other
// than 0.
bool fit_grid_codes(tf_ctx* c, const tf_fit_ops& f, std::vector<uint16_t>& ecode, std::vector<uint16_t>& fcode) {
    const char* forced = std::getenv("TETFIT_FIT_INDEXED");
    if (forced && *forced && std::strcmp(forced, "0") != 0) return false;
    if (!f.edge_sid || !f.face_sid) return false;
    const int32_t E = c->E, F = c->F;
    if (f.cells[0] <= 0 || f.cells[1] <= 0 || f.cells[2] <= 0) return false;
    const FitGrid g = make_fit_grid(f.cells[0], f.cells[1], f.cells[2]);
    // the divisions by invariant integers, checked against plain division
    for (const FastDiv* fd : {&g.dx, &g.dx1, &g.dy, &g.dy1})
        for (uint32_t n : {0u, 1u, fd->d - 1, fd->d, fd->d + 1, (uint32_t)E - 1, (uint32_t)F - 1, 0x7FFFFFFFu})
            if (fastdiv(*fd, n) != n / fd->d) throw std::runtime_error("fast division self-check failed");
    const int64_t nx = g.nx, ny = g.ny, nz = g.nz;
    if (nx <= 0 || ny <= 0 || nz <= 0) return false;
    if (nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz != E) return false;
    if ((int64_t)E >= ((int64_t)1 << 31) || (int64_t)F >= ((int64_t)1 << 31)) return false;
    if ((nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1) != F) return false;
    std::vector<int32_t> esid(f.edge_sid, f.edge_sid + E), fsid(f.face_sid, f.face_sid + F);
    {
        std::vector<uint8_t> seen_e(E, 0), seen_f(F, 0);
        for (int32_t v : esid) { if (v < 0 || v >= E || seen_e[v]) return false; seen_e[v] = 1; }
        for (int32_t v : fsid) { if (v < 0 || v >= F || seen_f[v]) return false; seen_f[v] = 1; }
    }
    auto map_of = [](const int32_t slot[4], const int64_t ref[4], uint32_t& m) {
        m = 0;
        for (int q = 0; q < 4; ++q) {
            int hit = -1;
            for (int k = 0; k < 4; ++k)
                if (slot[k] == ref[q]) { if (hit >= 0) return false; hit = k; }
            if (hit < 0) return false;
            m |= (uint32_t)hit << (2 * q);
        }
        return true;
    };
    ecode.assign(E, 0);
    for (int64_t q = 0; q < f.n_common; ++q) {
        const int32_t e = f.common_edges[q];
        if (f.ct_ptr[e + 1] - f.ct_ptr[e] != 4) return false;
        int32_t slot[4];
        int64_t ref[4];
        fit_edge_slots(g, (uint32_t)esid[e], slot);
        for (int k = 0; k < 4; ++k) ref[k] = fsid[f.ct_face[f.ct_ptr[e] + k]];
        uint32_t m;
        if (!map_of(slot, ref, m)) return false;
        const int d = f.dual_curl[e];
        const int sg[4] = {d, -d, d, -d};
        ecode[esid[e]] = (uint16_t)(m << 8 | code4(sg));
    }
    fcode.assign(F, 0);
    for (int64_t q = 0; q < f.n_tbc; ++q) {
        const int32_t fi = f.tbc_faces[q];
        int32_t slot[4];
        int64_t ref[4];
        fit_face_slots(g, (uint32_t)fsid[fi], slot);
        for (int k = 0; k < 4; ++k) ref[k] = esid[f.c_edge[4 * (int64_t)fi + k]];
        uint32_t m;
        if (!map_of(slot, ref, m)) return false;
        const int d = f.curl[fi];
        const int sg[4] = {d, -d, d, -d};
        fcode[fsid[fi]] = (uint16_t)(m << 8 | code4(sg));
    }
    c->grid = g;
    c->h_esid = std::move(esid);
    c->h_fsid = std::move(fsid);
    return true;
}

template <class T>
void create_fit(tf_ctx* c, const tf_problem* p) {
    const tf_fit_ops& f = p->fit;
    const int32_t E = c->E, F = c->F;
    std::vector<uint16_t> ecode, fcode;
    c->fit_grid = fit_grid_codes(c, f, ecode, fcode);
    // per-entity arrays in device order
    auto by_edge = [&](auto v) {
        if (!c->fit_grid) return v;
        auto o = v;
        for (int32_t e = 0; e < E; ++e) o[c->h_esid[e]] = v[e];
        return o;
    };
    auto by_face = [&](auto v) {
        if (!c->fit_grid) return v;
        auto o = v;
        for (int32_t q = 0; q < F; ++q) o[c->h_fsid[q]] = v[q];
        return o;
    };
    if (f.ep_vec && f.nu_vec) {
        upload(c->fit_ep, by_edge(std::vector<double>(f.ep_vec, f.ep_vec + E)));
        upload(c->fit_nu, by_face(std::vector<double>(f.nu_vec, f.nu_vec + F)));
    }
    const double dt = p->dt;
    std::vector<uint8_t> sg(E, 0);
    std::vector<T> ca(E, 0), cb(E, 0);
    std::vector<int4> f4(E, int4{0, 0, 0, 0});
    c->h_ca_common.assign(E, 0);
    c->h_cb_common.assign(E, 0);
    for (int64_t q = 0; q < f.n_common; ++q) {
        const int32_t e = f.common_edges[q];
        if (f.ct_ptr[e + 1] - f.ct_ptr[e] != 4) throw std::runtime_error("common edge without 4 faces");
        const int64_t b0 = f.ct_ptr[e];
        f4[e] = int4{f.ct_face[b0], f.ct_face[b0 + 1], f.ct_face[b0 + 2], f.ct_face[b0 + 3]};
        const int d = f.dual_curl[e];
        const int s[4] = {d, -d, d, -d};  // stored signs ignored for common edges (D:1873-1874)
        sg[e] = code4(s);
        const double up = f.upml[e];
        const double a = std::pow(up, 2) * f.M_h[e] * f.M_q[e];
        const double bb = std::pow(up, 2) * f.M_h[e] * up * dt;
        c->h_ca_common[e] = a;
        c->h_cb_common[e] = bb;
        ca[e] = (T)a;
        cb[e] = (T)bb;
    }
    c->h_fe_sg = sg;
    c->h_fe_f4 = f4;
    if (c->fit_grid) {
        c->h_fe_code = ecode;
        upload(c->fe_code, ecode);
        upload(c->fe_ca, by_edge(ca));
        upload(c->fe_cb, by_edge(cb));
    } else {
        upload(c->fe_sg, sg);
        upload(c->fe_ca, ca);
        upload(c->fe_cb, cb);
        upload(c->fe_f4, f4);
    }
    // special edges: uncommon (h source) and src edges (e / j)
    c->fit_uncommon.assign(f.uncommon_edges, f.uncommon_edges + f.n_uncommon);
    c->fit_src.assign(f.src_edges, f.src_edges + f.n_src);
    c->fit_sp_f4_u.clear();
    c->fit_sp_sg_u.clear();
    c->fit_sp_ca_u.clear();
    c->fit_sp_cb_u.clear();
    for (int32_t e : c->fit_uncommon) {
        int4 ff{0, 0, 0, 0};
        int s[4] = {0, 0, 0, 0};
        int n = 0;
        for (int64_t q = f.ct_ptr[e]; q < f.ct_ptr[e + 1] && n < 4; ++q, ++n) {
            (&ff.x)[n] = f.ct_face[q];
            s[n] = f.ct_sgn[q];
        }
        c->fit_sp_f4_u.push_back(ff);
        c->fit_sp_sg_u.push_back(code4(s));
        c->fit_sp_ca_u.push_back(f.M_h[e] * f.M_q[e]);
        c->fit_sp_cb_u.push_back(f.M_h[e] * dt);
    }
    // H phase over tbc faces
    std::vector<uint8_t> hs(F, 0);
    std::vector<T> cm(F, 0), cn(F, 0);
    std::vector<int4> e4(F, int4{0, 0, 0, 0});
    for (int64_t q = 0; q < f.n_tbc; ++q) {
        const int32_t fi = f.tbc_faces[q];
        const int32_t* ce = f.c_edge + 4 * (int64_t)fi;
        e4[fi] = int4{ce[0], ce[1], ce[2], ce[3]};
        const int d = f.curl[fi];
        const int s[4] = {d, -d, d, -d};
        hs[fi] = code4(s);
        cm[fi] = (T)(f.M_mu[fi] * f.fpml[fi]);
        cn[fi] = (T)(f.fpml[fi] * f.M_nu[fi]);
    }
    if (c->fit_grid) {
        upload(c->fh_code, fcode);
        upload(c->fh_cm, by_face(cm));
        upload(c->fh_cn, by_face(cn));
    } else {
        upload(c->fh_sg, hs);
        upload(c->fh_cm, cm);
        upload(c->fh_cn, cn);
        upload(c->fh_e4, e4);
    }
}

// K = (E - dt/2 SigMat)_f - [A 0]: row k is the row of free edge fe[k] of
// E - dt/2 SigMat (the operator Psi applies to U_old, D:1516) minus A's row
// perm[k] placed at the edges of its (free) columns, in fp64, exact zeros
// dropped. Then Psi's U_old terms minus A x are K U_old whenever x = U_old on
// the free edges, which gives the step's initial residual without the
// cancelling products (gf_rhs_init_kernel). With A = E_ff + dt/2 SigMat_ff,
// K holds E's and SigMat's Dirichlet columns and -dt SigMat_ff. Entries keep
// E's row order, A-only columns follow.
void residual_operator(const tf_dga_ops& d, double dt, const std::vector<int32_t>& perm,
                       const std::vector<int32_t>& fe, CsrBuf& o) {
    const bool lossy = d.Sig.nnz > 0;
    if (lossy && (d.Sig.nrows != d.E.nrows || d.Sig.nnz != d.E.nnz))
        throw std::runtime_error("SigMat must have E's pattern");
    o.rp.assign(1, 0);
    o.col.clear();
    o.val.clear();
    o.sg.clear();
    std::vector<std::pair<int32_t, double>> arow;
    std::vector<uint8_t> used;
    for (size_t k = 0; k < fe.size(); ++k) {
        const int32_t e = fe[k], a = perm[k];
        arow.clear();
        for (int64_t q = d.A.rowptr[a]; q < d.A.rowptr[a + 1]; ++q)
            arow.push_back({d.free_edges[d.A.col[q]], d.A.val[q]});
        used.assign(arow.size(), 0);
        for (int64_t q = d.E.rowptr[e]; q < d.E.rowptr[e + 1]; ++q) {
            double v = lossy ? d.E.val[q] - (0.5 * dt) * d.Sig.val[q] : d.E.val[q];
            for (size_t t = 0; t < arow.size(); ++t)
                if (!used[t] && arow[t].first == d.E.col[q]) {
                    v -= arow[t].second;
                    used[t] = 1;
                    break;
                }
            if (v != 0) {
                o.col.push_back(d.E.col[q]);
                o.val.push_back(v);
            }
        }
        for (size_t t = 0; t < arow.size(); ++t)
            if (!used[t] && arow[t].second != 0) {
                o.col.push_back(arow[t].first);
                o.val.push_back(-arow[t].second);
            }
        o.rp.push_back((int64_t)o.col.size());
    }
    o.v = tf_csr{};
    o.v.nrows = (int32_t)fe.size();
    o.v.ncols = d.E.ncols;
    o.v.nnz = o.rp.back();
    o.v.rowptr = o.rp.data();
    o.v.col = o.col.data();
    o.v.val = o.val.data();
}

template <class T>
void create_tet(tf_ctx* c, const tf_problem* p, const std::vector<int32_t>* batch_rows) {
    const tf_dga_ops& d = p->dga;
    const int32_t E = c->E, F = c->F;
    if (p->method == TF_METHOD_DGA) upload_sell<T>(c->sCt, d.Ct, false);
    if (d.C.nnz != 3 * (int64_t)F) throw std::runtime_error("C must have 3 entries per face");
    c->c_col.alloc(d.C.nnz * 4);
    TF_HIP(hipMemcpy(c->c_col.p, d.C.col, d.C.nnz * 4, hipMemcpyHostToDevice));
    c->c_fixed_signs = true;
    for (int64_t q = 0; q < d.C.nnz && c->c_fixed_signs; ++q)
        if (d.C.sgn[q] != (q % 3 == 1 ? -1 : 1)) c->c_fixed_signs = false;
    if (!c->c_fixed_signs) {
        c->c_sg.alloc(d.C.nnz);
        TF_HIP(hipMemcpy(c->c_sg.p, d.C.sgn, d.C.nnz, hipMemcpyHostToDevice));
    }
    upload_sell<T>(c->N, d.N, true);
    if (d.E.nrows == E && d.E.nnz > 0) upload_sell<T>(c->Efull, d.E, true);
    if ((p->method == TF_METHOD_GFDTD || p->method == TF_METHOD_FEM) && d.Sig.nnz > 0) upload_sell<T>(c->Sig, d.Sig, true);
    c->h_fixed_value.assign(d.fixed_value, d.fixed_value + E);
    zero_field<T>(c->curl_u, F, c->stream);
    zero_field<T>(c->curl_f, E, c->stream);
    if (p->method == TF_METHOD_DGA && d.frac.enabled) {
        // lossy media: the fractional operators in their own row order
        const tf_dga_frac& fr = d.frac;
        c->frac = true;
        c->nU = fr.Ku.nrows;
        c->nF = fr.Kf.nrows;
        c->q0 = fr.h_size + fr.p_size;
        c->q1 = c->q0 + fr.q_size;
        if (fr.Ru.nrows != c->nU || fr.Rf.nrows != c->nF || fr.Mu.nrows != E || fr.Tf.nrows != F)
            throw std::runtime_error("inconsistent fractional operators");
        upload_sell<T>(c->fKu, fr.Ku, true);
        upload_sell<T>(c->fRu, fr.Ru, true);
        upload_sell<T>(c->fKf, fr.Kf, true);
        upload_sell<T>(c->fRf, fr.Rf, true);
        auto two = [](const tf_csr& m, int32_t ncols, std::vector<int32_t>& a, std::vector<int32_t>& b) {
            a.assign(m.nrows, -1);
            b.assign(m.nrows, -1);
            for (int32_t r = 0; r < m.nrows; ++r) {
                const int64_t k0 = m.rowptr[r], k1 = m.rowptr[r + 1];
                if (k1 - k0 < 1 || k1 - k0 > 2) throw std::runtime_error("a Mu / Tf row needs one or two entries");
                for (int64_t k = k0; k < k1; ++k) {
                    if (m.col[k] < 0 || m.col[k] >= ncols || m.val[k] != 1.0)
                        throw std::runtime_error("Mu / Tf entries must be unit entries in range");
                }
                a[r] = m.col[k0];
                if (k1 - k0 == 2) b[r] = m.col[k0 + 1];
            }
        };
        two(fr.Mu, c->nU, c->h_mu_a, c->h_mu_b);
        two(fr.Tf, c->nF, c->h_tf_a, c->h_tf_b);
        upload(c->mu_a, c->h_mu_a);
        upload(c->mu_b, c->h_mu_b);
        upload(c->tf_a, c->h_tf_a);
        upload(c->tf_b, c->h_tf_b);
        std::vector<int32_t> sl(fr.slot, fr.slot + E);
        for (int32_t v : sl)
            if (v < -1 || v >= c->nU) throw std::runtime_error("fractional slot out of range");
        upload(c->fslot, sl);
        // U_d holds the class-4 values (ComputeEdgeBC at t = 0, D:762-773)
        std::vector<T> u0(std::max(c->nU, 1), (T)0);
        for (int32_t e = 0; e < E; ++e)
            if (d.edge_class[e] == 4 && sl[e] >= 0) u0[sl[e]] = (T)c->h_fixed_value[e];
        for (int k = 0; k < 2; ++k) {
            upload(c->uf[k], u0);
            zero_field<T>(c->ff[k], std::max(c->nF, 1), c->stream);
        }
        c->fcur = 0;
    } else if (p->method == TF_METHOD_DGA) {
        // H rows in SELL-C-sigma order; h_edge maps a SELL row to its edge
        // (a relabeled problem's H rows are already in its edges' locality order)
        if (c->tet_relabel) {
            c->h_perm.resize(d.H.nrows);
            for (int32_t r = 0; r < d.H.nrows; ++r) c->h_perm[r] = r;
        } else {
            c->h_perm = sigma_order(d.H);
        }
        CsrBuf hs;
        select_rows(d.H, c->h_perm.data(), d.H.nrows, nullptr, hs);
        upload_sell<T>(c->H, hs.v, true);
        std::vector<int32_t> row_edge(d.H.nrows, -1);
        for (int32_t e = 0; e < E; ++e)
            if (d.h_row[e] >= 0) row_edge[d.h_row[e]] = e;
        std::vector<int32_t> he(d.H.nrows);
        for (int32_t r = 0; r < d.H.nrows; ++r) {
            he[r] = row_edge[c->h_perm[r]];
            if (he[r] < 0) throw std::runtime_error("H row without an edge");
        }
        upload(c->h_edge, he);
    } else {
        // The CG runs over the free rows in the order of free_row_order: row k of
        // every free-row operator (A, and K, RHSmat1 and Ct for the initial
        // residual) is free edge free_edges[perm[k]], and A's columns are
        // relabeled to match, so all CG vectors use that order.
        if (d.A.nrows != d.n_free) throw std::runtime_error("A must have one row per free edge");
        const std::vector<int32_t> perm = batch_rows ? batch_row_order(d, *batch_rows) : free_row_order(d);
        std::vector<int32_t> inv(d.n_free);
        for (int32_t k = 0; k < d.n_free; ++k) inv[perm[k]] = k;
        CsrBuf a;
        select_rows(d.A, perm.data(), d.n_free, inv.data(), a);
        upload_sell<T>(c->A, a.v, true);
        // the CG kernels stream A's (and G's) slices assuming width >= 1:
        // every free row holds its diagonal
        if (c->A.nslices > 0 && c->A.min_width < 1) throw std::runtime_error("A has a 64-row slice without entries");
        c->hA_ptr = a.rp;
        c->hA_col = a.col;
        c->hA_val = a.val;
        c->hA_sptr = sell_layout(a.v);
        std::vector<int32_t> fe(d.n_free);
        for (int32_t k = 0; k < d.n_free; ++k) fe[k] = d.free_edges[perm[k]];
        if (p->method == TF_METHOD_FEM) {
            // fem residual: E and SigMat on the free rows (fem_w_kernel)
            c->fem = true;
            CsrBuf m;
            select_rows(d.E, fe.data(), d.n_free, nullptr, m);
            upload_sell<T>(c->Kf, m.v, true);
            c->k_dir_only = false;
            if (d.Sig.nnz > 0) {
                select_rows(d.Sig, fe.data(), d.n_free, nullptr, m);
            } else {
                m = CsrBuf{};
                m.rp.assign(d.n_free + 1, 0);
                m.v = tf_csr{};
                m.v.nrows = d.n_free;
                m.v.ncols = E;
                m.v.rowptr = m.rp.data();
                m.v.col = m.col.data();
                m.v.val = m.val.data();
            }
            upload_sell<T>(c->R1f, m.v, true);
            select_rows(d.Ct, fe.data(), d.n_free, nullptr, m);
            upload_sell<T>(c->Ctf, m.v, false);
            for (DevBuf* b : {&c->U_older, &c->fw, &c->fw2}) zero_field<T>(*b, E, c->stream);
        } else {
            CsrBuf m;
            residual_operator(d, p->dt, perm, fe, m);
            upload_sell<T>(c->Kf, m.v, true);
            std::vector<uint8_t> free_col(E, 0);
            for (int32_t e : fe) free_col[e] = 1;
            c->k_dir_only = true;
            for (int32_t col : m.col)
                if (free_col[col]) c->k_dir_only = false;
            select_rows(d.RHS1, fe.data(), d.n_free, nullptr, m);
            upload_sell<T>(c->R1f, m.v, true);
            select_rows(d.Ct, fe.data(), d.n_free, nullptr, m);
            upload_sell<T>(c->Ctf, m.v, false);
        }
        c->n_free = d.n_free;
        upload(c->free_edge, fe);
        std::vector<int32_t> er(E, -1);
        for (int32_t k = 0; k < d.n_free; ++k) er[fe[k]] = k;
        upload(c->edge_row, er);
        std::vector<uint8_t> is_free(E, 0);
        for (int32_t e : fe) is_free[e] = 1;
        c->h_dir_edge.clear();
        for (int32_t e = 0; e < E; ++e)
            if (!is_free[e]) c->h_dir_edge.push_back(e);
        c->n_dir = (int32_t)c->h_dir_edge.size();
        upload(c->dir_edge, c->h_dir_edge);
        // dir_new holds the Dirichlet values (ComputeEdgeBC for pec edges)
        std::vector<T> dn(E, 0);
        for (int32_t e : c->h_dir_edge) dn[e] = (T)c->h_fixed_value[e];
        upload(c->dir_new, dn);
        zero_field<T>(c->fb, E, c->stream);
        // Jacobi preconditioner: 1/A_jj, or 1 if missing / zero (CG.hpp:119-128)
        std::vector<T> pre(d.n_free, (T)1);
        for (int32_t i = 0; i < d.n_free; ++i)
            for (int64_t q = a.rp[i]; q < a.rp[i + 1]; ++q)
                if (a.col[q] == i && a.val[q] != 0) pre[i] = (T)(1.0 / a.val[q]);
        upload(c->pre, pre);
        const int64_t n = std::max<int32_t>(d.n_free, 1);
        for (DevBuf* b : {&c->x, &c->r, &c->z, &c->p0, &c->p1, &c->q, &c->r0}) zero_field<T>(*b, n, c->stream);
        for (DevBuf* b : {&c->rs0, &c->rs1}) zero_field<RS<T>>(*b, n, c->stream);
        c->cg_part.alloc(sizeof(double) * (kCgPartA + 3 * kCgPartB + 2 * kCgPartC));
        // two CG states (device and pinned host copies): the solve of step k
        // uses slot k & 1, so a step's kernels can run ahead of the host's
        // check of the previous solve (run_gfdtd)
        c->cg_state.alloc(2 * sizeof(CgState));
        TF_HIP(hipMemsetAsync(c->cg_state.p, 0, 2 * sizeof(CgState), c->stream));
        TF_HIP(hipHostMalloc((void**)&c->h_state, 2 * sizeof(CgState), hipHostMallocDefault));
        std::memset((void*)c->h_state, 0, 2 * sizeof(CgState));
    }
    // wide-mesh kernels when some operator needs windowed columns, or when
    // the operators a step streams repeatedly cannot stay in the 256 MB
    // Infinity Cache: dga reads H, N and Ct once per step, one step after
    // another; gfdtd re-reads A (and the learned G, 4 bytes per A entry) in
    // every CG pass of a step
    auto bytes = [](const Sell& m) {
        return m.sptr.bytes + (m.c16 ? m.col16.bytes : m.col.bytes) + m.val.bytes + m.sg.bytes;
    };
    c->wide = false;
    for (const Sell* m : {&c->sCt, &c->H, &c->N, &c->A, &c->Kf, &c->R1f, &c->Ctf, &c->Efull, &c->Sig, &c->fKu,
                          &c->fRu, &c->fKf, &c->fRf})
        c->wide = c->wide || m->cwin;
    const size_t streamed = p->method == TF_METHOD_DGA
                                ? bytes(c->sCt) + bytes(c->H) + bytes(c->N) + bytes(c->fKu) + bytes(c->fRu) +
                                      bytes(c->fKf) + bytes(c->fRf)
                                : bytes(c->A) + 4 * (size_t)c->A.entries;
    c->wide = c->wide || streamed > ((size_t)256 << 20);
    // $TETFIT_SELL_WIDE (non-empty, not 0) selects the wide-mesh kernels on
    // any mesh, to exercise them with one-base columns on small meshes
    if (const char* v = std::getenv("TETFIT_SELL_WIDE"))
        if (*v && std::strcmp(v, "0") != 0) c->wide = true;
    // initial U: fixed (class-4) values
    std::vector<T> u0(E, 0);
    for (int32_t e = 0; e < E; ++e)
        if (d.edge_class[e] == 4) u0[e] = (T)c->h_fixed_value[e];
    TF_HIP(hipStreamSynchronize(c->stream));
    TF_HIP(hipMemcpy(c->U.p, u0.data(), E * sizeof(T), hipMemcpyHostToDevice));
}

// Special (source) list of the FIT E phase from the current terms.
template <class T>
void build_fit_special(tf_ctx* c, const std::vector<std::vector<std::pair<int32_t, double>>>& by_edge_terms,
                       const std::vector<int>& edge_field) {
    std::vector<int32_t> edge, tb, te, tsrc;
    std::vector<int4> f4;
    std::vector<uint8_t> sg;
    std::vector<double> ca, cb, tw;
    std::vector<int8_t> kind;
    std::vector<uint8_t> main_sg = c->h_fe_sg;
    std::vector<uint16_t> main_code = c->h_fe_code;
    auto off_main = [&](int32_t e) {  // the special item updates it
        if (c->fit_grid) main_code[c->h_esid[e]] = 0;
        else main_sg[e] = 0;
    };
    auto dev_f4 = [&](int4 q) {
        return c->fit_grid ? int4{dev_face(c, q.x), dev_face(c, q.y), dev_face(c, q.z), dev_face(c, q.w)} : q;
    };
    auto add_terms = [&](int32_t e) {
        tb.push_back((int32_t)tsrc.size());
        auto it = by_edge_terms.size() > (size_t)e ? &by_edge_terms[e] : nullptr;
        if (it)
            for (auto& pr : *it) { tsrc.push_back(pr.first); tw.push_back(pr.second); }
        te.push_back((int32_t)tsrc.size());
    };
    for (size_t i = 0; i < c->fit_uncommon.size(); ++i) {
        const int32_t e = c->fit_uncommon[i];
        edge.push_back(dev_edge(c, e));
        f4.push_back(dev_f4(c->fit_sp_f4_u[i]));
        sg.push_back(c->fit_sp_sg_u[i]);
        ca.push_back(c->fit_sp_ca_u[i]);
        cb.push_back(c->fit_sp_cb_u[i]);
        kind.push_back(0);
        add_terms(e);
    }
    for (int32_t e : c->fit_src) {
        const int fld = edge_field.size() > (size_t)e ? edge_field[e] : TF_FIELD_E;
        edge.push_back(dev_edge(c, e));
        if (fld == TF_FIELD_J) {
            f4.push_back(dev_f4(c->h_fe_f4[e]));
            sg.push_back(c->h_fe_sg[e]);
            ca.push_back(c->h_ca_common[e]);
            cb.push_back(c->h_cb_common[e]);
            kind.push_back(2);
        } else {
            f4.push_back(int4{0, 0, 0, 0});
            sg.push_back(0);
            ca.push_back(0);
            cb.push_back(0);
            kind.push_back(1);
        }
        off_main(e);
        add_terms(e);
    }
    c->n_sp = (int32_t)edge.size();
    upload(c->sp_edge, edge);
    upload(c->sp_f4, f4);
    upload(c->sp_sg, sg);
    upload(c->sp_ca, ca);
    upload(c->sp_cb, cb);
    upload(c->sp_kind, kind);
    upload(c->sp_tb, tb);
    upload(c->sp_te, te);
    if (tsrc.empty()) { tsrc.push_back(0); tw.push_back(0); }
    upload(c->term_src, tsrc);
    upload(c->term_w, tw);
    if (c->fit_grid) upload(c->fe_code, main_code);
    else upload(c->fe_sg, main_sg);
}

void build_tet_inject(tf_ctx* c, const std::vector<std::vector<std::pair<int32_t, double>>>& by_edge_terms,
                      const std::vector<int>& edge_field) {
    std::vector<int32_t> edge, tb, te, tsrc;
    std::vector<int8_t> kind;
    std::vector<double> tw;
    for (int32_t e = 0; e < c->E; ++e) {
        if (by_edge_terms[e].empty()) continue;
        // a d edge (neither h, j nor e) is set to its e-source sum, 0, by the
        // dga loop before the H update (D:835-842); gfdtd overwrites it with
        // the solve, so nothing is injected there
        const bool solved = c->method == TF_METHOD_GFDTD || c->method == TF_METHOD_FEM;
        if (edge_field[e] == TF_FIELD_D && solved) continue;
        edge.push_back(e);
        kind.push_back(edge_field[e] == TF_FIELD_H ? 0 : edge_field[e] == TF_FIELD_J ? 2 : 1);
        if (edge_field[e] == TF_FIELD_J && !solved)
            throw std::runtime_error("j sources are not supported for dga");
        tb.push_back((int32_t)tsrc.size());
        for (auto& pr : by_edge_terms[e]) { tsrc.push_back(pr.first); tw.push_back(pr.second); }
        te.push_back((int32_t)tsrc.size());
    }
    c->n_inj = (int32_t)edge.size();
    c->dir_zero_static = c->k_dir_only && std::find(kind.begin(), kind.end(), (int8_t)1) == kind.end();
    for (int32_t e : c->h_dir_edge)
        if (c->h_fixed_value[e] != 0) c->dir_zero_static = false;
    upload(c->inj_edge, edge);
    upload(c->inj_kind, kind);
    upload(c->inj_tb, tb);
    upload(c->inj_te, te);
    if (tsrc.empty()) { tsrc.push_back(0); tw.push_back(0); }
    upload(c->term_src, tsrc);
    upload(c->term_w, tw);
}

// ---------------------------------------------------------------- run

bool allow_print(tf_ctx* c, double t) {
    if (!c->out_enabled) return false;
    if (t >= c->next_print) {
        c->next_print += c->period;
        return true;
    }
    return false;
}

// Output schedule: AllowPrint for single simulations; a fixed step period
// for batches (each instance's time is step * its own dt).
bool want_sample(tf_ctx* c, int64_t i, double t) {
    if (c->batch) {
        if (c->batch->every <= 0 || i % c->batch->every != 0) return false;
        c->batch->sample_steps.push_back(i);
        return true;
    }
    return allow_print(c, t);
}

// Host bookkeeping of one output sample at time t: its slot in the sample
// buffers (probe values, energies; grown when full), or -1 when the sample
// records nothing on the device (no probes, no energy output).
int64_t sample_reserve(tf_ctx* c, double t) {
    c->sample_times.push_back(t);
    const int64_t slot = c->n_samples++;
    if (c->n_probes == 0 && !c->energy_out) return -1;
    if (c->n_samples > c->sample_cap) {
        const int64_t cap = std::max<int64_t>(1024, c->sample_cap * 2);
        auto grow = [&](DevBuf& buf, int64_t per) {
            DevBuf nb;
            nb.alloc(cap * per * sizeof(double));
            if (slot && buf.p)
                TF_HIP(hipMemcpyAsync(nb.p, buf.p, slot * per * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            TF_HIP(hipStreamSynchronize(c->stream));
            std::swap(nb.p, buf.p);
            std::swap(nb.bytes, buf.bytes);
        };
        if (c->n_probes > 0) grow(c->samples, (int64_t)c->n_probes * 6);
        if (c->energy_out) grow(c->eser, 2);
        c->sample_cap = cap;
    }
    return slot;
}

// Partials per energy reduction (tf_energy uses the same grids, so a sample
// equals tf_energy of the same state bit for bit).
constexpr int kEnergyParts = 1024;

// Launches the two energy reductions of the current state into epart and
// returns their partial counts (ge, gn): dga / gfdtd / fem We = 1/2 U.E U,
// Wh = 1/2 B.N B (D:2773-2775), fdtd with the ep_vec / nu_vec weights
// (D:3468-3470).
template <class T>
void energy_partials(tf_ctx* c, const CgState* guard, int& ge, int& gn) {
    double* pp = c->epart.as<double>();
    if (c->method != TF_METHOD_FDTD) {
        ge = sell_blocks(c->Efull, kEnergyParts);
        gn = sell_blocks(c->N, kEnergyParts);
        TF_LAUNCH_NB(c->wide, 8, quad_form_kernel, T, ge, c->stream, c->Efull.vals<T>(), c->U.as<T>(), pp, guard);
        TF_LAUNCH_NB(c->wide, 8, quad_form_kernel, T, gn, c->stream, c->N.vals<T>(), c->B.as<T>(),
                     pp + kEnergyParts, guard);
    } else {
        ge = gn = kEnergyParts;
        weighted_sq_kernel<T><<<kEnergyParts, kBlock, 0, c->stream>>>(c->E, c->U.as<T>(), c->fit_ep.as<double>(), pp);
        weighted_sq_kernel<T><<<kEnergyParts, kBlock, 0, c->stream>>>(c->F, c->B.as<T>(), c->fit_nu.as<double>(),
                                                                      pp + kEnergyParts);
    }
}

template <class T>
void sample_launch(tf_ctx* c, int64_t slot, const CgState* guard) {
    if (slot < 0) return;
    if (c->n_probes > 0) {
        ev_begin(c, "probe_kernel");
        probe_kernel<T><<<blocks_for(c->n_probes * 6, 64), 64, 0, c->stream>>>(
            c->n_probes, c->probes.as<ProbeDev>(), c->U.as<T>(), c->B.as<T>(),
            c->samples.as<double>() + slot * c->n_probes * 6, guard);
        ev_end(c);
    }
    if (c->energy_out) {
        ev_begin(c, "energy_kernels");
        int ge = 0, gn = 0;
        energy_partials<T>(c, guard, ge, gn);
        energy_sum_kernel<<<1, 64, 0, c->stream>>>(c->epart.as<double>(), ge, c->epart.as<double>() + kEnergyParts,
                                                   gn, c->eser.as<double>() + 2 * slot, guard);
        ev_end(c);
    }
}

template <class T>
void sample(tf_ctx* c, double t) {
    sample_launch<T>(c, sample_reserve(c, t), nullptr);
}

void ensure_carr(tf_ctx* c, int64_t first, int64_t n) {
    // carriers for steps first-1 .. first+n-1 (row 0 is the previous step)
    const size_t ns = std::max<size_t>(c->src.size(), 1);
    std::vector<double> tab((size_t)(n + 1) * ns, 0.0);
    const bool per_src = c->src_dt.size() == c->src.size();
    for (int64_t k = 0; k <= n; ++k) {
        const int64_t i = first - 1 + k;
        for (size_t s = 0; s < c->src.size(); ++s) {
            const tf_source& so = c->src[s];
            const double dt = per_src ? c->src_dt[s] : c->dt;
            const double t = (double)i * dt;
            const double ts = so.field == TF_FIELD_H ? t - 0.5 * dt : t;
            tab[k * ns + s] = (i < 0) ? 0.0 : carrier_value(so, ts);
        }
    }
    if ((int64_t)tab.size() > c->carr_cap) {
        c->carr.alloc(tab.size() * sizeof(double));
        c->carr_cap = (int64_t)tab.size();
    }
    TF_HIP(hipMemcpyAsync(c->carr.p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    TF_HIP(hipStreamSynchronize(c->stream));
}

constexpr int kGraphSteps = 64;

void drop_graph(tf_ctx* c) {
    if (c->sg.exec) (void)hipGraphExecDestroy(c->sg.exec);
    if (c->sg.graph) (void)hipGraphDestroy(c->sg.graph);
    c->sg.exec = nullptr;
    c->sg.graph = nullptr;
}

// Graph replay is used when no launch of the chunk depends on the host: no
// probe sampling and no per-kernel event timing.
bool graph_steps(const tf_ctx* c) {
    if (!c->use_graphs || c->profiling) return false;
    return c->batch ? c->batch->every <= 0 : !c->out_enabled;
}

// Captures kGraphSteps calls of step(local step, carrier row, Poynting slot)
// into the context's graph (once), then runs steps [k0, k0 + m) of the
// current tf_run in chunks of kGraphSteps by replay; per_step = Poynting
// partials per step. Returns the number of steps run (a multiple of
// kGraphSteps); the caller runs the rest with ordinary launches.
template <class Step>
int64_t run_graph_chunks(tf_ctx* c, int64_t m, int per_step, Step step) {
    const int64_t chunks = m / kGraphSteps;
    if (chunks == 0) return 0;
    const size_t ns = std::max<size_t>(c->src.size(), 1);
    if (!c->sg.exec) {
        c->sg.carr.alloc((kGraphSteps + 1) * ns * sizeof(double));
        c->sg.poy.alloc(std::max(1, kGraphSteps * per_step) * sizeof(double));
        TF_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        try {
            for (int k = 0; k < kGraphSteps; ++k)
                step(k, c->sg.carr.as<double>() + (size_t)(k + 1) * ns, c->sg.poy.as<double>() + (size_t)k * per_step);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(c->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        TF_HIP(hipStreamEndCapture(c->stream, &c->sg.graph));
        TF_HIP(hipGraphInstantiate(&c->sg.exec, c->sg.graph, nullptr, nullptr, 0));
    }
    for (int64_t q = 0; q < chunks; ++q) {
        const int64_t k0 = q * kGraphSteps;
        // carrier rows k0 .. k0 + kGraphSteps of this run (row 0 = the step before)
        TF_HIP(hipMemcpyAsync(c->sg.carr.p, c->carr.as<double>() + (size_t)k0 * ns,
                              (kGraphSteps + 1) * ns * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        TF_HIP(hipGraphLaunch(c->sg.exec, c->stream));
        if (per_step)
            TF_HIP(hipMemcpyAsync(c->poy.as<double>() + (size_t)k0 * per_step, c->sg.poy.p,
                                  (size_t)kGraphSteps * per_step * sizeof(double), hipMemcpyDeviceToDevice,
                                  c->stream));
    }
    return chunks * kGraphSteps;
}

template <class T>
void run_fit(tf_ctx* c, int64_t first, int64_t n, tf_stats* st) {
    uint32_t ecnt[3], fcnt[3];
    fit_counts(c->grid, false, ecnt);
    fit_counts(c->grid, true, fcnt);
    const int main_blocks = c->fit_grid ? slab_grid(ecnt) : blocks_for(c->E);
    const int h_blocks = c->fit_grid ? slab_grid(fcnt) : blocks_for(c->F);
    const int sp_blocks = c->n_sp ? blocks_for(c->n_sp) : 0;
    const size_t ns = std::max<size_t>(c->src.size(), 1);
    if (sp_blocks) {
        if (c->poy_cap < n * sp_blocks) {
            c->poy.alloc(n * sp_blocks * sizeof(double));
            c->poy_cap = n * sp_blocks;
        }
    }
    // one step: E phase (with sources and the Poynting partials), H phase
    auto step = [&](const double* cs, double* poy) {
        launch_timed(c, "fit_e_kernel", fit_e_kernel<T>, main_blocks + sp_blocks,
            c->E, sp_blocks, c->fe_sg.as<uint8_t>(), c->fe_ca.as<T>(), c->fe_cb.as<T>(), c->fe_f4.as<int4>(),
            c->fit_grid ? c->fe_code.as<uint16_t>() : nullptr, c->grid, c->Fv.as<T>(), c->U.as<T>(), c->n_sp, c->sp_edge.as<int32_t>(), c->sp_f4.as<int4>(),
            c->sp_sg.as<uint8_t>(), c->sp_ca.as<double>(), c->sp_cb.as<double>(), c->sp_kind.as<int8_t>(),
            c->sp_tb.as<int32_t>(), c->sp_te.as<int32_t>(), c->term_src.as<int32_t>(), c->term_w.as<double>(), cs,
            sp_blocks ? poy : nullptr);
        launch_timed(c, "fit_h_kernel", fit_h_kernel<T>, h_blocks, c->F, c->fh_sg.as<uint8_t>(),
                     c->fh_cm.as<T>(), c->fh_cn.as<T>(), c->fh_e4.as<int4>(),
                     c->fit_grid ? c->fh_code.as<uint16_t>() : nullptr, c->grid, c->U.as<T>(), c->Fv.as<T>(),
                     c->B.as<T>(), c->dt);
    };
    const int64_t g = graph_steps(c) ? run_graph_chunks(c, n, sp_blocks, [&](int, const double* cs, double* poy) {
        step(cs, poy);
    }) : 0;
    for (int64_t k = g; k < n; ++k) {
        const int64_t i = first + k;
        const double t = (double)i * c->dt;
        step(c->carr.as<double>() + (k + 1) * ns, c->poy.as<double>() + k * sp_blocks);
        if (allow_print(c, t)) sample<T>(c, t);
        if (c->profiling && (k & 63) == 63) ev_flush(c);
    }
    TF_HIP(hipGetLastError());
    // Poynting per step
    std::vector<double> part((size_t)n * std::max(sp_blocks, 1), 0.0);
    if (sp_blocks)
        TF_HIP(hipMemcpyAsync(part.data(), c->poy.p, n * sp_blocks * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
    TF_HIP(hipStreamSynchronize(c->stream));
    for (int64_t k = 0; k < n; ++k) {
        double s = 0;
        for (int b = 0; b < sp_blocks; ++b) s += part[k * sp_blocks + b];
        c->poynting.push_back(s);
    }
    st->steps_done += n;
}

// Lossy dga (D:810-939 with edge classes 2/3 and face classes 2/3): one
// step is curl_f = Ct F with the h sources, the boundary values and e
// sources into U_frac, U_frac' = Ru U_frac + dt Ku curl_f, U = Mu U_frac',
// the Poynting partials, curl_u = C U with the B update, F_frac' = Rf F_frac
// - dt Kf curl_u, F = Tf F_frac'; the two U_frac (F_frac) buffers swap
// roles every step. Launched from the host step by step (no graph).
template <class T>
void run_dga_frac(tf_ctx* c, int64_t first, int64_t n, tf_stats* st) {
    const int inj_blocks = c->n_inj ? blocks_for(c->n_inj) : 0;
    const size_t ns = std::max<size_t>(c->src.size(), 1);
    if (inj_blocks && c->poy_cap < 2 * n * inj_blocks) {
        c->poy.alloc(2 * n * inj_blocks * sizeof(double));
        c->poy_cap = 2 * n * inj_blocks;
    }
    auto inject = [&](const double* cs, double* poy, int init) {
        if (!c->n_inj) return;
        launch_timed(c, "frac_inject_kernel", frac_inject_kernel<T>, inj_blocks, c->n_inj, c->inj_edge.as<int32_t>(),
                     c->inj_kind.as<int8_t>(), c->inj_tb.as<int32_t>(), c->inj_te.as<int32_t>(),
                     c->term_src.as<int32_t>(), c->term_w.as<double>(), cs, c->curl_f.as<T>(),
                     (const T*)c->U.as<T>(), c->uf[c->fcur].as<T>(), c->fslot.as<int32_t>(),
                     c->mu_a.as<int32_t>(), c->mu_b.as<int32_t>(), c->q0, c->q1, poy, init);
    };
    auto gather_u = [&]() {
        launch_timed(c, "frac_gather_kernel", frac_gather_kernel<T>, blocks_for(c->E), c->E, c->mu_a.as<int32_t>(),
                     c->mu_b.as<int32_t>(), (const T*)c->uf[c->fcur].as<T>(), c->U.as<T>());
    };
    if (!c->initial_done && first == 1) {
        // the pass before the loop (D:726-790), then the export at t = 0
        inject(c->carr.as<double>(), c->poy.as<double>(), 1);
        gather_u();
        if (want_sample(c, 0, 0.0)) sample<T>(c, 0.0);
    }
    c->initial_done = true;
    const int gb_ct = tuned_blocks(c, TK_CT, c->sCt, kSpmvCap);
    const int gb_u = sell_blocks(c->fKu, kSpmvCap), gb_f = sell_blocks(c->fKf, kSpmvCap);
    for (int64_t k = 0; k < n; ++k) {
        const int64_t i = first + k;
        const double t = (double)i * c->dt;
        const double* cs = c->carr.as<double>() + (k + 1) * ns;
        double* poy = c->poy.as<double>() + 2 * k * inj_blocks;
        TF_LAUNCH_TIMED(c, "sell_ct_kernel", c->tune[TK_CT].unroll, sell_ct_kernel, T, gb_ct,
                        tuned_view(c, TK_CT, c->sCt.signs()), c->Fv.as<T>(), c->curl_f.as<T>());
        inject(cs, poy, 0);
        const int nx = c->fcur ^ 1;
        if (c->nU > 0)
            TF_LAUNCH_TIMED(c, "sell_frac_kernel", 8, sell_frac_kernel, T, gb_u, c->fKu.vals<T>(), c->fRu.vals<T>(),
                            (const T*)c->curl_f.as<T>(), (const T*)c->uf[c->fcur].as<T>(), c->uf[nx].as<T>(), c->dt);
        launch_timed(c, "frac_gather_kernel", frac_gather_kernel<T>, blocks_for(c->E), c->E, c->mu_a.as<int32_t>(),
                     c->mu_b.as<int32_t>(), (const T*)c->uf[nx].as<T>(), c->U.as<T>());
        if (c->n_inj)
            launch_timed(c, "dga_poynting_kernel", dga_poynting_kernel<T>, inj_blocks, c->n_inj,
                         c->inj_edge.as<int32_t>(), c->inj_kind.as<int8_t>(), c->inj_tb.as<int32_t>(),
                         c->inj_te.as<int32_t>(), c->term_src.as<int32_t>(), c->term_w.as<double>(), cs,
                         (const T*)c->U.as<T>(), poy + inj_blocks);
        launch_timed(c, "dga_curl_kernel", dga_curl_kernel<T>, blocks_for(c->F), c->F, c->c_col.as<int32_t>(),
                     c->c_fixed_signs ? nullptr : c->c_sg.as<int8_t>(), (const T*)c->U.as<T>(), c->curl_u.as<T>(),
                     c->B.as<T>(), c->dt, (const double*)nullptr, (const CgState*)nullptr, CurlSources<T>{});
        if (c->nF > 0)
            TF_LAUNCH_TIMED(c, "sell_frac_kernel", 8, sell_frac_kernel, T, gb_f, c->fKf.vals<T>(), c->fRf.vals<T>(),
                            (const T*)c->curl_u.as<T>(), (const T*)c->ff[c->fcur].as<T>(), c->ff[nx].as<T>(), -c->dt);
        launch_timed(c, "frac_gather_kernel", frac_gather_kernel<T>, blocks_for(c->F), c->F, c->tf_a.as<int32_t>(),
                     c->tf_b.as<int32_t>(), (const T*)c->ff[nx].as<T>(), c->Fv.as<T>());
        c->fcur = nx;
        if (want_sample(c, i, t)) sample<T>(c, t);
        if (c->profiling && (k & 63) == 63) ev_flush(c);
    }
    TF_HIP(hipGetLastError());
    std::vector<double> part((size_t)2 * n * std::max(inj_blocks, 1), 0.0);
    if (inj_blocks)
        TF_HIP(hipMemcpyAsync(part.data(), c->poy.p, 2 * n * inj_blocks * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
    TF_HIP(hipStreamSynchronize(c->stream));
    for (int64_t k = 0; k < n; ++k) {
        double sum = 0;
        for (int b = 0; b < 2 * inj_blocks; ++b) sum += part[2 * k * inj_blocks + b];
        c->poynting.push_back(sum);
    }
    st->steps_done += n;
}

// The context's injection items as dga_curl_kernel's appended source work
// (mode 1: dga second Poynting half; mode 2: gfdtd / fem sources).
template <class T>
CurlSources<T> curl_sources(tf_ctx* c, int mode, const double* carr_step, const double* carr_prev, double* poy) {
    CurlSources<T> s;
    s.n_inj = c->n_inj;
    s.mode = mode;
    s.edge = c->inj_edge.as<int32_t>();
    s.kind = c->inj_kind.as<int8_t>();
    s.tb = c->inj_tb.as<int32_t>();
    s.te = c->inj_te.as<int32_t>();
    s.term_src = c->term_src.as<int32_t>();
    s.term_w = c->term_w.as<double>();
    s.carr_step = carr_step;
    s.carr_prev = carr_prev;
    s.fb = c->fb.p ? c->fb.as<T>() : nullptr;
    s.dir_new = c->dir_new.p ? c->dir_new.as<T>() : nullptr;
    s.delta = c->fem ? 1 : 0;
    s.poy_part = poy;
    return s;
}

template <class T>
void run_dga(tf_ctx* c, int64_t first, int64_t n, tf_stats* st) {
    if (c->frac) {
        run_dga_frac<T>(c, first, n, st);
        return;
    }
    const int inj_blocks = c->n_inj ? blocks_for(c->n_inj) : 0;
    const size_t ns = std::max<size_t>(c->src.size(), 1);
    const double* hdt = c->batch ? c->row_dt_h.as<double>() : nullptr;
    const double* fdt = c->batch ? c->row_dt_f.as<double>() : nullptr;
    const int gb_ct = tuned_blocks(c, TK_CT, c->sCt, kSpmvCap), gb_h = tuned_blocks(c, TK_H, c->H, kSpmvCap);
    const int gb_n = tuned_blocks(c, TK_N, c->N, kSpmvCap);
    // Poynting partials of a step: the injection kernel's workgroups (first
    // half, 0.5 U_old v), then the workgroups appended to dga_curl (second
    // half, 0.5 U_new v)
    const int per_step = c->n_inj ? 2 * inj_blocks : 0;
    if (per_step && c->poy_cap < (int64_t)per_step * n) {
        c->poy.alloc((size_t)per_step * n * sizeof(double));
        c->poy_cap = (int64_t)per_step * n;
    }
    if (!c->initial_done && first == 1) {
        // initial export at t = 0 (D:805-806), after the t = 0 e-source values
        if (c->n_inj) {
            const double* cs0 = c->carr.as<double>();  // row 0 = step first-1 = 0
            dga_inject_kernel<T><<<inj_blocks, kBlock, 0, c->stream>>>(
                c->n_inj, c->inj_edge.as<int32_t>(), c->inj_kind.as<int8_t>(), c->inj_tb.as<int32_t>(),
                c->inj_te.as<int32_t>(), c->term_src.as<int32_t>(), c->term_w.as<double>(), cs0, c->curl_f.as<T>(),
                c->U.as<T>(), c->poy.as<double>(), 1);
        }
        if (want_sample(c, 0, 0.0)) sample<T>(c, 0.0);
    }
    c->initial_done = true;
    // one step (D:810-939): curl_f = Ct F, the sources (h into curl_f, e into
    // U, the first Poynting half), U += dt H curl_f, curl_u = C U with the B
    // update and, in dga_curl's appended workgroups, the second Poynting half,
    // F -= dt N curl_u
    auto step = [&](const double* cs, double* poy) {
        TF_LAUNCH_TIMED(c, "sell_ct_kernel", c->tune[TK_CT].unroll, sell_ct_kernel, T, gb_ct,
                        tuned_view(c, TK_CT, c->sCt.signs()), c->Fv.as<T>(), c->curl_f.as<T>());
        if (c->n_inj)
            launch_timed(c, "dga_inject_kernel", dga_inject_kernel<T>, inj_blocks, c->n_inj, c->inj_edge.as<int32_t>(),
                         c->inj_kind.as<int8_t>(), c->inj_tb.as<int32_t>(), c->inj_te.as<int32_t>(),
                         c->term_src.as<int32_t>(), c->term_w.as<double>(), cs, c->curl_f.as<T>(), c->U.as<T>(), poy,
                         0);
        TF_LAUNCH_TIMED(c, "sell_h_kernel", c->tune[TK_H].unroll, sell_h_kernel, T, gb_h,
                        tuned_view(c, TK_H, c->H.vals<T>()), c->h_edge.as<int32_t>(), c->curl_f.as<T>(), c->U.as<T>(),
                        c->dt, hdt);
        const CurlSources<T> py = c->n_inj ? curl_sources<T>(c, 1, cs, nullptr, poy + inj_blocks) : CurlSources<T>{};
        launch_timed(c, "dga_curl_kernel", dga_curl_kernel<T>, blocks_for(c->F) + (c->n_inj ? inj_blocks : 0), c->F,
                     c->c_col.as<int32_t>(), c->c_fixed_signs ? nullptr : c->c_sg.as<int8_t>(), c->U.as<T>(),
                     c->curl_u.as<T>(), c->B.as<T>(), c->dt, fdt, (const CgState*)nullptr, py);
        TF_LAUNCH_TIMED(c, "sell_n_kernel", c->tune[TK_N].unroll, sell_n_kernel, T, gb_n,
                        tuned_view(c, TK_N, c->N.vals<T>()), c->curl_u.as<T>(), c->Fv.as<T>(), c->dt, 0, fdt, nullptr);
    };
    const int64_t g = graph_steps(c) ? run_graph_chunks(c, n, per_step, [&](int, const double* cs, double* poy) {
        step(cs, poy);
    }) : 0;
    for (int64_t k = g; k < n; ++k) {
        const int64_t i = first + k;
        const double t = (double)i * c->dt;
        step(c->carr.as<double>() + (k + 1) * ns, c->poy.as<double>() + (size_t)per_step * k);
        if (want_sample(c, i, t)) sample<T>(c, t);
        if (c->profiling && (k & 63) == 63) ev_flush(c);
    }
    TF_HIP(hipGetLastError());
    std::vector<double> part((size_t)std::max(per_step, 1) * n, 0.0);
    if (per_step)
        TF_HIP(hipMemcpyAsync(part.data(), c->poy.p, (size_t)per_step * n * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
    TF_HIP(hipStreamSynchronize(c->stream));
    for (int64_t k = 0; k < n; ++k) {
        double s = 0;
        for (int b = 0; b < per_step; ++b) s += part[(size_t)per_step * k + b];
        c->poynting.push_back(s);
    }
    st->steps_done += n;
}

// ---- CG solves on the device (single-reduction PCG, see sell.hpp) --------
// A solve is gf_rhs_init (or cg_restart for a redo), then passes it =
// 0, 1, ... of (cg_w, cg_upd) with Jacobi, or (cg_pg, cg_wu) with the
// learned preconditioner. Passes are launched in batches sized from the
// previous solve's count; kernels past the stop test return at once.
// `snaps` lists (iteration k, bank slot) captures of the iterate (r_k, x_k)
// (Jacobi solves), taken in stream order after the update that produces it.
//
// Vectors: u = z, w = q, p = p0, s = p1; the learned passes keep (r, s)
// pairs in rs0 (even passes) and rs1 (odd passes). Partials (fp64, one per
// block): [delta: kCgPartA][rr: kCgPartB][gamma ping, pong: 2 kCgPartB]
// [gamma of cg_pg: kCgPartC][rr of cg_pg: kCgPartC].
struct CgSolve {
    CgState init{};
    bool learned = false;
    const std::vector<std::pair<int, int>>* snaps = nullptr;
    int slot = 0;  // CG state slot (device cg_state[slot], pinned h_state[slot])
    bool dir_terms = true;  // gf_rhs_init: apply K U_old - RHS1 SrcFld (false: known to be zero)
    int it = 0;    // next pass
    std::vector<std::array<int64_t, 3>> tags;  // profiled launches: (record, pass, first kernel of a learned pass)
};

template <class T>
struct CgView {
    int32_t nf;
    int nb_w, nb_b, nb_g;
    double *part_del, *part_rr, *part_gj[2], *part_gl, *part_rl;
    T *x, *u, *w, *p, *sv, *rj;
    RS<T>* rs[2];
    SellView<T> A;
    SellView<uint16_t> G;
    const uint32_t* gpk;  // G packed with A's 16-bit columns, or null (G = bf16 values beside A's columns)
    CgView(tf_ctx* c, bool learned) {
        nf = c->n_free;
        nb_w = tuned_blocks(c, learned ? TK_WU : TK_W, c->A, kCgPartA);
        nb_b = std::max(1, std::min(blocks_for(nf), kCgPartB));
        nb_g = tuned_blocks(c, TK_G, c->A, kCgPartC);
        part_del = c->cg_part.as<double>();
        part_rr = part_del + kCgPartA;
        part_gj[0] = part_rr + kCgPartB;
        part_gj[1] = part_rr + 2 * kCgPartB;
        part_gl = part_rr + 3 * kCgPartB;
        part_rl = part_gl + kCgPartC;
        x = c->x.as<T>();
        u = c->z.as<T>();
        w = c->q.as<T>();
        p = c->p0.as<T>();
        sv = c->p1.as<T>();
        rj = c->r.as<T>();
        rs[0] = c->rs0.as<RS<T>>();
        rs[1] = c->rs1.as<RS<T>>();
        A = tuned_view(c, learned ? TK_WU : TK_W, c->A.vals<T>());
        G = tuned_view(c, TK_G, c->A.view<uint16_t>(c->lp.gval));
        gpk = c->lp.gpk.p ? c->lp.gpk.as<uint32_t>() : nullptr;
    }
};

// Copies the solve's initial state to its slot and launches the start of the
// solve: with_rhs, the step's initial residual (gf_rhs_init_kernel, guarded,
// which also keeps it in r0); otherwise a restart from r0 (cg_restart_kernel)
// with x back at the step's warm start.
// extrap (with_rhs only): -1 warm start; 0..2 the order of the extrapolated
// start (0: history shift only); launch_extrap false reuses the shift d of
// an earlier start of the same solve (x already holds the start).
template <class T>
void cg_start(tf_ctx* c, CgSolve& cs, bool with_rhs, const CgState* guard, int extrap = -1,
              bool launch_extrap = true) {
    const CgView<T> v(c, cs.learned);
    CgState* dst = c->cg_state.as<CgState>() + cs.slot;
    cs.init.need_conv = cs.learned ? 1 : 0;
    // the copy's source outlives the call: slot k & 1 is written again only
    // for the solve after next, after the host has waited on this one
    c->init_src[cs.slot] = cs.init;
    TF_HIP(hipMemcpyAsync(dst, &c->init_src[cs.slot], sizeof(CgState), hipMemcpyHostToDevice, c->stream));
    T* r0 = cs.learned ? &v.rs[0]->r : v.rj;
    const int r0_stride = cs.learned ? 2 : 1;
    if (with_rhs) {
        // gfdtd: r0 = (K U - RHS1 SrcFld) + dt Ct F + dt fb; fem: r0 = (E W /
        // dt^2 - 0.5/dt SigMat W2) - Ct (N C U_old) + fb / dt
        const double dt = c->dt;
        const RhsCoef co = c->fem ? RhsCoef{1 / (dt * dt), -0.5 / dt, -1.0, 1 / dt} : RhsCoef{1.0, -1.0, dt, dt};
        if (extrap >= 0 && launch_extrap)
            launch_timed(c, "cg_extrap_kernel", cg_extrap_kernel<T>, blocks_for(v.nf), v.nf, v.x, c->hx1.as<T>(),
                         c->hx2.as<T>(), c->dx.as<T>(), extrap, guard);
        const bool shifted = extrap > 0;
        TF_LAUNCH_TIMED(c, "gf_rhs_init_kernel", 8, gf_rhs_init_kernel, T, v.nb_b, c->Kf.vals<T>(),
                        c->R1f.vals<T>(), c->Ctf.signs(), (cs.dir_terms || c->fem) ? 1 : 0,
                        c->free_edge.as<int32_t>(), c->fem ? c->fw.as<T>() : c->U.as<T>(),
                        c->fem ? c->fw2.as<T>() : c->dir_new.as<T>(), c->Fv.as<T>(), c->fb.as<T>(),
                        c->r0.as<T>(), co, c->pre.as<T>(), r0, r0_stride, v.u, v.part_rr, v.part_gj[0], dst,
                        cs.learned ? 0 : 1, guard, v.A, shifted ? (const T*)c->dx.as<T>() : nullptr,
                        c->rr0part.as<double>(), CgBatch{}, (const double*)nullptr);
        if (shifted)
            launch_timed(c, "cg_set_rr0_kernel", cg_set_rr0_kernel, 1, c->rr0part.as<double>(), v.nb_b, dst, guard);
    } else {
        launch_timed(c, "cg_restart_kernel", cg_restart_kernel<T>, v.nb_b, v.nf, c->pre.as<T>(), c->r0.as<T>(), r0,
                     r0_stride, v.u, v.part_rr, v.part_gj[0], dst, cs.learned ? 0 : 1);
    }
    cs.it = 0;
    cs.tags.clear();
}

template <class T>
void cg_capture(tf_ctx* c, const CgSolve& cs, const CgView<T>& v, int k) {
    if (!cs.snaps) return;
    for (auto& sn : *cs.snaps)
        if (sn.first == k)
            bank_capture_kernel<T><<<blocks_for(c->lp.nw), kBlock, 0, c->stream>>>(c->lp.nw, sn.second,
                                                                                     v.x + c->lp.w0,
                                                                                     c->lp.Cb.as<float>());
}

// Launches passes [cs.it, min(cs.it + n, maxiter + 1)) and then an async copy
// of the state into the pinned slot.
template <class T>
void cg_passes(tf_ctx* c, CgSolve& cs, int n) {
    const CgView<T> v(c, cs.learned);
    CgState* dst = c->cg_state.as<CgState>() + cs.slot;
    auto tag = [&](int it, int first) {
        if (c->profiling) cs.tags.push_back({(int64_t)c->recs.size() - 1, it, first});
    };
    if (cs.it == 0) cg_capture<T>(c, cs, v, 0);
    const int stop = std::min(cs.it + n, cs.init.maxiter + 1);
    for (int& it = cs.it; it < stop; ++it) {
        if (cs.learned) {
            TF_LAUNCH_TIMED(c, "cg_pg_kernel", c->tune[TK_G].unroll, cg_pg_kernel, T, v.nb_g, it, v.G, v.gpk,
                            (const RS<T>*)v.rs[it == 0 ? 0 : (it - 1) & 1], v.rs[it & 1], (const T*)v.p, v.x,
                            v.u, (const double*)v.part_del, v.nb_w, v.part_gl, v.part_rl, dst);
            tag(it, 1);
            TF_LAUNCH_TIMED(c, "cg_wu_kernel", c->tune[TK_WU].unroll, cg_wu_kernel, T, v.nb_w, it, v.A,
                            (const T*)v.u, v.p, v.rs[it & 1], (const double*)v.part_rl, (const double*)v.part_gl,
                            v.nb_g, v.part_del, dst);
            tag(it, 0);
            continue;  // at it == maxiter the stop test of cg_wu ends the solve
        }
        TF_LAUNCH_TIMED(c, "cg_w_kernel", c->tune[TK_W].unroll, cg_w_kernel, T, v.nb_w, it, v.A, (const T*)v.u,
                        v.w, 1, (const double*)v.part_rr, v.nb_b, v.part_del, dst, CgBatch{});
        tag(it, 0);
        if (it == cs.init.maxiter) { ++it; break; }  // the pass at maxiter only runs the stop test
        launch_timed(c, "cg_upd_kernel", cg_upd_kernel<T>, v.nb_b, it, v.nf, (const T*)c->pre.as<T>(), v.x, v.rj,
                     v.u, (const T*)v.w, v.p, v.sv, (const double*)v.part_gj[it & 1], v.nb_b,
                     (const double*)v.part_del, v.nb_w, v.part_rr, v.part_gj[(it + 1) & 1], dst, CgBatch{});
        tag(it, 0);
        cg_capture<T>(c, cs, v, it + 1);
    }
    TF_HIP(hipMemcpyAsync(c->h_state + cs.slot, dst, sizeof(CgState), hipMemcpyDeviceToHost, c->stream));
}

// Waits for the state copied by the last cg_passes (unless the caller has
// already waited for it: synced), launching further batches of passes until
// the solve has stopped; the final state is in h_state[cs.slot].
template <class T>
void cg_finish(tf_ctx* c, CgSolve& cs, bool synced) {
    for (;;) {
        if (!synced) TF_HIP(hipStreamSynchronize(c->stream));
        synced = false;
        const CgState& h = c->h_state[cs.slot];
        if (h.done || cs.it > cs.init.maxiter) break;
        cg_passes<T>(c, cs, 4);
    }
    // Launches past the stop test did no work: timed under their own name,
    // so that per-kernel averages cover working launches only (the learned
    // pass that reaches the stop test still updates x and r in cg_pg).
    const CgState& h = c->h_state[cs.slot];
    if (c->profiling && h.done)
        for (auto& t : cs.tags)
            if (t[1] > h.iter || (t[1] == h.iter && !t[2])) c->recs[t[0]].name = "cg_stop_pass";
    (cs.learned ? c->cg_guess_pc : c->cg_guess) = h.iter;
}

// A whole solve restarted from the step's initial residual r0 (the Jacobi
// redo of a learned solve that missed the tolerance; x must be back at the
// step's warm start).
template <class T>
void cg_solve(tf_ctx* c, CgSolve& cs) {
    cg_start<T>(c, cs, false, nullptr);
    cg_passes<T>(c, cs, std::max(1, (cs.learned ? c->cg_guess_pc : c->cg_guess) + 1));
    cg_finish<T>(c, cs, false);
}

// ------------------------------------------------- learned preconditioner
//
// G is a symmetric sparse approximate inverse of A with A's pattern:
//   G_ej = (base_p + y_p) / sqrt(A_ee A_jj),  p = pair {e, j},
//   base_p = 1 (e == j), -Â_ej (e != j)   (y = 0: G = D^-1/2 (2I - Â) D^-1/2),
// y_p = MLP(features_p), features symmetric in (e, j). The network is trained
// on the device from correction pairs captured during the first Jacobi-PCG
// solves of the run, then frozen: G is a fixed linear operator for CG.

uint16_t bf16_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

constexpr int kCgSlots = 48;                      // bank slots for captured pairs
constexpr int kCaptureIters[4] = {0, 3, 7, 12};   // CG iterations captured per chosen step

// Host tables of the learned preconditioner: pair list, standardised
// features, SELL position -> pair map and Â on A's SELL pattern. Built on
// host threads while the context's first (collection) solves run on the
// device; nothing here touches the GPU.
std::unique_ptr<LearnedTables> learned_tables(const tf_ctx* c, std::vector<double> is, int32_t w0, int32_t w1) {
    auto T = std::make_unique<LearnedTables>();
    const int32_t n = c->n_free;
    const auto& rp = c->hA_ptr;
    const auto& col = c->hA_col;
    const auto& val = c->hA_val;
    const int64_t nnz = (int64_t)col.size();
    std::atomic<int> bad{0};  // 1 repeated column, 2 unsymmetric pattern
    // rows keep their CSR entry order (that of the SELL layout); ord lists
    // each row's positions by increasing column for lookups and merges
    std::vector<int64_t> ord(nnz);
    par_chunks(n, 4096, [&](int64_t lo, int64_t hi) {
        for (int64_t e = lo; e < hi; ++e) {
            for (int64_t q = rp[e]; q < rp[e + 1]; ++q) ord[q] = q;
            std::sort(ord.begin() + rp[e], ord.begin() + rp[e + 1],
                      [&](int64_t a, int64_t b) { return col[a] < col[b]; });
            for (int64_t q = rp[e] + 1; q < rp[e + 1]; ++q)
                if (col[ord[q]] == col[ord[q - 1]]) bad = 1;
        }
    });
    if (bad) throw std::runtime_error("A has a repeated column in a row");
    auto find = [&](int32_t row, int32_t cl) -> int64_t {
        const int64_t* b = ord.data() + rp[row];
        const int64_t* en = ord.data() + rp[row + 1];
        const int64_t* it = std::lower_bound(b, en, cl, [&](int64_t q, int32_t v) { return col[q] < v; });
        if (it == en || col[*it] != cl) { bad = 2; return rp[row]; }
        return *it;
    };
    auto sell_pos = [&](int32_t row, int64_t q) {
        return c->hA_sptr[row / kSlice] + (q - rp[row]) * kSlice + row % kSlice;
    };
    // pairs: upper triangle incl. diagonal, in CSR order (row-major prefix)
    std::vector<int64_t> pfirst((size_t)n + 1, 0);
    par_chunks(n, 65536, [&](int64_t lo, int64_t hi) {
        for (int64_t e = lo; e < hi; ++e) {
            int64_t k = 0;
            for (int64_t q = rp[e]; q < rp[e + 1]; ++q) k += col[q] >= e;
            pfirst[e + 1] = k;
        }
    });
    for (int32_t e = 0; e < n; ++e) pfirst[e + 1] += pfirst[e];
    const int64_t np64 = pfirst[n];
    if (np64 >= INT32_MAX) throw std::runtime_error("A has too many entry pairs for the learned preconditioner");
    const int32_t np = (int32_t)np64;
    std::vector<int32_t> pair_of(nnz, -1), pe(np), pj(np);
    par_chunks(n, 65536, [&](int64_t lo, int64_t hi) {
        for (int64_t e = lo; e < hi; ++e) {
            int64_t k = pfirst[e];
            for (int64_t q = rp[e]; q < rp[e + 1]; ++q)
                if (col[q] >= e) {
                    pair_of[q] = (int32_t)k;
                    pe[k] = (int32_t)e;
                    pj[k] = col[q];
                    ++k;
                }
        }
    });
    par_chunks(n, 65536, [&](int64_t lo, int64_t hi) {
        for (int64_t e = lo; e < hi; ++e)
            for (int64_t q = rp[e]; q < rp[e + 1]; ++q)
                if (col[q] < e) pair_of[q] = pair_of[find(col[q], (int32_t)e)];
    });
    if (bad) throw std::runtime_error("A pattern is not symmetric");
    T->n_pairs = np;
    T->n_pad = (np + 15) / 16 * 16;
    T->pair_w0 = pfirst[w0];
    T->pair_w1 = pfirst[w1];
    // row statistics of Â (off-diagonal part); means over fixed row chunks
    // summed in chunk order (independent of the thread count)
    std::vector<double> rho(n, 0), sig(n, 0), ssum(n, 0), nnzr(n, 0), ld(n, 0);
    constexpr int64_t kRowChunk = 65536;
    const int64_t nrc = (n + kRowChunk - 1) / kRowChunk;
    std::vector<double> cld(nrc, 0), cnz(nrc, 0);
    par_chunks(n, kRowChunk, [&](int64_t lo, int64_t hi) {
        double a_ld = 0, a_nz = 0;
        for (int64_t e = lo; e < hi; ++e) {
            double d = 0;
            for (int64_t q = rp[e]; q < rp[e + 1]; ++q) {
                if (col[q] == e) { d = val[q]; continue; }
                const double a = val[q] * is[e] * is[col[q]];
                rho[e] += std::fabs(a);
                sig[e] += a * a;
                ssum[e] += a;
            }
            nnzr[e] = (double)(rp[e + 1] - rp[e]);
            ld[e] = std::log(d);
            a_ld += ld[e];
            a_nz += nnzr[e];
        }
        cld[lo / kRowChunk] = a_ld;
        cnz[lo / kRowChunk] = a_nz;
    });
    double ldm = 0, nm = 0;
    for (int64_t k = 0; k < nrc; ++k) { ldm += cld[k]; nm += cnz[k]; }
    ldm /= n;
    nm /= n;
    par_chunks(n, kRowChunk, [&](int64_t lo, int64_t hi) {
        for (int64_t e = lo; e < hi; ++e) { ld[e] -= ldm; nnzr[e] /= nm; }
    });
    // features (16 per pair, symmetric in e, j), then standardised
    std::vector<float> F((size_t)np * kMlpIn);
    T->base.resize(np);
    T->scale.resize(np);
    T->pa.resize(np);
    T->pb.resize(np);
    constexpr int64_t kPairChunk = 65536;
    const int64_t npc = (np + kPairChunk - 1) / kPairChunk;
    std::vector<double> fsum((size_t)npc * kMlpIn, 0.0);
    par_chunks(np, kPairChunk, [&](int64_t lo, int64_t hi) {
        double* acc = &fsum[(size_t)(lo / kPairChunk) * kMlpIn];
        for (int64_t p = lo; p < hi; ++p) {
            const int32_t e = pe[p], j = pj[p];
            const bool dg = e == j;
            const int64_t qe = find(e, j);
            const double ah = val[qe] * is[e] * is[j];
            // (Â^2)_ej: merge of rows e and j
            double a2 = 0;
            for (int64_t x = rp[e], y = rp[j]; x < rp[e + 1] && y < rp[j + 1];) {
                const int64_t qx = ord[x], qy = ord[y];
                if (col[qx] < col[qy]) ++x;
                else if (col[qy] < col[qx]) ++y;
                else {
                    a2 += val[qx] * is[e] * is[col[qx]] * val[qy] * is[j] * is[col[qy]];
                    ++x;
                    ++y;
                }
            }
            const double id = dg ? 1.0 : 0.0;
            float* f = &F[(size_t)p * kMlpIn];
            const double mean2[5][2] = {{0.5 * (ld[e] + ld[j]), std::fabs(ld[e] - ld[j])},
                                        {0.5 * (rho[e] + rho[j]), std::fabs(rho[e] - rho[j])},
                                        {0.5 * (sig[e] + sig[j]), std::fabs(sig[e] - sig[j])},
                                        {0.5 * (ssum[e] + ssum[j]), std::fabs(ssum[e] - ssum[j])},
                                        {0.5 * (nnzr[e] + nnzr[j]), std::fabs(nnzr[e] - nnzr[j])}};
            f[0] = (float)id;
            f[1] = (float)(dg ? 0.0 : ah);
            f[2] = (float)a2;
            f[3] = (float)(dg ? 0.0 : std::fabs(ah));
            for (int k = 0; k < 5; ++k) { f[4 + 2 * k] = (float)mean2[k][0]; f[5 + 2 * k] = (float)mean2[k][1]; }
            f[14] = (float)(id * a2);
            f[15] = (float)((1 - id) * a2);
            for (int k = 0; k < kMlpIn; ++k) acc[k] += f[k];
            T->base[p] = (float)(dg ? 1.0 : -ah);
            T->scale[p] = (float)(is[e] * is[j]);
            T->pa[p] = sell_pos(e, qe);
            T->pb[p] = dg ? -1 : sell_pos(j, find(j, e));
        }
    });
    if (bad) throw std::runtime_error("A pattern is not symmetric");
    double mu[kMlpIn] = {}, sd[kMlpIn] = {};
    for (int64_t c2 = 0; c2 < npc; ++c2)
        for (int k = 0; k < kMlpIn; ++k) mu[k] += fsum[(size_t)c2 * kMlpIn + k];
    for (int k = 0; k < kMlpIn; ++k) mu[k] /= np;
    std::fill(fsum.begin(), fsum.end(), 0.0);
    par_chunks(np, kPairChunk, [&](int64_t lo, int64_t hi) {
        double* acc = &fsum[(size_t)(lo / kPairChunk) * kMlpIn];
        for (int64_t p = lo; p < hi; ++p)
            for (int k = 0; k < kMlpIn; ++k) {
                const double t = F[(size_t)p * kMlpIn + k] - mu[k];
                acc[k] += t * t;
            }
    });
    for (int64_t c2 = 0; c2 < npc; ++c2)
        for (int k = 0; k < kMlpIn; ++k) sd[k] += fsum[(size_t)c2 * kMlpIn + k];
    for (int k = 0; k < kMlpIn; ++k) {
        sd[k] = std::sqrt(sd[k] / np);
        if (!(sd[k] > 1e-12)) sd[k] = 1;
    }
    T->X.assign((size_t)T->n_pad * kMlpIn, 0);
    par_chunks(np, kPairChunk, [&](int64_t lo, int64_t hi) {
        for (int64_t p = lo; p < hi; ++p)
            for (int k = 0; k < kMlpIn; ++k)
                T->X[(size_t)p * kMlpIn + k] = bf16_bits((float)((F[(size_t)p * kMlpIn + k] - mu[k]) / sd[k]));
    });
    // SELL position -> pair, and Â on the SELL pattern
    const int64_t entries = c->A.entries;
    T->pos2pair.assign(entries, -1);
    T->ahat.assign(entries, 0.f);
    par_chunks(n, 65536, [&](int64_t lo, int64_t hi) {
        for (int64_t e = lo; e < hi; ++e)
            for (int64_t q = rp[e]; q < rp[e + 1]; ++q) {
                const int64_t o = sell_pos((int32_t)e, q);
                T->pos2pair[o] = pair_of[q];
                T->ahat[o] = (float)(val[q] * is[e] * is[col[q]]);
            }
    });
    return T;
}

// Parameter and bank set-up (once per context). The host tables of
// learned_tables() are built on host threads in the background and
// uploaded by learned_train() once the collection solves are done.
void learned_setup(tf_ctx* c, const tf_cg_opts* cg) {
    LearnedPc& L = c->lp;
    const int32_t n = c->n_free;
    if (n <= 0) throw std::runtime_error("learned preconditioner needs free edges");
    const auto& rp = c->hA_ptr;
    const auto& col = c->hA_col;
    const auto& val = c->hA_val;
    std::vector<double> is(n, 0.0);
    std::atomic<int> bad{0};
    par_chunks(n, 65536, [&](int64_t lo, int64_t hi) {
        for (int64_t e = lo; e < hi; ++e) {
            double d = 0;
            for (int64_t q = rp[e]; q < rp[e + 1]; ++q)
                if (col[q] == e) d = val[q];
            if (!(d > 0)) bad = 1;
            is[e] = 1.0 / std::sqrt(d);
        }
    });
    if (bad) throw std::runtime_error("A has a non-positive diagonal entry");
    // training window: rows [w0, w0 + nw) of the free-row order (whole
    // 64-row slices; the middle of the mesh when it has more rows)
    L.nw = n <= kTrainRows ? n : kTrainRows;
    L.w0 = n <= kTrainRows ? 0 : (n - L.nw) / 2 / kSlice * kSlice;
    std::vector<float> isd(n);
    for (int32_t e = 0; e < n; ++e) isd[e] = (float)is[e];
    upload(L.isd, isd);
    // parameters: He-uniform W1, W2 from a fixed LCG, zero biases, zero output layer
    std::vector<float> P(kMlpParams, 0.f);
    uint64_t st = 0x243F6A8885A308D3ULL;
    auto unif = [&] {
        st = st * 6364136223846793005ULL + 1442695040888963407ULL;
        return (double)(st >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int k = 0; k < kMlpHid * kMlpIn; ++k) P[kW1 + k] = (float)((2 * unif() - 1) * std::sqrt(6.0 / kMlpIn));
    for (int k = 0; k < kMlpHid * kMlpHid; ++k) P[kW2 + k] = (float)((2 * unif() - 1) * std::sqrt(6.0 / kMlpHid));
    upload(L.P, P);
    upload(L.Pbest, P);
    std::vector<float> zeros(kMlpParams, 0.f);
    upload(L.m, zeros);
    upload(L.v, zeros);
    const int nbb = 512;
    L.part.alloc((size_t)nbb * kMlpParams * 4);
    L.lpart.alloc(kSpmvCap * sizeof(double));
    L.cpart.alloc((size_t)kSpmvCap * kBank * sizeof(double));
    const size_t bank = (size_t)L.nw * kBank * 4;
    for (DevBuf* b : {&L.Rb, &L.Cb, &L.T1, &L.T2}) {
        b->alloc(bank);
        TF_HIP(hipMemsetAsync(b->p, 0, bank, c->stream));
    }
    L.slot.assign(kBank, 0);
    L.learn_steps = cg->learn_steps > 0 ? cg->learn_steps : 200;
    L.train_iters = cg->train_iters > 0 ? cg->train_iters : 80;
    L.lr = cg->lr > 0 ? cg->lr : 3e-3;
    L.state = LearnedPc::kCollecting;
    L.steps = 0;
    TF_HIP(hipStreamSynchronize(c->stream));
    const int32_t w0 = L.w0, w1 = L.w0 + L.nw;
    L.tables = std::async(std::launch::async, [c, is = std::move(is), w0, w1]() mutable {
        const auto t0 = std::chrono::steady_clock::now();
        auto T = learned_tables(c, std::move(is), w0, w1);
        if (std::getenv("TETFIT_HOST_TIMING"))
            std::fprintf(stderr, "[learned] host tables built in %.3f s (%d pairs)\n",
                         std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(), T->n_pairs);
        return T;
    });
}

// Waits for the host tables and uploads them (before the training).
void learned_upload_tables(tf_ctx* c) {
    LearnedPc& L = c->lp;
    std::unique_ptr<LearnedTables> T = L.tables.get();
    L.n_pairs = T->n_pairs;
    L.n_pad = T->n_pad;
    L.pair_w0 = T->pair_w0;
    L.pair_w1 = T->pair_w1;
    upload(L.X, T->X);
    upload(L.base, T->base);
    upload(L.scale, T->scale);
    upload(L.pa, T->pa);
    upload(L.pb, T->pb);
    upload(L.pos2pair, T->pos2pair);
    upload(L.ahat, T->ahat);
    T.reset();
    L.y.alloc((size_t)L.n_pad * 4);
    TF_HIP(hipMemsetAsync(L.y.p, 0, L.y.bytes, c->stream));
    L.gy.alloc((size_t)L.n_pad * 4);
    TF_HIP(hipMemsetAsync(L.gy.p, 0, L.gy.bytes, c->stream));
    const int64_t entries = c->A.entries;
    L.ghat.alloc(entries * 4);
    // positions outside the training window keep a zero gradient
    L.gpos.alloc(entries * 4);
    TF_HIP(hipMemsetAsync(L.gpos.p, 0, L.gpos.bytes, c->stream));
    // bf16 values: packed with A's 16-bit column offsets (cg_pg_kernel reads
    // one word per entry) when A has them, else beside A's int32 columns
    if (c->A.c16) L.gpk.alloc(entries * 4);
    else L.gval.alloc(entries * 2);
}

int row_blocks(int32_t n) { return std::max(1, std::min(kSpmvCap, (n + kBlock / kWave - 1) / (kBlock / kWave))); }

// Trains the MLP on the bank (Adam, train_iters steps), keeps the parameters
// of the lowest loss and writes G's values for CG. The loss is over the
// training window's rows: Â and Ĝ restricted to the window (columns outside
// it dropped), each sample's residual r~ = Â c~ formed from its correction,
// so only the pairs of the window's rows run through the MLP while training.
void learned_train(tf_ctx* c) {
    LearnedPc& L = c->lp;
    const auto t0 = std::chrono::steady_clock::now();
    // phase times on stderr with TETFIT_HOST_TIMING set (diagnostics)
    const bool timing = std::getenv("TETFIT_HOST_TIMING") != nullptr;
    auto since = [](std::chrono::steady_clock::time_point a) {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
    };
    L.tables.wait();
    const double t_wait = since(t0);
    learned_upload_tables(c);
    TF_HIP(hipStreamSynchronize(c->stream));
    const double t_upload = since(t0) - t_wait;
    const auto t1 = std::chrono::steady_clock::now();
    const int32_t nw = L.nw, w0 = L.w0;
    hipStream_t s = c->stream;
    const int64_t* sptr = c->A.sptr.as<int64_t>();
    const int32_t* col = c->A.col.as<int32_t>();
    const int rb = row_blocks(nw);
    const int64_t nb_elems = (int64_t)nw * kBank;
    // free slots become random probes c~ = unit-variance noise; every slot's
    // residual is r~ = Â c~ on the window
    uint64_t mask = 0;
    for (int k = 0; k < kBank; ++k)
        if (L.slot[k] == 0) { mask |= 1ULL << k; L.slot[k] = 2; }
    const int eb = (int)std::min<int64_t>((nb_elems + kBlock - 1) / kBlock, INT32_MAX);
    if (mask) bank_random_kernel<<<eb, kBlock, 0, s>>>(nw, mask, 0x5EEDULL, L.Cb.as<float>());
    bank_spmm_kernel<<<rb, kBlock, 0, s>>>(nw, w0, sptr, col, L.ahat.as<float>(), L.Cb.as<float>(), nullptr,
                                           L.T1.as<float>());
    bank_copy_cols_kernel<<<eb, kBlock, 0, s>>>(nw, ~0ULL, L.T1.as<float>(), L.Rb.as<float>());
    // weights: w_s = 1 / (S c~_s^T Â c~_s)
    bank_coldot_kernel<<<rb, kBlock, 0, s>>>(nw, L.Cb.as<float>(), L.T1.as<float>(), L.cpart.as<double>());
    std::vector<double> cp((size_t)rb * kBank);
    TF_HIP(hipMemcpyAsync(cp.data(), L.cpart.p, cp.size() * 8, hipMemcpyDeviceToHost, s));
    TF_HIP(hipStreamSynchronize(s));
    std::vector<float> w(kBank, 0.f);
    for (int k = 0; k < kBank; ++k) {
        double t = 0;
        for (int b = 0; b < rb; ++b) t += cp[(size_t)b * kBank + k];
        w[k] = (t > 0 && std::isfinite(t)) ? (float)(1.0 / (kBank * t)) : 0.f;
    }
    DevBuf wts;
    upload(wts, w);
    // the window's pairs (whole 16-pair tiles) and SELL positions
    const int64_t tile0 = L.pair_w0 / 16, tile1 = (L.pair_w1 + 15) / 16;
    const int64_t wtiles = std::max<int64_t>(tile1 - tile0, 0);
    const int64_t tiles = L.n_pad / 16;
    auto grid_of = [](int64_t t, int64_t cap) { return (int)std::min<int64_t>(cap, std::max<int64_t>(1, (t + 3) / 4)); };
    const int nbb = grid_of(wtiles, 512);
    const int64_t pos0 = c->hA_sptr[w0 / kSlice];
    const int64_t pos1 = c->hA_sptr[(w0 + nw + kSlice - 1) / kSlice];
    const int64_t entries = c->A.entries;
    const int xbw = (int)std::max<int64_t>(1, (pos1 - pos0 + kBlock - 1) / kBlock);
    const int64_t wpairs = L.pair_w1 - L.pair_w0;
    const int pb = blocks_for(std::max<int64_t>(wpairs, 1));
    const short* Xw = L.X.as<short>() + tile0 * 16 * kMlpIn;
    float* yw = L.y.as<float>() + tile0 * 16;
    float* gyw = L.gy.as<float>() + tile0 * 16;
    std::vector<double> lsum(rb);
    double best = INFINITY;
    for (int it = 0; it <= L.train_iters; ++it) {
        if (wtiles) mlp_forward_kernel<<<grid_of(wtiles, 4096), kBlock, 0, s>>>(wtiles, L.P.as<float>(), Xw, yw);
        expand_kernel<<<xbw, kBlock, 0, s>>>(pos1 - pos0, L.pos2pair.as<int32_t>() + pos0, L.base.as<float>(),
                                             L.scale.as<float>(), L.y.as<float>(), L.ghat.as<float>() + pos0,
                                             nullptr, nullptr, nullptr);
        bank_spmm_kernel<<<rb, kBlock, 0, s>>>(nw, w0, sptr, col, L.ghat.as<float>(), L.Rb.as<float>(),
                                               L.Cb.as<float>(), L.T1.as<float>());
        bank_spmm_kernel<<<rb, kBlock, 0, s>>>(nw, w0, sptr, col, L.ahat.as<float>(), L.T1.as<float>(), nullptr,
                                               L.T2.as<float>());
        bank_grad_kernel<<<rb, kBlock, 0, s>>>(nw, w0, sptr, col, L.Rb.as<float>(), L.T1.as<float>(),
                                               L.T2.as<float>(), wts.as<float>(), 2.f, L.gpos.as<float>(),
                                               L.lpart.as<double>());
        TF_HIP(hipMemcpyAsync(lsum.data(), L.lpart.p, rb * 8, hipMemcpyDeviceToHost, s));
        TF_HIP(hipStreamSynchronize(s));
        double loss = 0;
        for (double v : lsum) loss += v;
        if (it == 0) L.loss0 = loss;
        if (std::isfinite(loss) && loss < best) {
            best = loss;
            TF_HIP(hipMemcpyAsync(L.Pbest.p, L.P.p, kMlpParams * 4, hipMemcpyDeviceToDevice, s));
        }
        if (it == L.train_iters) break;
        if (wpairs > 0)
            pair_grad_kernel<<<pb, kBlock, 0, s>>>((int32_t)wpairs, L.pa.as<int64_t>() + L.pair_w0,
                                                   L.pb.as<int64_t>() + L.pair_w0, L.gpos.as<float>(),
                                                   L.gy.as<float>() + L.pair_w0);
        if (wtiles)
            mlp_backward_kernel<<<nbb, kBlock, 0, s>>>(wtiles, L.P.as<float>(), Xw, gyw, L.part.as<float>());
        mlp_adam_kernel<<<blocks_for(kMlpParams), kBlock, 0, s>>>(nbb, L.part.as<float>(), L.P.as<float>(),
                                                                  L.m.as<float>(), L.v.as<float>(), nullptr,
                                                                  (float)L.lr, 0.9f, 0.999f, 1e-8f, it + 1, 1.f);
    }
    L.loss = best;
    const double t_loop = since(t1);
    const auto t2 = std::chrono::steady_clock::now();
    // frozen G from the best parameters, over every pair
    mlp_forward_kernel<<<grid_of(tiles, 4096), kBlock, 0, s>>>(tiles, L.Pbest.as<float>(), L.X.as<short>(),
                                                               L.y.as<float>());
    const int xb = (int)((entries + kBlock - 1) / kBlock);
    expand_kernel<<<xb, kBlock, 0, s>>>(entries, L.pos2pair.as<int32_t>(), L.base.as<float>(), L.scale.as<float>(),
                                        L.y.as<float>(), nullptr, c->A.c16 ? nullptr : L.gval.as<uint16_t>(),
                                        c->A.col16.as<uint16_t>(), c->A.c16 ? L.gpk.as<uint32_t>() : nullptr);
    TF_HIP(hipGetLastError());
    TF_HIP(hipStreamSynchronize(s));
    // the training-only buffers are not needed once G is frozen
    for (DevBuf* b : {&L.Rb, &L.Cb, &L.T1, &L.T2, &L.ghat, &L.gpos, &L.X, &L.ahat, &L.pos2pair})
        b->release();
    L.train_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (timing)
        std::fprintf(stderr,
                     "[learned] tables wait %.3f s, upload %.3f s, training %.3f s (%d iterations, %d window rows), "
                     "G %.3f s\n",
                     t_wait, t_upload, t_loop, L.train_iters, L.nw, since(t2));
    L.state = LearnedPc::kActive;
}

// ---- the gfdtd step loop (D:1459-1565) -----------------------------------
// Step i: B -= dt C U (dga_curl), F = N B (sell_n), probe sample at t - dt,
// sources (gf_inject), rhs and r = b - A x (gf_rhs_init), the CG passes,
// then U = x on free edges and the Dirichlet values (gf_scatter).
//
// The host enqueues step k's scatter and step k+1's kernels up to
// gf_rhs_init right after solve k's first batch of passes, each guarded on
// solve k's state (cg_guard_ok), and only then waits for that batch: the GPU
// has that work queued while the host reads the state and enqueues the next
// passes. When solve k needs more passes, or a learned solve misses the
// tolerance and is redone with Jacobi, the guarded kernels have returned
// without effect and are enqueued again once solve k is complete.

struct GfPre {
    int64_t sample_slot = -1;  // probe sample slot of this step (-1: none)
    bool learned = false;      // preconditioner the step's gf_rhs_init was launched for
    int extrap = -1;           // start of the step's solve (cg_start's extrap)
};

// The learned preconditioner state the next solve will see: active, or
// about to become active when the collection ends with this solve.
bool predict_learned(const tf_ctx* c, bool want_learned) {
    const LearnedPc& L = c->lp;
    if (!want_learned) return false;
    if (L.state == LearnedPc::kActive) return true;
    return L.state == LearnedPc::kCollecting && L.steps + 1 >= L.learn_steps;
}

template <class T>
void run_gfdtd(tf_ctx* c, int64_t first, int64_t n, const tf_cg_opts* cg, tf_stats* st) {
    const int inj_blocks = c->n_inj ? blocks_for(c->n_inj) : 0;
    const size_t ns = std::max<size_t>(c->src.size(), 1);
    const int32_t nf = c->n_free;
    if (inj_blocks && c->poy_cap < n * inj_blocks) {
        c->poy.alloc(n * inj_blocks * sizeof(double));
        c->poy_cap = n * inj_blocks;
    }
    const bool lossy = c->Sig.nrows == c->E;
    const int loss_blocks = lossy ? sell_blocks(c->Sig, kSpmvCap) : 0;
    if (lossy && c->loss_cap < n * loss_blocks) {
        c->loss_part.alloc(n * loss_blocks * sizeof(double));
        c->loss_cap = n * loss_blocks;
    }
    CgState init{};
    init.tol = cg ? cg->tol : 1e-8;
    init.maxiter = cg ? cg->maxiter : 100;
    const bool want_learned = cg && cg->precond == TF_PRECOND_LEARNED;
    if (want_learned && c->lp.state == LearnedPc::kOff) learned_setup(c, cg);
    if (!c->solve_ev) TF_HIP(hipEventCreateWithFlags(&c->solve_ev, hipEventDisableTiming));
    // extrapolated starts: history buffers on first use
    const int guess = cg && !c->fem ? cg->guess : 0;
    if (guess && !c->dx.p) {
        const size_t nb = (size_t)std::max(nf, 1) * sizeof(T);
        for (DevBuf* b : {&c->hx1, &c->hx2, &c->dx}) b->alloc(nb);
        c->rr0part.alloc(kCgPartB * sizeof(double));
        c->hist = 0;
    }
    // the start of step i's solve: the history counts the solves of steps
    // i-1, i-2, ... just before it (a jump in steps or a state write clears
    // it). The learned preconditioner's collection solves start warm (their
    // captured pairs are warm-start data) while the history keeps filling.
    auto start_of = [&](int64_t i, bool learned_solve) {
        if (!guess) return -1;
        if (c->hist_next != i) c->hist = 0;
        const bool collecting = want_learned && !learned_solve && c->lp.state == LearnedPc::kCollecting;
        const int order = collecting ? 0 : std::min(guess, c->hist);
        c->hist = std::min(c->hist + 1, 2);
        c->hist_next = i + 1;
        return order;
    };
    LearnedPc& L = c->lp;
    const int gb_n = tuned_blocks(c, TK_N, c->N, kSpmvCap);
    CgState* states = c->cg_state.as<CgState>();

    // The Dirichlet terms of step k's residual vanish when every Dirichlet
    // value is zero: always for steps after the first of the run (the
    // previous scatter wrote the zero values), and for the first when U is
    // zero there now.
    auto dir_terms = [&](int64_t k) { return !(c->dir_zero_static && (k > 0 || c->ud_zero)); };
    // Kernels of step k up to the start of its solve (guard: the previous solve's state).
    auto pre = [&](int64_t k, GfPre& g, bool fresh, const CgState* guard) {
        const int64_t i = first + k;
        const double t = (double)i * c->dt;
        const double* cs = c->carr.as<double>() + (k + 1) * ns;
        const double* cp = c->carr.as<double>() + k * ns;
        // the step's sources (fb, dir_new, Poynting from U_old) ride along in
        // dga_curl's appended workgroups: nothing reads fb or dir_new before
        // the solve's initial residual, and U stays U_old until the scatter
        const CurlSources<T> inj = c->n_inj ? curl_sources<T>(c, 2, cs, i > 1 ? cp : nullptr,
                                                              c->poy.as<double>() + k * inj_blocks)
                                            : CurlSources<T>{};
        launch_timed(c, "dga_curl_kernel", dga_curl_kernel<T>, blocks_for(c->F) + (c->n_inj ? inj_blocks : 0), c->F,
                     c->c_col.as<int32_t>(), c->c_fixed_signs ? nullptr : c->c_sg.as<int8_t>(), c->U.as<T>(),
                     c->curl_u.as<T>(), c->B.as<T>(), c->dt, (const double*)nullptr, guard, inj);
        // gfdtd: F = N B (D:1468); fem: N C U_old for the curl-curl term (D:1729)
        TF_LAUNCH_TIMED(c, "sell_n_kernel", c->tune[TK_N].unroll, sell_n_kernel, T, gb_n,
                        tuned_view(c, TK_N, c->N.vals<T>()), c->fem ? c->curl_u.as<T>() : c->B.as<T>(), c->Fv.as<T>(), c->dt, 1, nullptr, guard);
        if (fresh) g.sample_slot = allow_print(c, t - c->dt) ? sample_reserve(c, t - c->dt) : -1;
        sample_launch<T>(c, g.sample_slot, guard);
        if (lossy) {
            // U_old.SigMat U_old of the losses (D:1519)
            TF_LAUNCH_TIMED(c, "quad_form_kernel", 8, quad_form_kernel, T, loss_blocks,
                c->Sig.vals<T>(), c->U.as<T>(), c->loss_part.as<double>() + k * loss_blocks, guard);
        }
        if (c->fem)
            launch_timed(c, "fem_w_kernel", fem_w_kernel<T>, blocks_for(c->E), c->E, c->edge_row.as<int32_t>(),
                         (const T*)c->U.as<T>(), (const T*)c->U_older.as<T>(), (const T*)c->dir_new.as<T>(),
                         c->fw.as<T>(), c->fw2.as<T>(), guard);
        CgSolve s0;
        s0.init = init;
        s0.learned = g.learned;
        s0.slot = (int)(k & 1);
        s0.dir_terms = dir_terms(k);
        if (fresh) g.extrap = start_of(i, g.learned);
        cg_start<T>(c, s0, true, guard, g.extrap);
    };
    auto scatter = [&](const CgState* guard) {
        launch_timed(c, "gf_scatter_kernel", gf_scatter_edges_kernel<T>, blocks_for(c->E), c->E,
                     c->edge_row.as<int32_t>(), (const T*)c->x.as<T>(), c->U.as<T>(), (const T*)c->dir_new.as<T>(),
                     guard, c->fem ? c->U_older.as<T>() : nullptr);
    };
    // collection captures of the solve of the current step
    auto collection_snaps = [&]() {
        std::vector<std::pair<int, int>> snaps;
        const int stride = std::max(1, L.learn_steps / (kCgSlots / 4));
        if (L.steps % stride == stride - 1) {
            int k = 0;
            for (int sl = 0; sl < kCgSlots && k < 4; ++sl)
                if (L.slot[sl] == 0) snaps.push_back({kCaptureIters[k++], sl});
            if (k < 4) snaps.clear();
        }
        return snaps;
    };

    // The pass before the loop (D:1434-1453), once, when the run starts at
    // step 1: U on the Dirichlet edges takes the boundary values and the e
    // sources at t = 0 (carrier row 0 is step first - 1 = 0). The fb and
    // Poynting slot it also writes are rewritten by step 1.
    if (first == 1) {
        if (c->n_inj)
            gf_inject_kernel<T><<<inj_blocks, kBlock, 0, c->stream>>>(
                c->n_inj, c->inj_edge.as<int32_t>(), c->inj_kind.as<int8_t>(), c->inj_tb.as<int32_t>(),
                c->inj_te.as<int32_t>(), c->term_src.as<int32_t>(), c->term_w.as<double>(), c->carr.as<double>(),
                nullptr, c->fb.as<T>(), c->dir_new.as<T>(), c->U.as<T>(), c->poy.as<double>(), nullptr, 0);
        // fem keeps U before this pass as U_older (D:1682-1683)
        if (c->n_dir > 0)
            gf_scatter_kernel<T><<<blocks_for(c->n_dir), kBlock, 0, c->stream>>>(
                0, c->free_edge.as<int32_t>(), c->x.as<T>(), c->U.as<T>(), c->n_dir, c->dir_edge.as<int32_t>(),
                c->dir_new.as<T>(), nullptr, c->fem ? c->U_older.as<T>() : nullptr);
        c->ud_zero = c->dir_zero_static;
    }
    GfPre cur;
    cur.learned = predict_learned(c, want_learned) && L.state == LearnedPc::kActive;
    pre(0, cur, true, nullptr);
    for (int64_t k = 0; k < n; ++k) {
        // the solve of step k: the mode it needs now, its start redone if the
        // start already launched was for the other preconditioner
        const bool learned = want_learned && L.state == LearnedPc::kActive;
        CgSolve sol;
        sol.init = init;
        sol.learned = learned;
        sol.slot = (int)(k & 1);
        std::vector<std::pair<int, int>> snaps;
        if (want_learned && L.state == LearnedPc::kCollecting) snaps = collection_snaps();
        sol.snaps = snaps.empty() ? nullptr : &snaps;
        if (learned != cur.learned) {
            cur.learned = learned;
            sol.dir_terms = dir_terms(k);
            cg_start<T>(c, sol, true, nullptr, cur.extrap, false);
        } else {
            sol.init.need_conv = learned ? 1 : 0;
            sol.it = 0;
        }
        cg_passes<T>(c, sol, std::max(1, (learned ? c->cg_guess_pc : c->cg_guess) + 1));
        TF_HIP(hipEventRecord(c->solve_ev, c->stream));
        // ahead of the check: this step's scatter and the next step's start
        const CgState* g = states + sol.slot;
        scatter(g);
        GfPre nxt;
        nxt.learned = predict_learned(c, want_learned);
        if (k + 1 < n) pre(k + 1, nxt, true, g);
        // the check of solve k's first batch
        TF_HIP(hipEventSynchronize(c->solve_ev));
        const CgState h0 = c->h_state[sol.slot];
        const bool ran_ahead = h0.done && (!learned || h0.relres <= init.tol);
        cg_finish<T>(c, sol, true);
        CgState h = c->h_state[sol.slot];
        bool ok = std::isfinite(h.relres) && h.relres <= init.tol;
        int64_t iters = h.iter;
        if (learned) {
            L.solves++;
            L.iters += h.iter;
            L.pc_iters.push_back(h.iter);
            if (!ok) {
                // Jacobi-PCG redo from the same warm start (U still holds x_old on free edges)
                st->cg_iters_total += h.iter;
                gather_free_kernel<T><<<blocks_for(nf), kBlock, 0, c->stream>>>(nf, c->free_edge.as<int32_t>(),
                                                                                c->U.as<T>(), c->x.as<T>());
                CgState* sh = c->h_state + sol.slot;
                CgSolve redo;
                redo.init = init;
                redo.slot = sol.slot;
                cg_solve<T>(c, redo);
                h = *sh;
                ok = std::isfinite(h.relres) && h.relres <= init.tol;
                iters = h.iter;
                if (++L.fallbacks > 3) L.state = LearnedPc::kDisabled;
            }
            // plain CG judges the gain: after 8 learned solves, no fewer
            // iterations than the collection's Jacobi solves turns G off
            if (L.state == LearnedPc::kActive && L.pc_iters.size() == 8 && L.jac_iters.size() >= 8) {
                double a = 0, b = 0;
                for (int q = 0; q < 8; ++q) { a += L.pc_iters[q]; b += L.jac_iters[L.jac_iters.size() - 8 + q]; }
                if (a >= b) L.state = LearnedPc::kDisabled;
            }
        } else if (want_learned && L.state == LearnedPc::kCollecting) {
            L.jac_iters.push_back(h.iter);
            if (ok)
                for (auto& sn : snaps)
                    if (sn.first < h.iter) {
                        bank_finalize_kernel<T><<<blocks_for(L.nw), kBlock, 0, c->stream>>>(
                            L.nw, sn.second, c->x.as<T>() + L.w0, L.isd.as<float>() + L.w0, L.Cb.as<float>());
                        L.slot[sn.second] = 1;
                    }
            if (++L.steps >= L.learn_steps) learned_train(c);
        }
        st->cg_iters_total += iters;
        st->cg_iters_max = std::max<int32_t>(st->cg_iters_max, (int32_t)iters);
        if (std::isfinite(h.relres)) st->cg_max_relres = std::max(st->cg_max_relres, h.relres);
        if (iters >= init.maxiter && !ok) st->cg_not_converged++;
        if (!ran_ahead) {
            // the guarded kernels returned without effect: run them now
            scatter(nullptr);
            if (k + 1 < n) pre(k + 1, nxt, false, nullptr);
        }
        cur = nxt;
        if (c->profiling) ev_flush(c);
    }
    if (n > 0) c->ud_zero = c->dir_zero_static;
    TF_HIP(hipGetLastError());
    std::vector<double> part((size_t)n * std::max(inj_blocks, 1), 0.0);
    if (inj_blocks)
        TF_HIP(hipMemcpyAsync(part.data(), c->poy.p, n * inj_blocks * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
    TF_HIP(hipStreamSynchronize(c->stream));
    for (int64_t k = 0; k < n; ++k) {
        double s = 0;
        for (int b = 0; b < inj_blocks; ++b) s += part[k * inj_blocks + b];
        c->poynting.push_back(s);
    }
    if (lossy) {
        std::vector<double> lp((size_t)n * loss_blocks);
        TF_HIP(hipMemcpy(lp.data(), c->loss_part.p, lp.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int64_t k = 0; k < n; ++k) {
            double s = 0;
            for (int b = 0; b < loss_blocks; ++b) s += lp[k * loss_blocks + b];
            c->loss_q.push_back(s);
        }
    } else {
        c->loss_q.insert(c->loss_q.end(), (size_t)n, 0.0);
    }
    st->steps_done += n;
    st->pc_state = L.state;
    st->pc_fallbacks = L.fallbacks;
    st->pc_learned_solves = L.solves;
    st->pc_learned_iters = L.iters;
    st->pc_train_seconds = L.train_seconds;
    st->pc_loss0 = L.loss0;
    st->pc_loss = L.loss;
    int np = 0;
    for (int v : L.slot) np += v == 1;
    st->pc_pairs = np;
    double jm = 0;
    for (int v : L.jac_iters) jm += v;
    st->pc_jacobi_mean_iters = L.jac_iters.empty() ? 0.0 : jm / L.jac_iters.size();
}

template <class F>
tf_status guarded(F f) {
    try {
        g_err.clear();
        f();
        return TF_OK;
    } catch (const HipError& e) {
        g_err = e.what();
        return (tf_status)e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return TF_ERR_NOMEM;
    } catch (const std::exception& e) {
        g_err = e.what();
        return TF_ERR_ARG;
    }
}

std::unique_ptr<tf_ctx> make_ctx(const tf_problem* p, int32_t prec,
                                 const std::vector<int32_t>* batch_rows = nullptr) {
    if (p->method < 0 || p->method > 3) throw std::runtime_error("unknown method");
    if (prec != TF_PREC_FP64 && prec != TF_PREC_FP32) throw std::runtime_error("unknown precision");
    if (p->n_edges <= 0 || p->n_faces <= 0) throw std::runtime_error("empty problem");
    int ndev = 0;
    TF_HIP(hipGetDeviceCount(&ndev));
    if (ndev < 1) throw HipError("no HIP device", TF_ERR_HIP);
    TF_HIP(hipSetDevice(0));
    auto c = std::make_unique<tf_ctx>();
    c->method = p->method;
    c->prec = prec;
    c->E = p->n_edges;
    c->F = p->n_faces;
    c->dt = p->dt;
    c->tsize = prec == TF_PREC_FP64 ? 8 : 4;
    TF_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->U.alloc(c->E * c->tsize);
    c->Fv.alloc(c->F * c->tsize);
    c->B.alloc(c->F * c->tsize);
    TF_HIP(hipMemsetAsync(c->U.p, 0, c->U.bytes, c->stream));
    TF_HIP(hipMemsetAsync(c->Fv.p, 0, c->Fv.bytes, c->stream));
    TF_HIP(hipMemsetAsync(c->B.p, 0, c->B.bytes, c->stream));
    if (p->method == TF_METHOD_FDTD) {
        if (prec == TF_PREC_FP64) create_fit<double>(c.get(), p);
        else create_fit<float>(c.get(), p);
        std::vector<std::vector<std::pair<int32_t, double>>> none;
        std::vector<int> fld;
        if (prec == TF_PREC_FP64) build_fit_special<double>(c.get(), none, fld);
        else build_fit_special<float>(c.get(), none, fld);
    } else {
        TetRelabel rl;
        const tf_problem* q = p;
        if (relabel_tet(p, rl)) {
            q = &rl.p;
            c->tet_relabel = true;
            c->h_esid = rl.ep;
            c->h_fsid = rl.fp;
        }
        if (prec == TF_PREC_FP64) create_tet<double>(c.get(), q, batch_rows);
        else create_tet<float>(c.get(), q, batch_rows);
        std::vector<std::vector<std::pair<int32_t, double>>> none(c->E);
        std::vector<int> fld(c->E, TF_FIELD_H);
        build_tet_inject(c.get(), none, fld);
    }
    ensure_carr(c.get(), 1, 0);
    TF_HIP(hipStreamSynchronize(c->stream));
    return c;
}

// Block-diagonal concatenation of n dga problems: edge, face and H-row
// indices of instance k are offset by the sizes of instances 0..k-1.
struct MergedDga {
    std::vector<int64_t> c_ptr, ct_ptr, h_ptr, n_ptr;
    std::vector<int32_t> c_col, ct_col, h_col, n_col, h_row, bnd_edge;
    std::vector<int8_t> c_sg, ct_sg, edge_class;
    std::vector<double> h_val, n_val, fixed_value, h_dt, f_dt;
    std::vector<int32_t> eoff, foff;
    tf_problem p{};
};

void append_csr(const tf_csr& m, int32_t col_off, std::vector<int64_t>& ptr, std::vector<int32_t>& col,
                std::vector<int8_t>* sg, std::vector<double>* val) {
    if (ptr.empty()) ptr.push_back(0);
    const int64_t base = ptr.back();
    for (int32_t r = 0; r < m.nrows; ++r) ptr.push_back(base + m.rowptr[r + 1]);
    for (int64_t q = 0; q < m.nnz; ++q) {
        const int64_t c64 = (int64_t)m.col[q] + col_off;
        if (c64 > INT32_MAX) throw std::runtime_error("batch too large for int32 indices");
        col.push_back((int32_t)c64);
    }
    if (sg) {
        if (!m.sgn) throw std::runtime_error("incidence without signs");
        sg->insert(sg->end(), m.sgn, m.sgn + m.nnz);
    }
    if (val) {
        if (!m.val) throw std::runtime_error("operator without values");
        val->insert(val->end(), m.val, m.val + m.nnz);
    }
}

tf_csr csr_view(int32_t nrows, int32_t ncols, const std::vector<int64_t>& ptr, const std::vector<int32_t>& col,
                const std::vector<int8_t>* sg, const std::vector<double>* val) {
    tf_csr o{};
    o.nrows = nrows;
    o.ncols = ncols;
    o.nnz = ptr.empty() ? 0 : ptr.back();
    o.rowptr = ptr.data();
    o.col = col.data();
    o.sgn = sg ? sg->data() : nullptr;
    o.val = val ? val->data() : nullptr;
    return o;
}

void merge_dga(const tf_problem* const* probs, int32_t n, MergedDga& m) {
    m.eoff.assign(1, 0);
    m.foff.assign(1, 0);
    int32_t hoff = 0, boff = 0;
    for (int32_t k = 0; k < n; ++k) {
        const tf_problem* p = probs[k];
        if (!p) throw std::runtime_error("null problem in batch");
        if (p->method != TF_METHOD_DGA) throw std::runtime_error("batches hold method dga problems");
        if (p->dga.frac.enabled) throw std::runtime_error("batches hold lossless dga problems only");
        if (p->n_edges <= 0 || p->n_faces <= 0) throw std::runtime_error("empty problem in batch");
        if (!(p->dt > 0)) throw std::runtime_error("batch instance without a time step");
        const tf_dga_ops& d = p->dga;
        const int32_t E = p->n_edges, F = p->n_faces, eo = m.eoff.back(), fo = m.foff.back();
        if ((int64_t)eo + E > INT32_MAX || (int64_t)fo + F > INT32_MAX) throw std::runtime_error("batch too large");
        if (d.C.nrows != F || d.Ct.nrows != E || d.N.nrows != F) throw std::runtime_error("inconsistent operators");
        append_csr(d.C, eo, m.c_ptr, m.c_col, &m.c_sg, nullptr);
        append_csr(d.Ct, fo, m.ct_ptr, m.ct_col, &m.ct_sg, nullptr);
        append_csr(d.H, eo, m.h_ptr, m.h_col, nullptr, &m.h_val);
        append_csr(d.N, fo, m.n_ptr, m.n_col, nullptr, &m.n_val);
        for (int32_t e = 0; e < E; ++e) {
            m.h_row.push_back(d.h_row[e] >= 0 ? d.h_row[e] + hoff : -1);
            m.bnd_edge.push_back(d.bnd_edge && d.bnd_edge[e] >= 0 ? d.bnd_edge[e] + boff : -1);
            m.edge_class.push_back(d.edge_class[e]);
            m.fixed_value.push_back(d.fixed_value[e]);
        }
        m.h_dt.insert(m.h_dt.end(), (size_t)d.H.nrows, p->dt);
        m.f_dt.insert(m.f_dt.end(), (size_t)F, p->dt);
        hoff += d.H.nrows;
        boff += d.n_bnd;
        m.eoff.push_back(eo + E);
        m.foff.push_back(fo + F);
    }
    const int32_t E = m.eoff.back(), F = m.foff.back();
    tf_problem& p = m.p;
    p.method = TF_METHOD_DGA;
    p.n_edges = E;
    p.n_faces = F;
    p.dt = probs[0]->dt;
    p.duration = probs[0]->duration;
    for (int32_t k = 0; k < n; ++k) p.n_vols += probs[k]->n_vols;
    tf_dga_ops& d = p.dga;
    d.C = csr_view(F, E, m.c_ptr, m.c_col, &m.c_sg, nullptr);
    d.Ct = csr_view(E, F, m.ct_ptr, m.ct_col, &m.ct_sg, nullptr);
    d.H = csr_view(hoff, E, m.h_ptr, m.h_col, nullptr, &m.h_val);
    d.N = csr_view(F, F, m.n_ptr, m.n_col, nullptr, &m.n_val);
    d.h_row = m.h_row.data();
    d.bnd_edge = m.bnd_edge.data();
    d.n_bnd = boff;
    d.edge_class = m.edge_class.data();
    d.fixed_value = m.fixed_value.data();
}

// Rebuilds the merged source and probe tables of a batch from its
// per-instance lists (edge, face and source indices offset per instance).
void batch_sources(tf_ctx* c) {
    auto& b = *c->batch;
    std::vector<std::vector<std::pair<int32_t, double>>> by_edge(c->E);
    std::vector<int> fld(c->E, -1);
    c->src.clear();
    c->src_dt.clear();
    for (int32_t k = 0; k < b.n; ++k) {
        const int32_t so = (int32_t)c->src.size();
        for (const tf_source& s : b.src[k]) {
            c->src.push_back(s);
            c->src_dt.push_back(b.dt[k]);
        }
        for (const tf_source_term& t : b.terms[k]) {
            const int32_t e = t.edge + b.eoff[k];
            by_edge[e].push_back({t.source + so, t.weight});
            if (fld[e] < 0) fld[e] = b.src[k][t.source].field;
        }
    }
    for (auto& f : fld)
        if (f < 0) f = TF_FIELD_E;
    build_tet_inject(c, by_edge, fld);
}

// ------------------------------------------------------------- autotuner
//
// For each SELL kernel of the context: candidate grid sizes (one slice per
// wave, 1/2, 1/4, 1/8 of that, and 256..8192 blocks) x gather batches
// {4, 8, 16} x slice mappings {chunk, band}; each candidate is launched twice to warm up and timed over
// five launches with HIP events; the fastest is kept. Results are cached per
// (kernel, precision, rows, entries) in the process and, when the
// environment names a file in TETFIT_TUNE_CACHE, in that file.
std::map<std::string, KTune> g_tune_cache;
bool g_tune_loaded = false;

std::string tune_key(const tf_ctx* c, int k, const Sell& m) {
    return std::string(kTuneName[k]) + "|" + (c->prec == TF_PREC_FP64 ? "f64" : "f32") + "|" +
           std::to_string(m.nrows) + "|" + std::to_string(m.entries);
}

void tune_cache_load() {
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    const char* path = std::getenv("TETFIT_TUNE_CACHE");
    if (!path) return;
    FILE* f = std::fopen(path, "r");
    if (!f) return;
    char key[256];
    int b = 0, u = 0, band = 0;
    while (std::fscanf(f, "%255s %d %d %d", key, &b, &u, &band) == 4) g_tune_cache[key] = KTune{b, u, band};
    std::fclose(f);
}

void tune_cache_store(const std::string& key, const KTune& t) {
    g_tune_cache[key] = t;
    const char* path = std::getenv("TETFIT_TUNE_CACHE");
    if (!path) return;
    FILE* f = std::fopen(path, "a");
    if (!f) return;
    std::fprintf(f, "%s %d %d %d\n", key.c_str(), t.blocks, t.unroll, t.band);
    std::fclose(f);
}

template <class L>
KTune tune_one(tf_ctx* c, int k, const Sell& m, int cap, L launch) {
    tune_cache_load();
    const std::string key = tune_key(c, k, m);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) return it->second;
    const int need = sell_blocks(m, INT32_MAX);
    std::vector<int> cand;
    // whole multiples of the 256 CUs up to 8 workgroups per CU (kernels whose
    // registers allow fewer resident waves run best at those), then larger
    for (int b : {need, (need + 1) / 2, (need + 3) / 4, (need + 7) / 8, 256, 512, 768, 1024, 1280, 1536, 1792, 2048,
                  4096, 8192})
        if (b >= 1 && b <= std::min(need, cap)) cand.push_back(b);
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    hipEvent_t e0, e1;
    TF_HIP(hipEventCreate(&e0));
    TF_HIP(hipEventCreate(&e1));
    KTune best{sell_blocks(m, cap), 8, 0};
    float best_ms = INFINITY;
    for (int b : cand)
        for (int u : {4, 8, 16})
            for (int band : {0, 1}) {
                const KTune t{b, u, band};
                launch(t);
                launch(t);
                TF_HIP(hipEventRecord(e0, c->stream));
                for (int r = 0; r < 5; ++r) launch(t);
                TF_HIP(hipEventRecord(e1, c->stream));
                TF_HIP(hipEventSynchronize(e1));
                float ms = 0;
                TF_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best_ms) { best_ms = ms; best = t; }
            }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    TF_HIP(hipGetLastError());
    tune_cache_store(key, best);
    return best;
}

// Launches used for tuning leave the simulation state as it was: Ct F only
// rewrites curl_f (recomputed first in every dga step), H and N updates run
// with dt = 0, F = N B (gfdtd) rewrites F with the value it already has,
// and the CG kernels run on a scratch state with the outputs in scratch
// vectors (w, u), which every solve recomputes before use.
template <class T>
void autotune(tf_ctx* c) {
    hipStream_t s = c->stream;
    if (c->method == TF_METHOD_DGA) {
        c->tune[TK_CT] = tune_one(c, TK_CT, c->sCt, kSpmvCap, [&](const KTune& t) {
            TF_LAUNCH_NB(c->wide, t.unroll, sell_ct_kernel, T, t.blocks, s, with_band(c->sCt.signs(), t),
                         c->Fv.as<T>(), c->curl_f.as<T>());
        });
        if (c->frac) return;  // lossy media: only sell_ct of the tuned kernels runs
        c->tune[TK_H] = tune_one(c, TK_H, c->H, kSpmvCap, [&](const KTune& t) {
            TF_LAUNCH_NB(c->wide, t.unroll, sell_h_kernel, T, t.blocks, s, with_band(c->H.vals<T>(), t),
                         c->h_edge.as<int32_t>(), c->curl_f.as<T>(), c->U.as<T>(), 0.0, nullptr);
        });
        c->tune[TK_N] = tune_one(c, TK_N, c->N, kSpmvCap, [&](const KTune& t) {
            TF_LAUNCH_NB(c->wide, t.unroll, sell_n_kernel, T, t.blocks, s, with_band(c->N.vals<T>(), t),
                         c->curl_u.as<T>(), c->Fv.as<T>(), 0.0, 0, nullptr, nullptr);
        });
    } else if (c->method == TF_METHOD_GFDTD || c->method == TF_METHOD_FEM) {
        c->tune[TK_N] = tune_one(c, TK_N, c->N, kSpmvCap, [&](const KTune& t) {
            TF_LAUNCH_NB(c->wide, t.unroll, sell_n_kernel, T, t.blocks, s, with_band(c->N.vals<T>(), t), c->B.as<T>(),
                         c->Fv.as<T>(), c->dt, 1, nullptr, nullptr);
        });
        // Tuning launches run pass 1 (the full work of a pass) on a scratch
        // CG state that never stops (tol 0, no iteration cap, rr0 = 1: the
        // ratio test sees a non-zero residual) with unit partials; that it
        // did not stop is checked after the last launch, since timings of
        // launches past a stop would be of empty passes. Their outputs go to vectors every solve rewrites before
        // use: u (z), w (q, also standing in for x), r, p (p0), s (p1) and
        // the (r, s) pairs (rs0, rs1). A pass kernel's grid sets how many
        // partials the next kernel sums, so each candidate is timed together
        // with the kernel that follows it in a pass, reading that many.
        CgState sc{};
        sc.tol = 0.0;
        sc.maxiter = INT32_MAX;
        sc.rr0 = 1.0;
        sc.gam[0] = sc.gam[1] = sc.alpha[0] = sc.alpha[1] = 1.0;
        DevBuf scratch, ones;
        upload(scratch, std::vector<CgState>{sc});
        upload(ones, std::vector<double>(std::max(kCgPartA, kCgPartC), 1.0));
        double* part = c->cg_part.as<double>();
        const double* one = ones.as<double>();
        CgState* st = scratch.as<CgState>();
        const int32_t nf = c->n_free;
        const int nbb = std::max(1, std::min(blocks_for(nf), kCgPartB));  // cg_upd's grid (CgView)
        // Jacobi pass: cg_w (stop test on cg_upd's partials) + the cg_upd that sums its b partials
        c->tune[TK_W] = tune_one(c, TK_W, c->A, kCgPartA, [&](const KTune& t) {
            TF_LAUNCH_NB(c->wide, t.unroll, cg_w_kernel, T, t.blocks, s, 1, with_band(c->A.vals<T>(), t), c->z.as<T>(),
                         c->q.as<T>(), 1, one, nbb, part, st, CgBatch{});
            cg_upd_kernel<T><<<nbb, kBlock, 0, s>>>(1, nf, c->pre.as<T>(), c->q.as<T>(), c->r.as<T>(), c->z.as<T>(),
                                                    c->q.as<T>(), c->p0.as<T>(), c->p1.as<T>(), one, nbb, one,
                                                    t.blocks, part, part, st, CgBatch{});
        });
        if (c->lp.state == LearnedPc::kActive) {
            const SellView<uint16_t> gv = c->A.view<uint16_t>(c->lp.gval);
            const uint32_t* gpk = c->lp.gpk.p ? c->lp.gpk.as<uint32_t>() : nullptr;
            auto pg = [&](const KTune& t, int nparts) {
                TF_LAUNCH_NB(c->wide, t.unroll, cg_pg_kernel, T, t.blocks, s, 1, with_band(gv, t), gpk,
                             c->rs0.as<RS<T>>(), c->rs1.as<RS<T>>(), c->p0.as<T>(), c->q.as<T>(), c->z.as<T>(), one,
                             nparts, part, part, st);
            };
            auto wu = [&](const KTune& t, int nparts) {
                TF_LAUNCH_NB(c->wide, t.unroll, cg_wu_kernel, T, t.blocks, s, 1, with_band(c->A.vals<T>(), t),
                             c->z.as<T>(), c->p0.as<T>(), c->rs1.as<RS<T>>(), one, one, nparts, part, st);
            };
            // cg_pg (summing cg_wu's partials) + the cg_wu that sums its b partials, then
            // cg_wu (summing the chosen cg_pg's partials) + the cg_pg that sums its b
            KTune W0 = c->tune[TK_WU];
            W0.blocks = tuned_blocks(c, TK_WU, c->A, kCgPartA);
            c->tune[TK_G] = tune_one(c, TK_G, c->A, kCgPartC, [&](const KTune& t) {
                pg(t, W0.blocks);
                wu(W0, t.blocks);
            });
            KTune G = c->tune[TK_G];
            G.blocks = tuned_blocks(c, TK_G, c->A, kCgPartC);
            c->tune[TK_WU] = tune_one(c, TK_WU, c->A, kCgPartA, [&](const KTune& t) {
                wu(t, G.blocks);
                pg(G, t.blocks);
            });
        }
        CgState back{};
        TF_HIP(hipMemcpyAsync(&back, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
        TF_HIP(hipStreamSynchronize(s));
        if (back.done) throw std::runtime_error("autotune: the scratch CG state stopped during tuning");
    }
}

// Block-diagonal concatenation of n lossless gfdtd problems for a batched
// CG (run_gfdtd_batch). Instance k's free edges are padded to a multiple of
// 64 with solver-only edges appended after its own edges: no face touches
// them (empty C columns and Ct rows), their rows of A and E are the diagonal
// 1, so their CG rows stay 0 (no right-hand side) and every 64-row slice of
// the CG belongs to one instance. row0[k] is instance k's first CG row.
struct MergedGf {
    std::vector<int64_t> c_ptr, ct_ptr, n_ptr, e_ptr, a_ptr, r1_ptr;
    std::vector<int32_t> c_col, ct_col, n_col, e_col, a_col, r1_col, bnd_edge, free_edges;
    std::vector<int8_t> c_sg, ct_sg, edge_class;
    std::vector<double> n_val, e_val, a_val, r1_val, fixed_value, f_dt, cg_dt;
    std::vector<int32_t> eoff, foff, ecnt, row0;
    tf_problem p{};
};

void merge_gfdtd(const tf_problem* const* probs, int32_t n, MergedGf& m) {
    m.eoff.assign(1, 0);
    m.foff.assign(1, 0);
    m.row0.assign(1, 0);
    int32_t boff = 0;
    auto empty_rows = [](std::vector<int64_t>& ptr, int32_t k) {
        if (ptr.empty()) ptr.push_back(0);
        for (int32_t r = 0; r < k; ++r) ptr.push_back(ptr.back());
    };
    auto unit_rows = [](std::vector<int64_t>& ptr, std::vector<int32_t>& col, std::vector<double>& val,
                        int32_t first_col, int32_t k) {
        if (ptr.empty()) ptr.push_back(0);
        for (int32_t r = 0; r < k; ++r) {
            col.push_back(first_col + r);
            val.push_back(1.0);
            ptr.push_back((int64_t)col.size());
        }
    };
    for (int32_t k = 0; k < n; ++k) {
        const tf_problem* p = probs[k];
        if (!p) throw std::runtime_error("null problem in batch");
        if (p->method != TF_METHOD_GFDTD) throw std::runtime_error("a batch holds one method (dga or gfdtd)");
        const tf_dga_ops& d = p->dga;
        if (d.Sig.nnz > 0) throw std::runtime_error("gfdtd batches hold lossless problems only");
        if (!(p->dt > 0)) throw std::runtime_error("batch instance without a time step");
        const int32_t E = p->n_edges, F = p->n_faces, nf = d.n_free;
        if (d.C.nrows != F || d.Ct.nrows != E || d.N.nrows != F || d.A.nrows != nf || d.RHS1.nrows != E ||
            d.E.nrows != E || !d.free_edges)
            throw std::runtime_error("inconsistent operators");
        const int32_t pad = (kSlice - nf % kSlice) % kSlice;
        const int64_t eo = m.eoff.back(), fo = m.foff.back(), ro = m.row0.back();
        if (eo + E + pad > INT32_MAX || fo + F > INT32_MAX) throw std::runtime_error("batch too large");
        append_csr(d.C, (int32_t)eo, m.c_ptr, m.c_col, &m.c_sg, nullptr);
        append_csr(d.Ct, (int32_t)fo, m.ct_ptr, m.ct_col, &m.ct_sg, nullptr);
        empty_rows(m.ct_ptr, pad);
        append_csr(d.N, (int32_t)fo, m.n_ptr, m.n_col, nullptr, &m.n_val);
        append_csr(d.E, (int32_t)eo, m.e_ptr, m.e_col, nullptr, &m.e_val);
        unit_rows(m.e_ptr, m.e_col, m.e_val, (int32_t)(eo + E), pad);
        append_csr(d.RHS1, (int32_t)eo, m.r1_ptr, m.r1_col, nullptr, &m.r1_val);
        empty_rows(m.r1_ptr, pad);
        append_csr(d.A, (int32_t)ro, m.a_ptr, m.a_col, nullptr, &m.a_val);
        unit_rows(m.a_ptr, m.a_col, m.a_val, (int32_t)(ro + nf), pad);
        for (int32_t i = 0; i < nf; ++i) m.free_edges.push_back((int32_t)(eo + d.free_edges[i]));
        for (int32_t j = 0; j < pad; ++j) m.free_edges.push_back((int32_t)(eo + E + j));
        for (int32_t e = 0; e < E + pad; ++e) {
            const bool own = e < E;
            m.bnd_edge.push_back(own && d.bnd_edge && d.bnd_edge[e] >= 0 ? d.bnd_edge[e] + boff : -1);
            m.edge_class.push_back(own ? d.edge_class[e] : (int8_t)1);
            m.fixed_value.push_back(own ? d.fixed_value[e] : 0.0);
        }
        m.f_dt.insert(m.f_dt.end(), (size_t)F, p->dt);
        m.cg_dt.insert(m.cg_dt.end(), (size_t)(nf + pad), p->dt);
        boff += d.n_bnd;
        m.ecnt.push_back(E);
        m.eoff.push_back((int32_t)(eo + E + pad));
        m.foff.push_back((int32_t)(fo + F));
        m.row0.push_back((int32_t)(ro + nf + pad));
    }
    const int32_t E = m.eoff.back(), F = m.foff.back(), NF = m.row0.back();
    tf_problem& p = m.p;
    p.method = TF_METHOD_GFDTD;
    p.n_edges = E;
    p.n_faces = F;
    p.dt = probs[0]->dt;
    p.duration = probs[0]->duration;
    for (int32_t k = 0; k < n; ++k) p.n_vols += probs[k]->n_vols;
    tf_dga_ops& d = p.dga;
    d.C = csr_view(F, E, m.c_ptr, m.c_col, &m.c_sg, nullptr);
    d.Ct = csr_view(E, F, m.ct_ptr, m.ct_col, &m.ct_sg, nullptr);
    d.N = csr_view(F, F, m.n_ptr, m.n_col, nullptr, &m.n_val);
    d.E = csr_view(E, E, m.e_ptr, m.e_col, nullptr, &m.e_val);
    d.A = csr_view(NF, NF, m.a_ptr, m.a_col, nullptr, &m.a_val);
    d.RHS1 = csr_view(E, E, m.r1_ptr, m.r1_col, nullptr, &m.r1_val);
    d.free_edges = m.free_edges.data();
    d.n_free = NF;
    d.bnd_edge = m.bnd_edge.data();
    d.n_bnd = boff;
    d.edge_class = m.edge_class.data();
    d.fixed_value = m.fixed_value.data();
}

// CG workgroup table of a gfdtd batch: each instance's slices in chunks of
// kBatchChunk (whole instances of up to 4 waves' worth of slices per
// workgroup), the workgroups of one instance consecutive.
constexpr int kBatchChunk = 8;
void build_batch_table(tf_ctx* c, const std::vector<int32_t>& row0) {
    std::vector<int32_t> inst, s0, s1, w0, w1;
    const int32_t n = (int32_t)row0.size() - 1;
    for (int32_t k = 0; k < n; ++k) {
        const int32_t a = row0[k] / kSlice, b = row0[k + 1] / kSlice;
        w0.push_back((int32_t)inst.size());
        for (int32_t q = a; q < b; q += kBatchChunk) {
            inst.push_back(k);
            s0.push_back(q);
            s1.push_back(std::min(b, q + kBatchChunk));
        }
        if ((int32_t)inst.size() - w0.back() > kMaxPart) throw std::runtime_error("batch instance too large");
        if (a == b) throw std::runtime_error("batch instance without free edges");
        w1.push_back((int32_t)inst.size());
    }
    c->bt_chunks = (int32_t)inst.size();
    upload(c->bt_inst, inst);
    upload(c->bt_s0, s0);
    upload(c->bt_s1, s1);
    upload(c->bt_w0, w0);
    upload(c->bt_w1, w1);
    // partials per workgroup: w.u, r.r and r.u (two), fp64 (at least the
    // single-solve layout, which the autotuner's trial launches use)
    c->cg_part.alloc(sizeof(double) *
                     std::max<size_t>(4 * (size_t)c->bt_chunks, kCgPartA + 3 * kCgPartB + 2 * kCgPartC));
    // two state slots per instance (slot 0 used), pinned host copies
    c->cg_state.alloc(2 * sizeof(CgState) * (size_t)n);
    TF_HIP(hipMemsetAsync(c->cg_state.p, 0, c->cg_state.bytes, c->stream));
    if (c->h_state) TF_HIP(hipHostFree(c->h_state));
    c->h_state = nullptr;
    TF_HIP(hipHostMalloc((void**)&c->h_state, 2 * sizeof(CgState) * (size_t)n, hipHostMallocDefault));
    std::memset((void*)c->h_state, 0, 2 * sizeof(CgState) * (size_t)n);
}

// n steps of a gfdtd batch (D:1459-1565 for every instance): the loop of
// run_gfdtd with plain Jacobi CG, one solve per instance per step (each on
// its own scalars and stop test, CgBatch), the host waiting on all of them
// before the next step.
template <class T>
void run_gfdtd_batch(tf_ctx* c, int64_t first, int64_t n, const tf_cg_opts* cg, tf_stats* st) {
    if (cg && cg->precond != TF_PRECOND_JACOBI) throw std::runtime_error("gfdtd batches solve with Jacobi CG only");
    if (cg && cg->guess) throw std::runtime_error("gfdtd batches start each solve from U_n (guess 0)");
    auto& b = *c->batch;
    const int32_t ni = b.n;
    const int inj_blocks = c->n_inj ? blocks_for(c->n_inj) : 0;
    const size_t ns = std::max<size_t>(c->src.size(), 1);
    const int32_t nf = c->n_free;
    if (inj_blocks && c->poy_cap < n * inj_blocks) {
        c->poy.alloc(n * inj_blocks * sizeof(double));
        c->poy_cap = n * inj_blocks;
    }
    CgState init{};
    init.tol = cg ? cg->tol : 1e-8;
    init.maxiter = cg ? cg->maxiter : 100;
    std::vector<CgState> inits((size_t)2 * ni, init);
    const CgBatch bt{c->bt_inst.as<int32_t>(), c->bt_s0.as<int32_t>(), c->bt_s1.as<int32_t>(),
                     c->bt_w0.as<int32_t>(), c->bt_w1.as<int32_t>()};
    const int nb = c->bt_chunks;
    double* part_del = c->cg_part.as<double>();
    double* part_rr = part_del + nb;
    double* part_gj[2] = {part_rr + nb, part_rr + 2 * nb};
    CgState* states = c->cg_state.as<CgState>();
    const int gb_n = tuned_blocks(c, TK_N, c->N, kSpmvCap);
    const double* fdt = c->row_dt_f.as<double>();
    const SellView<T> A = c->A.vals<T>();
    auto dir_terms = [&](int64_t k) { return !(c->dir_zero_static && (k > 0 || c->ud_zero)); };
    // the pass before the loop (D:1434-1453), when the run starts at step 1
    if (first == 1) {
        if (c->n_inj)
            gf_inject_kernel<T><<<inj_blocks, kBlock, 0, c->stream>>>(
                c->n_inj, c->inj_edge.as<int32_t>(), c->inj_kind.as<int8_t>(), c->inj_tb.as<int32_t>(),
                c->inj_te.as<int32_t>(), c->term_src.as<int32_t>(), c->term_w.as<double>(), c->carr.as<double>(),
                nullptr, c->fb.as<T>(), c->dir_new.as<T>(), c->U.as<T>(), c->poy.as<double>(), nullptr, 0);
        if (c->n_dir > 0)
            gf_scatter_kernel<T><<<blocks_for(c->n_dir), kBlock, 0, c->stream>>>(
                0, c->free_edge.as<int32_t>(), c->x.as<T>(), c->U.as<T>(), c->n_dir, c->dir_edge.as<int32_t>(),
                c->dir_new.as<T>(), nullptr, nullptr);
        c->ud_zero = c->dir_zero_static;
    }
    for (int64_t k = 0; k < n; ++k) {
        const int64_t i = first + k;
        const double t = (double)i * c->dt;
        const double* cs = c->carr.as<double>() + (k + 1) * ns;
        const double* cp = c->carr.as<double>() + k * ns;
        const CurlSources<T> inj = c->n_inj ? curl_sources<T>(c, 2, cs, i > 1 ? cp : nullptr,
                                                              c->poy.as<double>() + k * inj_blocks)
                                            : CurlSources<T>{};
        launch_timed(c, "dga_curl_kernel", dga_curl_kernel<T>, blocks_for(c->F) + (c->n_inj ? inj_blocks : 0), c->F,
                     c->c_col.as<int32_t>(), c->c_fixed_signs ? nullptr : c->c_sg.as<int8_t>(), c->U.as<T>(),
                     c->curl_u.as<T>(), c->B.as<T>(), c->dt, fdt, (const CgState*)nullptr, inj);
        TF_LAUNCH_TIMED(c, "sell_n_kernel", c->tune[TK_N].unroll, sell_n_kernel, T, gb_n,
                        tuned_view(c, TK_N, c->N.vals<T>()), c->B.as<T>(), c->Fv.as<T>(), c->dt, 1, nullptr,
                        nullptr);
        // the state before the solve, at t - dt (as run_gfdtd samples it)
        if (want_sample(c, i - 1, t - c->dt)) sample<T>(c, t - c->dt);
        TF_HIP(hipMemcpyAsync(states, inits.data(), inits.size() * sizeof(CgState), hipMemcpyHostToDevice,
                              c->stream));
        // x = U_old on the free rows is the warm start (x keeps the last solve)
        TF_LAUNCH_TIMED(c, "gf_rhs_init_kernel", 8, gf_rhs_init_kernel, T, nb, c->Kf.vals<T>(), c->R1f.vals<T>(),
                        c->Ctf.signs(), dir_terms(k) ? 1 : 0, c->free_edge.as<int32_t>(), c->U.as<T>(),
                        c->dir_new.as<T>(), c->Fv.as<T>(), c->fb.as<T>(), c->r0.as<T>(), RhsCoef{1.0, -1.0, 0.0, 0.0},
                        c->pre.as<T>(), c->r.as<T>(), 1, c->z.as<T>(), part_rr, part_gj[0], states, 1,
                        (const CgState*)nullptr, A, (const T*)nullptr, (double*)nullptr, bt,
                        c->row_dt_cg.as<double>());
        int it = 0;
        auto passes = [&](int m) {
            const int stop = std::min(it + m, init.maxiter + 1);
            for (; it < stop; ++it) {
                TF_LAUNCH_TIMED(c, "cg_w_kernel", c->tune[TK_W].unroll, cg_w_kernel, T, nb, it, A,
                                (const T*)c->z.as<T>(), c->q.as<T>(), 1, (const double*)part_rr, nb, part_del, states,
                                bt);
                if (it == init.maxiter) { ++it; break; }
                launch_timed(c, "cg_upd_kernel", cg_upd_kernel<T>, nb, it, nf, (const T*)c->pre.as<T>(),
                             c->x.as<T>(), c->r.as<T>(), c->z.as<T>(), (const T*)c->q.as<T>(), c->p0.as<T>(),
                             c->p1.as<T>(), (const double*)part_gj[it & 1], nb, (const double*)part_del, nb, part_rr,
                             part_gj[(it + 1) & 1], states, bt);
            }
            TF_HIP(hipMemcpyAsync(c->h_state, states, 2 * sizeof(CgState) * (size_t)ni, hipMemcpyDeviceToHost,
                                  c->stream));
            TF_HIP(hipStreamSynchronize(c->stream));
        };
        passes(std::max(1, c->cg_guess + 1));
        for (;;) {
            bool all = true;
            for (int32_t q = 0; q < ni && all; ++q) all = c->h_state[2 * q].done != 0;
            if (all || it > init.maxiter) break;
            passes(4);
        }
        int32_t it_max = 0;
        for (int32_t q = 0; q < ni; ++q) {
            const CgState& h = c->h_state[2 * q];
            const int64_t iters = h.done ? h.iter : init.maxiter;
            it_max = std::max<int32_t>(it_max, (int32_t)iters);
            st->cg_iters_total += iters;
            st->cg_iters_max = std::max<int32_t>(st->cg_iters_max, (int32_t)iters);
            if (std::isfinite(h.relres)) st->cg_max_relres = std::max(st->cg_max_relres, h.relres);
            if (!h.done || !(h.relres <= init.tol)) st->cg_not_converged++;
        }
        c->cg_guess = it_max;
        launch_timed(c, "gf_scatter_kernel", gf_scatter_edges_kernel<T>, blocks_for(c->E), c->E,
                     c->edge_row.as<int32_t>(), (const T*)c->x.as<T>(), c->U.as<T>(), (const T*)c->dir_new.as<T>(),
                     (const CgState*)nullptr, (T*)nullptr);
        if (c->profiling) ev_flush(c);
    }
    if (n > 0) c->ud_zero = c->dir_zero_static;
    TF_HIP(hipGetLastError());
    std::vector<double> part((size_t)n * std::max(inj_blocks, 1), 0.0);
    if (inj_blocks)
        TF_HIP(hipMemcpyAsync(part.data(), c->poy.p, n * inj_blocks * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
    TF_HIP(hipStreamSynchronize(c->stream));
    for (int64_t k = 0; k < n; ++k) {
        double s = 0;
        for (int q = 0; q < inj_blocks; ++q) s += part[k * inj_blocks + q];
        c->poynting.push_back(s);
    }
    st->steps_done += n;
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

const char* tf_last_error(void) { return g_err.c_str(); }

int32_t tf_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

tf_status tf_create(const tf_problem* p, int32_t prec, tf_ctx** out) {
    if (!p || !out) { g_err = "null argument"; return TF_ERR_ARG; }
    *out = nullptr;
    return guarded([&] { *out = make_ctx(p, prec).release(); });
}

tf_status tf_create_batch(const tf_problem* const* probs, int32_t n, int32_t prec, tf_ctx** out) {
    if (!probs || !out || n < 1) { g_err = "bad argument"; return TF_ERR_ARG; }
    *out = nullptr;
    return guarded([&] {
        if (!probs[0]) throw std::runtime_error("null problem in batch");
        if (probs[0]->method == TF_METHOD_GFDTD) {
            // CG-solved instances: block-diagonal system, one Jacobi CG per
            // instance per step (run_gfdtd_batch)
            MergedGf m;
            merge_gfdtd(probs, n, m);
            auto c = make_ctx(&m.p, prec, &m.row0);
            build_batch_table(c.get(), m.row0);
            // per CG row dt: the CG rows permute only inside each instance's
            // block (batch_row_order), so the merged order's values hold
            if ((size_t)c->n_free != m.cg_dt.size()) throw std::runtime_error("batch CG rows");
            upload(c->row_dt_cg, m.cg_dt);
            upload(c->row_dt_f, m.f_dt);
            auto bt = std::make_unique<tf_ctx::Batch>();
            bt->n = n;
            bt->gfdtd = true;
            bt->eoff = m.eoff;
            bt->foff = m.foff;
            bt->ecnt = m.ecnt;
            bt->poff.assign(n + 1, 0);
            bt->src.resize(n);
            bt->terms.resize(n);
            bt->probes.resize(n);
            for (int32_t k = 0; k < n; ++k) bt->dt.push_back(probs[k]->dt);
            c->batch = std::move(bt);
            *out = c.release();
            return;
        }
        MergedDga m;
        merge_dga(probs, n, m);
        auto c = make_ctx(&m.p, prec);
        // the instances' operators are block-diagonal: banded slice order
        // (measured 2-3% faster on 256 instances of 8^3-32^3)
        for (Sell* s : {&c->sCt, &c->H, &c->N}) s->band = true;
        auto bt = std::make_unique<tf_ctx::Batch>();
        bt->n = n;
        bt->eoff = m.eoff;
        bt->foff = m.foff;
        for (int32_t k = 0; k < n; ++k) bt->ecnt.push_back(m.eoff[k + 1] - m.eoff[k]);
        bt->poff.assign(n + 1, 0);
        bt->src.resize(n);
        bt->terms.resize(n);
        bt->probes.resize(n);
        for (int32_t k = 0; k < n; ++k) bt->dt.push_back(probs[k]->dt);
        std::vector<double> hdt(m.h_dt.size());
        for (size_t r = 0; r < hdt.size(); ++r) hdt[r] = m.h_dt[c->h_perm[r]];  // H rows are in SELL order
        upload(c->row_dt_h, hdt);
        upload(c->row_dt_f, m.f_dt);
        c->batch = std::move(bt);
        *out = c.release();
    });
}

tf_status tf_set_sources(tf_ctx* c, const tf_source* src, int32_t n_src, const tf_source_term* terms,
                         int64_t n_terms) {
    if (!c || (n_src && !src) || (n_terms && !terms)) { g_err = "null argument"; return TF_ERR_ARG; }
    return guarded([&] {
        drop_graph(c);
        c->src.assign(src, src + n_src);
        std::vector<std::vector<std::pair<int32_t, double>>> by_edge(c->E);
        std::vector<int> fld(c->E, -1);
        for (int64_t i = 0; i < n_terms; ++i) {
            const tf_source_term& t = terms[i];
            if (t.edge < 0 || t.edge >= c->E || t.source < 0 || t.source >= n_src)
                throw std::runtime_error("source term out of range");
            // FIT maps to grid positions in build_fit_special; tets here
            const int32_t e = c->method == TF_METHOD_FDTD ? t.edge : dev_edge(c, t.edge);
            by_edge[e].push_back({t.source, t.weight});
            if (fld[e] < 0) fld[e] = src[t.source].field;
        }
        for (auto& f : fld)
            if (f < 0) f = TF_FIELD_E;
        if (c->method == TF_METHOD_FDTD) {
            for (int32_t e : c->fit_uncommon) fld[e] = TF_FIELD_H;
            if (c->prec == TF_PREC_FP64) build_fit_special<double>(c, by_edge, fld);
            else build_fit_special<float>(c, by_edge, fld);
        } else {
            build_tet_inject(c, by_edge, fld);
        }
    });
}

tf_status tf_set_probes(tf_ctx* c, const tf_probe* probes, int32_t n) {
    if (!c || (n && !probes)) { g_err = "null argument"; return TF_ERR_ARG; }
    return guarded([&] {
        std::vector<ProbeDev> v(n);
        for (int32_t i = 0; i < n; ++i) {
            const tf_probe& p = probes[i];
            if (p.ne < 0 || p.ne > 12 || p.nh < 0 || p.nh > 6) throw std::runtime_error("probe term count");
            ProbeDev& d = v[i];
            std::memset(&d, 0, sizeof(d));
            d.ne = p.ne;
            d.nh = p.nh;
            for (int k = 0; k < p.ne; ++k) {
                if (p.e_idx[k] < 0 || p.e_idx[k] >= c->E) throw std::runtime_error("probe edge out of range");
                d.e_idx[k] = dev_edge(c, p.e_idx[k]);
                for (int q = 0; q < 3; ++q) d.e_w[k][q] = p.e_w[k][q];
            }
            for (int k = 0; k < p.nh; ++k) {
                if (p.h_idx[k] < 0 || p.h_idx[k] >= c->F) throw std::runtime_error("probe face out of range");
                d.h_idx[k] = dev_face(c, p.h_idx[k]);
                for (int q = 0; q < 3; ++q) d.h_w[k][q] = p.h_w[k][q];
            }
        }
        c->n_probes = n;
        upload(c->probes, v);
        c->n_samples = 0;
        c->sample_times.clear();
        c->sample_cap = 0;
        c->samples.release();
    });
}

tf_status tf_set_output(tf_ctx* c, double delay, double period) {
    if (!c) { g_err = "null argument"; return TF_ERR_ARG; }
    c->out_enabled = true;
    c->next_print = delay;  // Output::Initialize (Output.hpp:54-57)
    c->period = period;
    return TF_OK;
}

tf_status tf_run(tf_ctx* c, int64_t first, int64_t n, const tf_cg_opts* cg, tf_stats* st) {
    if (!c || n < 0) { g_err = "bad argument"; return TF_ERR_ARG; }
    tf_stats local{};
    if (!st) st = &local;
    *st = tf_stats{};
    return guarded([&] {
        if (n == 0) return;
        if (cg && cg->precond != TF_PRECOND_JACOBI && cg->precond != TF_PRECOND_LEARNED)
            throw std::runtime_error("unknown preconditioner");
        if (cg && (cg->maxiter < 0 || !(cg->tol >= 0) || cg->guess < 0 || cg->guess > 2))
            throw std::runtime_error("bad CG options");
        ensure_carr(c, first, n);
        c->ktimes.clear();
        hipEvent_t e0, e1;
        TF_HIP(hipEventCreate(&e0));
        TF_HIP(hipEventCreate(&e1));
        TF_HIP(hipEventRecord(e0, c->stream));
        const bool f64 = c->prec == TF_PREC_FP64;
        if (c->method == TF_METHOD_FDTD) {
            if (f64) run_fit<double>(c, first, n, st); else run_fit<float>(c, first, n, st);
        } else if (c->method == TF_METHOD_DGA) {
            if (f64) run_dga<double>(c, first, n, st); else run_dga<float>(c, first, n, st);
        } else if (c->batch && c->batch->gfdtd) {
            if (f64) run_gfdtd_batch<double>(c, first, n, cg, st); else run_gfdtd_batch<float>(c, first, n, cg, st);
        } else {
            if (f64) run_gfdtd<double>(c, first, n, cg, st); else run_gfdtd<float>(c, first, n, cg, st);
        }
        TF_HIP(hipEventRecord(e1, c->stream));
        TF_HIP(hipEventSynchronize(e1));
        float ms = 0;
        TF_HIP(hipEventElapsedTime(&ms, e0, e1));
        st->seconds = ms * 1e-3;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        ev_flush(c);
    });
}

tf_status tf_read_state(tf_ctx* c, double* U, double* F, double* B) {
    if (!c) { g_err = "null argument"; return TF_ERR_ARG; }
    return guarded([&] {
        TF_HIP(hipStreamSynchronize(c->stream));
        auto rd = [&](const DevBuf& b, double* dst, int64_t n) {
            if (!dst) return;
            if (c->prec == TF_PREC_FP64) {
                TF_HIP(hipMemcpy(dst, b.p, n * 8, hipMemcpyDeviceToHost));
            } else {
                std::vector<float> tmp(n);
                TF_HIP(hipMemcpy(tmp.data(), b.p, n * 4, hipMemcpyDeviceToHost));
                for (int64_t i = 0; i < n; ++i) dst[i] = tmp[i];
            }
        };
        // FIT grid order or tet locality order: entity i sits at device position pos[i]
        auto rd_perm = [&](const DevBuf& b, double* dst, int64_t n, const std::vector<int32_t>& pos) {
            if (!dst) return;
            std::vector<double> tmp(n);
            rd(b, tmp.data(), n);
            for (int64_t i = 0; i < n; ++i) dst[i] = tmp[pos[i]];
        };
        if (c->fit_grid || c->tet_relabel) {
            rd_perm(c->U, U, c->E, c->h_esid);
            if (!c->fem) rd_perm(c->Fv, F, c->F, c->h_fsid);
            rd_perm(c->B, B, c->F, c->h_fsid);
        } else {
            rd(c->U, U, c->E);
            if (!c->fem) rd(c->Fv, F, c->F);
            rd(c->B, B, c->F);
        }
        // fem has no F field (its buffer holds N C U_old of the last step)
        if (c->fem && F) std::fill(F, F + c->F, 0.0);
    });
}

tf_status tf_write_state(tf_ctx* c, const double* U, const double* F, const double* B) {
    if (!c) { g_err = "null argument"; return TF_ERR_ARG; }
    return guarded([&] {
        TF_HIP(hipStreamSynchronize(c->stream));
        c->hist = 0;  // extrapolated starts begin again from the written state
        c->hist_next = -1;
        auto wr = [&](DevBuf& b, const double* src, int64_t n) {
            if (!src) return;
            if (c->prec == TF_PREC_FP64) {
                TF_HIP(hipMemcpy(b.p, src, n * 8, hipMemcpyHostToDevice));
            } else {
                std::vector<float> tmp(src, src + n);
                TF_HIP(hipMemcpy(b.p, tmp.data(), n * 4, hipMemcpyHostToDevice));
            }
        };
        if (c->fit_grid || c->tet_relabel) {
            auto wr_perm = [&](DevBuf& b, const double* src, int64_t n, const std::vector<int32_t>& pos) {
                if (!src) return;
                std::vector<double> tmp(n);
                for (int64_t i = 0; i < n; ++i) tmp[pos[i]] = src[i];
                wr(b, tmp.data(), n);
            };
            wr_perm(c->U, U, c->E, c->h_esid);
            wr_perm(c->Fv, F, c->F, c->h_fsid);
            wr_perm(c->B, B, c->F, c->h_fsid);
        } else {
            wr(c->U, U, c->E);
            wr(c->Fv, F, c->F);
            wr(c->B, B, c->F);
        }
        if (U) {
            c->ud_zero = true;
            for (int32_t e : c->h_dir_edge)
                if (U[e] != 0) c->ud_zero = false;
        }
        // lossy dga: each edge's (face's) value split evenly over its
        // fractional entries, the entries no edge (face) reads set to zero
        if (c->frac) {
            auto split = [&](DevBuf& dst, int32_t nx, const double* src, int32_t n, const std::vector<int32_t>& a,
                             const std::vector<int32_t>& b) {
                if (!src) return;
                std::vector<double> v(std::max(nx, 1), 0.0);
                for (int32_t i = 0; i < n; ++i) {
                    const double w = b[i] >= 0 ? 0.5 * src[i] : src[i];
                    v[a[i]] = w;
                    if (b[i] >= 0) v[b[i]] = w;
                }
                wr(dst, v.data(), (int64_t)v.size());
            };
            split(c->uf[c->fcur], c->nU, U, c->E, c->h_mu_a, c->h_mu_b);
            split(c->ff[c->fcur], c->nF, F, c->F, c->h_tf_a, c->h_tf_b);
        }
        // gfdtd and fem: the next solve's warm start is the written field on
        // the free edges, as after every step (the solution vector is U's
        // free part); fem also takes it as the step before (U_older)
        if (U && c->fem) TF_HIP(hipMemcpy(c->U_older.p, c->U.p, c->U.bytes, hipMemcpyDeviceToDevice));
        if (U && (c->method == TF_METHOD_GFDTD || c->method == TF_METHOD_FEM) && c->n_free > 0) {
            if (c->prec == TF_PREC_FP64)
                gather_free_kernel<double><<<blocks_for(c->n_free), kBlock, 0, c->stream>>>(
                    c->n_free, c->free_edge.as<int32_t>(), c->U.as<double>(), c->x.as<double>());
            else
                gather_free_kernel<float><<<blocks_for(c->n_free), kBlock, 0, c->stream>>>(
                    c->n_free, c->free_edge.as<int32_t>(), c->U.as<float>(), c->x.as<float>());
            TF_HIP(hipGetLastError());
            TF_HIP(hipStreamSynchronize(c->stream));
        }
    });
}

tf_status tf_read_probes(tf_ctx* c, double* times, double* vals, int64_t cap, int64_t* n) {
    if (!c || !n) { g_err = "null argument"; return TF_ERR_ARG; }
    return guarded([&] {
        TF_HIP(hipStreamSynchronize(c->stream));
        *n = c->n_samples;
        const int64_t m = std::min(cap, c->n_samples);
        if (times)
            for (int64_t k = 0; k < m; ++k) times[k] = c->sample_times[k];
        if (vals && m > 0 && c->n_probes > 0)
            TF_HIP(hipMemcpy(vals, c->samples.p, m * c->n_probes * 6 * sizeof(double), hipMemcpyDeviceToHost));
    });
}

tf_status tf_read_poynting(tf_ctx* c, double* vals, int64_t cap, int64_t* n) {
    if (!c || !n) { g_err = "null argument"; return TF_ERR_ARG; }
    *n = (int64_t)c->poynting.size();
    if (vals)
        for (int64_t k = 0; k < std::min(cap, *n); ++k) vals[k] = c->poynting[k];
    return TF_OK;
}

tf_status tf_read_losses(tf_ctx* c, double* vals, int64_t cap, int64_t* n) {
    if (!c || !n) { g_err = "null argument"; return TF_ERR_ARG; }
    // the losses column of poynting.dat, one value per stepped step: gfdtd
    // pushes 1/2 U_old.SigMat U_old before each solve and, from step 2 on,
    // U.SigMat U after it (D:1519, D:1543-1545), and the file pairs the first
    // of those values with the steps (D:1573-1576); dga / fdtd: 0 (lossless
    // dga here; the FIT Joule weights exist only in l2norm mode, D:5593-5597)
    *n = (int64_t)c->poynting.size();
    std::vector<double> col;
    if (c->method == TF_METHOD_GFDTD || c->method == TF_METHOD_FEM) {  // fem: the same pushes, D:1748-1768
        const int64_t ns = (int64_t)c->loss_q.size();  // q[i-1] = U_old(i).SigMat U_old(i)
        for (int64_t i = 1; i <= ns && (int64_t)col.size() < *n; ++i) {
            col.push_back(0.5 * c->loss_q[i - 1]);
            if (i > 1 && i < ns) col.push_back(c->loss_q[i]);  // U(i) = U_old(i+1)
        }
    }
    col.resize(*n, 0.0);
    if (vals)
        for (int64_t k = 0; k < std::min(cap, *n); ++k) vals[k] = col[k];
    return TF_OK;
}

// Energy matrices of the context: E and N (tet methods), the ep_vec / nu_vec
// weights (fdtd); batches of dga problems carry none.
static bool has_energy(const tf_ctx* c) {
    return c->method == TF_METHOD_FDTD ? c->fit_ep.p != nullptr : c->Efull.nrows == c->E;
}

tf_status tf_energy(tf_ctx* c, double* we, double* wh) {
    if (!c) { g_err = "null argument"; return TF_ERR_ARG; }
    if (!has_energy(c)) {
        g_err = "no energy matrices in this context (batches of dga problems carry none)";
        return TF_ERR_STATE;
    }
    return guarded([&] {
        if (!c->epart.p) c->epart.alloc(2 * kEnergyParts * sizeof(double));
        int ge = 0, gn = 0;
        if (c->prec == TF_PREC_FP64) energy_partials<double>(c, nullptr, ge, gn);
        else energy_partials<float>(c, nullptr, ge, gn);
        std::vector<double> h(2 * kEnergyParts);
        TF_HIP(hipMemcpyAsync(h.data(), c->epart.p, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
        TF_HIP(hipStreamSynchronize(c->stream));
        double a = 0, b = 0;
        for (int k = 0; k < ge; ++k) a += h[k];
        for (int k = 0; k < gn; ++k) b += h[kEnergyParts + k];
        if (we) *we = 0.5 * a;
        if (wh) *wh = 0.5 * b;
    });
}

tf_status tf_set_energy_output(tf_ctx* c, int32_t on) {
    if (!c) { g_err = "null argument"; return TF_ERR_ARG; }
    if (on && !has_energy(c)) {
        g_err = "no energy matrices in this context (batches of dga problems carry none)";
        return TF_ERR_STATE;
    }
    return guarded([&] {
        drop_graph(c);
        if (on && !c->epart.p) c->epart.alloc(2 * kEnergyParts * sizeof(double));
        c->energy_out = on != 0;
        // a new series: samples taken so far are dropped
        c->n_samples = 0;
        c->sample_times.clear();
        c->sample_cap = 0;
        c->samples.release();
        c->eser.release();
    });
}

tf_status tf_read_energy(tf_ctx* c, double* times, double* vals, int64_t cap, int64_t* n) {
    if (!c || !n) { g_err = "null argument"; return TF_ERR_ARG; }
    if (!c->energy_out) { g_err = "energy output is off (tf_set_energy_output)"; return TF_ERR_STATE; }
    return guarded([&] {
        TF_HIP(hipStreamSynchronize(c->stream));
        *n = c->n_samples;
        const int64_t m = std::min(cap, c->n_samples);
        if (times)
            for (int64_t k = 0; k < m; ++k) times[k] = c->sample_times[k];
        if (vals && m > 0) TF_HIP(hipMemcpy(vals, c->eser.p, m * 2 * sizeof(double), hipMemcpyDeviceToHost));
    });
}

tf_status tf_kernel_times(tf_ctx* c, char* names, int32_t cap_names, double* ms_avg, int64_t* launches, int32_t cap,
                          int32_t* n) {
    if (!c || !n) { g_err = "null argument"; return TF_ERR_ARG; }
    int32_t k = 0;
    int32_t off = 0;
    for (auto& kv : c->ktimes) {
        if (k >= cap) break;
        const std::string& nm = kv.first;
        if (names && off + (int32_t)nm.size() + 1 <= cap_names) {
            std::memcpy(names + off, nm.c_str(), nm.size() + 1);
            off += (int32_t)nm.size() + 1;
        }
        if (ms_avg) ms_avg[k] = kv.second.first / std::max<int64_t>(kv.second.second, 1);
        if (launches) launches[k] = kv.second.second;
        ++k;
    }
    *n = k;
    return TF_OK;
}

tf_status tf_set_profiling(tf_ctx* c, int32_t on) {
    if (!c) { g_err = "null argument"; return TF_ERR_ARG; }
    c->profiling = on != 0;
    return TF_OK;
}

tf_status tf_device_bytes(tf_ctx* c, int64_t* bytes) {
    if (!c || !bytes) { g_err = "null argument"; return TF_ERR_ARG; }
    *bytes = (int64_t)c->device_bytes();
    return TF_OK;
}

tf_status tf_destroy(tf_ctx* c) {
    delete c;
    return TF_OK;
}

tf_status tf_mlp_grad(const float* params, const float* x, const float* gy, int64_t n, float* y, float* grad) {
    if (!params || n < 0 || (n && (!x || !gy))) { g_err = "bad argument"; return TF_ERR_ARG; }
    static_assert(kMlpParams == TF_MLP_PARAMS && kMlpIn == TF_MLP_IN && kMlpHid == TF_MLP_HIDDEN,
                  "MLP layout mismatch");
    return guarded([&] {
        int ndev = 0;
        TF_HIP(hipGetDeviceCount(&ndev));
        if (ndev < 1) throw HipError("no HIP device", TF_ERR_HIP);
        TF_HIP(hipSetDevice(0));
        const int64_t npad = std::max<int64_t>(16, (n + 15) / 16 * 16);
        std::vector<uint16_t> xb((size_t)npad * kMlpIn, 0);
        for (int64_t k = 0; k < n * kMlpIn; ++k) xb[k] = bf16_bits(x[k]);
        std::vector<float> g(npad, 0.f);
        std::copy(gy, gy + n, g.begin());
        std::vector<float> P(params, params + kMlpParams);
        DevBuf dP, dX, dG, dY, dPart, dGrad;
        upload(dP, P);
        upload(dX, xb);
        upload(dG, g);
        dY.alloc(npad * 4);
        const int64_t tiles = npad / 16;
        const int nbb = (int)std::min<int64_t>(512, (tiles + 3) / 4);
        dPart.alloc((size_t)nbb * kMlpParams * 4);
        dGrad.alloc(kMlpParams * 4);
        mlp_forward_kernel<<<(int)std::min<int64_t>(4096, (tiles + 3) / 4), kBlock>>>(tiles, dP.as<float>(),
                                                                                     dX.as<short>(), dY.as<float>());
        mlp_backward_kernel<<<nbb, kBlock>>>(tiles, dP.as<float>(), dX.as<short>(), dG.as<float>(), dPart.as<float>());
        mlp_adam_kernel<<<blocks_for(kMlpParams), kBlock>>>(nbb, dPart.as<float>(), dP.as<float>(), nullptr, nullptr,
                                                            dGrad.as<float>(), 0.f, 0.9f, 0.999f, 1e-8f, 1, 1.f);
        TF_HIP(hipGetLastError());
        TF_HIP(hipDeviceSynchronize());
        if (y && n) TF_HIP(hipMemcpy(y, dY.p, n * 4, hipMemcpyDeviceToHost));
        if (grad) TF_HIP(hipMemcpy(grad, dGrad.p, kMlpParams * 4, hipMemcpyDeviceToHost));
    });
}

tf_status tf_autotune(tf_ctx* c) {
    if (!c) { g_err = "null argument"; return TF_ERR_ARG; }
    return guarded([&] {
        drop_graph(c);
        if (c->prec == TF_PREC_FP64) autotune<double>(c);
        else autotune<float>(c);
    });
}

tf_status tf_fit_layout(tf_ctx* c, int32_t* grid_order) {
    if (!c || !grid_order) { g_err = "null argument"; return TF_ERR_ARG; }
    *grid_order = c->fit_grid ? 1 : 0;
    return TF_OK;
}

tf_status tf_sell_info(tf_ctx* c, int32_t which, int64_t* info) {
    if (!c || !info || which < 0 || which > 6) { g_err = "bad argument"; return TF_ERR_ARG; }
    const Sell* m[7] = {&c->sCt, &c->H, &c->N, &c->A, &c->Kf, &c->R1f, &c->Ctf};
    const Sell& s = *m[which];
    info[0] = s.nrows;
    info[1] = s.nslices;
    info[2] = s.nnz;
    info[3] = s.entries;
    info[4] = s.c16 ? 2 : 4;
    info[5] = s.c16 ? (s.cwin ? 4 : 1) : 0;
    return TF_OK;
}

tf_status tf_set_graphs(tf_ctx* c, int32_t on) {
    if (!c) { g_err = "null argument"; return TF_ERR_ARG; }
    c->use_graphs = on != 0;
    return TF_OK;
}

tf_status tf_set_tuning(tf_ctx* c, const int32_t* blocks, const int32_t* unroll, const int32_t* band, int32_t n) {
    if (!c || !blocks || !unroll || n < 0 || n > TK_COUNT) { g_err = "bad argument"; return TF_ERR_ARG; }
    for (int k = 0; k < n; ++k)
        if (blocks[k] < 0 || (unroll[k] != 4 && unroll[k] != 8 && unroll[k] != 16) ||
            (band && band[k] != 0 && band[k] != 1)) {
            g_err = "tuning: blocks >= 0 (0 = default), unroll 4, 8 or 16, band 0 or 1";
            return TF_ERR_ARG;
        }
    drop_graph(c);
    for (int k = 0; k < n; ++k) c->tune[k] = KTune{blocks[k], unroll[k], band ? band[k] : 0};
    return TF_OK;
}

tf_status tf_get_tuning(tf_ctx* c, int32_t* blocks, int32_t* unroll, int32_t* band, int32_t cap, int32_t* n) {
    if (!c || !n) { g_err = "null argument"; return TF_ERR_ARG; }
    *n = TK_COUNT;
    for (int k = 0; k < std::min<int32_t>(cap, TK_COUNT); ++k) {
        if (blocks) blocks[k] = c->tune[k].blocks;
        if (unroll) unroll[k] = c->tune[k].unroll;
        if (band) band[k] = c->tune[k].band;
    }
    return TF_OK;
}

tf_status tf_batch_set_sources(tf_ctx* c, int32_t inst, const tf_source* src, int32_t n_src,
                               const tf_source_term* terms, int64_t n_terms) {
    if (!c || !c->batch || inst < 0 || inst >= c->batch->n || n_src < 0 || n_terms < 0 || (n_src && !src) ||
        (n_terms && !terms)) {
        g_err = "bad argument";
        return TF_ERR_ARG;
    }
    return guarded([&] {
        drop_graph(c);
        auto& b = *c->batch;
        const int32_t E = b.ecnt[inst];
        for (int64_t i = 0; i < n_terms; ++i)
            if (terms[i].edge < 0 || terms[i].edge >= E || terms[i].source < 0 || terms[i].source >= n_src)
                throw std::runtime_error("source term out of range");
        b.src[inst].assign(src, src + n_src);
        b.terms[inst].assign(terms, terms + n_terms);
        batch_sources(c);
    });
}

tf_status tf_batch_set_probes(tf_ctx* c, int32_t inst, const tf_probe* probes, int32_t n) {
    if (!c || !c->batch || inst < 0 || inst >= c->batch->n || n < 0 || (n && !probes)) {
        g_err = "bad argument";
        return TF_ERR_ARG;
    }
    return guarded([&] {
        auto& b = *c->batch;
        const int32_t eo = b.eoff[inst], fo = b.foff[inst];
        const int32_t E = b.ecnt[inst], F = b.foff[inst + 1] - fo;
        std::vector<tf_probe> v(probes, probes + n);
        for (tf_probe& p : v) {
            if (p.ne < 0 || p.ne > TF_PROBE_MAX_E || p.nh < 0 || p.nh > TF_PROBE_MAX_H)
                throw std::runtime_error("probe term count");
            for (int k = 0; k < p.ne; ++k) {
                if (p.e_idx[k] < 0 || p.e_idx[k] >= E) throw std::runtime_error("probe edge out of range");
                p.e_idx[k] += eo;
            }
            for (int k = 0; k < p.nh; ++k) {
                if (p.h_idx[k] < 0 || p.h_idx[k] >= F) throw std::runtime_error("probe face out of range");
                p.h_idx[k] += fo;
            }
        }
        b.probes[inst] = std::move(v);
        std::vector<tf_probe> all;
        b.poff.assign(1, 0);
        for (int32_t k = 0; k < b.n; ++k) {
            all.insert(all.end(), b.probes[k].begin(), b.probes[k].end());
            b.poff.push_back((int32_t)all.size());
        }
        b.sample_steps.clear();
        const tf_status rc = tf_set_probes(c, all.data(), (int32_t)all.size());
        if (rc != TF_OK) throw std::runtime_error(g_err);
    });
}

tf_status tf_batch_set_output(tf_ctx* c, int64_t every) {
    if (!c || !c->batch || every < 0) { g_err = "bad argument"; return TF_ERR_ARG; }
    c->batch->every = every;
    return TF_OK;
}

tf_status tf_batch_read_probes(tf_ctx* c, int32_t inst, double* times, double* vals, int64_t cap, int64_t* n) {
    if (!c || !c->batch || !n || inst < 0 || inst >= c->batch->n) { g_err = "bad argument"; return TF_ERR_ARG; }
    return guarded([&] {
        auto& b = *c->batch;
        TF_HIP(hipStreamSynchronize(c->stream));
        *n = c->n_samples;
        const int64_t m = std::min(cap, c->n_samples);
        if (times)
            for (int64_t k = 0; k < m; ++k) times[k] = (double)b.sample_steps[k] * b.dt[inst];
        const int32_t P = c->n_probes, p0 = b.poff[inst], np = b.poff[inst + 1] - p0;
        if (vals && m > 0 && np > 0) {
            std::vector<double> all((size_t)m * P * 6);
            TF_HIP(hipMemcpy(all.data(), c->samples.p, all.size() * sizeof(double), hipMemcpyDeviceToHost));
            for (int64_t k = 0; k < m; ++k)
                std::memcpy(vals + (size_t)k * np * 6, all.data() + ((size_t)k * P + p0) * 6, (size_t)np * 6 * 8);
        }
    });
}

tf_status tf_batch_read_state(tf_ctx* c, int32_t inst, double* U, double* F, double* B) {
    if (!c || !c->batch || inst < 0 || inst >= c->batch->n) { g_err = "bad argument"; return TF_ERR_ARG; }
    return guarded([&] {
        auto& b = *c->batch;
        std::vector<double> u(c->E), f(c->F), bb(c->F);
        const tf_status rc = tf_read_state(c, u.data(), f.data(), bb.data());
        if (rc != TF_OK) throw std::runtime_error(g_err);
        const int32_t eo = b.eoff[inst], fo = b.foff[inst];
        if (U) std::copy(u.begin() + eo, u.begin() + eo + b.ecnt[inst], U);
        if (F) std::copy(f.begin() + fo, f.begin() + b.foff[inst + 1], F);
        if (B) std::copy(bb.begin() + fo, bb.begin() + b.foff[inst + 1], B);
    });
}

}  // extern "C"

#include "noalloc.hpp"
