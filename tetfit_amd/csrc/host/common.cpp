// common.cpp - CSR assembly helpers, source evaluation split, dt power
// iteration and output schedule shared by the FIT and tetrahedral builders.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <mutex>
#include <numeric>

#include "build.hpp"

namespace tfh {

static double wall_seconds() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

PhaseClock::PhaseClock() : on(std::getenv("TETFIT_HOST_TIMING") != nullptr), t0(wall_seconds()), t(t0) {}

void PhaseClock::mark(const char* phase) {
    if (!on) return;
    const double now = wall_seconds();
    std::fprintf(stderr, "[host] %-28s %8.3f s (total %8.3f s)\n", phase, now - t, now - t0);
    t = now;
}

tf_csr CsrStore::view() const {
    tf_csr c{};
    c.nrows = nrows;
    c.ncols = ncols;
    c.nnz = rowptr.empty() ? 0 : rowptr.back();
    c.rowptr = rowptr.empty() ? nullptr : rowptr.data();
    c.col = col.empty() ? nullptr : col.data();
    c.val = val.empty() ? nullptr : val.data();
    c.sgn = sgn.empty() ? nullptr : sgn.data();
    return c;
}

CsrStore Triplets::to_csr(int32_t nrows, int32_t ncols, bool overwrite) const {
    CsrStore m;
    m.nrows = nrows;
    m.ncols = ncols;
    const size_t n = r.size();
    // Counting sort by row keeps insertion order within a row.
    std::vector<int64_t> cnt(nrows + 1, 0);
    for (size_t k = 0; k < n; ++k) cnt[r[k] + 1]++;
    for (int32_t i = 0; i < nrows; ++i) cnt[i + 1] += cnt[i];
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    std::vector<int32_t> order(n);
    for (size_t k = 0; k < n; ++k) order[pos[r[k]]++] = (int32_t)k;
    m.rowptr.assign(nrows + 1, 0);
    m.col.reserve(n);
    m.val.reserve(n);
    std::vector<int32_t> idx;
    for (int32_t i = 0; i < nrows; ++i) {
        idx.assign(order.begin() + cnt[i], order.begin() + cnt[i + 1]);
        std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) { return c[a] < c[b]; });
        for (size_t q = 0; q < idx.size();) {
            int32_t j = c[idx[q]];
            double acc = v[idx[q]];
            size_t w = q + 1;
            for (; w < idx.size() && c[idx[w]] == j; ++w) acc = overwrite ? v[idx[w]] : acc + v[idx[w]];
            m.col.push_back(j);
            m.val.push_back(acc);
            q = w;
        }
        m.rowptr[i + 1] = (int64_t)m.col.size();
    }
    return m;
}

// ---------------------------------------------------------------- sources

double SourceEval::carrier(const SourceDef& s, double t) {
    if (t < 0) return 0;  // heaviside of Compute*Source (D:3909, D:4091)
    double ret = s.amp;
    if (s.st == "j") ret *= s.impedance;
    const long double arg = t - s.delay;
    if (s.carrier == "sin") ret *= std::sin(2 * PI_L * s.freq * arg);
    else if (s.carrier == "cos") ret *= s.amp * std::cos(2 * PI_L * s.freq * arg);
    else if (s.carrier == "gaussian") ret *= std::exp(-std::pow(t - s.delay, 2) / std::pow(s.width, 2));
    else if (s.carrier == "ricker") {
        long double a = PI_L * s.freq * arg;
        ret *= (1 - 2 * a * a) * std::exp(-(a * a));
    } else if (s.carrier == "gaussmod") {
        ret *= std::exp(-std::pow(t - s.delay, 2) / std::pow(s.width, 2));
        ret *= std::sin(2 * PI_L * s.freq * arg);
    }
    return ret;
}

// One mode factor of the wave profile (Source.hpp:362-417): sin / cos of
// 2 pi k x in long double like the reference (its PI is long double), the
// GSL Bessel and Legendre forms through the C++17 special functions
// (J_n and y_l of negative arguments by parity, as GSL evaluates them).
static long double mode_factor(const std::string& f, double k, double x) {
    if (f == "sin") return std::sin(2 * PI_L * k * (long double)x);
    if (f == "cos") return std::cos(2 * PI_L * k * (long double)x);
    if (f.empty()) return 1.0;
    const char c = f[0];
    if (c != 'J' && c != 'Y' && c != 'j' && c != 'y' && c != 'P') return 1.0;
    const int l = std::stoi(f.substr(1));
    const double a = k * x;
    const double sgn = (l % 2 == 0) ? 1.0 : -1.0;
    switch (c) {
        case 'J': return a < 0 ? sgn * std::cyl_bessel_j((double)l, -a) : std::cyl_bessel_j((double)l, a);
        case 'Y': return std::cyl_neumann((double)l, a);
        case 'j': return a < 0 ? sgn * std::sph_bessel((unsigned)l, -a) : std::sph_bessel((unsigned)l, a);
        case 'y': return a < 0 ? -sgn * std::sph_neumann((unsigned)l, -a) : std::sph_neumann((unsigned)l, a);
        default: return std::legendre((unsigned)l, std::cos(a));
    }
}

Vec3 SourceEval::space(const SourceDef& s, const Vec3& p) {
    double ret = 1.0;
    if (s.decay != 0)
        ret *= std::pow(std::pow(p[0] - s.center[0], 2) + std::pow(p[1] - s.center[1], 2) +
                            std::pow(p[2] - s.center[2], 2),
                        -0.5 * s.decay);
    auto gauss = [&]() {
        double e = -((std::pow(p[0] - s.center[0], 2) + std::pow(p[1] - s.center[1], 2) +
                      std::pow(p[2] - s.center[2], 2)) /
                     std::pow(s.radius, 2));
        ret *= std::exp(e);
    };
    const std::array<double, 3>& c = s.center;
    if (s.dir == "rho" || s.dir == "theta" || s.dir == "phi") {
        // spherical coordinates of the point and the centre (Source.hpp:334-349)
        double r1[3] = {0, 0, 0}, r2[3] = {0, 0, 0};
        if (p[0] != 0 || p[1] != 0 || p[2] != 0) {
            r1[0] = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
            r1[1] = std::acos(p[2] / r1[0]);
            r1[2] = std::atan2(p[1], p[0]);
        }
        if (c[0] != 0 || c[1] != 0 || c[2] != 0) {
            r2[0] = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            r2[1] = std::acos(c[2] / r2[0]);
            r2[2] = std::atan2(c[1], c[0]);
        }
        if (s.prof == "gaussian") gauss();
        else if (s.prof == "wave")
            for (int j = 0; j < 3; ++j) ret *= mode_factor(s.bfuncs[j], s.kvec[j], r1[j] - r2[j]);
        const double st = std::sin(r1[1]), ct = std::cos(r1[1]), sp = std::sin(r1[2]), cp = std::cos(r1[2]);
        if (s.dir == "rho") return Vec3(st * cp * ret, st * sp * ret, ct * ret);
        // the reference's theta vector has -r sin(phi) as its z component (Source.hpp:426)
        if (s.dir == "theta") return Vec3(r1[0] * ct * cp * ret, r1[0] * ct * sp * ret, -r1[0] * sp * ret);
        return Vec3(-r1[0] * st * sp * ret, r1[0] * st * cp * ret, 0);
    }
    if (s.dir == "r" || s.dir == "psi") {
        // cylindrical coordinates (Source.hpp:430-457); the centre's angle
        // branch follows the sign of the point's x, as in the reference
        double r1[3] = {0, 0, p[2]}, r2[3] = {0, 0, c[2]};
        if (p[0] != 0 || p[1] != 0) {
            r1[0] = std::sqrt(p[0] * p[0] + p[1] * p[1]);
            r1[1] = p[0] >= 0 ? std::asin(p[1] / r1[0]) : (double)(-std::asin(p[1] / r1[0]) + PI_L);
        }
        if (c[0] != 0 || c[1] != 0 || c[2] != 0) {
            r2[0] = std::sqrt(c[0] * c[0] + c[1] * c[1]);
            if (c[0] == 0 && c[1] == 0) r2[1] = 0;
            else r2[1] = p[0] >= 0 ? std::asin(c[1] / r2[0]) : (double)(-std::asin(c[1] / r2[0]) + PI_L);
        }
        if (s.prof == "gaussian") gauss();
        else if (s.prof == "wave")
            for (int j = 0; j < 3; ++j) ret *= mode_factor(s.bfuncs[j], s.kvec[j], r1[j] - r2[j]);
        if (s.dir == "r") return Vec3(std::cos(r1[1]) * ret, std::sin(r1[1]) * ret, 0);
        return Vec3(-r1[0] * std::sin(r1[1]) * ret, r1[0] * std::cos(r1[1]) * ret, 0);
    }
    if (s.prof == "gaussian") gauss();
    else if (s.prof == "wave") {
        for (int j = 0; j < 3; ++j) {
            const std::string& f = s.bfuncs[j];
            if (f == "sin") ret *= std::sin(2 * PI_L * s.kvec[j] * (long double)(p[j] - s.center[j]));
            else if (f == "cos") ret *= std::cos(2 * PI_L * s.kvec[j] * (long double)(p[j] - s.center[j]));
        }
    }
    if (s.dir == "x") return Vec3(ret, 0, 0);
    if (s.dir == "y") return Vec3(0, ret, 0);
    return Vec3(0, 0, ret);
}

static int carrier_kind(const std::string& c) {
    if (c == "sin") return TF_CARRIER_SIN;
    if (c == "cos") return TF_CARRIER_COS;
    if (c == "gaussian") return TF_CARRIER_GAUSSIAN;
    if (c == "ricker") return TF_CARRIER_RICKER;
    if (c == "gaussmod") return TF_CARRIER_GAUSSMOD;
    return TF_CARRIER_DC;
}

static int field_kind(const std::string& st) {
    if (st == "e") return TF_FIELD_E;
    if (st == "j") return TF_FIELD_J;
    if (st == "h") return TF_FIELD_H;
    return TF_FIELD_D;
}

tf_source make_source(const SourceDef& s) {
    tf_source r{};
    r.carrier = carrier_kind(s.carrier);
    r.field = field_kind(s.st);
    r.amp = s.amp;
    r.freq = s.freq;
    r.delay = s.delay;
    r.width = s.width;
    r.impedance = s.impedance;
    return r;
}

tf_source_desc describe_source(const SourceDef& s) {
    tf_source_desc d{};
    d.type = field_kind(s.st);
    d.carrier = carrier_kind(s.carrier);
    d.profile = s.prof == "wave" ? 1 : (s.prof == "gaussian" ? 2 : 0);
    static const char* const kDirs[] = {"x", "y", "z", "rho", "theta", "phi", "r", "psi"};
    d.dir = 2;
    for (int k = 0; k < 8; ++k)
        if (s.dir == kDirs[k]) d.dir = k;
    for (int j = 0; j < 3; ++j) {
        const std::string& f = s.bfuncs[j];
        d.bfunc[j] = 0;
        d.border[j] = 0;
        if (f == "sin") d.bfunc[j] = 1;
        else if (f == "cos") d.bfunc[j] = 2;
        else if (!f.empty() && std::string("JYjyP").find(f[0]) != std::string::npos) {
            d.bfunc[j] = 3 + (int32_t)std::string("JYjyP").find(f[0]);
            d.border[j] = std::stoi(f.substr(1));
        }
        d.kvec[j] = s.kvec[j];
        d.center[j] = s.center[j];
    }
    d.amp = s.amp;
    d.freq = s.freq;
    d.delay = s.delay;
    d.width = s.width;
    d.impedance = s.impedance;
    d.radius = s.radius;
    d.decay = s.decay;
    return d;
}

// ---------------------------------------------------------------- rows

CsrStore assemble_rows(int32_t nrows, int32_t ncols,
                       const std::function<void(int32_t, std::vector<RowEntry>&)>& row) {
    constexpr int32_t kBlk = 4096;
    const int32_t nb = (nrows + kBlk - 1) / kBlk;
    std::vector<std::vector<int32_t>> bcol(nb);
    std::vector<std::vector<double>> bval(nb);
    CsrStore m;
    m.nrows = nrows;
    m.ncols = ncols;
    m.rowptr.assign((size_t)nrows + 1, 0);
    std::atomic<bool> failed{false};
    std::string what;
    std::mutex mu;
#pragma omp parallel
    {
        std::vector<RowEntry> ent;
#pragma omp for schedule(dynamic, 1)
        for (int32_t blk = 0; blk < nb; ++blk) {
            if (failed.load(std::memory_order_relaxed)) continue;
            try {
                const int32_t r0 = blk * kBlk, r1 = std::min(nrows, r0 + kBlk);
                for (int32_t r = r0; r < r1; ++r) {
                    ent.clear();
                    row(r, ent);
                    // stable insertion sort by column: equal columns keep their order
                    for (size_t i = 1; i < ent.size(); ++i) {
                        RowEntry x = ent[i];
                        size_t j = i;
                        while (j > 0 && ent[j - 1].col > x.col) { ent[j] = ent[j - 1]; --j; }
                        ent[j] = x;
                    }
                    int64_t cnt = 0;
                    for (size_t i = 0; i < ent.size();) {
                        double s = 0.0;
                        size_t j = i;
                        for (; j < ent.size() && ent[j].col == ent[i].col; ++j) s += ent[j].val;
                        bcol[blk].push_back(ent[i].col);
                        bval[blk].push_back(s);
                        ++cnt;
                        i = j;
                    }
                    m.rowptr[(size_t)r + 1] = cnt;
                }
            } catch (const std::exception& e) {
                std::lock_guard<std::mutex> g(mu);
                if (!failed.exchange(true)) what = e.what();
            }
        }
    }
    if (failed) throw Error(what);
    for (int32_t r = 0; r < nrows; ++r) m.rowptr[(size_t)r + 1] += m.rowptr[r];
    m.col.resize((size_t)m.rowptr[nrows]);
    m.val.resize((size_t)m.rowptr[nrows]);
#pragma omp parallel for schedule(static)
    for (int32_t blk = 0; blk < nb; ++blk) {
        const int64_t at = m.rowptr[(size_t)blk * kBlk];
        std::copy(bcol[blk].begin(), bcol[blk].end(), m.col.begin() + at);
        std::copy(bval[blk].begin(), bval[blk].end(), m.val.begin() + at);
        std::vector<int32_t>().swap(bcol[blk]);
        std::vector<double>().swap(bval[blk]);
    }
    return m;
}

// ---------------------------------------------------------------- dt

double par_dot(const double* a, const double* b, size_t n) {
    constexpr size_t kBlk = 32768;
    const int64_t nb = (int64_t)((n + kBlk - 1) / kBlk);
    std::vector<double> part(nb);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nb; ++i) {
        const size_t lo = (size_t)i * kBlk, hi = std::min(n, lo + kBlk);
        double s = 0;
        for (size_t j = lo; j < hi; ++j) s += a[j] * b[j];
        part[i] = s;
    }
    double s = 0;
    for (double v : part) s += v;
    return s;
}

double power_lambda_max(size_t n,
                        const std::function<void(const std::vector<double>&, std::vector<double>&)>& apply,
                        int iters, const std::vector<double>* start) {
    // Deterministic start vector with all modes present (or the caller's).
    std::vector<double> x(n), y(n);
    if (start) {
        if (start->size() != n) throw Error("power iteration: start vector size");
        x = *start;
    } else {
        uint64_t st = 0x9E3779B97F4A7C15ull;
        for (size_t i = 0; i < n; ++i) {
            st ^= st << 13; st ^= st >> 7; st ^= st << 17;
            x[i] = (double)(st >> 11) * (1.0 / 9007199254740992.0) - 0.5;
        }
    }
    double lam = 0;
    for (int it = 0; it < iters; ++it) {
        double nx = std::sqrt(par_dot(x.data(), x.data(), n));
        if (nx == 0) return 0;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)n; ++i) x[i] /= nx;
        apply(x, y);
        lam = par_dot(x.data(), y.data(), n);
        std::swap(x, y);
    }
    return lam;
}

// ---------------------------------------------------------------- output

void set_output(const Model& m, const SimulationDef& sim, Build& b) {
    auto it = m.outputs.find(sim.output);
    if (it == m.outputs.end()) {
        b.out_mode = 0;
        b.out_name = "simulation";
        return;
    }
    const OutputDef& o = it->second;
    const std::string& md = o.mode;
    b.out_mode = md == "probepoint" ? 1 : md == "maxerror" ? 2 : md == "l2norm" ? 3 : md == "silo" ? 4 : 0;
    b.out_delay = o.delay;
    b.out_period = o.period;
    b.out_frame = (int32_t)o.ref_sys;
    b.out_name = o.name;
}

void Build::finalize_views() {
    prob.n_edges = (int32_t)(method == "fdtd" ? M_h.size() : edge_class.size());
    tf_fit_ops& f = prob.fit;
    f = tf_fit_ops{};
    if (method == "fdtd") {
        f.ct_ptr = ct_ptr.data();
        f.ct_face = ct_face.data();
        f.ct_sgn = ct_sgn.data();
        f.c_edge = c_edge.data();
        f.curl = curl.data();
        f.dual_curl = dual_curl.data();
        f.M_h = M_h.data();
        f.M_q = M_q.data();
        f.upml = upml.data();
        f.M_mu = M_mu.data();
        f.M_nu = M_nu.data();
        f.fpml = fpml.data();
        f.common_edges = common_edges.data();
        f.n_common = (int64_t)common_edges.size();
        f.uncommon_edges = uncommon_edges.data();
        f.n_uncommon = (int64_t)uncommon_edges.size();
        f.bc_edges = bc_edges.data();
        f.n_bc = (int64_t)bc_edges.size();
        f.bc_value = bc_value.data();
        f.src_edges = src_edges.data();
        f.n_src = (int64_t)src_edges.size();
        f.tbc_faces = tbc_faces.data();
        f.n_tbc = (int64_t)tbc_faces.size();
        f.ep_vec = M_e.data();
        f.nu_vec = Nu_vec.data();
        if (!edge_sid.empty() && !face_sid.empty()) {
            f.edge_sid = edge_sid.data();
            f.face_sid = face_sid.data();
            for (int a = 0; a < 3; ++a) f.cells[a] = cells[a];
        }
    } else {
        tf_dga_ops& d = prob.dga;
        d = tf_dga_ops{};
        d.C = C.view();
        d.Ct = Ct.view();
        d.bnd_edge = bnd_edge.data();
        d.n_bnd = n_bnd;
        d.edge_class = edge_class.data();
        d.fixed_value = fixed_value.data();
        d.H = H.view();
        d.h_row = h_row.empty() ? nullptr : h_row.data();
        d.N = N.view();
        d.E = E.view();
        d.A = A.view();
        d.RHS1 = RHS1.view();
        d.free_edges = free_edges.empty() ? nullptr : free_edges.data();
        d.n_free = (int32_t)free_edges.size();
        d.Sig = Sig.view();
        d.edge_rank = edge_rank.empty() ? nullptr : edge_rank.data();
        d.face_rank = face_rank.empty() ? nullptr : face_rank.data();
        if (!frac_slot.empty()) {
            tf_dga_frac& fr = d.frac;
            fr.enabled = 1;
            fr.h_size = frac_sizes[0];
            fr.p_size = frac_sizes[1];
            fr.q_size = frac_sizes[2];
            fr.b_size = frac_sizes[3];
            fr.n_size = frac_sizes[4];
            fr.r_size = frac_sizes[5];
            fr.s_size = frac_sizes[6];
            fr.Ku = frac_Ku.view();
            fr.Ru = frac_Ru.view();
            fr.Mu = frac_M.view();
            fr.Kf = frac_Kf.view();
            fr.Rf = frac_Rf.view();
            fr.Tf = frac_T.view();
            fr.slot = frac_slot.data();
        }
    }
}

}  // namespace tfh
