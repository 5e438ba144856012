// oracle.cpp - CPU checker (test infrastructure, see oracle.h).
//
// Single-threaded fp64 restatement of the reference loops; every function
// names the reference lines it follows (D: = src/DGA/Discretization.hpp).
// the C++ special functions (cyl_bessel_j & co.) under -std=c++14
#define __STDCPP_WANT_MATH_SPEC_FUNCS__ 1
#include "oracle.h"

#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

const long double PI_L = 3.141592653589793238L;  // Utilities.hpp:92

struct V3 {
    double x = 0, y = 0, z = 0;
};

// Source::Compute, Source.hpp:313-327 (carrier), :328-331 (decay) and
// :520-567 (cartesian directions, gaussian / wave profiles).
V3 source_compute(const tf_source_desc& s, const double p[3], double t) {
    double ret = s.amp;
    if (s.type == TF_FIELD_J) ret *= s.impedance;
    switch (s.carrier) {
        case TF_CARRIER_SIN: ret *= std::sin(2 * PI_L * s.freq * (long double)(t - s.delay)); break;
        case TF_CARRIER_COS: ret *= s.amp * std::cos(2 * PI_L * s.freq * (long double)(t - s.delay)); break;
        case TF_CARRIER_GAUSSIAN: ret *= std::exp(-std::pow(t - s.delay, 2) / std::pow(s.width, 2)); break;
        case TF_CARRIER_RICKER:
            ret *= (1 - 2 * std::pow(PI_L * s.freq * (long double)(t - s.delay), 2)) *
                   std::exp(-std::pow(PI_L * s.freq * (long double)(t - s.delay), 2));
            break;
        case TF_CARRIER_GAUSSMOD:  // extension carrier: gaussian envelope times sin
            ret *= std::exp(-std::pow(t - s.delay, 2) / std::pow(s.width, 2));
            ret *= std::sin(2 * PI_L * s.freq * (long double)(t - s.delay));
            break;
        default: break;
    }
    if (s.decay != 0)
        ret *= std::pow(std::pow(p[0] - s.center[0], 2) + std::pow(p[1] - s.center[1], 2) +
                            std::pow(p[2] - s.center[2], 2),
                        -0.5 * s.decay);
    const double* c = s.center;
    // coordinates of the profile: cartesian, spherical (rho, theta, phi,
    // Source.hpp:334-349) or cylindrical (r, psi, z; the centre's angle
    // branch takes the sign of the point's x, Source.hpp:430-457)
    double q1[3] = {p[0], p[1], p[2]}, q2[3] = {c[0], c[1], c[2]};
    const bool sph = s.dir >= 3 && s.dir <= 5, cyl = s.dir >= 6;
    if (sph) {
        q1[0] = q1[1] = q1[2] = q2[0] = q2[1] = q2[2] = 0;
        if (p[0] != 0 || p[1] != 0 || p[2] != 0) {
            q1[0] = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
            q1[1] = std::acos(p[2] / q1[0]);
            q1[2] = std::atan2(p[1], p[0]);
        }
        if (c[0] != 0 || c[1] != 0 || c[2] != 0) {
            q2[0] = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            q2[1] = std::acos(c[2] / q2[0]);
            q2[2] = std::atan2(c[1], c[0]);
        }
    } else if (cyl) {
        q1[0] = q1[1] = q2[0] = q2[1] = 0;
        if (p[0] != 0 || p[1] != 0) {
            q1[0] = std::sqrt(p[0] * p[0] + p[1] * p[1]);
            if (p[0] >= 0) q1[1] = std::asin(p[1] / q1[0]);
            else q1[1] = -std::asin(p[1] / q1[0]) + PI_L;
        }
        if (c[0] != 0 || c[1] != 0 || c[2] != 0) {
            q2[0] = std::sqrt(c[0] * c[0] + c[1] * c[1]);
            if (c[0] == 0 && c[1] == 0) q2[1] = 0;
            else if (p[0] >= 0) q2[1] = std::asin(c[1] / q2[0]);
            else q2[1] = -std::asin(c[1] / q2[0]) + PI_L;
        }
    }
    if (s.profile == 2) {
        double e = -((std::pow(p[0] - s.center[0], 2) + std::pow(p[1] - s.center[1], 2) +
                      std::pow(p[2] - s.center[2], 2)) /
                     std::pow(s.radius, 2));
        ret *= std::exp(e);
    } else if (s.profile == 1) {
        for (int j = 0; j < 3; ++j) {
            const double x = q1[j] - q2[j];
            const int n = s.border[j];
            if (s.bfunc[j] == 1) ret *= std::sin(2 * PI_L * s.kvec[j] * (long double)x);
            else if (s.bfunc[j] == 2) ret *= std::cos(2 * PI_L * s.kvec[j] * (long double)x);
            else if (s.bfunc[j] >= 3 && (sph || cyl)) {
                // GSL's Jn, Yn, jl, yl, Pl (Source.hpp:362-404); x < 0 by parity
                const double a = s.kvec[j] * x;
                const double par = (n % 2 == 0) ? 1.0 : -1.0;
                switch (s.bfunc[j]) {
                    case 3: ret *= a < 0 ? par * std::cyl_bessel_j((double)n, -a) : std::cyl_bessel_j((double)n, a); break;
                    case 4: ret *= std::cyl_neumann((double)n, a); break;
                    case 5: ret *= a < 0 ? par * std::sph_bessel((unsigned)n, -a) : std::sph_bessel((unsigned)n, a); break;
                    case 6: ret *= a < 0 ? -par * std::sph_neumann((unsigned)n, -a) : std::sph_neumann((unsigned)n, a); break;
                    default: ret *= std::legendre((unsigned)n, std::cos(a)); break;
                }
            }
        }
    }
    V3 v;
    switch (s.dir) {
        case 0: v.x = ret; break;
        case 1: v.y = ret; break;
        case 3: v.x = std::sin(q1[1]) * std::cos(q1[2]) * ret; v.y = std::sin(q1[1]) * std::sin(q1[2]) * ret;
                v.z = std::cos(q1[1]) * ret; break;
        case 4: v.x = q1[0] * std::cos(q1[1]) * std::cos(q1[2]) * ret;
                v.y = q1[0] * std::cos(q1[1]) * std::sin(q1[2]) * ret;
                v.z = -q1[0] * std::sin(q1[2]) * ret; break;  // sin(phi), as Source.hpp:426 has it
        case 5: v.x = -q1[0] * std::sin(q1[1]) * std::sin(q1[2]) * ret;
                v.y = q1[0] * std::sin(q1[1]) * std::cos(q1[2]) * ret; break;
        case 6: v.x = std::cos(q1[1]) * ret; v.y = std::sin(q1[1]) * ret; break;
        case 7: v.x = -q1[0] * std::sin(q1[1]) * ret; v.y = q1[0] * std::cos(q1[1]) * ret; break;
        default: v.z = ret; break;
    }
    return v;
}

struct SourceTable {
    std::vector<tf_source_desc> sd;
    std::vector<tf_source_edge> se;
    std::vector<int64_t> by_edge;  // edge -> index into se, or -1
};

// ComputeHfieldSource (D:4088-4202): sum of h-type sources at the edge
// barycentre, dotted with the boundary dual edge vector; 0 for t < 0.
double h_source(const SourceTable& st, int64_t k, double t) {
    if (t < 0) return 0;
    const tf_source_edge& e = st.se[k];
    V3 acc;
    for (int i = 0; i < e.n_src; ++i) {
        const tf_source_desc& s = st.sd[e.src[i]];
        if (s.type != TF_FIELD_H) continue;
        V3 v = source_compute(s, e.bary, t);
        acc.x += v.x; acc.y += v.y; acc.z += v.z;
    }
    return acc.x * e.vec[0] + acc.y * e.vec[1] + acc.z * e.vec[2];
}

// ComputeEfieldSource (D:3907-3940): e-type sources dotted with the edge vector.
double e_source(const SourceTable& st, int64_t k, double t) {
    if (t < 0) return 0;
    const tf_source_edge& e = st.se[k];
    V3 acc;
    for (int i = 0; i < e.n_src; ++i) {
        const tf_source_desc& s = st.sd[e.src[i]];
        if (s.type != TF_FIELD_E) continue;
        V3 v = source_compute(s, e.bary, t);
        acc.x += v.x; acc.y += v.y; acc.z += v.z;
    }
    return acc.x * e.vec[0] + acc.y * e.vec[1] + acc.z * e.vec[2];
}

// ComputeCurrentSource (D:3994-4028): j-type sources dotted with the unit edge vector.
double j_source(const SourceTable& st, int64_t k, double t) {
    if (t < 0) return 0;
    const tf_source_edge& e = st.se[k];
    V3 acc;
    for (int i = 0; i < e.n_src; ++i) {
        const tf_source_desc& s = st.sd[e.src[i]];
        if (s.type != TF_FIELD_J) continue;
        V3 v = source_compute(s, e.bary, t);
        acc.x += v.x; acc.y += v.y; acc.z += v.z;
    }
    const double nrm = std::sqrt(e.vec[0] * e.vec[0] + e.vec[1] * e.vec[1] + e.vec[2] * e.vec[2]);
    return (1.0 / nrm) * (acc.x * e.vec[0] + acc.y * e.vec[1] + acc.z * e.vec[2]);
}

// 3x3 inverse by cofactors (what Eigen's Matrix3d::inverse computes).
void inv3(const double m[9], double r[9]) {
    const double c00 = m[4] * m[8] - m[5] * m[7], c01 = m[5] * m[6] - m[3] * m[8], c02 = m[3] * m[7] - m[4] * m[6];
    const double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
    r[0] = c00 / det;
    r[1] = (m[2] * m[7] - m[1] * m[8]) / det;
    r[2] = (m[1] * m[5] - m[2] * m[4]) / det;
    r[3] = c01 / det;
    r[4] = (m[0] * m[8] - m[2] * m[6]) / det;
    r[5] = (m[2] * m[3] - m[0] * m[5]) / det;
    r[6] = c02 / det;
    r[7] = (m[1] * m[6] - m[0] * m[7]) / det;
    r[8] = (m[0] * m[4] - m[1] * m[3]) / det;
}

// Output::AllowPrint (Output.hpp:318-328)
struct Printer {
    bool on;
    double next, period;
    bool allow(double t) {
        if (!on) return false;
        if (t >= next) { next += period; return true; }
        return false;
    }
};

// 1/2 x.(M x) over a CSR matrix, rows in order (the energy diagnostic).
double half_quad(const tf_csr& m, const std::vector<double>& x) {
    double s = 0;
    for (int32_t r = 0; r < m.nrows; ++r) {
        double mx = 0;
        for (int64_t q = m.rowptr[r]; q < m.rowptr[r + 1]; ++q) mx += m.val[q] * x[m.col[q]];
        s += x[r] * mx;
    }
    return 0.5 * s;
}

double half_weighted(const double* w, const std::vector<double>& x) {
    double s = 0;
    for (size_t i = 0; i < x.size(); ++i) s += w[i] * x[i] * x[i];
    return 0.5 * s;
}

// We, Wh of a state: ExportFields l2norm (D:2773-2775) for the tet methods,
// ExportFitFields (D:3468-3470) for fdtd.
void energies(const tf_problem* p, const std::vector<double>& U, const std::vector<double>& B, double& we,
              double& wh) {
    if (p->method == TF_METHOD_FDTD) {
        we = half_weighted(p->fit.ep_vec, U);
        wh = half_weighted(p->fit.nu_vec, B);
    } else {
        we = half_quad(p->dga.E, U);
        wh = half_quad(p->dga.N, B);
    }
}

struct Sink {
    oracle_out* o;
    int32_t np;
    const tf_problem* p;
    // one output sample: the probes' six values and, when asked for, the
    // energies of the same state (U, B)
    void put(double t, const std::vector<double>& six, const std::vector<double>& U, const std::vector<double>& B) {
        if (o->n_samples < o->cap_samples) {
            if (o->times) o->times[o->n_samples] = t;
            if (o->vals)
                std::memcpy(o->vals + o->n_samples * np * 6, six.data(), sizeof(double) * six.size());
            if (o->energy) energies(p, U, B, o->energy[2 * o->n_samples], o->energy[2 * o->n_samples + 1]);
        }
        o->n_samples++;
    }
    void poy(double v) {
        if (o->poynting && o->n_poynting < o->cap_poynting) o->poynting[o->n_poynting] = v;
        o->n_poynting++;
    }
};

// GetFitElectricField D:3555-3568 and GetFitMagneticField D:3570-3588.
void export_fit(const std::vector<double>& U, const std::vector<double>& B, const tf_probe_geom* pg, int32_t np,
                std::vector<double>& six) {
    static const int axis[12] = {0, 1, 2, 1, 2, 0, 2, 2, 0, 1, 1, 0};
    static const int faxis[6] = {2, 1, 0, 0, 1, 2};
    six.assign((size_t)np * 6, 0.0);
    for (int p = 0; p < np; ++p) {
        const tf_probe_geom& g = pg[p];
        const double Lx = g.L[0], Ly = g.L[1], Lz = g.L[2];
        const double volume = Lx * Ly * Lz;
        const double da[3] = {Ly * Lz / 4, Lx * Lz / 4, Lx * Ly / 4};
        double e[3] = {0, 0, 0};
        for (int i = 0; i < 12; ++i) {
            const double u = U[g.edges[i]];
            for (int c = 0; c < 3; ++c) e[c] = e[c] + (c == axis[i] ? u * da[c] : u * 0.0);
        }
        for (int c = 0; c < 3; ++c) six[p * 6 + c] = e[c] / volume;
        const double inc[3] = {0.5 * Lx, 0.5 * Ly, 0.5 * Lz};
        double h[3] = {0, 0, 0};
        for (int i = 0; i < 6; ++i) {
            const double u = B[g.faces[i]];
            for (int c = 0; c < 3; ++c) h[c] = h[c] + (c == faxis[i] ? u * inc[c] : u * 0.0);
        }
        double mi[9];
        inv3(g.mu, mi);
        for (int c = 0; c < 3; ++c)
            six[p * 6 + 3 + c] = (mi[c * 3] * h[0] + mi[c * 3 + 1] * h[1] + mi[c * 3 + 2] * h[2]) / volume;
    }
}

// Inverse of a 4x4 matrix by Gauss-Jordan with partial pivoting.
void inv4(const double a[16], double r[16]) {
    double m[4][8];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 8; ++j) m[i][j] = j < 4 ? a[i * 4 + j] : (j - 4 == i ? 1.0 : 0.0);
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        for (int i = c + 1; i < 4; ++i)
            if (std::fabs(m[i][c]) > std::fabs(m[piv][c])) piv = i;
        for (int j = 0; j < 8; ++j) std::swap(m[c][j], m[piv][j]);
        const double d = m[c][c];
        for (int j = 0; j < 8; ++j) m[c][j] /= d;
        for (int i = 0; i < 4; ++i) {
            if (i == c) continue;
            const double f = m[i][c];
            for (int j = 0; j < 8; ++j) m[i][j] -= f * m[c][j];
        }
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) r[i * 4 + j] = m[i][j + 4];
}

// GetWhitneyElectricField D:3768-3875 (nodal functions as the inverse of the
// [x y z 1] node matrix) and GetMagneticField D:3877-3897.
void export_tet(const std::vector<double>& U, const std::vector<double>& B, const tf_probe_geom* pg, int32_t np,
                std::vector<double>& six) {
    six.assign((size_t)np * 6, 0.0);
    static const int le[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};
    for (int p = 0; p < np; ++p) {
        const tf_probe_geom& g = pg[p];
        double node[16];
        for (int i = 0; i < 4; ++i) {
            for (int c = 0; c < 3; ++c) node[i * 4 + c] = g.nodes[i][c];
            node[i * 4 + 3] = 1;
        }
        double wn[16];  // wn = node^-1: row r of [p 1] * wn gives the nodal values
        inv4(node, wn);
        double w[4];
        for (int i = 0; i < 4; ++i)
            w[i] = g.point[0] * wn[0 * 4 + i] + g.point[1] * wn[1 * 4 + i] + g.point[2] * wn[2 * 4 + i] + wn[3 * 4 + i];
        double e[3] = {0, 0, 0};
        for (int k = 0; k < 6; ++k) {
            const int i = le[k][0], j = le[k][1];
            const double u = U[g.edges[k]];
            for (int c = 0; c < 3; ++c) e[c] += (w[i] * wn[c * 4 + j] - w[j] * wn[c * 4 + i]) * u;
        }
        for (int c = 0; c < 3; ++c) six[p * 6 + c] = e[c];
        double pev[6][3];
        for (int k = 0; k < 6; ++k)
            for (int c = 0; c < 3; ++c) pev[k][c] = g.nodes[le[k][1]][c] - g.nodes[le[k][0]][c];
        double fq[4];
        for (int i = 0; i < 4; ++i) fq[i] = B[g.faces[i]];
        const double v = 3 * g.volume;
        double ret[3] = {0, 0, 0};
        for (int c = 0; c < 3; ++c) {
            ret[c] += (pev[0][c] * fq[1] - pev[2][c] * fq[2] + pev[3][c] * fq[3]) / v;
            ret[c] += (pev[0][c] * fq[0] - pev[1][c] * fq[2] + pev[4][c] * fq[3]) / v;
            ret[c] += (-pev[1][c] * fq[1] + pev[2][c] * fq[0] + pev[5][c] * fq[3]) / v;
            ret[c] += (pev[3][c] * fq[0] - pev[4][c] * fq[1] + pev[5][c] * fq[2]) / v;
        }
        double mi[9];
        inv3(g.mu, mi);
        for (int c = 0; c < 3; ++c)
            six[p * 6 + 3 + c] = 0.25 * (mi[c * 3] * ret[0] + mi[c * 3 + 1] * ret[1] + mi[c * 3 + 2] * ret[2]);
    }
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------- fdtd
// D:1840-1933
void run_fdtd(const tf_problem* p, const SourceTable& st, const tf_probe_geom* pg, int32_t np, Printer pr,
              int64_t first, int64_t n, bool faithful, const double* U0, const double* F0, const double* B0,
              oracle_out* out) {
    const tf_fit_ops& f = p->fit;
    const int32_t E = p->n_edges, Fn = p->n_faces;
    const double dt = p->dt;
    std::vector<double> U(E, 0.0), F(Fn, 0.0), B(Fn, 0.0), I(E, 0.0), U_old, F_old, curl_fb(E, 0.0);
    if (U0) U.assign(U0, U0 + E);
    if (F0) F.assign(F0, F0 + Fn);
    if (B0) B.assign(B0, B0 + Fn);
    // Ct_vec as the reference keeps it: vector of vector of signed indices
    struct Sgn { uint32_t val; int32_t sgn; };
    std::vector<std::vector<Sgn>> Ct_vec;
    if (faithful) {
        Ct_vec.resize(E);
        for (int32_t e = 0; e < E; ++e)
            for (int64_t q = f.ct_ptr[e]; q < f.ct_ptr[e + 1]; ++q)
                Ct_vec[e].push_back(Sgn{(uint32_t)f.ct_face[q], f.ct_sgn[q]});
    }
    Sink sink{out, np, p};
    std::vector<double> six;
    double loop_s = 0;
    for (int64_t i = first; i < first + n; ++i) {
        const double t0 = now_s();
        const double t = (double)i * dt;
        U_old = U;
        for (int64_t q = 0; q < f.n_bc; ++q) U[f.bc_edges[q]] = t < 0 ? 0.0 : f.bc_value[q];  // ComputeEdgeBC
        for (int64_t q = 0; q < f.n_src; ++q) {
            const int32_t j = f.src_edges[q];
            const int64_t k = st.by_edge[j];
            if (k < 0) continue;
            if (st.se[k].field == TF_FIELD_J) I[j] = j_source(st, k, t);
            else U[j] = e_source(st, k, t);
        }
        for (int64_t q = 0; q < f.n_common; ++q) {
            const int32_t j = f.common_edges[q];
            if (faithful) {
                auto abs_ct_vec = std::vector<uint32_t>(
                    {Ct_vec[j][0].val, Ct_vec[j][1].val, Ct_vec[j][2].val, Ct_vec[j][3].val});
                U[j] = std::pow(f.upml[j], 2) * f.M_h[j] *
                       (f.M_q[j] * U_old[j] +
                        f.upml[j] * dt *
                            (f.dual_curl[j] * (F[abs_ct_vec[0]] - F[abs_ct_vec[1]] + F[abs_ct_vec[2]] - F[abs_ct_vec[3]]) -
                             I[j]));
            } else {
                const int32_t* c = f.ct_face + f.ct_ptr[j];
                U[j] = std::pow(f.upml[j], 2) * f.M_h[j] *
                       (f.M_q[j] * U_old[j] +
                        f.upml[j] * dt * (f.dual_curl[j] * (F[c[0]] - F[c[1]] + F[c[2]] - F[c[3]]) - I[j]));
            }
        }
        for (int64_t q = 0; q < f.n_uncommon; ++q) {
            const int32_t j = f.uncommon_edges[q];
            const int64_t k = st.by_edge[j];
            curl_fb[j] = k >= 0 ? h_source(st, k, t - 0.5 * dt) : 0.0;
            double val[3] = {0, 0, 0}, sg[3] = {0, 0, 0};
            int m = 0;
            for (int64_t r = f.ct_ptr[j]; r < f.ct_ptr[j + 1] && m < 3; ++r, ++m) {
                val[m] = F[f.ct_face[r]];
                sg[m] = f.ct_sgn[r] < 0 ? -1 : 1;
            }
            const double d = val[0] * sg[0] + val[1] * sg[1] + val[2] * sg[2];
            U[j] = f.M_h[j] * (f.M_q[j] * U_old[j] + dt * (d + curl_fb[j] - I[j]));
        }
        double poy = 0;
        for (int32_t e = 0; e < E; ++e) poy += (U_old[e] + U[e]) * curl_fb[e];
        sink.poy(0.5 * poy);
        F_old = F;
        for (int64_t q = 0; q < f.n_tbc; ++q) {
            const int32_t j = f.tbc_faces[q];
            const int32_t* c = f.c_edge + 4 * (int64_t)j;
            B[j] = f.M_mu[j] * (f.fpml[j] * F_old[j]) - dt * f.curl[j] * (U[c[0]] - U[c[1]] + U[c[2]] - U[c[3]]);
            F[j] = f.fpml[j] * f.M_nu[j] * B[j];
        }
        loop_s += now_s() - t0;
        if (pr.allow(t)) {
            export_fit(U, B, pg, np, six);
            sink.put(t, six, U, B);
        }
    }
    out->seconds = loop_s;
    if (out->U) std::memcpy(out->U, U.data(), sizeof(double) * E);
    if (out->F) std::memcpy(out->F, F.data(), sizeof(double) * Fn);
    if (out->B) std::memcpy(out->B, B.data(), sizeof(double) * Fn);
}

// y = M x over CSR rows, sequential in increasing column order
void csr_mv(const tf_csr& M, const std::vector<double>& x, std::vector<double>& y, double scale_vals = 1.0) {
    y.assign(M.nrows, 0.0);
    for (int32_t r = 0; r < M.nrows; ++r) {
        double s = 0;
        for (int64_t q = M.rowptr[r]; q < M.rowptr[r + 1]; ++q) {
            const double v = M.val ? M.val[q] : (double)M.sgn[q];
            s += (scale_vals == 1.0 ? v : scale_vals * v) * x[M.col[q]];
        }
        y[r] = s;
    }
}

// ---------------------------------------------------------------- dga
// D:538-939 for sigma = chi = 0 (free edges class 1, faces class 1)
void run_dga(const tf_problem* p, const SourceTable& st, const tf_probe_geom* pg, int32_t np, Printer pr,
             int64_t first, int64_t n, const double* U0, const double* F0, const double* B0, oracle_out* out) {
    const tf_dga_ops& d = p->dga;
    const int32_t E = p->n_edges, Fn = p->n_faces;
    const double dt = p->dt;
    std::vector<double> U(E, 0.0), F(Fn, 0.0), B(Fn, 0.0), I(E, 0.0);
    std::vector<double> U_a(d.H.nrows, 0.0), U_d(E, 0.0);  // U_d indexed by edge for class-4
    if (U0) U.assign(U0, U0 + E);
    if (F0) F.assign(F0, F0 + Fn);
    if (B0) B.assign(B0, B0 + Fn);
    for (int32_t e = 0; e < E; ++e) {
        if (d.h_row[e] >= 0) U_a[d.h_row[e]] = U[e];
        if (d.edge_class[e] == 4) U_d[e] = d.fixed_value[e];  // ComputeEdgeBC(ee,0)
    }
    // initial e-sources at t = 0 (D:762-790)
    // before the loop every source edge, h edges included, takes
    // ComputeEfieldSource(ee, 0), the sum of its "e" sources (D:746-765)
    for (int64_t k = 0; k < (int64_t)st.se.size(); ++k) {
        const int32_t e = st.se[k].edge;
        if (st.se[k].field == TF_FIELD_J) continue;  // rejected for dga (A-7)
        if (d.edge_class[e] == 4) U_d[e] = e_source(st, k, 0.0);
        else if (d.h_row[e] >= 0) U_a[d.h_row[e]] = e_source(st, k, 0.0);
    }
    auto assemble_U = [&]() {  // U = M*U_frac
        for (int32_t e = 0; e < E; ++e) {
            if (d.h_row[e] >= 0) U[e] = U_a[d.h_row[e]];
            else if (d.edge_class[e] == 4) U[e] = U_d[e];
        }
    };
    assemble_U();
    Sink sink{out, np, p};
    std::vector<double> six, U_old, curl_b(E, 0.0), cf, hc, cu, nc;
    if (first == 1 && pr.allow(0.0)) {  // initial export at t = 0 (D:805-806, A-6)
        export_tet(U, B, pg, np, six);
        sink.put(0.0, six, U, B);
    }
    double loop_s = 0;
    for (int64_t i = first; i < first + n; ++i) {
        const double t0 = now_s();
        const double t = (double)i * dt;
        std::fill(curl_b.begin(), curl_b.end(), 0.0);  // Fb = 0
        U_old = U;
        for (int32_t e = 0; e < E; ++e)
            if (d.edge_class[e] == 4 && st.by_edge[e] < 0) U_d[e] = t < 0 ? 0.0 : d.fixed_value[e];
        for (int64_t k = 0; k < (int64_t)st.se.size(); ++k) {
            const tf_source_edge& s = st.se[k];
            if (s.field == TF_FIELD_H) curl_b[s.edge] = h_source(st, k, t - 0.5 * dt);  // Ctb*Fb
            else if (s.field == TF_FIELD_E && d.edge_class[s.edge] == 4) U_d[s.edge] = e_source(st, k, t);
            else if (s.field != TF_FIELD_J && d.h_row[s.edge] >= 0)
                U_a[d.h_row[s.edge]] = e_source(st, k, t);  // a d edge: U_a = 0 before the H update (D:835-848)
        }
        // curl_f = C^T F + curl_b - I
        csr_mv(d.Ct, F, cf);
        for (int32_t e = 0; e < E; ++e) cf[e] = cf[e] + curl_b[e] - I[e];
        // U_a = U_a + dt*(H*curl_f)
        csr_mv(d.H, cf, hc);
        for (int32_t r = 0; r < d.H.nrows; ++r) U_a[r] = U_a[r] + dt * hc[r];
        assemble_U();
        double poy = 0;
        for (int32_t e = 0; e < E; ++e) poy += (U_old[e] + U[e]) * curl_b[e];
        sink.poy(0.5 * poy);
        // curl_u = C U ; B -= dt curl_u ; F_a -= (dt N) curl_u
        csr_mv(d.C, U, cu);
        for (int32_t f = 0; f < Fn; ++f) B[f] = B[f] - dt * cu[f];
        csr_mv(d.N, cu, nc, dt);
        for (int32_t f = 0; f < Fn; ++f) F[f] = F[f] - nc[f];
        loop_s += now_s() - t0;
        if (pr.allow(t)) {
            export_tet(U, B, pg, np, six);
            sink.put(t, six, U, B);
        }
    }
    out->seconds = loop_s;
    if (out->U) std::memcpy(out->U, U.data(), sizeof(double) * E);
    if (out->F) std::memcpy(out->F, F.data(), sizeof(double) * Fn);
    if (out->B) std::memcpy(out->B, B.data(), sizeof(double) * Fn);
}

// ---------------------------------------------------------------- dga, lossy
// D:538-939 with edge classes 2/3 and face classes 2/3: the fractional
// unknowns U_frac = [U_a | U_b | U_c | U_d], F_frac = [F_a | F_b | F_c]
// (the stacked operators of tf_dga_frac: Ku = [H; Mp; Mq], Ru = [I; diag
// P_p; Q; I], Kf = [N; Tr; Ts], Rf = [I; diag R_r; S]). Deviation: the
// reference maps Old_F_c onto the old-U buffer (D:586), reading the wrong
// (and possibly out-of-range) entries; F_c here reads the old F_c.
void run_dga_frac(const tf_problem* p, const SourceTable& st, const tf_probe_geom* pg, int32_t np, Printer pr,
                  int64_t first, int64_t n, const double* U0, const double* F0, const double* B0, oracle_out* out) {
    const tf_dga_ops& d = p->dga;
    const tf_dga_frac& fr = d.frac;
    const int32_t E = p->n_edges, Fn = p->n_faces;
    const int32_t h = fr.h_size, pp = fr.p_size, q = fr.q_size;
    const int32_t nU = fr.Ku.nrows, nF = fr.Kf.nrows;
    const int32_t nn = fr.n_size, nr = fr.r_size;
    const double dt = p->dt;
    std::vector<double> U(E, 0.0), F(Fn, 0.0), B(Fn, 0.0), I(E, 0.0), Uf(nU, 0.0), Ff(nF, 0.0);
    // a written state: each edge's (face's) value split evenly over its
    // fractional entries
    if (U0) {
        U.assign(U0, U0 + E);
        for (int32_t e = 0; e < E; ++e) {
            const int64_t a = fr.Mu.rowptr[e], c = fr.Mu.rowptr[e + 1];
            for (int64_t k = a; k < c; ++k) Uf[fr.Mu.col[k]] = U[e] / (double)(c - a);
        }
    }
    if (F0) {
        F.assign(F0, F0 + Fn);
        for (int32_t f = 0; f < Fn; ++f) {
            const int64_t a = fr.Tf.rowptr[f], c = fr.Tf.rowptr[f + 1];
            for (int64_t k = a; k < c; ++k) Ff[fr.Tf.col[k]] = F[f] / (double)(c - a);
        }
    }
    if (B0) B.assign(B0, B0 + Fn);
    // before the loop (D:726-790): boundary edges take ComputeEdgeBC(ee,0),
    // every source edge ComputeEfieldSource(ee,0) in each of its fractional
    // entries (half of it in U_c)
    for (int32_t e = 0; e < E; ++e)
        if (d.edge_class[e] == 4 && st.by_edge[e] < 0) Uf[fr.slot[e]] = d.fixed_value[e];
    auto put_all = [&](int32_t e, double v) {
        for (int64_t k = fr.Mu.rowptr[e]; k < fr.Mu.rowptr[e + 1]; ++k) {
            const int32_t c = fr.Mu.col[k];
            Uf[c] = (c >= h + pp && c < h + pp + q) ? 0.5 * v : v;
        }
    };
    for (int64_t k = 0; k < (int64_t)st.se.size(); ++k) {
        if (st.se[k].field == TF_FIELD_J) continue;  // rejected for dga (A-7)
        put_all(st.se[k].edge, e_source(st, k, 0.0));
    }
    auto mv_add = [](const tf_csr& M, const std::vector<double>& x, int32_t r) {
        double s = 0;
        for (int64_t k = M.rowptr[r]; k < M.rowptr[r + 1]; ++k) s += M.val[k] * x[M.col[k]];
        return s;
    };
    auto gather = [](const tf_csr& M, const std::vector<double>& x, std::vector<double>& y) {
        for (int32_t r = 0; r < M.nrows; ++r) {
            double s = 0;
            for (int64_t k = M.rowptr[r]; k < M.rowptr[r + 1]; ++k) s += x[M.col[k]];
            y[r] = s;
        }
    };
    gather(fr.Mu, Uf, U);  // U = M*U_frac
    Sink sink{out, np, p};
    std::vector<double> six, U_old, Uf_old, Ff_old, curl_b(E, 0.0), cf, kc, cu, kf;
    if (first == 1 && pr.allow(0.0)) {
        export_tet(U, B, pg, np, six);
        sink.put(0.0, six, U, B);
    }
    double loop_s = 0;
    for (int64_t i = first; i < first + n; ++i) {
        const double t0 = now_s();
        const double t = (double)i * dt;
        std::fill(curl_b.begin(), curl_b.end(), 0.0);
        U_old = U;
        Uf_old = Uf;
        for (int32_t e = 0; e < E; ++e)
            if (d.edge_class[e] == 4 && st.by_edge[e] < 0) Uf[fr.slot[e]] = t < 0 ? 0.0 : d.fixed_value[e];
        for (int64_t k = 0; k < (int64_t)st.se.size(); ++k) {
            const tf_source_edge& s = st.se[k];
            if (s.field == TF_FIELD_H) curl_b[s.edge] = h_source(st, k, t - 0.5 * dt);
            else if (s.field != TF_FIELD_J) put_all(s.edge, e_source(st, k, t));
        }
        csr_mv(d.Ct, F, cf);
        for (int32_t e = 0; e < E; ++e) cf[e] = cf[e] + curl_b[e] - I[e];
        csr_mv(fr.Ku, cf, kc);
        // U_a = U_a + dt*(H curl_f); U_b = P_p .* Old_U_b + dt*(Mp curl_f);
        // U_c = Q Old_U_c + dt*(Mq curl_f)  (D:888-890)
        for (int32_t r = 0; r < h; ++r) Uf[r] = Uf[r] + dt * kc[r];
        for (int32_t r = h; r < h + pp + q; ++r) Uf[r] = mv_add(fr.Ru, Uf_old, r) + dt * kc[r];
        gather(fr.Mu, Uf, U);
        double poy = 0;
        for (int32_t e = 0; e < E; ++e) poy += (U_old[e] + U[e]) * curl_b[e];
        sink.poy(0.5 * poy);
        Ff_old = Ff;
        csr_mv(d.C, U, cu);
        for (int32_t f = 0; f < Fn; ++f) B[f] = B[f] - dt * cu[f];
        // F_a -= (dt N) curl_u; F_b = R_r .* Old_F_b - (dt Tr) curl_u;
        // F_c = S Old_F_c - (dt Ts) curl_u  (D:916-919)
        csr_mv(fr.Kf, cu, kf, dt);
        for (int32_t r = 0; r < nn; ++r) Ff[r] = Ff[r] - kf[r];
        for (int32_t r = nn; r < nF; ++r) Ff[r] = mv_add(fr.Rf, Ff_old, r) - kf[r];
        (void)nr;
        gather(fr.Tf, Ff, F);
        loop_s += now_s() - t0;
        if (pr.allow(t)) {
            export_tet(U, B, pg, np, six);
            sink.put(t, six, U, B);
        }
    }
    out->seconds = loop_s;
    if (out->U) std::memcpy(out->U, U.data(), sizeof(double) * E);
    if (out->F) std::memcpy(out->F, F.data(), sizeof(double) * Fn);
    if (out->B) std::memcpy(out->B, B.data(), sizeof(double) * Fn);
}

// ConjugateGradientSolver, ConjugateGradientSolver.hpp:52-128
struct Cg {
    int32_t n;
    const int64_t* ptr;
    const int32_t* col;
    const double* val;
    std::vector<double> pre;
    double tol;
    int32_t maxiter;
    int32_t iter = 0;
    double final_error = 0;
    void compute() {
        pre.assign(n, 0.0);
        for (int32_t j = 0; j < n; ++j) {
            double dj = 0;
            bool found = false;
            for (int64_t q = ptr[j]; q < ptr[j + 1]; ++q)
                if (col[q] == j) { dj = val[q]; found = true; break; }
            pre[j] = (found && dj != 0) ? 1.0 / dj : 1.0;
        }
    }
    void Ax(const std::vector<double>& x, std::vector<double>& y) const {
        for (int32_t i = 0; i < n; ++i) {
            double s = 0;
            for (int64_t q = ptr[i]; q < ptr[i + 1]; ++q) s += val[q] * x[col[q]];
            y[i] = s;
        }
    }
    static double dotp(const std::vector<double>& a, const std::vector<double>& b) {
        double s = 0;
        for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
        return s;
    }
    std::vector<double> solve(const std::vector<double>& b, const std::vector<double>& x0) {
        double rhoA = 0, rho_p;
        iter = 0;
        std::vector<double> r(n), z(n), p(n), q(n), x = x0, ax(n);
        Ax(x, ax);
        for (int32_t i = 0; i < n; ++i) r[i] = b[i] - ax[i];
        const double residual0 = std::sqrt(dotp(r, r));
        double residual = residual0;
        while (residual / residual0 > tol && iter < maxiter) {
            for (int32_t i = 0; i < n; ++i) z[i] = pre[i] * r[i];
            rho_p = rhoA;
            rhoA = dotp(r, z);
            if (iter == 0) p = z;
            else {
                const double beta = rhoA / rho_p;
                for (int32_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
            }
            Ax(p, q);
            const double a = rhoA / dotp(p, q);
            for (int32_t i = 0; i < n; ++i) { x[i] = x[i] + a * p[i]; r[i] = r[i] - a * q[i]; }
            residual = std::sqrt(dotp(r, r));
            iter++;
        }
        final_error = residual / residual0;
        return x;
    }
};

// ---------------------------------------------------------------- gfdtd
// D:1360-1565
void run_gfdtd(const tf_problem* p, const SourceTable& st, const tf_probe_geom* pg, int32_t np, Printer pr,
               int64_t first, int64_t n, double tol, int32_t maxiter, const double* U0, const double* F0,
               const double* B0, oracle_out* out) {
    const tf_dga_ops& d = p->dga;
    const int32_t E = p->n_edges, Fn = p->n_faces, nf = d.n_free;
    const double dt = p->dt;
    std::vector<double> U(E, 0.0), F(Fn, 0.0), B(Fn, 0.0), I(E, 0.0), SrcFld(E, 0.0), Fb(E, 0.0), Fb_old;
    if (U0) U.assign(U0, U0 + E);
    if (F0) F.assign(F0, F0 + Fn);
    if (B0) B.assign(B0, B0 + Fn);
    std::vector<uint8_t> is_free(E, 0);
    for (int32_t k = 0; k < nf; ++k) is_free[d.free_edges[k]] = 1;
    // Dirichlet edges: class-4 values (ComputeEdgeBC), e sources
    auto bc_pass = [&](double t, bool h_too) {
        for (int32_t e = 0; e < E; ++e)
            if (!is_free[e] && st.by_edge[e] < 0) { U[e] = t < 0 ? 0.0 : d.fixed_value[e]; SrcFld[e] = U[e]; }
        for (int64_t k = 0; k < (int64_t)st.se.size(); ++k) {
            const tf_source_edge& s = st.se[k];
            if (s.field == TF_FIELD_H) {
                if (h_too) Fb[s.edge] = h_source(st, k, t - 0.5 * dt);
            } else if (s.field == TF_FIELD_J) {
                if (h_too) I[s.edge] = j_source(st, k, t);  // D:1491-1492
                SrcFld[s.edge] = U[s.edge];
            } else {  // e, and d edges (whose "e" sum is 0; free, so the solve overwrites U)
                U[s.edge] = e_source(st, k, t);
                SrcFld[s.edge] = U[s.edge];
            }
        }
    };
    bc_pass(0.0, false);  // D:1431-1453
    std::vector<double> U_old = U;
    Cg cg{nf, d.A.rowptr, d.A.col, d.A.val, {}, tol, maxiter};
    cg.compute();
    std::vector<double> solution(nf, 0.0), rhs(nf, 0.0), cu, eu, r1, ctf, su;
    for (int32_t k = 0; k < nf; ++k) solution[k] = U[d.free_edges[k]];
    Sink sink{out, np, p};
    std::vector<double> six, sx;
    // U.(SigMat U) and the Losses vector (zero when every tet is lossless)
    auto quad_sig = [&](const std::vector<double>& x) {
        if (d.Sig.nnz == 0) return 0.0;
        csr_mv(d.Sig, x, sx);
        double q = 0;
        for (int32_t e = 0; e < E; ++e) q += x[e] * sx[e];
        return q;
    };
    auto loss = [&](double v) {
        if (out->losses && out->n_losses < out->cap_losses) out->losses[out->n_losses] = v;
        out->n_losses++;
    };
    double loop_s = 0;
    for (int64_t i = first; i < first + n; ++i) {
        double t0 = now_s();
        const double t = (double)i * dt;
        csr_mv(d.C, U, cu);
        for (int32_t f = 0; f < Fn; ++f) B[f] -= dt * cu[f];
        csr_mv(d.N, B, F);
        loop_s += now_s() - t0;
        if (pr.allow(t - dt)) {
            export_tet(U, B, pg, np, six);
            sink.put(t - dt, six, U, B);
        }
        t0 = now_s();
        Fb_old = Fb;
        bc_pass(t, true);
        // Psi = E U_old - dt/2 SigMat U_old - RHSmat1 SrcFld - dt I + dt C^T F + dt Ctb Fb  (D:1516)
        csr_mv(d.E, U_old, eu);
        if (d.Sig.nnz > 0) {
            csr_mv(d.Sig, U_old, su);
            for (int32_t e = 0; e < E; ++e) eu[e] -= (0.5 * dt) * su[e];  // - (dt/2) SigMat U_old
        }
        csr_mv(d.RHS1, SrcFld, r1);
        csr_mv(d.Ct, F, ctf);
        double poy = 0;
        for (int32_t e = 0; e < E; ++e) poy += U_old[e] * (0.5 * (Fb[e] + Fb_old[e]));
        sink.poy(poy);
        loss(0.5 * quad_sig(U_old));  // D:1519
        for (int32_t k = 0; k < nf; ++k) {
            const int32_t e = d.free_edges[k];
            rhs[k] = ((eu[e] - r1[e]) - dt * I[e]) + dt * ctf[e] + dt * Fb[e];
        }
        solution = cg.solve(rhs, solution);
        out->cg_iters_total += cg.iter;
        if (cg.iter > out->cg_iters_max) out->cg_iters_max = cg.iter;
        if (std::isfinite(cg.final_error) && cg.final_error > out->cg_max_relres) out->cg_max_relres = cg.final_error;
        if (cg.iter >= maxiter && cg.final_error > tol) out->cg_not_converged++;
        for (int32_t k = 0; k < nf; ++k) U[d.free_edges[k]] = solution[k];
        if (i > 1) loss(quad_sig(U));  // Joule_L, D:1543-1545
        U_old = U;
        loop_s += now_s() - t0;
    }
    if (first + n > 1) loss(quad_sig(U));  // after the loop, D:1566-1570
    out->seconds = loop_s;
    if (out->U) std::memcpy(out->U, U.data(), sizeof(double) * E);
    if (out->F) std::memcpy(out->F, F.data(), sizeof(double) * Fn);
    if (out->B) std::memcpy(out->B, B.data(), sizeof(double) * Fn);
}

// ---------------------------------------------------------------- fem
// D:1580-1790: Whitney edge elements, second order in time. Per step i
// (t = i dt): curl_u = C U, B -= dt curl_u, export at t - dt; sources at t;
//   Psi = 2/dt^2 E U_old - Ct N C U_old + 1/dt Ctb (Fb - Fb_old)
//         - 1/dt^2 E U_older + 0.5/dt SigMat U_older - RHSmat1 SrcFld - I/dt,
// A solution = Psi on the free edges (warm start: the previous solution,
// zero before the first solve), U_older = U_old, U_old = U.
void run_fem(const tf_problem* p, const SourceTable& st, const tf_probe_geom* pg, int32_t np, Printer pr,
             int64_t first, int64_t n, double tol, int32_t maxiter, const double* U0, const double* B0,
             oracle_out* out) {
    const tf_dga_ops& d = p->dga;
    const int32_t E = p->n_edges, Fn = p->n_faces, nf = d.n_free;
    const double dt = p->dt;
    std::vector<double> U(E, 0.0), B(Fn, 0.0), I(E, 0.0), SrcFld(E, 0.0), Fb(E, 0.0), Fb_old(E, 0.0);
    if (U0) U.assign(U0, U0 + E);
    if (B0) B.assign(B0, B0 + Fn);
    std::vector<uint8_t> is_free(E, 0);
    for (int32_t k = 0; k < nf; ++k) is_free[d.free_edges[k]] = 1;
    std::vector<double> U_old = U, U_older = U_old;
    // the pass before the loop (D:1660-1680): no h sources
    for (int32_t e = 0; e < E; ++e)
        if (!is_free[e] && st.by_edge[e] < 0) { U[e] = d.fixed_value[e]; SrcFld[e] = U[e]; }
    auto src_pass = [&](double t, bool h_too) {
        for (int64_t k = 0; k < (int64_t)st.se.size(); ++k) {
            const tf_source_edge& s = st.se[k];
            if (s.field == TF_FIELD_H) {
                if (h_too) Fb[s.edge] = h_source(st, k, t - 0.5 * dt);
            } else if (s.field == TF_FIELD_J) {
                I[s.edge] = j_source(st, k, t);
                SrcFld[s.edge] = U[s.edge];
            } else {
                U[s.edge] = e_source(st, k, t);
                SrcFld[s.edge] = U[s.edge];
            }
        }
    };
    src_pass(0.0, false);
    U_older = U_old;
    U_old = U;
    Cg cg{nf, d.A.rowptr, d.A.col, d.A.val, {}, tol, maxiter};
    cg.compute();
    std::vector<double> solution(nf, 0.0), rhs(nf, 0.0), cu, nc, cc, eu, eo, so, r1, sx;
    Sink sink{out, np, p};
    std::vector<double> six;
    auto quad_sig = [&](const std::vector<double>& x) {
        if (d.Sig.nnz == 0) return 0.0;
        csr_mv(d.Sig, x, sx);
        double q = 0;
        for (int32_t e = 0; e < E; ++e) q += x[e] * sx[e];
        return q;
    };
    auto loss = [&](double v) {
        if (out->losses && out->n_losses < out->cap_losses) out->losses[out->n_losses] = v;
        out->n_losses++;
    };
    const double s2 = 1 / (dt * dt);
    double loop_s = 0;
    for (int64_t i = first; i < first + n; ++i) {
        double t0 = now_s();
        const double t = (double)i * dt;
        csr_mv(d.C, U, cu);
        for (int32_t f = 0; f < Fn; ++f) B[f] -= dt * cu[f];
        loop_s += now_s() - t0;
        if (pr.allow(t - dt)) {
            export_tet(U, B, pg, np, six);
            sink.put(t - dt, six, U, B);
        }
        t0 = now_s();
        Fb_old = Fb;
        for (int32_t e = 0; e < E; ++e)
            if (!is_free[e] && st.by_edge[e] < 0) { U[e] = t < 0 ? 0.0 : d.fixed_value[e]; SrcFld[e] = U[e]; }
        src_pass(t, true);
        csr_mv(d.C, U_old, cu);
        csr_mv(d.N, cu, nc);
        csr_mv(d.Ct, nc, cc);
        csr_mv(d.E, U_old, eu);
        csr_mv(d.E, U_older, eo);
        const bool sig = d.Sig.nnz > 0;
        if (sig) csr_mv(d.Sig, U_older, so);
        csr_mv(d.RHS1, SrcFld, r1);
        double poy = 0;
        for (int32_t e = 0; e < E; ++e) poy += U_old[e] * (0.5 * (Fb[e] + Fb_old[e]));
        sink.poy(poy);
        loss(0.5 * quad_sig(U_old));
        for (int32_t k = 0; k < nf; ++k) {
            const int32_t e = d.free_edges[k];
            double psi = 2.0 * eu[e] * s2 - cc[e];
            psi = psi + (1 / dt) * (Fb[e] - Fb_old[e]);
            psi = psi - eo[e] * s2;
            if (sig) psi = psi + (0.5 / dt) * so[e];
            psi = psi - r1[e];
            psi = psi - (1 / dt) * I[e];
            rhs[k] = psi;
        }
        solution = cg.solve(rhs, solution);
        out->cg_iters_total += cg.iter;
        if (cg.iter > out->cg_iters_max) out->cg_iters_max = cg.iter;
        if (std::isfinite(cg.final_error) && cg.final_error > out->cg_max_relres) out->cg_max_relres = cg.final_error;
        if (cg.iter >= maxiter && cg.final_error > tol) out->cg_not_converged++;
        for (int32_t k = 0; k < nf; ++k) U[d.free_edges[k]] = solution[k];
        if (i > 1) loss(quad_sig(U));
        U_older = U_old;
        U_old = U;
        loop_s += now_s() - t0;
    }
    if (first + n > 1) loss(quad_sig(U));
    out->seconds = loop_s;
    if (out->U) std::memcpy(out->U, U.data(), sizeof(double) * E);
    if (out->F) std::fill(out->F, out->F + Fn, 0.0);
    if (out->B) std::memcpy(out->B, B.data(), sizeof(double) * Fn);
}

}  // namespace

extern "C" {

const char* oracle_last_error(void) { return g_err.c_str(); }

void oracle_export(int32_t fdtd, const tf_probe_geom* pg, int32_t n_probes, int32_t n_edges, int32_t n_faces,
                   const double* U, const double* B, double* six) {
    std::vector<double> u(U, U + n_edges), b(B, B + n_faces), out;
    if (fdtd) export_fit(u, b, pg, n_probes, out);
    else export_tet(u, b, pg, n_probes, out);
    std::memcpy(six, out.data(), sizeof(double) * out.size());
}

void oracle_energy(const tf_problem* p, const double* U, const double* B, double* we, double* wh) {
    std::vector<double> u(U, U + p->n_edges), b(B, B + p->n_faces);
    energies(p, u, b, *we, *wh);
}

void oracle_source_compute(const tf_source_desc* s, const double p[3], double t, double out3[3]) {
    V3 v = source_compute(*s, p, t);
    out3[0] = v.x;
    out3[1] = v.y;
    out3[2] = v.z;
}

int32_t oracle_cg_solve(int32_t n, const int64_t* rowptr, const int32_t* col, const double* val, const double* b,
                        const double* x0, double tol, int32_t maxiter, double* x, double* err) {
    Cg cg{n, rowptr, col, val, {}, tol, maxiter};
    cg.compute();
    std::vector<double> bb(b, b + n), xx(x0, x0 + n);
    std::vector<double> r = cg.solve(bb, xx);
    std::memcpy(x, r.data(), sizeof(double) * n);
    if (err) *err = cg.final_error;
    return cg.iter;
}

int32_t oracle_run(const tf_problem* p, const tf_source_desc* sd, int32_t n_sd, const tf_source_edge* se,
                   int64_t n_se, const tf_probe_geom* pg, int32_t n_probes, int32_t out_enabled, double out_delay,
                   double out_period, int64_t first, int64_t n, int32_t faithful, double cg_tol, int32_t cg_maxiter,
                   const double* U0, const double* F0, const double* B0, oracle_out* out) {
    try {
        g_err.clear();
        out->n_samples = 0;
        out->n_poynting = 0;
        out->n_losses = 0;
        out->cg_iters_total = 0;
        out->cg_iters_max = 0;
        out->cg_not_converged = 0;
        out->cg_max_relres = 0;
        SourceTable st;
        st.sd.assign(sd, sd + n_sd);
        st.se.assign(se, se + n_se);
        st.by_edge.assign(p->n_edges, -1);
        for (int64_t k = 0; k < n_se; ++k) st.by_edge[se[k].edge] = k;
        Printer pr{out_enabled != 0, out_delay, out_period};
        if (p->method == TF_METHOD_FDTD)
            run_fdtd(p, st, pg, n_probes, pr, first, n, faithful != 0, U0, F0, B0, out);
        else if (p->method == TF_METHOD_DGA && p->dga.frac.enabled)
            run_dga_frac(p, st, pg, n_probes, pr, first, n, U0, F0, B0, out);
        else if (p->method == TF_METHOD_DGA)
            run_dga(p, st, pg, n_probes, pr, first, n, U0, F0, B0, out);
        else if (p->method == TF_METHOD_FEM)
            run_fem(p, st, pg, n_probes, pr, first, n, cg_tol, cg_maxiter, U0, B0, out);
        else
            run_gfdtd(p, st, pg, n_probes, pr, first, n, cg_tol, cg_maxiter, U0, F0, B0, out);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// Preconditioner MLP in fp64 (no reference counterpart: the reference has no
// learned component). Parameter layout W1[64][16], b1[64], W2[64][64], b2[64],
// w3[64], b3. y[i] = w3 . relu(W2 relu(W1 x_i + b1) + b2) + b3 and
// grad = sum_i gy[i] dy[i]/dparams, by hand-written backpropagation.
void oracle_mlp(const float* P, const float* X, const double* gy, int64_t n, double* y, double* grad) {
    const int I = 16, H = 64;
    const int oW1 = 0, oB1 = oW1 + H * I, oW2 = oB1 + H, oB2 = oW2 + H * H, oW3 = oB2 + H, oB3 = oW3 + H;
    if (grad)
        for (int k = 0; k <= oB3; ++k) grad[k] = 0;
    std::vector<double> z1(H), h1(H), z2(H), h2(H), g2(H), g1(H);
    for (int64_t i = 0; i < n; ++i) {
        const float* x = X + i * I;
        for (int a = 0; a < H; ++a) {
            double t = P[oB1 + a];
            for (int k = 0; k < I; ++k) t += (double)P[oW1 + a * I + k] * x[k];
            z1[a] = t;
            h1[a] = t > 0 ? t : 0;
        }
        double out = P[oB3];
        for (int a = 0; a < H; ++a) {
            double t = P[oB2 + a];
            for (int b = 0; b < H; ++b) t += (double)P[oW2 + a * H + b] * h1[b];
            z2[a] = t;
            h2[a] = t > 0 ? t : 0;
            out += (double)P[oW3 + a] * h2[a];
        }
        if (y) y[i] = out;
        if (!grad) continue;
        const double g = gy[i];
        grad[oB3] += g;
        for (int a = 0; a < H; ++a) {
            grad[oW3 + a] += g * h2[a];
            g2[a] = z2[a] > 0 ? g * P[oW3 + a] : 0;
            grad[oB2 + a] += g2[a];
            for (int b = 0; b < H; ++b) grad[oW2 + a * H + b] += g2[a] * h1[b];
        }
        for (int b = 0; b < H; ++b) {
            double t = 0;
            for (int a = 0; a < H; ++a) t += (double)P[oW2 + a * H + b] * g2[a];
            g1[b] = z1[b] > 0 ? t : 0;
            grad[oB1 + b] += g1[b];
            for (int k = 0; k < I; ++k) grad[oW1 + b * I + k] += g1[b] * x[k];
        }
    }
}

}  // extern "C"
