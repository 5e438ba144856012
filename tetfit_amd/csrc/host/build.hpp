// build.hpp - one simulation of a Model, meshed and assembled into the plain
// tables of tf_problem (plus sources, probes and checker-facing geometry).
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "model.hpp"
#include "tetfit_host.h"
#include "tetfit_problem.h"

namespace tfh {

// Wall time of the set-up phases, printed to stderr when TETFIT_HOST_TIMING
// is set (for finding where host set-up time goes).
struct PhaseClock {
    bool on;
    double t0, t;
    PhaseClock();
    void mark(const char* phase);
};

// Owning storage behind a tf_csr.
struct CsrStore {
    int32_t nrows = 0, ncols = 0;
    std::vector<int64_t> rowptr;
    std::vector<int32_t> col;
    std::vector<double> val;
    std::vector<int8_t> sgn;
    tf_csr view() const;
};

// (row, col, value) accumulation into a CSR with duplicates summed in
// insertion order (what Eigen's setFromTriplets with a summing functor does).
struct Triplets {
    std::vector<int32_t> r, c;
    std::vector<double> v;
    void add(int32_t i, int32_t j, double x) { r.push_back(i); c.push_back(j); v.push_back(x); }
    size_t size() const { return r.size(); }
    // overwrite=true keeps the last value instead of the sum (overwrite_to_sparse).
    CsrStore to_csr(int32_t nrows, int32_t ncols, bool overwrite = false) const;
};

// One (column, value) contribution to a row under assembly.
struct RowEntry {
    int32_t col;
    double val;
};

// A CSR matrix assembled row by row on all host threads. `row(r, out)`
// appends the contributions of row r in summation order; equal columns are
// summed in that order from +0.0, so the result is bitwise the one of a
// sequential loop that adds the same contributions in the same per-entry
// order (zero contributions still enter the pattern). Exceptions thrown by
// `row` are rethrown on the calling thread.
CsrStore assemble_rows(int32_t nrows, int32_t ncols,
                       const std::function<void(int32_t, std::vector<RowEntry>&)>& row);

struct Build {
    tf_problem prob{};
    std::string method;
    bool possibly_unstable = false;
    double stable_dt = 0;
    int64_t n_steps = 0, first_step = 0;
    double cg_tol = 1e-8;
    int32_t out_mode = 0;  // 0 none, 1 probepoint, 2 maxerror, 3 l2norm, 4 silo
    double out_delay = 0, out_period = 0;
    int32_t out_frame = 0;  // Output::ReferenceFrame: 0 cartesian, 1 cylindrical, 2 spherical
    std::string out_name;

    // ---- FIT tables
    std::vector<int64_t> ct_ptr;
    std::vector<int32_t> ct_face;
    std::vector<int8_t> ct_sgn;
    std::vector<int32_t> c_edge;
    std::vector<int8_t> curl, dual_curl;
    std::vector<double> M_h, M_q, upml, M_mu, M_nu, fpml, M_e, Nu_vec;
    std::vector<int32_t> common_edges, uncommon_edges, bc_edges, src_edges, tbc_faces;
    std::vector<double> bc_value;
    std::vector<int32_t> e_cluster, d_faces, g_nodes;  // 12/cube, 6/cube, 2/edge
    std::vector<int32_t> edge_sid, face_sid;           // grid positions (tf_fit_ops)
    int32_t cells[3] = {0, 0, 0};
    std::vector<double> pts;                            // 3/node
    std::vector<int32_t> boundary_face;
    int32_t n_cubes = 0;
    double L[3] = {0, 0, 0};

    // ---- DGA tables
    CsrStore C, Ct, H, N, E, A, RHS1, Sig;
    // lossy dga (tf_dga_frac): stacked fractional operators, the U_frac entry
    // each edge's boundary / source value goes to, and the block sizes
    // h, p, q, b, n, r, s
    CsrStore frac_Ku, frac_Ru, frac_Kf, frac_Rf, frac_M, frac_T;
    std::vector<int32_t> frac_slot;
    int32_t frac_sizes[7] = {0, 0, 0, 0, 0, 0, 0};
    std::vector<int32_t> bnd_edge, h_row, free_edges;
    std::vector<int32_t> edge_rank, face_rank;  // locality ranks (tf_dga_ops.edge_rank, face_rank)
    std::vector<int8_t> edge_class;
    std::vector<double> fixed_value;
    int32_t n_bnd = 0;
    std::vector<int32_t> vols, vtf, edge_nodes, face_nodes, vol_material;
    std::vector<double> cell_volumes;
    int32_t n_pts = 0;

    // ---- sources and probes
    std::vector<tf_source> sources;
    std::vector<tf_source_term> terms;
    std::vector<tf_source_desc> src_desc;
    std::vector<tf_source_edge> src_edge_desc;
    std::vector<tf_probe> probes;
    std::vector<double> probe_points;
    std::vector<tf_probe_geom> probe_geom;

    void finalize_views();
};

// Source.hpp:313-567 split into time and space parts.
struct SourceEval {
    // Time part: amp (x impedance for j) x carrier(t), evaluated in long double
    // like the reference (PI is long double there), 0 for t < 0.
    static double carrier(const SourceDef& s, double t);
    // Space part: decay(p) x profile(p) x direction vector; throws for the
    // Bessel/Legendre modes and curvilinear directions (GSL-backed).
    static Vec3 space(const SourceDef& s, const Vec3& p);
};

tf_source_desc describe_source(const SourceDef& s);
tf_source make_source(const SourceDef& s);

// Power-iteration estimate of lambda_max for a symmetric-positive operator
// given through `apply(x, y)` on vectors of size n (replaces Spectra's
// GenEigsSolver call in ComputeFDTDTimeStep / ComputeDGATimeStep).
// Dot product over fixed blocks of 32,768 entries summed in block order:
// the same result for any number of threads.
double par_dot(const double* a, const double* b, size_t n);

double power_lambda_max(size_t n,
                        const std::function<void(const std::vector<double>&, std::vector<double>&)>& apply,
                        int iters, const std::vector<double>* start = nullptr);

void build_fit(const Model& m, const SimulationDef& sim, Build& b);
void build_tet(const Model& m, const SimulationDef& sim, Build& b);

// Output::AllowPrint-style schedule parameters from the simulation's output.
void set_output(const Model& m, const SimulationDef& sim, Build& b);

// Uniform [0, 1) from (seed, index): one SplitMix64 finalisation of
// seed * golden + index (order-independent per-cell draws).
inline double cell_uniform(uint64_t seed, uint64_t index) {
    uint64_t z = seed * 0x9E3779B97F4A7C15ull + index + 0x632BE59BD9B4E019ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace tfh
