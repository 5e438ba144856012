// capi.cpp - extern "C" surface of libtetfit_host.so (declared in
// include/tetfit_host.h). No C++ exception crosses it.
#include <cstring>
#include <memory>
#include <string>

#include "build.hpp"

namespace tfh {
int32_t write_kuhn_mesh(const char* path, int32_t n, double jitter, uint64_t seed, int32_t block,
                        double* min_vol_ratio);
}

struct tf_model {
    tfh::Model m;
};
struct tf_build {
    tfh::Build b;
};

static thread_local std::string g_err;

template <class F>
static int32_t guarded(F f) {
    try {
        g_err.clear();
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    } catch (...) {
        g_err = "unknown error";
        return 1;
    }
}

extern "C" {

const char* tfh_last_error(void) { return g_err.c_str(); }

int32_t tf_model_load(const char* path, tf_model** out) {
    if (!path || !out) { g_err = "null argument"; return 1; }
    *out = nullptr;
    return guarded([&] {
        auto m = std::make_unique<tf_model>();
        m->m.load(path);
        *out = m.release();
    });
}

int32_t tf_model_n_simulations(const tf_model* m) { return m ? (int32_t)m->m.simulations.size() : 0; }

int32_t tf_model_simulation_label(const tf_model* m, int32_t i) {
    if (!m || i < 0 || i >= (int32_t)m->m.simulations.size()) return -1;
    auto it = m->m.simulations.begin();
    std::advance(it, i);
    return (int32_t)it->first;
}

void tf_model_free(tf_model* m) { delete m; }

int32_t tf_build_simulation(const tf_model* m, int32_t label, tf_build** out) {
    if (!m || !out) { g_err = "null argument"; return 1; }
    *out = nullptr;
    return guarded([&] {
        auto it = m->m.simulations.find((uint32_t)label);
        if (it == m->m.simulations.end()) throw tfh::Error("no simulation with that label");
        const tfh::SimulationDef& sim = it->second;
        auto mi = m->m.meshes.find(sim.mesh);
        if (mi == m->m.meshes.end()) throw tfh::Error("simulation refers to an undefined mesh");
        const std::string& type = mi->second.type;
        if (type == "tetrahedral" && sim.method == "fdtd")
            throw tfh::Error("Input file error at line " + std::to_string(m->m.method_line) +
                             ": If mesh type is tetrahedral, method must be fem or dga. If mesh type is cartesian, method must be fdtd");
        if (type == "cartesian" && sim.method != "fdtd")
            throw tfh::Error("Input file error at line " + std::to_string(m->m.method_line) +
                             ": If mesh type is tetrahedral, method must be fem or dga. If mesh type is cartesian, method must be fdtd");
        auto b = std::make_unique<tf_build>();
        b->b.method = sim.method;
        if (sim.method == "fdtd") tfh::build_fit(m->m, sim, b->b);
        else if (sim.method == "dga" || sim.method == "gfdtd" || sim.method == "fem") tfh::build_tet(m->m, sim, b->b);
        else throw tfh::Error("method '" + sim.method + "' is outside this build's hot path (fdtd, dga, gfdtd, fem)");
        if ((sim.method == "gfdtd" || sim.method == "fem") && sim.solver == "agmg")
            throw tfh::Error("solver agmg needs the Fortran AGMG library, not available; use solver cg");
        b->b.finalize_views();
        *out = b.release();
    });
}

void tf_build_free(tf_build* b) { delete b; }

const tf_problem* tf_build_problem(const tf_build* b) { return b ? &b->b.prob : nullptr; }
int32_t tf_build_possibly_unstable(const tf_build* b) { return b ? (int32_t)b->b.possibly_unstable : 0; }
double tf_build_stable_dt(const tf_build* b) { return b ? b->b.stable_dt : 0; }
int64_t tf_build_n_steps(const tf_build* b) { return b ? b->b.n_steps : 0; }
int64_t tf_build_first_step(const tf_build* b) { return b ? b->b.first_step : 0; }
double tf_build_cg_tol(const tf_build* b) { return b ? b->b.cg_tol : 0; }
int32_t tf_build_output_mode(const tf_build* b) { return b ? b->b.out_mode : 0; }
double tf_build_output_delay(const tf_build* b) { return b ? b->b.out_delay : 0; }
double tf_build_output_period(const tf_build* b) { return b ? b->b.out_period : 0; }
const char* tf_build_output_name(const tf_build* b) { return b ? b->b.out_name.c_str() : ""; }
int32_t tf_build_output_frame(const tf_build* b) { return b ? b->b.out_frame : 0; }

void tf_build_sources(const tf_build* b, const tf_source** src, int32_t* n_src, const tf_source_term** terms,
                      int64_t* n_terms) {
    *src = b->b.sources.data();
    *n_src = (int32_t)b->b.sources.size();
    *terms = b->b.terms.data();
    *n_terms = (int64_t)b->b.terms.size();
}

void tf_build_probes(const tf_build* b, const tf_probe** probes, const double** points, int32_t* n) {
    *probes = b->b.probes.data();
    *points = b->b.probe_points.data();
    *n = (int32_t)b->b.probes.size();
}

void tf_build_source_desc(const tf_build* b, const tf_source_desc** d, int32_t* n, const tf_source_edge** e,
                          int64_t* ne) {
    *d = b->b.src_desc.data();
    *n = (int32_t)b->b.src_desc.size();
    *e = b->b.src_edge_desc.data();
    *ne = (int64_t)b->b.src_edge_desc.size();
}

void tf_build_probe_geom(const tf_build* b, const tf_probe_geom** g, int32_t* n) {
    *g = b->b.probe_geom.data();
    *n = (int32_t)b->b.probe_geom.size();
}

void tf_build_fit_tables(const tf_build* b, const int32_t** e_cluster, const int32_t** d_faces, const int32_t** g_nodes,
                         const double** pts, int32_t* n_cubes, int32_t* n_pts, const int32_t** boundary_face,
                         double* L3) {
    const tfh::Build& x = b->b;
    *e_cluster = x.e_cluster.data();
    *d_faces = x.d_faces.data();
    *g_nodes = x.g_nodes.data();
    *pts = x.pts.data();
    *n_cubes = x.n_cubes;
    *n_pts = x.n_pts;
    *boundary_face = x.boundary_face.data();
    L3[0] = x.L[0]; L3[1] = x.L[1]; L3[2] = x.L[2];
}

void tf_build_tet_tables(const tf_build* b, const int32_t** vols, const double** pts, const int32_t** vtf,
                         const int32_t** edge_nodes, const int32_t** face_nodes, int32_t* n_vols, int32_t* n_pts,
                         const int32_t** boundary_face, const int32_t** vol_material) {
    const tfh::Build& x = b->b;
    *vols = x.vols.data();
    *pts = x.pts.data();
    *vtf = x.vtf.data();
    *edge_nodes = x.edge_nodes.data();
    *face_nodes = x.face_nodes.data();
    *n_vols = (int32_t)(x.vols.size() / 4);
    *n_pts = x.n_pts;
    *boundary_face = x.boundary_face.data();
    *vol_material = x.vol_material.data();
}

int32_t tf_write_kuhn_mesh(const char* path, int32_t n, double jitter, uint64_t seed, int32_t block,
                           double* min_vol_ratio) {
    int32_t r = -1;
    int32_t st = guarded([&] { r = tfh::write_kuhn_mesh(path, n, jitter, seed, block, min_vol_ratio); });
    if (st) return -2;
    if (r < 0) g_err = "kuhn mesh: a jittered tetrahedron is inverted";
    return r;
}

}  // extern "C"
