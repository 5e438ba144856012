/*
 * tetfit_host.h - C ABI of libtetfit_host.so: the host side that keeps
 * TetFIT's .fdtd input language, its mesh / material / source / output
 * primitives and the per-method operator producers, and hands an assembled
 * tf_problem to a stepping backend (libtetfit_hip.so, or the CPU oracle in
 * the tests).
 *
 * Reference counterparts (src/DGA/..., D: = Discretization.hpp):
 *   tf_model_load        Discretization(std::string) parser, D:45-211
 *   tf_build             RunSimulation set-up, D:289-800 (ReadMesh D:4204,
 *                        ReadStructuredMesh D:4347, ReadUnstructuredMesh D:5868,
 *                        ConstructCodecasaMaterialMatrices D:8694,
 *                        ConstructGFDTDMaterialMatrices D:8132, FindProbe
 *                        D:3645, FindFitProbe D:3590)
 *   tf_write_kuhn_mesh   no counterpart: writes synthetic netgen neutral
 *                        meshes in the format ReadUnstructuredMesh reads.
 */
#ifndef TETFIT_HOST_H
#define TETFIT_HOST_H

#include <stdint.h>
#include "tetfit_problem.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tf_model tf_model;   /* a parsed .fdtd input                 */
typedef struct tf_build tf_build;   /* one simulation, meshed and assembled */

const char* tfh_last_error(void);

/* Parses an .fdtd input; 0 on success. Mesh files named in it are resolved
 * against the current directory first, then the input file's directory. */
int32_t tf_model_load(const char* fdtd_path, tf_model** out);
int32_t tf_model_n_simulations(const tf_model* m);
int32_t tf_model_simulation_label(const tf_model* m, int32_t i);
void    tf_model_free(tf_model* m);

/* Meshes and assembles simulation `sim_label` of the model. dt policy:
 * `SET step` is used when it is not above the host's stability estimate,
 * otherwise the estimate times courant is used and the build is flagged
 * possibly-unstable (D:5476-5483, D:9146-9153). */
int32_t tf_build_simulation(const tf_model* m, int32_t sim_label, tf_build** out);
void    tf_build_free(tf_build* b);

const tf_problem* tf_build_problem(const tf_build* b);
int32_t tf_build_possibly_unstable(const tf_build* b);
double  tf_build_stable_dt(const tf_build* b);   /* host estimate            */
int64_t tf_build_n_steps(const tf_build* b);     /* loop trip count          */
int64_t tf_build_first_step(const tf_build* b);  /* 0 fdtd, 1 dga/gfdtd      */
double  tf_build_cg_tol(const tf_build* b);
int32_t tf_build_output_mode(const tf_build* b); /* 0 none,1 probepoint,...  */
double  tf_build_output_delay(const tf_build* b);
double  tf_build_output_period(const tf_build* b);
const char* tf_build_output_name(const tf_build* b);
/* Output::ReferenceFrame (Output.hpp:331): 0 cartesian, 1 cylindrical, 2 spherical. */
int32_t tf_build_output_frame(const tf_build* b);

void tf_build_sources(const tf_build* b, const tf_source** src, int32_t* n_src,
                      const tf_source_term** terms, int64_t* n_terms);
void tf_build_probes(const tf_build* b, const tf_probe** probes,
                     const double** points /* 3 per probe */, int32_t* n);

/* Reference-native description of the sources, for a checker that restates
 * Source::Compute (Source.hpp:313-567) and the Compute*Source functions. */
typedef struct {
    int32_t type;          /* tf_field_kind                                   */
    int32_t carrier;       /* tf_carrier_kind                                 */
    int32_t profile;       /* 0 dc/const, 1 wave, 2 gaussian                  */
    int32_t dir;           /* 0 x, 1 y, 2 z, 3 rho, 4 theta, 5 phi (spherical),*/
                           /* 6 r, 7 psi (cylindrical)                        */
    int32_t bfunc[3];      /* 0 none, 1 sin, 2 cos, 3 J_n, 4 Y_n, 5 j_l, 6 y_l, */
                           /* 7 P_l (Source.hpp:362-417)                      */
    double  amp, freq, delay, width, impedance, radius, decay;
    double  kvec[3], center[3];
    int32_t border[3];     /* order n / l of the Bessel / Legendre modes       */
} tf_source_desc;

typedef struct {
    int32_t edge;
    int32_t field;         /* tf_field_kind of the edge's first source        */
    int32_t n_src;         /* number of sources on the edge (<= 4)            */
    int32_t src[4];        /* indices into the tf_source_desc array           */
    double  bary[3];       /* edge barycentre                                 */
    double  vec[3];        /* h: boundary dual edge vector; e/j: edge vector  */
} tf_source_edge;

void tf_build_source_desc(const tf_build* b, const tf_source_desc** d,
                          int32_t* n, const tf_source_edge** e, int64_t* ne);

/* Reference-native geometry of each probe's element, for a checker that
 * restates GetFitElectricField / GetWhitneyElectricField & co. */
typedef struct {
    int32_t elem;          /* cube or tet id                                  */
    int32_t edges[12];     /* FIT: E_cluster; tet: local edge order (6)       */
    int32_t faces[6];      /* FIT: D; tet: vtf order (4)                      */
    double  point[3];
    double  nodes[4][3];   /* tet nodes (sorted ids)                          */
    double  L[3];          /* FIT cell steps                                  */
    double  mu[9];         /* element permeability tensor, row-major          */
    double  volume;        /* tet: signed CellVolumes                         */
} tf_probe_geom;

void tf_build_probe_geom(const tf_build* b, const tf_probe_geom** g, int32_t* n);

/* Cartesian (FIT) mesh tables for checkers: E_cluster (12/cube), D (6/cube),
 * edge end nodes G (2/edge), node coordinates, boundary_face labels. */
void tf_build_fit_tables(const tf_build* b, const int32_t** e_cluster,
                         const int32_t** d_faces, const int32_t** g_nodes,
                         const double** pts, int32_t* n_cubes, int32_t* n_pts,
                         const int32_t** boundary_face, double* L3);

/* Tet mesh tables for checkers: volumes (4 sorted node ids / tet), node
 * coordinates, vtf (4 face ids / tet), edge end nodes (2 / edge), face nodes
 * (3 / face), boundary_face labels. */
void tf_build_tet_tables(const tf_build* b, const int32_t** vols,
                         const double** pts, const int32_t** vtf,
                         const int32_t** edge_nodes, const int32_t** face_nodes,
                         int32_t* n_vols, int32_t* n_pts,
                         const int32_t** boundary_face, const int32_t** vol_material);

/* Writes an n^3-cube Kuhn tetrahedral mesh of [0,1]^3 in netgen neutral
 * format. Interior vertices are jittered by U[-jitter*h, +jitter*h] per
 * coordinate (splitmix64 stream seeded by `seed`); walls are labelled z-min=1,
 * x-min=2, y-max=3, x-max=4, y-min=5, z-max=6 (the cartesian mesher's labels).
 * Material labels: block (bi,bj,bk) of `block`^3 cubes gets label
 * 1 + linear block index, skipping 7 (label-7 tets are dropped by the reader).
 * Returns the number of material labels used, or -1 (an inverted tet). */
int32_t tf_write_kuhn_mesh(const char* path, int32_t n, double jitter,
                           uint64_t seed, int32_t block, double* min_vol_ratio);

#ifdef __cplusplus
}
#endif

#endif /* TETFIT_HOST_H */
