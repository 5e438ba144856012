/*
 * tetfit_problem.h - plain-C description of one assembled TetFIT time-stepping
 * problem, as handed from the host side (parser, meshers, operator producers)
 * to a stepping backend (the HIP library or the CPU oracle).
 *
 * Every array is owned by the producer; consumers copy what they need.
 * Index type is int32, incidence signs are int8 (+1 / -1), values are fp64.
 *
 * The tables mirror what the reference's Discretization object holds when its
 * time loop starts (src/DGA/Discretization.hpp, abbreviated D: below):
 *   FIT/FDTD : Ct_vec, C_vec, curl, dual_curl, M_h, M_q, upmlcoeff1, M_mu,
 *              M_nu, fpmlcoeff1, common_edges, uncommon_edges, bc_edges,
 *              tbc_surfaces (D:10826-10834, loop D:1840-1933)
 *   DGA      : C, Ctb (as bnd_edge), H with h_index, N, edge classes
 *              (D:10779-10780, loop D:810-939)
 *   GFDTD    : C, Ctb, N, E, A, RHSmat1, compressed_dirichlet
 *              (loop D:1459-1565)
 */
#ifndef TETFIT_PROBLEM_H
#define TETFIT_PROBLEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    TF_METHOD_FDTD  = 0,   /* "method fdtd"  on a cartesian mesher grid       */
    TF_METHOD_DGA   = 1,   /* "method dga"   explicit, local-inverse Hodge     */
    TF_METHOD_GFDTD = 2,   /* "method gfdtd" consistent eps mass, CG solve    */
    TF_METHOD_FEM   = 3    /* "method fem" Whitney edge elements, second order */
                           /* in time, CG solve (D:1580-1790)                 */
} tf_method;

/* Compressed sparse rows. rowptr has nrows+1 entries (int64 so that row
 * pointers of the largest meshes fit). For incidence matrices val is NULL and
 * sgn holds +1/-1 per entry; for material matrices sgn is NULL. */
typedef struct {
    int32_t        nrows;
    int32_t        ncols;
    int64_t        nnz;
    const int64_t* rowptr;
    const int32_t* col;
    const double*  val;
    const int8_t*  sgn;
} tf_csr;

/* FIT / FDTD operators (method fdtd). */
typedef struct {
    /* Ct_vec: signed faces around each edge, in the reference's push order */
    const int64_t* ct_ptr;        /* n_edges+1                               */
    const int32_t* ct_face;       /* ct_ptr[n_edges]                          */
    const int8_t*  ct_sgn;
    const int32_t* c_edge;        /* C_vec: 4 edges per face, 4*n_faces       */
    const int8_t*  curl;          /* per face, +1/-1                          */
    const int8_t*  dual_curl;     /* per edge, +1/-1                          */
    const double*  M_h;           /* per edge                                 */
    const double*  M_q;
    const double*  upml;          /* upmlcoeff1, 1.0 without PML              */
    const double*  M_mu;          /* per face                                 */
    const double*  M_nu;
    const double*  fpml;          /* fpmlcoeff1                               */
    const int32_t* common_edges;   int64_t n_common;
    const int32_t* uncommon_edges; int64_t n_uncommon;  /* h-source edges     */
    const int32_t* bc_edges;       int64_t n_bc;        /* pec edges          */
    const double*  bc_value;      /* per bc edge: BCs[...].GetValue()         */
    const int32_t* src_edges;      int64_t n_src;       /* e / j source edges */
    const int32_t* tbc_faces;      int64_t n_tbc;       /* tbc_surfaces       */
    /* energy diagnostic weights (D:3468-3470): We = 1/2 U.(ep_vec U),
     * Wh = 1/2 B.(nu_vec B); ep_vec = eps_bar/L (D:5596), nu_vec = nu_bar/A (D:5507) */
    const double*  ep_vec;
    const double*  nu_vec;
    /* structured position of every edge and face on the nx x ny x nz cell
     * grid (or NULL): x edges (i<nx, j<=ny, k<=nz) at ((k(ny+1)+j)nx+i), then
     * y edges (i<=nx, j<ny, k<=nz) at ((k ny+j)(nx+1)+i), then z edges
     * (i<=nx, j<=ny, k<nz) at ((k(ny+1)+j)(nx+1)+i); x faces (i<=nx, j<ny,
     * k<nz) at ((k ny+j)(nx+1)+i), then y faces (i<nx, j<=ny, k<nz) at
     * ((k(ny+1)+j)nx+i), then z faces (i<nx, j<ny, k<=nz) at ((k ny+j)nx+i).
     * Lets the device store the fields in grid order and find an edge's
     * faces (and a face's edges) from its position instead of index tables. */
    const int32_t* edge_sid;
    const int32_t* face_sid;
    int32_t        cells[3];
} tf_fit_ops;

/* Fractional unknowns of method dga with lossy (sigma) or magnetic-loss
 * (chi) media, ConstructCodecasaMaterialMatrices D:9180-9481 and the loop
 * D:883-920. Edge classes 1..4 and face classes 1..3 split the unknowns into
 *   U_frac = [U_a (h_size) | U_b (p_size) | U_c (q_size) | U_d (b_size)]
 *   F_frac = [F_a (n_size) | F_b (r_size) | F_c (s_size)]
 * and each step does
 *   U_frac <- Ru U_frac + dt Ku curl_f,   U = Mu U_frac,
 *   F_frac <- Rf F_frac - dt Kf curl_u,   F = Tf F_frac,
 * with Ku = [H; Mp; Mq; 0], Ru = [I; diag P_p; Q; I], Kf = [N; Tr; Ts],
 * Rf = [I; diag R_r; S] (all rows of these are in that stacked order). The
 * boundary values and e sources of an edge are written to U_frac entry
 * slot[edge] (U_d for class 4, U_a for class 1; -1: no entry) before the
 * update. enabled = 0 (and everything else empty) for lossless media, which
 * step with H and N directly. */
typedef struct {
    int32_t        enabled;
    int32_t        h_size, p_size, q_size, b_size;
    int32_t        n_size, r_size, s_size;
    tf_csr         Ku;            /* nU x n_edges, nU = h+p+q+b               */
    tf_csr         Ru;            /* nU x nU                                  */
    tf_csr         Mu;            /* n_edges x nU, entries 1 (1 or 2 per row)  */
    tf_csr         Kf;            /* nF x n_faces, nF = n+r+s                 */
    tf_csr         Rf;            /* nF x nF                                  */
    tf_csr         Tf;            /* n_faces x nF, entries 1 (1 or 2 per row)  */
    const int32_t* slot;          /* per edge                                 */
} tf_dga_frac;

/* DGA operators (methods dga and gfdtd). */
typedef struct {
    tf_csr         C;             /* faces x edges, signs                     */
    tf_csr         Ct;            /* edges x faces, transpose of C            */
    const int32_t* bnd_edge;      /* per edge: column of Ctb, or -1           */
    int32_t        n_bnd;         /* columns of Ctb                           */
    const int8_t*  edge_class;    /* classify_edges: 1..4                     */
    const double*  fixed_value;   /* per edge, value of class-4 edges         */
    /* dga */
    tf_csr         H;             /* H_size x n_edges                         */
    const int32_t* h_row;         /* per edge: h_index for class 1, else -1   */
    /* both */
    tf_csr         N;             /* n_faces x n_faces                        */
    /* gfdtd and fem */
    tf_csr         E;             /* n_edges x n_edges eps mass (fem: the     */
                                  /* Whitney element mass, D:7917-8058)       */
    tf_csr         A;             /* n_free x n_free system matrix (fem:      */
                                  /* E/dt^2 + SigMat/(2 dt), D:8086-8112)     */
    tf_csr         RHS1;          /* RHSmat1: n_edges x n_edges               */
    const int32_t* free_edges;    /* compressed_dirichlet                      */
    int32_t        n_free;
    tf_csr         Sig;           /* SigMat: n_edges x n_edges conductivity    */
                                  /* mass on E's pattern (gfdtd); nnz 0 when   */
                                  /* no tet is lossy                           */
    /* dga with lossy media: the fractional operators (H is then empty and N
     * is the lossless all-face N of the energy diagnostic) */
    tf_dga_frac    frac;
    /* optional (NULL: none; given together): locality ranks of every edge
     * and face, used by a stepping backend only to lay out its rows and
     * fields: entities of one shape class on one grid line get consecutive
     * ranks, so the rows of a 64-row slice gather neighbouring entries from a
     * few cache lines. Any permutations are valid; results do not depend on
     * them beyond the rounding of reductions. */
    const int32_t* edge_rank;
    const int32_t* face_rank;
} tf_dga_ops;

typedef struct {
    int32_t    method;            /* tf_method                                 */
    int32_t    n_edges;
    int32_t    n_faces;
    int32_t    n_vols;
    double     dt;                /* t_step                                    */
    double     duration;          /* simulation_time                           */
    tf_fit_ops fit;
    tf_dga_ops dga;
} tf_problem;

/* ---- sources ---------------------------------------------------------- */

typedef enum {
    TF_CARRIER_SIN      = 0,
    TF_CARRIER_COS      = 1,   /* multiplies amp twice, Source.hpp:322       */
    TF_CARRIER_GAUSSIAN = 2,
    TF_CARRIER_RICKER   = 3,
    TF_CARRIER_DC       = 4,
    TF_CARRIER_GAUSSMOD = 5    /* extension: amp*exp(-(t-d)^2/w^2)*sin(2pi f (t-d)) */
} tf_carrier_kind;

typedef enum {
    TF_FIELD_H = 0,            /* boundary magnetic source (Fb / curl_fb)    */
    TF_FIELD_E = 1,            /* imposed edge voltage U(e)                   */
    TF_FIELD_J = 2,            /* impressed current I(e)                      */
    TF_FIELD_D = 3             /* any other type on an edge (d): the loops of */
                               /* D:827-842 and D:1485-1500 route such an     */
                               /* edge to ComputeEfieldSource, which sums "e" */
                               /* sources only (D:3928) and so gives 0: dga   */
                               /* resets the edge's U to 0 before the H       */
                               /* update, gfdtd overwrites it with the solve  */
} tf_field_kind;

/* Time part of one DEFINE source (Source::Compute's carrier, Source.hpp:316-327). */
typedef struct {
    int32_t carrier;           /* tf_carrier_kind                             */
    int32_t field;             /* tf_field_kind                               */
    double  amp;
    double  freq;
    double  delay;
    double  width;
    double  impedance;         /* j sources multiply amp by it                */
} tf_source;

/* Space part of one source on one edge: the value injected at edge `edge` at
 * time t is sum over its terms of carrier(source, t) * weight. */
typedef struct {
    int32_t edge;
    int32_t source;            /* index into the tf_source array              */
    double  weight;
} tf_source_term;

/* ---- probes ----------------------------------------------------------- */

/* A probe samples E as sum_k U[e_idx[k]] * e_w[k] and H as sum_k B[h_idx[k]] *
 * h_w[k] (3-vectors). FIT probes use 12 edges / 6 faces of the cube that holds
 * the point (D:3555-3588); DGA probes use the Whitney forms of the 6 tet edges
 * and the 4 tet faces (D:3768-3897). */
#define TF_PROBE_MAX_E 12
#define TF_PROBE_MAX_H 6
typedef struct {
    int32_t ne, nh;
    int32_t e_idx[TF_PROBE_MAX_E];
    double  e_w[TF_PROBE_MAX_E][3];
    int32_t h_idx[TF_PROBE_MAX_H];
    double  h_w[TF_PROBE_MAX_H][3];
} tf_probe;

#ifdef __cplusplus
}
#endif

#endif /* TETFIT_PROBLEM_H */
