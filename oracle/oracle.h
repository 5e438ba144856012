/*
 * oracle.h - CPU checker for the TetFIT time-stepping hot path.
 *
 * TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load liboracle.so; it is the checker, never the thing
 * measured as the product or shipped.
 *
 * A single-threaded fp64 restatement of the reference loops
 * (src/DGA/Discretization.hpp, "D:"):
 *   fdtd   D:1840-1933  with ComputeHfieldSource D:4088-4202,
 *                       ComputeEfieldSource D:3907, ComputeCurrentSource D:3994,
 *                       ExportFitFields D:3093-3158 -> GetFitElectricField
 *                       D:3555, GetFitMagneticField D:3570
 *   dga    D:810-939    with ExportFields D:2400-2470 -> GetWhitneyElectricField
 *                       D:3768-3875, GetMagneticField D:3877-3897
 *   gfdtd  D:1459-1565  with ConjugateGradientSolver.hpp:52-103
 *   Source::Compute     Source.hpp:313-567 (cartesian directions; x87 long
 *                       double where the reference's PI promotes to it)
 * It consumes the same assembled operators as the device path (tf_problem)
 * plus the reference-native source and probe descriptions, and evaluates
 * sources and probes itself.
 */
#ifndef TETFIT_ORACLE_H
#define TETFIT_ORACLE_H

#include <stdint.h>

#include "../include/tetfit_host.h"
#include "../include/tetfit_problem.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    /* probe samples: times[k], vals[(k*n_probes + p)*6 + c] */
    double* times;
    double* vals;
    int64_t cap_samples;
    int64_t n_samples;
    /* per-step Poynting scalar */
    double* poynting;
    int64_t cap_poynting;
    int64_t n_poynting;
    /* final state (any may be NULL) */
    double* U;
    double* F;
    double* B;
    /* CG statistics (gfdtd) */
    int64_t cg_iters_total;
    int32_t cg_iters_max;
    int32_t cg_not_converged;
    double cg_max_relres;
    /* steady-clock seconds of the step loop, probe export excluded (D:812/926) */
    double seconds;
    /* gfdtd Losses vector in push order (D:1519, D:1543-1545, D:1566-1570) */
    double* losses;
    int64_t cap_losses;
    int64_t n_losses;
    /* energies at each probe sample (may be NULL): energy[2k], energy[2k+1]
     * = We, Wh of the state the sample's probes read (ExportFields l2norm
     * D:2773-2775: 1/2 U.(E U), 1/2 B.(N B); ExportFitFields D:3468-3470:
     * 1/2 U.(ep_vec U), 1/2 B.(nu_vec B)); cap = cap_samples */
    double* energy;
} oracle_out;

/* Runs steps [first, first+n) of the problem's method. faithful=1 keeps the
 * reference's containers in the fdtd loop (vector-of-vector Ct_vec, a heap
 * vector per common edge per step, D:1873); faithful=0 uses flat arrays with
 * the same arithmetic. U0/F0/B0 (may be NULL) set the initial state.
 * Returns 0, or -1 with oracle_last_error(). */
int32_t oracle_run(const tf_problem* p, const tf_source_desc* sd, int32_t n_sd, const tf_source_edge* se,
                   int64_t n_se, const tf_probe_geom* pg, int32_t n_probes, int32_t out_enabled,
                   double out_delay, double out_period, int64_t first, int64_t n, int32_t faithful,
                   double cg_tol, int32_t cg_maxiter, const double* U0, const double* F0, const double* B0,
                   oracle_out* out);

/* We, Wh of the state (U, B) as the energy series records them. */
void oracle_energy(const tf_problem* p, const double* U, const double* B, double* we, double* wh);

/* Source::Compute(p, t) for one source (3-vector into out3). */
void oracle_source_compute(const tf_source_desc* s, const double p[3], double t, double out3[3]);

/* ConjugateGradientSolver::solveWithGuess on a CSR matrix with the Jacobi
 * preconditioner of compute(); returns iterations, writes x and final error. */
int32_t oracle_cg_solve(int32_t n, const int64_t* rowptr, const int32_t* col, const double* val, const double* b,
                        const double* x0, double tol, int32_t maxiter, double* x, double* err);

const char* oracle_last_error(void);

/* One probe sample of the state (U, B): fdtd=1 GetFitElectricField /
 * GetFitMagneticField (D:3555-3588), else GetWhitneyElectricField /
 * GetMagneticField (D:3768-3897); six[p*6 + c], c = Ex,Ey,Ez,Hx,Hy,Hz. */
void oracle_export(int32_t fdtd, const tf_probe_geom* pg, int32_t n_probes, int32_t n_edges, int32_t n_faces,
                   const double* U, const double* B, double* six);

/* Preconditioner MLP (16 -> 64 -> 64 -> 1, ReLU) forward and backward in
 * fp64: y[i] and grad = sum_i gy[i] dy[i]/dparams. */
void oracle_mlp(const float* params, const float* x, const double* gy, int64_t n, double* y, double* grad);

#ifdef __cplusplus
}
#endif

#endif
