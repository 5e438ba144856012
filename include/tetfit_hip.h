/*
 * tetfit_hip.h - C ABI of libtetfit_hip.so, the MI355X (gfx950) stepping
 * backend for TetFIT's leapfrog time loops.
 *
 * The reference has no plugin or FFI seam: the loops live inside
 * Discretization::RunSimulation (src/DGA/Discretization.hpp, "D:" below).
 * These entry points replace exactly the loop bodies
 *     dga    D:810-939     (explicit DGA leapfrog)
 *     gfdtd  D:1459-1565   (DGA leapfrog with a CG solve per step)
 *     fdtd   D:1840-1933   (FIT / Yee leapfrog)
 * including the per-step BC/source evaluation (ComputeEdgeBC D:3899,
 * ComputeHfieldSource D:4088, ComputeEfieldSource D:3907,
 * ComputeCurrentSource D:3994), the probe sampling of ExportFields /
 * ExportFitFields (D:2400-2470, D:3093-3158) and the Poynting scalar that feeds
 * poynting.dat (D:894, D:1516-1518, D:1901).
 *
 * Conventions: no exception crosses this boundary; every call returns a
 * tf_status and tf_last_error() gives the text of the last failure on the
 * calling thread. Host buffers are copied in and out; the context owns its
 * device memory and one HIP stream. Calls on one context must be serialised;
 * separate contexts may be driven from separate threads.
 */
#ifndef TETFIT_HIP_H
#define TETFIT_HIP_H

#include <stdint.h>
#include "tetfit_problem.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    TF_OK              = 0,
    TF_ERR_ARG         = 1,   /* invalid argument / inconsistent problem      */
    TF_ERR_HIP         = 2,   /* HIP runtime failure                          */
    TF_ERR_NOMEM       = 3,   /* device allocation failed                     */
    TF_ERR_STATE       = 4,   /* call not valid in the context's state        */
    TF_ERR_NOT_CONVERGED = 5  /* CG hit maxiter (reported, not fatal)         */
} tf_status;

typedef enum {
    TF_PREC_FP64 = 0,          /* plain path: fp64 fields and operators        */
    TF_PREC_FP32 = 1           /* fp32 fields and operators, fp64 reductions   */
} tf_precision;

typedef enum {
    TF_PRECOND_JACOBI  = 0,    /* 1/diag(A), ConjugateGradientSolver.hpp:119  */
    TF_PRECOND_LEARNED = 1     /* per-edge diagonal predicted by the MLP      */
} tf_precond_kind;

typedef struct {
    double  tol;               /* relative residual vs warm-start residual    */
    int32_t maxiter;           /* 100 in the reference callers (D:1388)       */
    int32_t precond;           /* tf_precond_kind                             */
} tf_cg_opts;

typedef struct {
    int64_t steps_done;
    int64_t cg_iters_total;    /* sum over steps of CG iterations              */
    int32_t cg_iters_max;
    int32_t cg_not_converged;  /* steps that stopped at maxiter                */
    double  cg_max_relres;     /* max over steps of final ||r||/||r0||          */
    double  seconds;           /* wall time of the step loop (hip events)      */
} tf_stats;

typedef struct tf_ctx tf_ctx;

/* Text of the last error raised on this thread ("" if none). */
const char* tf_last_error(void);

/* Number of visible HIP devices (0 when no GPU is present). */
int32_t tf_device_count(void);

/* Copies the operators of `prob` to device 0 in the layout the kernels use.
 * `prec` selects fp64 or fp32 storage for fields and operator values. */
tf_status tf_create(const tf_problem* prob, int32_t prec, tf_ctx** out);

/* Sources used by the simulation (Simulation::Src(), Simulation.hpp:40-57). */
tf_status tf_set_sources(tf_ctx* ctx, const tf_source* src, int32_t n_src,
                         const tf_source_term* terms, int64_t n_terms);

/* Probe points, already located and weighted on the host. */
tf_status tf_set_probes(tf_ctx* ctx, const tf_probe* probes, int32_t n_probes);

/* Output schedule of Output::AllowPrint (Output.hpp:318-328): a sample is taken
 * whenever t >= next, then next += period (period 0 = every step). */
tf_status tf_set_output(tf_ctx* ctx, double delay, double period);

/* Advances steps [first_step, first_step+n_steps). Step i is at time i*dt.
 * fdtd starts at i = 0, dga and gfdtd at i = 1 (D:1840, D:810, D:1459).
 * cg may be NULL for explicit methods. */
tf_status tf_run(tf_ctx* ctx, int64_t first_step, int64_t n_steps,
                 const tf_cg_opts* cg, tf_stats* stats);

/* Field state. Any pointer may be NULL. Sizes: U n_edges, F and B n_faces. */
tf_status tf_read_state(tf_ctx* ctx, double* U, double* F, double* B);
tf_status tf_write_state(tf_ctx* ctx, const double* U, const double* F,
                         const double* B);

/* Probe samples taken so far: times[k] and vals[(k*n_probes + p)*6 + c]
 * with c = Ex,Ey,Ez,Hx,Hy,Hz. *n receives the number of samples. */
tf_status tf_read_probes(tf_ctx* ctx, double* times, double* vals,
                         int64_t cap_samples, int64_t* n);

/* Per-step Poynting scalar (the first column of poynting.dat). */
tf_status tf_read_poynting(tf_ctx* ctx, double* vals, int64_t cap, int64_t* n);

/* Discrete energy of the current state in fp64 (We, Wh); fdtd and dga. */
tf_status tf_energy(tf_ctx* ctx, double* we, double* wh);

/* Average duration of each kernel launched by the last tf_run, measured with
 * HIP events on the context's stream: names[i] (NUL-separated, packed into
 * `names` of capacity cap_names), ms_avg[i], launches[i]. */
tf_status tf_kernel_times(tf_ctx* ctx, char* names, int32_t cap_names,
                          double* ms_avg, int64_t* launches, int32_t cap,
                          int32_t* n);

/* Enables per-kernel event timing for subsequent runs (costs launch overhead). */
tf_status tf_set_profiling(tf_ctx* ctx, int32_t on);

/* Device bytes held by the context. */
tf_status tf_device_bytes(tf_ctx* ctx, int64_t* bytes);

tf_status tf_destroy(tf_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* TETFIT_HIP_H */
