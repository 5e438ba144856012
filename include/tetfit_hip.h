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
    TF_PRECOND_LEARNED = 1     /* learned sparse approximate inverse, below   */
} tf_precond_kind;

/* TF_PRECOND_LEARNED (gfdtd only; it replaces the diagonal preconditioner of
 * ConjugateGradientSolver.hpp:119-128 and the AGMG option of agmg.hpp):
 * G is symmetric with the pattern of A,
 *     G_ej = (base_ej + y_ej) / sqrt(A_ee A_jj),
 *     base = 1 on the diagonal, -A_ej / sqrt(A_ee A_jj) off it,
 * and y_ej is the output of a 16 -> 64 -> 64 -> 1 ReLU MLP on 16 features of
 * the entry pair {e, j} (symmetric in e, j). The first `learn_steps` solves
 * of the context run Jacobi-PCG and capture (residual, correction) pairs
 * (r_k, x* - x_k) at CG iterations 0, 3, 7 and 12 of a dozen of them; the
 * network is then trained on the device (bf16 MFMA forward/backward, Adam)
 * to minimise the A-norm of the error left by one step x += G r, and frozen.
 * Later solves use G; a solve that misses the tolerance is redone with
 * Jacobi-PCG from the same start, and plain CG also judges the gain (no fewer
 * iterations than Jacobi over the first 8 learned solves turns G off). The
 * stopping test is the reference's (||r|| / ||r0|| <= tol) in every case. */
typedef struct {
    double  tol;               /* relative residual vs warm-start residual    */
    int32_t maxiter;           /* 100 in the reference callers (D:1388)       */
    int32_t precond;           /* tf_precond_kind                             */
    int32_t learn_steps;       /* LEARNED: solves that supply training pairs (0: 200) */
    int32_t train_iters;       /* LEARNED: Adam steps (0: 80)                 */
    double  lr;                /* LEARNED: Adam learning rate (0: 3e-3)       */
    int32_t guess;             /* gfdtd start of each solve: 0 the warm start
                                  U_n (the reference); 1 / 2 linear / quadratic
                                  extrapolation 2U_n - U_{n-1} /
                                  3U_n - 3U_{n-1} + U_{n-2} once the history
                                  holds enough steps (LEARNED: the collection
                                  solves start warm). The stop test stays
                                  ||r|| <= tol ||b - A U_n|| (the reference's
                                  warm-start residual). No reference
                                  counterpart. */
} tf_cg_opts;

typedef struct {
    int64_t steps_done;
    int64_t cg_iters_total;    /* sum over steps of CG iterations              */
    int32_t cg_iters_max;
    int32_t cg_not_converged;  /* steps that stopped at maxiter                */
    double  cg_max_relres;     /* max over steps of final ||r||/||r0||          */
    double  seconds;           /* wall time of the step loop (hip events)      */
    /* learned preconditioner, cumulative since the context was created */
    int32_t pc_state;          /* 0 unused, 1 collecting, 2 active, 3 turned off */
    int32_t pc_pairs;          /* captured (residual, correction) pairs        */
    int64_t pc_fallbacks;      /* learned solves redone with Jacobi-PCG        */
    int64_t pc_learned_solves;
    int64_t pc_learned_iters;  /* CG iterations of the learned solves          */
    double  pc_jacobi_mean_iters; /* mean iterations of the collection solves  */
    double  pc_train_seconds;
    double  pc_loss0, pc_loss; /* training loss at the start / best            */
} tf_stats;

/* Parameter vector of the preconditioner MLP (fp32): W1[64][16], b1[64],
 * W2[64][64], b2[64], w3[64], b3. */
#define TF_MLP_IN 16
#define TF_MLP_HIDDEN 64
#define TF_MLP_PARAMS 5313

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

/* The losses column of poynting.dat, one value per stepped step, as the
 * reference writes it: gfdtd pushes 1/2 U_old.SigMat U_old before each solve
 * and, from step 2 on, U.SigMat U after it (D:1519, D:1543-1545), and the
 * file pairs the first n of those values with the n steps (D:1573-1576);
 * dga and fdtd give 0 (dga runs lossless media only; the FIT Joule weights
 * are filled in l2norm output mode only, D:5593-5597). Replaces the Losses
 * vector of Discretization::RunSimulation. */
tf_status tf_read_losses(tf_ctx* ctx, double* vals, int64_t cap, int64_t* n);

/* Discrete energy of the current state, fp64 reductions: fdtd
 * We = 1/2 U.(ep_vec U), Wh = 1/2 B.(nu_vec B) (D:3468-3470); dga and
 * gfdtd We = 1/2 U.(E U), Wh = 1/2 B.(N B) with E the problem's eps mass
 * matrix (D:2773-2775). TF_ERR_STATE for contexts without these matrices
 * (batches). */
tf_status tf_energy(tf_ctx* ctx, double* we, double* wh);

/* Energy time series (the discrete-energy part of the l2norm output,
 * ExportFields D:2250-2284 / D:2773-2775, ExportFitFields D:3468-3470):
 * with energy output on, every output sample (the probe schedule of
 * tf_set_output, AllowPrint) also records We and Wh of the state the probes
 * see, with the reductions of tf_energy (the same values bit for bit).
 * Switching it resets the sample series. TF_ERR_STATE for contexts without
 * the energy matrices. The Poynting flux and losses columns come from
 * tf_read_poynting / tf_read_losses (one value per step). */
tf_status tf_set_energy_output(tf_ctx* ctx, int32_t on);
/* times[k] and vals[2k], vals[2k+1] = (We, Wh) of sample k, k < min(cap, n);
 * n = samples so far. TF_ERR_STATE when energy output is off. */
tf_status tf_read_energy(tf_ctx* ctx, double* times, double* vals, int64_t cap, int64_t* n);

/* Average duration of each kernel launched by the last tf_run, measured with
 * HIP events on the context's stream: names[i] (NUL-separated, packed into
 * `names` of capacity cap_names), ms_avg[i], launches[i]. */
tf_status tf_kernel_times(tf_ctx* ctx, char* names, int32_t cap_names,
                          double* ms_avg, int64_t* launches, int32_t cap,
                          int32_t* n);

/* Enables per-kernel event timing for subsequent runs (costs launch overhead). */
tf_status tf_set_profiling(tf_ctx* ctx, int32_t on);

/* fdtd and dga contexts step in HIP graphs when no launch depends on the host
 * (no probe sampling, no profiling): 64 steps of launches are captured once
 * and replayed; on (default) or off (every step launched from the host).
 * Results are identical either way. No reference counterpart. */
tf_status tf_set_graphs(tf_ctx* ctx, int32_t on);

/* Device bytes held by the context. */
tf_status tf_device_bytes(tf_ctx* ctx, int64_t* bytes);

tf_status tf_destroy(tf_ctx* ctx);

/* Autotuner: times the context's SELL kernels (Ct, H, N for dga; N, the
 * Jacobi pass's A product and, once the learned preconditioner is active,
 * its two passes (G and A products) for gfdtd) over grid sizes, gather
 * batch depths and slice-to-wave mappings (band 0: one contiguous chunk of
 * slices per workgroup; 1: an XCD's waves take its slices round-robin) and
 * keeps the fastest for later tf_run calls. Every row is
 * summed by one lane in entry order for each choice, so explicit fields are
 * identical; the CG's dot products are summed per workgroup, so a different
 * grid changes their rounding (gfdtd / fem fields agree to rounding). Choices are cached per (kernel, precision, rows, entries) in the
 * process and in the file named by $TETFIT_TUNE_CACHE if set. No reference
 * counterpart. tf_get_tuning reports the choice per kernel (order:
 * sell_ct, sell_h, sell_n, cg_w, cg_pg, cg_wu; blocks 0 = default). */
tf_status tf_autotune(tf_ctx* ctx);
tf_status tf_get_tuning(tf_ctx* ctx, int32_t* blocks, int32_t* unroll, int32_t* band, int32_t cap, int32_t* n);
/* Sets the launch shape of the first n kernels in tf_get_tuning's order
 * (blocks 0 = default grid; unroll 4, 8 or 16; band 0 or 1, null = all 0),
 * e.g. to replay a choice. */
tf_status tf_set_tuning(tf_ctx* ctx, const int32_t* blocks, const int32_t* unroll, const int32_t* band,
                        int32_t n);

/* Storage of one device operator, for byte accounting: which = 0 Ct (dga),
 * 1 H, 2 N, 3 A (gfdtd free rows), 4 K = E_f - [A 0] (gfdtd free rows: the
 * couplings left once the warm start cancels, used for the initial residual),
 * 5 RHSmat1, 6 Ct (gfdtd free rows);
 * info[0..5] = rows, 64-row slices, non-zeros, stored entries (with
 * padding), bytes per stored column index (2: 16-bit code, 4: int32),
 * column bases per slice (1: the code is the offset from the slice's base;
 * 4: the code is a 2-bit window and a 14-bit offset from that window's
 * base; 0 with int32 columns). No reference counterpart. */
tf_status tf_sell_info(tf_ctx* ctx, int32_t which, int64_t* info);

/* FIT storage of a context: *grid_order = 1 when fields and coefficients sit
 * at the grid positions of their edges and faces (the problem carries
 * edge_sid / face_sid and every common edge and tbc face has exactly its
 * grid neighbours; the kernels then need no face / edge index tables), 0 for
 * the index-table layout. For byte accounting. No reference counterpart. */
tf_status tf_fit_layout(tf_ctx* ctx, int32_t* grid_order);

/* ---- batches ----------------------------------------------------------
 * Many independent dga simulations (uneven meshes, materials, sources and
 * time steps) stepped together: the problems are concatenated into one
 * block-diagonal context so that each phase of the leapfrog is one launch
 * over every instance; rows and faces carry their own simulation's dt and
 * each source its own time base (t = i * dt_instance). No reference
 * counterpart: the reference runs simulations one after another
 * (dga_main.cpp:58-62). tf_run then advances all instances by the same
 * number of steps (cg may be NULL); Poynting values are summed over the
 * batch. Edge, face and source indices in the per-instance calls are local
 * to the instance. */
tf_status tf_create_batch(const tf_problem* const* probs, int32_t n, int32_t prec, tf_ctx** out);
tf_status tf_batch_set_sources(tf_ctx* ctx, int32_t inst, const tf_source* src, int32_t n_src,
                               const tf_source_term* terms, int64_t n_terms);
tf_status tf_batch_set_probes(tf_ctx* ctx, int32_t inst, const tf_probe* probes, int32_t n);
/* Probe samples every `every` steps (and at t = 0), 0 = none. */
tf_status tf_batch_set_output(tf_ctx* ctx, int64_t every);
tf_status tf_batch_read_probes(tf_ctx* ctx, int32_t inst, double* times, double* vals,
                               int64_t cap_samples, int64_t* n);
tf_status tf_batch_read_state(tf_ctx* ctx, int32_t inst, double* U, double* F, double* B);

/* One forward and backward pass of the preconditioner MLP on the device
 * (the kernels training uses), for gradient checks: y[i] = MLP(x[i]) and
 * grad = sum_i gy[i] dy[i]/dparams, for n feature rows x (n x TF_MLP_IN,
 * rounded to bf16 on upload). No reference counterpart: the reference has no
 * learned component. */
tf_status tf_mlp_grad(const float* params, const float* x, const float* gy,
                      int64_t n, float* y, float* grad);

/* tf_mlp_grad without any device allocation, for callers that cannot
 * allocate (code under HIP stream capture, serving loops, harnesses that
 * forbid hipMalloc). No reference counterpart. All pointers are device
 * memory owned by the caller: params (TF_MLP_PARAMS floats), x (n x
 * TF_MLP_IN bf16 bit patterns), gy (n), y (n, out), grad (TF_MLP_PARAMS,
 * out); work happens on `stream` (a hipStream_t, null = the null stream)
 * and is not waited for. y equals tf_mlp_grad's bit for bit; grad sums the
 * same per-row terms over another partition (fp32 rounding differs). The
 * work buffers live in one fixed device array per GPU (about 2.7 MB, kept
 * until the process ends) whatever n is; n above 2^31 rows is rejected.
 * Calls that overlap on one GPU (other streams, other threads) share that
 * array and race; calls in stream order do not. */
tf_status tf_mlp_grad_noalloc(const float* params, const uint16_t* x, const float* gy,
                              int64_t n, float* y, float* grad, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TETFIT_HIP_H */
