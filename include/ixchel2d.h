/*
 * ixchel2d.h -- C-ABI of the MI355X (gfx950) SIMPLE-iteration library.
 *
 * Drop-in boundary for the time-step loop body of Ixchel2D
 * (source/main.f90:391-1271).  Upstream has no FFI: its operator API is a set
 * of Fortran module procedures called per cell (ensambla_*, residuo_u) or per
 * line (tridiagonal) from loop nests in main.f90, with all state in module
 * arrays inside an OpenACC data region (main.f90:356-389).  Calls per cell
 * across an FFI are not workable, so each entry point below replaces one
 * main.f90 loop nest or one inner-loop pass; the comment on each names what
 * it replaces.  The Fortran binding (ixchel2d_amd/fortran/mod_ixchel_hip.f90)
 * and the Python binding (ixchel2d_amd/_lib.py) bind exactly these symbols.
 *
 * Conventions
 *  - Every call returns 0 on success or a negative IXH_ERR_* code; the text
 *    of the last error on the calling thread is in ixh_last_error().
 *  - Arrays cross the boundary in upstream's Fortran layout: element (i,j),
 *    1-based, of an array with leading dimension ld sits at (i-1)+(j-1)*ld.
 *    Extents: u(mi,nj+1), v(mi+1,nj), T/p/p'/b_o(mi+1,nj+1), a_u(mi,nj+1),
 *    a_v(mi+1,nj) (source/mod_ec_momento.f90:27-41, mod_ec_energia.f90:23-24).
 *  - Host pointers in, host pointers out; device residency is owned by the
 *    state object.  One HIP stream per state; calls on one state are not
 *    thread-safe, separate states are independent.
 */
#ifndef IXCHEL2D_H
#define IXCHEL2D_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IXH_OK 0
#define IXH_ERR_ARG (-1)      /* invalid argument or state */
#define IXH_ERR_HIP (-2)      /* a HIP runtime call failed */
#define IXH_ERR_NOMEM (-3)    /* device allocation failed */
#define IXH_ERR_UNSUPP (-4)   /* size or option outside what is built */

/* fields (ixh_upload_field / ixh_download_field) */
enum {
  IXH_U = 0, IXH_V = 1, IXH_T = 2, IXH_P = 3,
  IXH_U_ANT = 4, IXH_V_ANT = 5, IXH_T_ANT = 6,
  IXH_PCORR = 7, IXH_BO = 8, IXH_AU = 9, IXH_AV = 10, IXH_RESU = 11,
  IXH_NFIELDS = 12
};
/* variables for boundary conditions and sources */
enum { IXH_VAR_U = 0, IXH_VAR_V = 1, IXH_VAR_T = 2 };
/* sides: a = west (x=0), b = south, c = east, d = north */
enum { IXH_SIDE_A = 0, IXH_SIDE_B = 1, IXH_SIDE_C = 2, IXH_SIDE_D = 3 };
/* boundary node types (tipo_cond_front % tipo_condi) */
enum { IXH_BC_NONE = 0, IXH_BC_DIRI = 1, IXH_BC_NEUM = 2 };
/* sweep directions: Y = lines along y (the `_y` assemblies), X = along x */
enum { IXH_DIR_Y = 0, IXH_DIR_X = 1 };
/* line solvers */
enum {
  IXH_SOLVER_PCR = 0,      /* PCR in LDS, division-free assembly (default) */
  IXH_SOLVER_THOMAS = 1,   /* upstream's Thomas and assembly order: bit-faithful */
  IXH_SOLVER_PCR_EXACT = 2 /* PCR with upstream-order assembly */
};
/* arithmetic of the sweeps.
 * FP64: every field and sweep in FP64 (the drop-in).
 * FP32_MIXED: u, v, T, p' stored and swept in FP32; p, mass residual, norms
 *   and Nusselt integrals in FP64.  Stalls about 6e-8 relative above the
 *   FP64 steady state.
 * FP32_REFINED: every field stored in FP64; each sweep assembles its rows in
 *   FP32 and solves for the FP32 increment of the line against the line's
 *   FP64 residual, so the iteration converges to the FP64 steady state.
 * FP32_STOCHASTIC: the storage of FP32_MIXED and the increment sweeps of
 *   FP32_REFINED (the residual formed in FP64 from the FP32 fields); each new
 *   u, v, T (sweeps and correction) is phi + delta formed in FP64 and stored
 *   with stochastic rounding, so increments below the FP32 spacing are kept
 *   on average instead of dropped.  Runs are deterministic: the rounding
 *   bits are a hash of the value and its node. */
enum { IXH_FP64 = 0, IXH_FP32_MIXED = 1, IXH_FP32_REFINED = 2, IXH_FP32_STOCHASTIC = 3 };

typedef struct ixh_state ixh_state;

/* mesh: node coordinates as read by lectura_mallas_escalonadas
 * (source/mod_malla.f90:62-162): xu(mi), yv(nj), xp(mi+1), yp(nj+1) */
typedef struct {
  int32_t mi, nj;
  const double *xu, *yv, *xp, *yp;
} ixh_grid;

/* the scalars the loop body reads (source/main.f90:164-244) */
typedef struct {
  double dt, rel_pres, rel_vel, rel_ener;
  double gamma_momen, gamma_energ; /* constant diffusivities (main.f90:221-244) */
  double ri_x, ri_y;               /* Ri (u equation) and Riy (v equation) */
  double conv_u, conv_t, conv_p, conv_resi, conv_paso;
  int32_t iter_ecuaci_max, iter_simple_max;
  int32_t solver;                  /* IXH_SOLVER_* */
  int32_t precision;               /* IXH_FP64 | IXH_FP32_MIXED | IXH_FP32_REFINED | IXH_FP32_STOCHASTIC */
} ixh_params;

/* one record per SIMPLE iteration: sqrt of the squared-difference sums of
 * the momentum (v only), p' and energy passes, sqrt(sum b_o^2) and the u
 * residual norm (main.f90:716-725, 884-906, 1128-1136, 1192-1199); pass
 * counts; the SIMPLE counter after the exit test and whether upstream writes
 * the line to fort.102 (main.f90:1202-1215) */
typedef struct {
  double err_v, err_p, maxbo, err_t, residuo;
  int32_t passes_mom, passes_p, passes_t, simple_iter, logged, pad_;
} ixh_hist;

const char *ixh_last_error(void);
const char *ixh_version(void);

/* Create the device state: geometry factors (mod_malla.f90:167-217), field
 * buffers, a_u = a_v = 1e40 (main.f90:309-310), everything else 0.
 * Replaces the `!$acc data` region (main.f90:356-389). */
int ixh_create(const ixh_grid *grid, const ixh_params *params, ixh_state **out);
int ixh_destroy(ixh_state *s);
int ixh_set_params(ixh_state *s, const ixh_params *params);
/* run the state's work on an existing HIP stream (NULL = its own stream) */
int ixh_set_stream(ixh_state *s, void *hip_stream);
int ixh_synchronize(ixh_state *s);

/* field transfer, Fortran layout with leading dimension ld (>= the field's
 * first extent).  Replaces the data-region copy/copyin/update clauses. */
int ixh_upload_field(ixh_state *s, int32_t field, const double *src, int32_t ld);
int ixh_download_field(ixh_state *s, int32_t field, double *dst, int32_t ld);

/* Boundary conditions per node of one side, flattened from tipo_cond_front
 * (source/mod_cond_frontera.f90:25-38) by the lectura_cond_frontera index
 * rule (:125-149).  type/value have one entry per node along the side: u sides
 * a/c nj+1, b/d mi; v sides a/c nj, b/d mi+1; T sides a/c nj+1, b/d mi+1.
 * Nodes 2..n-1 must all be typed. */
int ixh_set_bc(ixh_state *s, int32_t var, int32_t side, const int32_t *type, const double *value);

/* Source terms S = Sc + Sp*phi of one variable in its own extent (NULL,NULL
 * clears them).  Immersed bodies enter here as definir_cuerpo's large-source
 * penalty (source/mod_front_inmersa.f90:54-179). */
int ixh_set_sources(ixh_state *s, int32_t var, const double *sc, const double *sp);

/* ---- one directional sweep: assembly + boundary rows + line solves +
 * copy-back.  dir = IXH_DIR_Y replaces main.f90:420-484 (u), 570-640 (v);
 * dir = IXH_DIR_X replaces 492-556 (u), 647-711 (v).  The Y sweep of v takes
 * the snapshot of v (main.f90:559-564); the X sweep of v returns the momentum
 * error sqrt(sum (v-fv)^2) (main.f90:716-725) in *err_v when err_v != NULL. */
int ixh_momentum_u(ixh_state *s, int32_t dir);
int ixh_momentum_v(ixh_state *s, int32_t dir, double *err_v);
/* p' = 0 (main.f90:752-758) */
int ixh_pcorr_zero(ixh_state *s);
/* p' sweeps (main.f90:762-879); the X sweep stores b_o and returns the p'
 * error and maxbo = sqrt(sum b_o^2) (main.f90:884-906) when non-NULL */
int ixh_pcorr(ixh_state *s, int32_t dir, double *err_p, double *maxbo);
/* p += 0.6 p'; u, v corrections (main.f90:931-962) */
int ixh_correct(ixh_state *s);
/* T sweeps (main.f90:974-1121); the X sweep returns the energy error
 * (main.f90:1128-1136) when err_t != NULL */
int ixh_energy(ixh_state *s, int32_t dir, double *err_t);
/* residuo_u over the interior and its norm (mod_ec_momento.f90:1107-1266,
 * main.f90:1169-1199) */
int ixh_residual_u(ixh_state *s, double *residuo);
/* lagged copies T_ant = T, u_ant = u, v_ant = v (main.f90:1251-1270) */
int ixh_lag(ixh_state *s);

/* nsteps time steps of main.f90:391-1270 with upstream's inner-loop and
 * SIMPLE exit rules.  When every loop is single-pass (iter_ecuaci_max < 0 and
 * iter_simple_max < 0) the steps run as 6 launches per step on the state's
 * stream (independent phases of consecutive steps share a launch) with no
 * host synchronisation; otherwise the errors are read back after every pass.
 * One record per SIMPLE iteration goes to hist (up to cap); *nrec receives
 * the count (may be NULL). */
int ixh_simple_iterations(ixh_state *s, int32_t nsteps, ixh_hist *hist, int32_t cap, int32_t *nrec);

/* nusselt_promedio_y on the current T (source/mod_postproceso.f90:100-215):
 * integrals over ii = i0..i1 (step 2) on the walls j=1 and j=nj+1 */
int ixh_nusselt_y(ixh_state *s, int32_t i0, int32_t i1, double *nu0, double *nu1);
/* mean Nusselt number on the vertical walls (no upstream counterpart; the
 * definition is in ixh_nusselt_x's implementation and in the oracle) */
int ixh_nusselt_x(ixh_state *s, double *nu_west, double *nu_east);

/* tridiagonal(a,b,c,r,n) (source/mod_solucionador.f90:25-47) for nsys
 * independent systems stored one after another (stride n), solved on the
 * device; b and r are overwritten as upstream does (r = solution). */
int ixh_tridiagonal(const double *a, double *b, const double *c, double *r, int32_t n,
                    int32_t nsys, int32_t solver);
/* ixh_tridiagonal without any device allocation: no upstream counterpart;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, per-line calls such as the Fortran
 * tridiagonal wrapper).  The systems go through a device scratch array of
 * 4 x 262144 doubles (8 MiB per GPU, held until the process ends), so
 * n * nsys must be at most 262144 (IXH_ERR_UNSUPP otherwise).  Same
 * arguments, results and errors as ixh_tridiagonal.  Every call on one GPU
 * uses that GPU's one array: calls that overlap on one GPU (other host
 * threads) race. */
int ixh_tridiagonal_noalloc(const double *a, double *b, const double *c, double *r, int32_t n,
                            int32_t nsys, int32_t solver);

/* per-kernel timing of the ixh_simple_iterations calls made with profiling
 * on (ixh_set_profiling resets it): for each kernel id the total milliseconds
 * and the launch count, measured with HIP events on the state's stream.  Ids
 * 0-10 are the single kernels of the sequential step (events around each
 * launch); ids 11-16 the 6 launches of the fused step, whose events carry
 * the kernels' own start and end (hipExtLaunchKernelGGL) and are read back
 * from a ring without stalling the launches. */
int ixh_set_profiling(ixh_state *s, int32_t on);
int ixh_kernel_times(ixh_state *s, double *ms, int32_t *launches, int32_t cap, int32_t *nkernels);
const char *ixh_kernel_name(int32_t id);

/* ---- batches: many independent cavities (any sizes up to 1023 x 1023,
 * their own parameters, boundary conditions and sources, one common solver)
 * advanced together.  Each phase of a time step is one launch over the lines
 * of every cavity; a cavity's arithmetic and phase order are those of
 * ixh_simple_iterations' sequential step, so its fields equal a run of that
 * cavity alone.  No upstream counterpart: upstream runs one cavity per
 * process (main.f90:348-1272). */
typedef struct ixh_batch ixh_batch;
/* workgroup orderings of a batch launch */
enum {
  IXH_ORDER_CAVITY = 0,  /* cavity by cavity, lines in order */
  IXH_ORDER_LARGEST = 1, /* largest cavities first */
  IXH_ORDER_XCD = 2      /* whole cavities binned onto the 8 XCDs (default) */
};
/* the states stay owned by the caller and must outlive the batch */
int ixh_batch_create(ixh_state *const *states, int32_t n, ixh_batch **out);
int ixh_batch_destroy(ixh_batch *b);
/* ordering, and whether each step runs as one captured HIP graph */
int ixh_batch_set_schedule(ixh_batch *b, int32_t order, int32_t graph);
/* threads per line of the sweep and residual launches: 64 (one wave per line
 * with ceil(n/64) rows per lane; the pcr solver, every line at most 320
 * nodes, IXH_ERR_UNSUPP otherwise) or 256 (rows per thread from the longest
 * line); lines per workgroup in the sweeps: 1, or IXH_BATCH_LINE_GROUP
 * consecutive lines of a cavity, one wave each (64 threads per line only),
 * whose copies in the other layout are stored as runs of neighbouring values
 * through LDS; and whether the u residual rows run in the T x-sweep's launch
 * (1; one line per workgroup only) or in a launch of their own (0).  A step in
 * which some cavity's u equation reads T (ri_x != 0) runs the residual on its
 * own either way.  Defaults: 64 threads and groups of IXH_BATCH_LINE_GROUP
 * lines where allowed, else 256 threads; the residual in its own launch.
 * Every shape gives the same fields and norms bit for bit. */
#define IXH_BATCH_LINE_GROUP 4
int ixh_batch_set_shape(ixh_batch *b, int32_t line_threads, int32_t lines_per_group, int32_t residual_with_t);
/* nsteps single-pass time steps of every cavity (each needs
 * iter_ecuaci_max < 0 and iter_simple_max < 0); the records of the first
 * min(nsteps, cap) steps of cavity c go to hist[c*cap + k] */
int ixh_batch_iterations(ixh_batch *b, int32_t nsteps, ixh_hist *hist, int32_t cap);
/* time trial_steps steps under every schedule (ordering, graph or direct
 * launches, threads per line, lines per workgroup, residual placement) the
 * batch allows, keep the fastest and return it.  ms_per_step (NULL or
 * IXH_BATCH_SCHEDULES entries, index 16*order + 8*graph +
 * 4*(line_threads == 256) + 2*(lines_per_group > 1) + residual_with_t)
 * receives the timings, -1 where a schedule was not allowed.  The trial steps
 * are real time steps of every cavity (trial_steps + 4 per schedule tried),
 * identical under every schedule. */
#define IXH_BATCH_SCHEDULES 48
int ixh_batch_autotune(ixh_batch *b, int32_t trial_steps, int32_t *order, int32_t *graph, int32_t *line_threads,
                       int32_t *lines_per_group, int32_t *residual_with_t, double *ms_per_step);

/* ---- discrete adjoint of the energy transport.  No upstream counterpart.
 * ixh_adjoint_record runs nsteps single-pass time steps from the current
 * state (the state advances as under ixh_simple_iterations) and keeps a
 * checkpoint of the fields every `every` steps.  ixh_adjoint_energy then runs
 * the adjoint of the recorded energy steps backward, recomputing each
 * segment's velocities from its checkpoint: given tbar = dJ/dT of the final
 * temperature ((mi+1) x (nj+1), leading dimension ld) it returns dJ/d(value)
 * of every T boundary node of sides a, b, c, d (NULL = not wanted; lengths as
 * in ixh_set_bc) and dJ/dT of the initial temperature (with T_ant taken equal
 * to it).  The adjoint holds the velocities of each step fixed: it is the
 * exact gradient when the flow does not depend on T (ri_x = ri_y = 0), and
 * the frozen-flow part of it under buoyancy.  Afterwards u, v, T, p and the
 * lagged u_ant, v_ant, T_ant are those at the end of the recorded run, which
 * is all a later time step reads; p' reads as zero, and a_u, a_v, b_o and
 * Res_u hold values of the recomputation (every step recomputes them before
 * use).  FP64 path only. */
int ixh_adjoint_record(ixh_state *s, int32_t nsteps, int32_t every);
int ixh_adjoint_energy(ixh_state *s, const double *tbar, int32_t ld, double *grad_a, double *grad_b,
                       double *grad_c, double *grad_d, double *grad_t0, int32_t ld0);
/* The same outputs through the adjoint of the whole step (u, v sweeps with
 * their coefficient dependence on the velocities, p' sweeps, correction,
 * energy): the exact gradient of the recorded K-step map under buoyancy.
 * Needs ri_x = 0 (upstream's u equation never reads T).  dJ/dT(initial) is
 * taken with u, v, p of the start held fixed and T_ant equal to T. */
int ixh_adjoint_coupled(ixh_state *s, const double *tbar, int32_t ld, double *grad_a, double *grad_b,
                        double *grad_c, double *grad_d, double *grad_t0, int32_t ld0);

#ifdef __cplusplus
}
#endif
#endif /* IXCHEL2D_H */
