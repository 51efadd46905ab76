/*
 * giraffe_amd.h -- C ABI of the MI355X-native Newton-step hot path of GIRAFFE's
 * static analysis (Beam_1 / Shell_1 element evaluation, global scatter, tangent
 * solve, Newton update / commit / rollback and the Static load-step loop).
 *
 * Plain C types only: pointers, sizes, doubles, ints.  Every entry point returns
 * an int status (GF_OK == 0, negative on error) and never throws.  Host arrays
 * passed in are borrowed for the duration of the call only.
 *
 * Reference interfaces each group replaces (GIRAFFE, alfredogneto/GIRAFFE):
 *   model set-up      Database::PreCalc src/Database.cpp:704-757,
 *                     Solution::DOFsActive / SetGlobalDOFs src/Solution.cpp:121-224, 40-118
 *   gf_begin_step     Solution::Zeros src/Solution.cpp:798-814 and the load-time
 *                     bookkeeping of Static::Solve src/Static.cpp:178-191
 *   gf_assemble       Solution::Clear / MountLocal / MountElementLoads / MountLoads /
 *                     MountGlobal / MountSparse src/Solution.cpp:227-387, 833-863,
 *                     Element::Mount / MountGlobal src/Element.h:53-76,
 *                     the sign flip src/Static.cpp:210 and the max-norm scans of
 *                     ConvergenceCriteria::EstablishResidualCriteria /
 *                     CheckResidualConvergence src/ConvergenceCriteria.cpp:187-288, 460-598
 *   gf_solve          Matrix sparsesystem(SparseMatrix&, Matrix&, int* info_fail, int, int)
 *                     src/SparseMatrix.h:28, src/SparseMatrix.cpp:74-166
 *   gf_update         Solution::UpdateDisps src/Solution.cpp:390-423 and the scans of
 *                     ConvergenceCriteria::CheckGLConvergence src/ConvergenceCriteria.cpp:289-459
 *   gf_commit         Solution::SaveConfiguration src/Solution.cpp:426-529
 *                     (Node::SaveConfiguration src/Node.cpp:325-384,
 *                      Beam_1::SaveLagrange src/Beam_1.cpp:1494-1506,
 *                      Shell_1::SaveLagrange src/Shell_1.cpp:1650-1664)
 *   gf_rollback       Solution::RestoreConfiguration src/Solution.cpp:532-574
 *   gf_static_solve   bool Static::Solve() src/Static.cpp:140-311
 */
#ifndef GIRAFFE_AMD_H
#define GIRAFFE_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define GF_OK            0
#define GF_E_ARG        -1   /* bad argument / inconsistent model description   */
#define GF_E_DEVICE     -2   /* HIP runtime error                               */
#define GF_E_NOMEM      -3   /* device or host allocation failed                */
#define GF_E_SINGULAR   -4   /* zero or non-finite pivot (sparsesystem info_fail) */
#define GF_E_STATE      -5   /* call out of sequence                            */
#define GF_E_UNSUPPORTED -6  /* feature not built                               */
#define GF_E_CAPACITY   -7   /* request exceeds a fixed capacity                */

/* ---- precision of the tangent path -------------------------------------- */
#define GF_PREC_FP64   0     /* tangent assembled and factored in FP64          */
#define GF_PREC_MIXED  1     /* FP32 tangent + FP32 factor, FP64 refinement     */

/* ---- gravity factor of a solution step (BoolTable::GetLinearFactorAtCurrentTime,
 * src/BoolTable.cpp:84-109: the gravity BoolTable's entries for this and the
 * previous solution step) -------------------------------------------------- */
#define GF_GRAVITY_OFF        0  /* off now and before: factor 0                   */
#define GF_GRAVITY_RAMP_UP    1  /* on now, off before: 0 -> 1 across the window   */
#define GF_GRAVITY_ON         2  /* on now and before: factor 1                    */
#define GF_GRAVITY_RAMP_DOWN  3  /* off now, on before: 1 -> 0 across the window   */

/* ---- load kinds (src/NodalLoad.cpp, src/NodalFollowerLoad.cpp, src/ShellLoad.cpp) */
#define GF_LOAD_NODAL           0  /* dead force + pseudo-moment (NodalLoad)      */
#define GF_LOAD_NODAL_FOLLOWER  1  /* follower force / moment (NodalFollowerLoad) */
#define GF_LOAD_SHELL_PRESSURE  2  /* follower pressure on Shell_1 (ShellLoad)    */

/* Per-element property records. */
#define GF_BEAM_NPROP   9   /* Beam_1: E, nu, rho, A, I11, I22, I12, It, T0 (pre-tension);
                               Pipe_1: EA, EI, GJ, GA, Rho (mass per length), 0, 0, 0, 0 */
#define GF_BEAM_KIND_BEAM1 0 /* Beam_1 (src/Beam_1.cpp)                         */
#define GF_BEAM_KIND_PIPE1 1 /* Pipe_1 (src/Pipe_1.cpp): Beam_1 kinematics, the pipe
                               section's D = diag(GA, GA, EA, EI, EI, GJ), buoyancy */
#define GF_SHELL_NPROP  4   /* E, nu, rho, thickness                              */

/* Model description: the hot-path subset of a GIRAFFE deck, already resolved
 * (node sets expanded, sections and materials folded into per-element records).
 * Node and element indices are 0-based. */
typedef struct gf_model_desc {
    int n_nodes;
    const double *node_xyz;          /* [n_nodes][3] reference coordinates          */
    const unsigned char *node_fix;   /* [n_nodes][6] 1 = DOF prescribed to zero     */

    int n_beams;
    const int *beam_conn;            /* [n_beams][3] nodes 1,2,3 of Beam_1            */
    const double *beam_cs;           /* [n_beams][9] element CS rows E1,E2,E3 (unit)  */
    const double *beam_prop;         /* [n_beams][GF_BEAM_NPROP]                      */

    int n_shells;
    const int *shell_conn;           /* [n_shells][6] corners 1-3, mid-sides 4-6      */
    const double *shell_prop;        /* [n_shells][GF_SHELL_NPROP]                    */

    int n_loads;
    const int *load_kind;            /* [n_loads] GF_LOAD_*                           */
    const int *load_set_off;         /* [n_loads+1] offsets into load_set_items       */
    const int *load_set_items;       /* node ids (nodal) or shell element ids (pressure) */
    const double *load_cs;           /* [n_loads][9] CS rows E1,E2,E3 of the load     */
    const int *load_flags;           /* [n_loads] ShellLoad AreaUpdate (0/1)          */
    const int *load_tab_off;         /* [n_loads+1] row offsets into tab_time/tab_val */
    const double *tab_time;          /* [rows] table instants                         */
    const double *tab_val;           /* [rows][6] FX FY FZ MX MY MZ (pressure: col 0) */

    double gravity[3];               /* GravityData G                                 */
    int gravity_on;                  /* 1: BoolTable ramp 0->1 over the solution     */

    /* Element numbers as in the deck (Element::number; together a permutation of
     * 1..n_beams+n_shells), used by the results writer.  NULL: Beam_1 numbered
     * 1..n_beams, then Shell_1 n_beams+1..n_beams+n_shells. */
    const int *beam_number;          /* [n_beams] or NULL                             */
    const int *shell_number;         /* [n_shells] or NULL                            */

    /* Pipe_1 and OceanData (src/Pipe_1.cpp:1278-1340, src/Environment.cpp:384-470).
     * beam_kind[e]: GF_BEAM_KIND_* (NULL: all Beam_1).  beam_buoy[e]: buoyant mass
     * per length rho_fluid * pi De^2 / 4 of a Pipe_1 (NULL: none); at a Gauss point
     * below the surface (depth = (x - surface) . G / |G| > 0) the self-weight uses
     * Rho - beam_buoy instead of Rho.  ocean_on: OceanData present; sea current
     * (Morison) loads are not on this path. */
    const int *beam_kind;            /* [n_beams] or NULL                             */
    const double *beam_buoy;         /* [n_beams] or NULL                             */
    int ocean_on;
    double ocean_surface[3];         /* SurfacePosition                               */
    double ocean_rho_fluid;          /* RhoFluid (kept for the deck writer)           */
} gf_model_desc;

/* Static solution-step and ConvergenceCriteria parameters, in the order of
 * Static::Read src/Static.cpp:43-130 and ConvergenceCriteria::Read
 * src/ConvergenceCriteria.cpp:64-165. */
typedef struct gf_static_params {
    double start_time, end_time, time_step, max_time_step, min_time_step;
    int max_it, min_it, conv_increase;
    double inc_factor;
    int sample;
    double force_tol, moment_tol, force_min, moment_min, constraint_min;
    double disp_tol, rot_tol, lag_tol, disp_min, rot_min, lag_min;
    double divergence_ref;
} gf_static_params;

/* Outcome of one Static::Solve.  step_iters (optional, caller-owned) receives
 * the Newton iteration count of every attempted load step, negated for steps
 * that diverged and were bisected. */
typedef struct gf_static_report {
    int aborted;                 /* Static::Solve returned false                  */
    int converged_steps;
    int diverged_steps;
    int total_iterations;        /* Newton iterations (assemblies)                */
    int total_solves;            /* tangent solves                                */
    double final_time;
    double max_rel_residual;     /* max over converged steps of the final
                                    max(|r_f|/ref_f, |r_m|/ref_m)                 */
    int *step_iters;             /* [step_iters_cap] or NULL                      */
    int step_iters_cap;
    int n_step_iters;            /* entries written                               */
    double t_assemble_ms, t_solve_ms, t_total_ms;
} gf_static_report;

/* Device options for a model. */
typedef struct gf_options {
    int precision;               /* GF_PREC_FP64 or GF_PREC_MIXED                 */
    int device;                  /* HIP device ordinal                            */
    void *stream;                /* hipStream_t to run on, or NULL for a new one  */
    int refine_max;              /* max FP64 refinement sweeps per solve (mixed)  */
    double refine_tol;           /* refinement ends when max|correction| <= refine_tol max|x|
                                    (default 1e-11: the increment is then accurate far below the
                                    Newton loop's own tolerances) */
    int refine_reuse;            /* mixed: keep the FP32 factor of an earlier tangent while the
                                    refinement against the current FP64 tangent converges fast
                                    enough; it is refactored when it does not (default 1) */
    int solver;                  /* GF_SOLVER_*: tangent solver of a model whose band is too
                                    wide for the block-tridiagonal form (default AUTO) */
} gf_options;

/* Tangent solvers.  Narrow bands always take the block-tridiagonal cyclic
 * reduction (except under FRONTS); wider ones the band LU or the multifrontal
 * factorisation over a nested-dissection ordering (PARDISO's own strategy,
 * src/SparseMatrix.cpp:115-117).  AUTO picks the multifrontal one when its
 * factorisation takes less than half the band LU's flops and fewer launches. */
#define GF_SOLVER_AUTO   0
#define GF_SOLVER_BAND   1
#define GF_SOLVER_FRONTS 2

typedef struct gf_model gf_model;

/* ---- library ------------------------------------------------------------- */
int gf_version(void);                                    /* 100*major + minor  */
const char *gf_last_error(void);                          /* thread-local text */
int gf_default_options(gf_options *opt);

/* ---- model lifetime: Database::PreCalc + DOFsActive + SetGlobalDOFs ------- */
int gf_model_create(const gf_model_desc *desc, const gf_options *opt, gf_model **out);
int gf_model_destroy(gf_model *m);
/* n_free, half-bandwidth, BTD block size (0 for the band layout), block count or
 * padded order, colours, bytes of the tangent storage, layout (0 BTD, 1 band, 2 band
 * assembly with the multifrontal solve), precision. */
int gf_model_info(const gf_model *m, long long *info_out /* [8] */);
/* Tangent solves so far, FP64 refinement sweeps (mixed precision), solves that
 * fell back to the FP64 factorisation, solves factored in FP64 directly by the
 * mixed path's back-off (after two fallbacks in a row the next 16 solves; a retry
 * that falls back again doubles the next back-off, up to 1,024; a converged FP32
 * solve resets it; a refinement that does not contract at all starts the 1,024
 * at once; with the multifrontal solver, a converged FP32 solve that cost more
 * than an FP64 factorisation and solve starts a back-off as two fallbacks do),
 * FP32 factorisations. */
int gf_model_stats(const gf_model *m, long long *stats_out /* [5] */);
/* Device bytes gf_model_create would allocate for this description (host-side
 * planning only; no device call).  Used for admission by the request server. */
int gf_model_footprint(const gf_model_desc *desc, const gf_options *opt, unsigned long long *bytes);
/* Host-side planning only (no device call; no counterpart in the reference, whose
 * ordering is PARDISO's): info_out = n_free, half-bandwidth, BTD block size (0 for
 * the band layout), 1 when the DOFs are numbered in reverse Cuthill-McKee node
 * order (chosen when it narrows the band of the caller's node order). */
int gf_model_plan(const gf_model_desc *desc, long long *info_out /* [4] */);
/* Host-side planning only (no device call; no counterpart in the reference, whose
 * supernodes are PARDISO's): the multifrontal plan the description gets under
 * GF_SOLVER_FRONTS.  info_out: fronts, levels, row-map entries, front matrix
 * elements, diagonal-inverse elements, algorithmic flops of one factorisation,
 * factorisation launches, solve launches.  front_out (NULL or [fronts][6]): own
 * DOFs k, padded own rows kp, padded order mp, level, parent front (-1 for a
 * root), offset of the row map; rows_out (NULL or [row-map entries]): band row of
 * each front row (-1 for padding); dof_out (NULL or [n_nodes][6]): band row of
 * each node DOF (-1 when not free).  Fronts are numbered level by level. */
int gf_front_plan(const gf_model_desc *desc, long long *info_out /* [8] */, int *front_out, int *rows_out,
                  int *dof_out);

/* ---- Newton step pieces --------------------------------------------------- */
/* Time window of the current solution step; the gravity BoolTable ramp
 * (BoolTable::GetLinearFactorAtCurrentTime src/BoolTable.cpp:84-109) reads it.
 * gf_static_solve sets it from its parameters. */
int gf_set_solution_window(gf_model *m, double start_time, double end_time);
/* Gravity factor mode of the next solution step (GF_GRAVITY_*); a model starts
 * in GF_GRAVITY_RAMP_UP when its description has gravity on, the first
 * solution step of a deck whose gravity table is on. */
int gf_set_gravity_mode(gf_model *m, int mode);
/* New load step: incremental displacements := 0, load time = t_last + dt. */
int gf_begin_step(gf_model *m, double t_last, double dt);
/* Assemble K and r = f_ext - f_int at the current state; res_out[0..2] =
 * max|r| over free translational DOFs, over free rotational DOFs, and 1.0 if
 * any entry is non-finite (>=1e300 or NaN) else 0.0. */
int gf_assemble(gf_model *m, double *res_out /* [3] */);
/* x = K^-1 r.  *info_fail = 1 when the factorisation meets a zero or non-finite
 * pivot (the caller then bisects), as sparsesystem reports PARDISO errors. */
int gf_solve(gf_model *m, int *info_fail);
/* d += x; inc_out = max|x_u|, max|x_rot|, max|d_u|, max|d_rot|, non-finite flag. */
int gf_update(gf_model *m, double *inc_out /* [5] */);
int gf_commit(gf_model *m);
int gf_rollback(gf_model *m);

/* Whole Static::Solve loop (adaptive steps, bisection, convergence criteria). */
int gf_static_solve(gf_model *m, const gf_static_params *p, gf_static_report *rep);
/* The solution-step sequence of the reference's main loop (src/main.cpp:53-62):
 * steps[i] with gravity mode gravity_modes[i] in turn on the same state,
 * stopping after a step that aborts; reps[i] for each step run, *n_run of them. */
int gf_static_solve_steps(gf_model *m, int n, const gf_static_params *steps, const int *gravity_modes,
                          gf_static_report *reps, int *n_run);

/* Return the model to the state right after gf_model_create (reference
 * configuration, identity Gauss-point rotations, no mixed-precision back-off),
 * keeping the device buffers, so the same resident model can be solved again. */
int gf_reset_state(gf_model *m);

/* ---- device timing ------------------------------------------------------------
 * HIP events on the model's stream around each phase; accumulated until read. */
#define GF_PH_ASSEMBLE 0   /* element kernels (all colours)            */
#define GF_PH_FACTOR   1   /* tangent factorisation                    */
#define GF_PH_SOLVE    2   /* forward / back substitution, refinement  */
#define GF_NPHASE      3
int gf_timing_enable(gf_model *m, int on);
/* ms_out[GF_NPHASE] accumulated milliseconds, launches_out[GF_NPHASE] kernel launches. */
int gf_timing_read(gf_model *m, double *ms_out, long long *launches_out, int reset);

/* Per-kernel timing (no counterpart in the reference; a measurement aid).  While
 * timing is enabled, every launch of a kernel class whose bit is set in `mask`
 * is bracketed by HIP events on the model's stream.  Read per class:
 * accumulated milliseconds, launches, and work units (GF_KC_* comments). */
#define GF_KC_BTD_LEVEL      0  /* btd_level_kernel; unit: block row eliminated     */
#define GF_KC_BTD_TAIL       1  /* btd_tail_kernel; unit: block row eliminated      */
#define GF_KC_BAND_PANEL     2  /* band_batch_inv_kernel; unit: flop of the batch inverses */
#define GF_KC_BAND_UPDATE    3  /* band_step_kernel; unit: band column eliminated   */
#define GF_KC_BAND_SWEEP     4  /* band_batch_apply_kernel, band_couple_kernel; unit: row
                                   solved (forward or backward; counted on the apply
                                   launches)                                        */
#define GF_KC_BEAM_ASSEMBLE  5  /* beam_assemble_kernel; unit: element              */
#define GF_KC_SHELL_ASSEMBLE 6  /* shell_element_kernel; unit: element             */
#define GF_KC_FRONT_STEP     7  /* front_step_kernel; unit: algorithmic flop of the partial
                                   factorisations, sum over eliminated columns c of
                                   2 (m - c - 1)^2 for a front of order m            */
#define GF_KC_FRONT_ASSEMBLE 8  /* front_assemble_kernel; unit: front entry written */
#define GF_KC_FRONT_SOLVE    9  /* front_fwd_own/fwd_upd/bwd_upd/bwd_own_kernel; unit: factor
                                   byte read (own triangle with the diagonal inverses,
                                   or the update panel)                              */
#define GF_NKC              10
int gf_kernel_probe(gf_model *m, unsigned mask);
/* ms_out, launches_out, units_out: GF_NKC entries each (any may be NULL). */
int gf_kernel_probe_read(gf_model *m, double *ms_out, long long *launches_out, double *units_out, int reset);

/* ---- results --------------------------------------------------------------- */
/* Per node: X Y Z of the last converged configuration and the Euler rotation
 * vector (Node::WriteResults src/Node.cpp:208-219). */
int gf_read_nodes(gf_model *m, double *out /* [n_nodes][6] */);
/* Per element, Gauss-point averages of the back-rotated generalised stresses and
 * strains of the last converged configuration, in description order:
 *   Beam_1  [n_beams][12]:  sigma_r (n_r, m_r), epsilon_r (eta_r, kappa_r), each
 *           the mean of the two Gauss points (Beam_1::WriteResults
 *           src/Beam_1.cpp:444-458);
 *   Shell_1 [n_shells][24]: n_r1, m_r1, n_r2, m_r2, eta_r1, kappa_r1, eta_r2,
 *           kappa_r2, each the mean of the three Gauss points (Shell_1::WriteResults
 *           src/Shell_1.cpp:624-666).
 * The reference keeps the values of its last Mount, which runs at the converged
 * configuration; here they are evaluated from the committed Gauss-point state,
 * which holds the same configuration.  Either pointer may be NULL. */
int gf_read_element_results(gf_model *m, double *beam_out, double *shell_out);
/* Configuration results text file of the last converged configuration, as
 * PostFiles::WriteConfigurationResults writes it (src/PostFiles.cpp:1429-1464):
 * one Node::WriteResults line per node (src/Node.cpp:208-219), then one
 * Beam_1 / Shell_1 WriteResults line per element in element-number order. */
int gf_write_results(gf_model *m, const char *path);
/* Per node and DOF: incremental displacement d, or residual r (0 where not free). */
int gf_read_disp(gf_model *m, double *out /* [n_nodes][6] */);
int gf_read_residual(gf_model *m, double *out /* [n_nodes][6] */);
/* Last solve's increment x per node and DOF (0 where not free). */
int gf_read_increment(gf_model *m, double *out /* [n_nodes][6] */);
/* Overwrite the incremental displacements (testing of assembly at chosen states). */
int gf_write_disp(gf_model *m, const double *in /* [n_nodes][6] */);
/* Dense copy of the free-free tangent K_AA in node-DOF space for small models:
 * out[(6*i+a)*(6*n_nodes) + 6*j+b]; entries of non-free DOFs are 0. */
int gf_read_tangent_dense(gf_model *m, double *out, long long cap);

/* ---- element-level parity entry (Beam_1::Mount on the device) -------------
 * For `count` independent elements: x[9] node coordinates of nodes 1,2,3,
 * cs[9], prop[GF_BEAM_NPROP], d[18] incremental node displacements (global),
 * gp[2][15] Gauss-point state (Q_i row-major 9, kappa_i 3, z'_i 3).
 * Outputs the global-frame 18x18 tangent (row-major) and internal force 18. */
int gf_beam_eval(int count, const double *x, const double *cs, const double *prop,
                 const double *d, const double *gp, double *K_out, double *f_out);

/* ---- standalone band solve -------------------------------------------------
 * Replaces bandsystem (src/BandMatrix.cpp:257-287, LAPACK dgbtrf/dgbtrs) with
 * the device band solver (block LU over 64 x 64 blocks, each diagonal block
 * inverted without interchanges) for one right-hand side.
 * ab_rows[i*(2hb+1) + (c - i + hb)] = A(i, c) for |i - c| <= hb (entries with c
 * outside [0, n) are ignored).  A pivot smaller than 1e-13 of its diagonal
 * block's largest entry is replaced by that size with its sign and the solve
 * goes on (PARDISO's static pivoting, as sparsesystem's mtype 11 defaults);
 * *info_fail = 1 on a non-finite pivot or an all-zero diagonal block. */
int gf_bandsystem(int n, int hb, const double *ab_rows, const double *b, double *x, int *info_fail);
/* Block-tridiagonal solve with the cyclic-reduction solver the narrow-band
 * models use (no counterpart in the reference; a test entry point): nblk block
 * rows of m x m row-major blocks, A[i] coupling row i to i-1 (A[0] ignored),
 * B[i] the diagonal, C[i] coupling to i+1 (C[nblk-1] ignored), 1 <= m <= 32.
 * info_fail = 1 on a zero or non-finite pivot. */
int gf_btdsystem(int m, int nblk, const double *A, const double *B, const double *C, const double *b, double *x,
                 int *info_fail);
/* gf_btdsystem on the caller's DEVICE buffers, allocating nothing.  No
 * counterpart in the reference; it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  A, B, C, b, x
 * and info_fail are device pointers (info_fail: one int, written 0 or 1); the
 * work is enqueued on `stream` (a hipStream_t, NULL for the null stream) and
 * the call returns without waiting.  The factor workspace lives in a fixed
 * device array of the library (9.4 MB on each GPU it runs on, held until the
 * process ends): nblk * m * m <= 131072, else GF_E_CAPACITY before any device
 * work.  Calls that overlap on one GPU share that array and race. */
int gf_btdsystem_noalloc(int m, int nblk, const double *A, const double *B, const double *C, const double *b,
                         double *x, int *info_fail, void *stream);

/* ---- element-level parity entry (Shell_1::Mount on the device) ------------
 * x[18] reference coordinates of nodes 1..6, prop[GF_SHELL_NPROP], d[36]
 * incremental displacements of the 6 nodes (6 per node; corner rotations are
 * ignored), gp[3][21] Gauss-point state (Q_i row-major 9, kappa_r1_i 3,
 * kappa_r2_i 3, z_x1_i 3, z_x2_i 3, local frame).  Outputs the global-frame
 * 27x27 tangent (row-major, DOF order of Shell_1::MountGlobal: u of nodes
 * 1..6, then alpha of nodes 4..6) and internal force 27. */
int gf_shell_eval(int count, const double *x, const double *prop, const double *d,
                  const double *gp, double *K_out, double *f_out);

/* ---- input decks (src/IO.cpp:193-770 dispatch, hot-path subset) -------------
 * gf_deck_read parses a GIRAFFE .inp deck restricted to what this path runs:
 * Nodes, NodeSets, ElementSets, Elements (Beam_1, Shell_1, Pipe_1), Materials
 * (Hooke), Sections (Tube, Rectangle), PipeSections, ShellSections
 * (Homogeneous), CoordinateSystems, Loads (NodalLoad, NodalFollowerLoad,
 * ShellLoad with NTimes tables), Constraints (NodalConstraint, the same in every
 * solution step), Environment (GravityData; OceanData whose sea current is off
 * or still), SolutionSteps (Static steps, run in sequence) and
 * ConvergenceCriteria; Monitor, PostFiles, SolverOptions, ExecutionData and
 * ConfigurationSave are skipped; anything else is refused with a message.
 * gf_deck_desc returns the resolved description (arrays owned by the deck) and
 * the first step's Static parameters; gf_deck_steps / gf_deck_step give every
 * step (start time = the previous step's end time,
 * Database::EvaluateStartEndTimes src/Database.cpp:760-770) with its gravity mode (GF_GRAVITY_*) for
 * gf_static_solve_steps.  gf_deck_write emits a description and one step as
 * such a deck. */
typedef struct gf_deck gf_deck;
int gf_deck_read(const char *path, gf_deck **out);
int gf_deck_desc(const gf_deck *deck, gf_model_desc *desc_out, gf_static_params *params_out);
int gf_deck_steps(const gf_deck *deck, int *n_steps);
int gf_deck_step(const gf_deck *deck, int step, gf_static_params *params_out, int *gravity_mode);
int gf_deck_destroy(gf_deck *deck);
int gf_deck_write(const char *path, const gf_model_desc *desc, const gf_static_params *params);

/* ---- request server ----------------------------------------------------------
 * No counterpart in the reference (it runs one Static::Solve per process); this is
 * the serving loop around the same Static drop-in: many independent solves
 * resident on one GPU, each with its own Newton state, tangent storage and
 * stream, admitted FIFO while fewer than max_inflight are resident and their
 * footprints fit under mem_cap, and advanced one Newton step at a time as their
 * device work completes.  One host thread drives a server. */
typedef struct gf_server gf_server;
typedef struct gf_server_options {
    int precision;                 /* GF_PREC_FP64 or GF_PREC_MIXED                  */
    int device;                    /* HIP device ordinal                             */
    int max_inflight;              /* resident requests at once (1..256)             */
    unsigned long long mem_cap;    /* device bytes the resident requests may hold    */
} gf_server_options;
int gf_default_server_options(gf_server_options *opt);
int gf_server_create(const gf_server_options *opt, gf_server **out);
int gf_server_destroy(gf_server *s);
/* Copies the description and parameters; *id_out identifies the request. */
int gf_server_submit(gf_server *s, const gf_model_desc *desc, const gf_static_params *p, long long *id_out);
/* Admits and advances requests for up to wait_us microseconds (returns earlier
 * when nothing is queued or resident).  Returns the number of requests that
 * finished during the call, or a negative status. */
int gf_server_pump(gf_server *s, int wait_us);
/* queued, resident, finished-not-taken, device bytes held by resident requests */
int gf_server_counts(const gf_server *s, long long *out /* [4] */);
/* Results of a finished request (then forgotten): report (step_iters filled up to
 * its cap), node results [n_nodes][6] as gf_read_nodes (may be NULL), and
 * times_ms[3] = submit->admit, submit->finish, admit->finish.  Returns the
 * request's status (GF_OK or the error that ended it). */
int gf_server_take(gf_server *s, long long id, gf_static_report *rep, double *nodes_out, double *times_ms);

#ifdef __cplusplus
}
#endif
#endif /* GIRAFFE_AMD_H */
