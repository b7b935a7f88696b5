/*
 * hsc.h -- C ABI of the MI355X HSC (Heterogeneous Subspace Correction) linear
 * solver.  Plain C: device pointers, sizes and status codes; no C++ or torch
 * types cross this boundary, no exceptions escape it.
 *
 * It replaces the reference's libuipc linear-solver plugin
 *   class LinearHSC final : public IterativeSolver
 *   (reproduction/hsc_backend/linear_hsc.cu:153-438), whose contract is
 *   IterativeSolver::do_build / do_solve
 *   (reproduction/hsc_backend/upstream/src/backends/cuda/linear_system/iterative_solver.h:19-21).
 *
 * Threading: one handle = one host thread = one HIP stream at a time.  Calls
 * on one handle are not thread-safe.  All device buffers are caller-owned
 * unless stated.  hsc_setup allocates the handle's workspaces (re-used, grown
 * only when a bigger system arrives), picks the persistent-PCG launch shape
 * (timed once per size class and kept in the tune table, see hsc_tune_file)
 * and sizes that launch's scratch; hsc_solve then allocates nothing and runs
 * no tuning (hsc_alloc_count lets a caller check this).
 */
#ifndef HSC_MI355X_H
#define HSC_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HSC_ABI_VERSION 3

/* Status codes: one per failure mode of the reference plugin. */
enum hsc_status {
  HSC_OK = 0,
  HSC_ERR_INVALID_ARGUMENT = 1,   /* bad config / DOF layout: linear_hsc.cu:188-189,199-200 */
  HSC_ERR_FEM_DIAGONAL = 2,       /* non-positive FEM 3x3 diagonal block: linear_hsc.cu:73-78,208 */
  HSC_ERR_UNEXPECTED_COUPLING = 3,/* coupling block not ABD<->FEM: linear_hsc.cu:230 */
  HSC_ERR_ABD_NOT_SPD = 4,        /* non-SPD 12x12 ABD block: linear_hsc.cu:268-269 */
  HSC_ERR_ABD_INVERSE = 5,        /* ABD Neumann doubling diverged / tol missed: linear_hsc.cu:292-297 */
  HSC_ERR_NONFINITE_FACTORS = 6,  /* non-finite P or C (or C not SPD): linear_hsc.cu:300 */
  HSC_ERR_BREAKDOWN = 7,          /* PCG pSp<=0 or rz<=0 or non-finite: linear_hsc.cu:370-371 */
  HSC_ERR_NOT_CONVERGED = 8,      /* max_iter reached: linear_hsc.cu:394 */
  HSC_ERR_NOT_DESCENT = 9,        /* b^T x <= 0: linear_hsc.cu:400-402 */
  HSC_ERR_FULL_RESIDUAL = 10,     /* full residual > full_residual_limit: linear_hsc.cu:433-435 */
  HSC_ERR_HIP = 11,               /* HIP runtime error: linear_hsc.cu:34-37 */
  HSC_ERR_NOT_SETUP = 12,         /* solve before setup (no reference counterpart) */
  HSC_ERR_DEVICE_TIMEOUT = 13,    /* a persistent kernel's bounded wait expired (no counterpart) */
  HSC_ERR_CAPACITY = 14,          /* a system exceeds a compiled-in limit (no counterpart) */
  HSC_ERR_STALE_BATCH = 15        /* a captured batch's handle was set up or reconfigured after
                                     the capture (no counterpart) */
};

/* Arithmetic of the bulk (matrix/vector) traffic. */
enum hsc_precision {
  HSC_PRECISION_FP64 = 0,         /* everything FP64 (the reference) */
  HSC_PRECISION_MIXED = 1         /* FP32 SpMV / low-rank / Neumann terms; FP64 dots, alpha/beta,
                                     x, r and a true FP64 residual every residual_replace_every */
};

/* How A^-1 (the ABD block inverse) is formed. */
enum hsc_abd_inverse {
  HSC_ABD_EXACT = 0,              /* FP64 Cholesky of A (core.py:149-150 exact_affine=True) */
  HSC_ABD_NEUMANN = 1,            /* the plugin's inverse (default): block Neumann polynomial of
                                     neumann_order terms and weight omega (core.py:74-92;
                                     linear_hsc.cu:264-275), doubled to abd_inverse_tol when
                                     adaptive_abd_inverse = 1 (linear_hsc.cu:276-297) */
  HSC_ABD_NEUMANN_ADAPTIVE = 2    /* polynomial + doubling to abd_inverse_tol (linear_hsc.cu:276-297) */
};

/* Preconditioner. */
enum hsc_preconditioner {
  HSC_PRECOND_NEUMANN = 0,        /* block-Jacobi Neumann series of D (linear_hsc.cu:330-343) */
  HSC_PRECOND_LEARNED = 1,        /* the CHEBYSHEV passes on the interval ratio a graph MLP picks
                                     from D's node features at setup; needs hsc_set_learned_weights
                                     (no counterpart) */
  HSC_PRECOND_CHEBYSHEV = 2       /* same passes, per-pass weights from the Chebyshev roots on
                                     [b/poly_interval_ratio, b], b = 1.2 x a power-iteration estimate
                                     of lambda_max(Dinv D) (no counterpart) */
};

/* Name of a preconditioner / precision as written to the JSONL row
 * ("neumann", "learned", "chebyshev"; "fp64", "mixed"); "unknown" otherwise. */
const char* hsc_preconditioner_name(int32_t preconditioner);
const char* hsc_precision_name(int32_t precision);

/* Solver configuration.  The first eight fields are the reference's scene
 * config keys and defaults (integration.patch:34-55; linear_system/tol_rate
 * default 1e-3 in scene_default_config.cpp:146-151); the rest are additions. */
typedef struct hsc_config {
  int32_t max_rank;              /* linear_system/hsc/max_rank            (128)  */
  int32_t neumann_order;         /* linear_system/hsc/neumann_order       (5)    */
  int32_t max_iter;              /* linear_system/hsc/max_iter            (2000) */
  int32_t adaptive_abd_inverse;  /* linear_system/hsc/adaptive_abd_inverse(1)    */
  double abd_inverse_tol;        /* linear_system/hsc/abd_inverse_tol     (1e-5) */
  double full_residual_limit;    /* linear_system/hsc/full_residual_limit (0.1)  */
  double omega;                  /* linear_system/hsc/omega               (0.7)  */
  double tol_rate;               /* linear_system/tol_rate                (1e-3) */
  /* --- additions --- */
  int32_t abd_inverse;           /* enum hsc_abd_inverse (HSC_ABD_NEUMANN)        */
  int32_t precision;             /* enum hsc_precision (HSC_PRECISION_FP64)       */
  int32_t residual_replace_every;/* mixed: FP64 true-residual period (25); 0 = off */
  int32_t true_residual_every_iteration; /* 1 = stop test on the true residual each
                                     iteration (core.py:120-132 schedule); 0 = recursive
                                     residual + one true check (linear_hsc.cu:374-387) */
  double aca_rtol;               /* ACA stop on ||E-UV^T||_F/||E||_F (core.py:22-50); 0 = none */
  int32_t lanes_per_row;         /* SpMV lanes per block row: 0 = autotuned, else 4/8/16 */
  int32_t wg_per_cu;             /* persistent workgroups per CU: 0 = autotuned, else 1..4 */
  int32_t preconditioner;        /* enum hsc_preconditioner (HSC_PRECOND_NEUMANN)  */
  int32_t check_acceptance;      /* 1 = fail on non-descent / full residual (reference); 0 = report only */
  int32_t schedule;              /* persistent-PCG work split: 0 = autotuned; 1 = latency (grid sized by the
                                    SpMV row tiles, 128-node elementwise tiles); 2 = throughput (grid covers
                                    row and low-rank tiles, split by tile count, 512-node tiles); 3 = latency
                                    on row tiles cut along D's connected components (at 4 lanes per row;
                                    otherwise as 1), so that no tile straddles two bodies */
  double poly_interval_ratio;    /* HSC_PRECOND_CHEBYSHEV: b / a of the fitted interval (100) */
  int32_t pass_sync;             /* latency schedule, order > 2: 1 = each row tile waits only for its
                                    neighbour tiles between preconditioner passes; 0 = grid barriers;
                                    2 (default) = pass pairs where they apply (one system, mixed precision,
                                    4 lanes per row, rows of at most 20 blocks): each row tile also computes
                                    its halo rows' pass in LDS, so it reads its neighbours' values once per
                                    two passes, as tagged words (no flags or waits), else as 1.
                                    Same arithmetic in every mode (bitwise-equal results) */
  int32_t pipelined;             /* 1 = the pipelined PCG loop where it applies (latency schedule, 4 lanes per
                                    row, Chebyshev/learned or Neumann passes of order >= 2, recursive-residual
                                    stopping): one grid barrier per iteration; q = D p kept by recurrence in
                                    FP64 (no counterpart; default 0 = the reference's loop) */
  int32_t operator_mapping;      /* kernels of the standalone operators (hsc_apply_*) and of the solve's
                                    acceptance D x: 0 (default) = by working set: the streaming kernels
                                    (64-block chunks per wavefront) when the operation reads 256 MiB or more
                                    (beyond the Infinity Cache), else the row mapping; 1 = row mapping
                                    (16 lanes per row; 8 for FP32 passes); 2 = streaming */
  int32_t reserved0;
} hsc_config;

/* Device view of the assembled Hessian: the stored triangle of 3x3 blocks
 * (libuipc CBCOOMatrixView<double,3> after ge2sym/convert: one block per
 * unordered block pair, diagonal blocks full).  Mirrors SolvingInfo::A()
 * (integration.patch:7,17-20). */
typedef struct hsc_bcoo3_view {
  const int32_t* row_indices;    /* [triplet_count] block row (device) */
  const int32_t* col_indices;    /* [triplet_count] block col (device) */
  const double* values;          /* [triplet_count*9] 3x3 column-major (device) */
  int64_t triplet_count;
  int64_t block_dim;             /* number of block rows = N/3 */
} hsc_bcoo3_view;

/* Per-solve statistics: the fields of the reference's hsc_stats.jsonl row
 * (linear_hsc.cu:421-431) plus additions. */
typedef struct hsc_stats {
  int32_t fem_dofs, abd_dofs, active_fem_dofs, rank;
  double aca_relative_error;
  int32_t pcg_iterations;
  double schur_relative_residual, full_relative_residual;
  double setup_ms, total_ms;
  int32_t active_mesh_contacts;  /* contact counters of the row (linear_hsc.cu:410-414): not known to the
                                    solver; a solve writes 0 and the caller fills them before
                                    hsc_append_jsonl */
  int32_t active_ground_contacts;
  int32_t adaptive_abd_inverse;  /* 1 when the adaptive doubling of the ABD inverse ran (abd_inverse =
                                    NEUMANN with adaptive_abd_inverse = 1), else 0 */
  int32_t neumann_doublings;
  double abd_inverse_residual, full_residual_limit;
  int32_t accepted, fallback;    /* fallback: learned preconditioner fell back to Neumann */
  int32_t status;                /* enum hsc_status of this solve */
  /* --- additions --- */
  int32_t precision, preconditioner, true_residual_checks, restarts;
  double pcg_ms;                 /* device time of the persistent PCG launch */
  double descent;                /* b^T x */
  double gbps;                   /* algorithmic bytes of the PCG launch / pcg_ms, GB/s (0 when not timed) */
  double roofline_fraction;      /* gbps / 8000 (MI355X HBM peak) */
  double fp64_oracle_error;      /* |x - x_ref| / |x_ref| against an FP64 reference solution: not known to
                                    the solver; a solve writes -1 (JSONL null) and a caller that has a
                                    reference fills it before hsc_append_jsonl */
  int32_t learned_used;          /* 1 when learned parameters shaped this solve's preconditioner */
  int32_t fallback_reason;       /* 0 none; 1 no learned weights loaded; 2 network output rejected
                                    (non-finite or outside its bounds): plain preconditioner used */
} hsc_stats;

typedef struct hsc_handle hsc_handle;
typedef void* hsc_stream;        /* hipStream_t (NULL = default stream) */

int32_t hsc_abi_version(void);
void hsc_config_default(hsc_config* cfg);

/* Create / destroy.  Mirrors do_build (linear_hsc.cu:172-193): validates the
 * config (HSC_ERR_INVALID_ARGUMENT like its UIPC_ASSERT at :188-189). */
int32_t hsc_create(const hsc_config* cfg, int32_t device, hsc_handle** out);
void hsc_destroy(hsc_handle* h);
const char* hsc_last_error(const hsc_handle* h);
int32_t hsc_set_config(hsc_handle* h, const hsc_config* cfg);

/* Setup for one Newton system (linear_hsc.cu:195-304): split A into D | E |
 * A_abd, invert the FEM 3x3 diagonal blocks, ACA(E) = U V^T, A^-1, C, all on
 * the device.  qo/qm: ABD DOF offset/count (qm % 12 == 0); fo/fn: FEM DOF
 * offset/count (fn % 3 == 0); qm + fn = 3*A->block_dim. */
int32_t hsc_setup(hsc_handle* h, const hsc_bcoo3_view* A, int32_t qo, int32_t qm,
                  int32_t fo, int32_t fn, hsc_stream stream);

/* Solve H x = b after hsc_setup (linear_hsc.cu:349-435): Schur PCG on the FEM
 * DOFs, back-substitution of the ABD DOFs, descent and full-residual checks.
 * b, x: device FP64 vectors of N = qm + fn; x is fully overwritten.
 * Synchronous (returns when x and *stats are final). */
int32_t hsc_solve(hsc_handle* h, const double* b, double* x, hsc_stats* stats,
                  hsc_stream stream);

/* do_solve in one call: hsc_setup + hsc_solve (setup_ms covers the setup). */
int32_t hsc_do_solve(hsc_handle* h, const hsc_bcoo3_view* A, int32_t qo, int32_t qm,
                     int32_t fo, int32_t fn, const double* b, double* x,
                     hsc_stats* stats, hsc_stream stream);

/* Append one JSONL row in the reference's hsc_stats.jsonl schema
 * (linear_hsc.cu:420-432) to `path`; frame/newton label the row.  After the
 * reference's fields come precision, preconditioner (by name), the learned
 * fields, status, pcg_ms, gbps, roofline_fraction and fp64_oracle_error
 * (null when negative). */
int32_t hsc_append_jsonl(const hsc_stats* s, const char* path, int32_t frame, int32_t newton);

/* ---- launch-shape tuning and allocation accounting (no reference counterpart) ----
 * The tune table maps a size class (nodes, blocks per row, rank, precision,
 * batch, preconditioner, order, pass sync, device) to the fastest persistent
 * PCG launch shape.  It is loaded from and appended to the file named here
 * (default: $HSC_TUNE_FILE, else $HOME/.cache/hsc_mi355x/tune_v4.txt; NULL or
 * "" = keep it in memory only).  hsc_tune runs the timing for h's current
 * system now (hsc_setup does it too when the class is new).  hsc_alloc_count
 * is the number of device and pinned-host allocations the library has made
 * in this process. */
int32_t hsc_tune_file(const char* path);
int32_t hsc_tune(hsc_handle* h, hsc_stream stream);
int64_t hsc_alloc_count(void);

/* ---- batched solves (no reference counterpart: many independent systems
 * of uneven size in ONE persistent launch) ----
 * handles[i] must each have been hsc_setup; all share one config precision
 * and preconditioner. b[i], x[i]: device vectors of system i.  The first call
 * for a batch size class times its launch shapes (the tune table), and the
 * launch scratch lives in handles[0], grown when a larger batch arrives: a
 * caller that cannot allocate captures the batch once (hsc_batch_create) and
 * replays it. */
int32_t hsc_solve_batched(hsc_handle* const* handles, int32_t count,
                          const double* const* b, double* const* x,
                          hsc_stats* stats, hsc_stream stream);

/* ---- operator access for parity tests (device vectors of fn FEM DOFs) ---- */
int32_t hsc_apply_D(hsc_handle* h, const double* x, double* y, hsc_stream stream);
int32_t hsc_apply_schur(hsc_handle* h, const double* x, double* y, hsc_stream stream);
int32_t hsc_apply_precond(hsc_handle* h, const double* r, double* z, hsc_stream stream);
/* One block-Jacobi pass y = x + w Dinv (r - D x) over the FEM part, with the
 * streaming pass kernel (no reference counterpart: it is one Horner step of
 * the series that linear_hsc.cu:329-342 sums term by term).  precision
 * HSC_PRECISION_FP64: x and y are double, with the FP64 values and inverse
 * blocks; HSC_PRECISION_MIXED: x and y are float, with the FP32 copies (r is
 * double in both).  x and y must not overlap. */
int32_t hsc_apply_pass(hsc_handle* h, int32_t precision, double w, const void* x, const double* r, void* y,
                       hsc_stream stream);

/* Setup summary and host copies of the small factors (any pointer may be NULL).
 * U: qm x rank column-major; V: active x rank column-major; C: rank x rank;
 * P: qm x qm column-major (A^-1 approximation); active: active FEM DOFs
 * (relative to fo); aca_history: rank+1 entries. */
typedef struct hsc_setup_info {
  int32_t fem_dofs, abd_dofs, fem_nodes, nnzb, active_fem_dofs, rank;
  int32_t lanes_per_row, workgroups, neumann_doublings, abd_block_diagonal;
  double aca_relative_error, abd_inverse_residual;
  int32_t schedule;              /* lanes_per_row, workgroups (per CU), schedule: the last solve's launch */
  int32_t pass_sync;             /* 1 when the last solve's preconditioner passes synchronised per tile */
  double lambda_max;             /* power-iteration estimate of lambda_max(Dinv D) (Chebyshev or learned; else 0) */
  double poly_ratio;             /* interval ratio b/a of the polynomial passes (Chebyshev: the config's; learned:
                                    the network's); 0 for the Neumann series or the learned fallback */
  double learned_out[3];         /* learned: the network's node-mean outputs (log10(b/a) - 2.5 at degrees 8, 12, 16) */
  int32_t learned_used;          /* 1 when the network's interval is in use */
  int32_t fallback_reason;       /* 2 when the network's output was rejected (Neumann series used) */
  int32_t pipelined;             /* 1 when the last solve's launch ran the pipelined PCG loop */
  int32_t pass_pairs;            /* 1 when the last solve ran its preconditioner passes in pairs (pass_sync 2,
                                    and every row tile's halo fit: otherwise one wait per pass) */
  int32_t streaming_operators;   /* 1 when operator_mapping selects the streaming kernels for D x on this
                                    system */
  int32_t row_tiles;             /* row tiles of schedule 3, cut along D's connected components (a component
                                    of 128 rows or more starts a tile; tiles of at most 128 rows); 0 = none
                                    (too many rows or components: schedule 3 runs as 1) */
} hsc_setup_info;
int32_t hsc_get_setup_info(const hsc_handle* h, hsc_setup_info* out);

/* ---- FEM Hessian generator (before the path) ----
 * The per-tet log-Neo-Hookean Hessian of the reference's FEM constitution
 * (log_neo_hookean_function.h:93-150, log_neo_hookean_3d.cu:62-135 of the
 * overlays): F = Ds Dm^-1 (Ds = [x1-x0, x2-x0, x3-x0]), the 9x9 d2Psi/dF2 with
 * vec(F) column-major, projected SPD by bounded cyclic Jacobi when
 * project_spd = 1, scaled by volume * dt^2 and pulled back to the 10 node
 * blocks a <= b of each tet.  Writes 10 * n_tets general triplets (row = the
 * smaller node index, 3x3 column-major) for hsc_assemble_bcoo.  tets:
 * [n_tets][4]; x: [nodes][3]; Dm_inv: [n_tets][9] column-major; mu, lambda,
 * volumes: [n_tets].  Device memory; synchronises `stream`. */
int32_t hsc_fem_hessian(const int32_t* tets, const double* x, const double* Dm_inv, const double* mu,
                        const double* lambda, const double* volumes, double dt, int64_t n_tets, int32_t project_spd,
                        int32_t* out_rows, int32_t* out_cols, double* out_values, hsc_stream stream);

/* ---- ABD-FEM coupling blocks (before the path) ----
 * The reference's off-diagonal subsystem assembly
 * (abd_fem_linear_subsystem.cu:18-61): each contact block H (3x3 column-major,
 * rows on ABD vertex abd_vertex[i], columns on FEM vertex fem_vertex[i])
 * becomes J^T H (12x3) with J = d x / d q the affine Jacobian at the vertex's
 * rest position xbar (x = p + A xbar, q = (p, rows of A)), written as 4
 * triplets (abd_block_offset + 4 body + k, fem_block_offset + fem vertex),
 * values zero when body_fixed[body] or vertex_fixed[fem vertex] is nonzero.
 * v2b, xbar: [n_abd_vertices](x3); body_fixed: [n_bodies]; vertex_fixed:
 * [n_fem_vertices]; outputs: 4 * count triplets, the E blocks of the plugin's
 * input for hsc_assemble_bcoo.  An index out of range returns
 * HSC_ERR_INVALID_ARGUMENT.  Device memory; synchronises `stream`; calls in
 * one process are serialised (they share a per-device error word). */
int32_t hsc_abd_fem_coupling(const int32_t* abd_vertex, const int32_t* fem_vertex, const double* hess,
                             int64_t count, const int32_t* v2b, const double* xbar, int32_t n_abd_vertices,
                             const int32_t* body_fixed, int32_t n_bodies, const int32_t* vertex_fixed,
                             int32_t n_fem_vertices, int32_t abd_block_offset, int32_t fem_block_offset,
                             int32_t* out_rows, int32_t* out_cols, double* out_values, hsc_stream stream);

/* ---- assembly before the path ----
 * Replaces the caller-side conversion that produces the plugin's input
 * (libuipc MatrixConverter::ge2sym + convert, global_linear_system.cu:232-234):
 * `count` general 3x3-block triplets (block row, block column, 9 values
 * column-major; duplicates allowed) -> the upper-triangle BCOO of unique
 * blocks the plugin receives, sorted by (row, column), duplicates summed in
 * input order.  A lower block (row > column) is folded onto (column, row)
 * transposed when fold_lower = 1, or dropped when 0 (the input holds both
 * triangles of a symmetric H).  All arrays are device memory; the outputs
 * need `count` entries; *out_count receives the number of unique blocks.
 * Synchronises `stream`. */
int32_t hsc_assemble_bcoo(const int32_t* rows, const int32_t* cols, const double* values, int64_t count,
                          int32_t fold_lower, int32_t* out_rows, int32_t* out_cols, double* out_values,
                          int64_t* out_count, hsc_stream stream);

/* ---- assembly without allocation (no counterpart in the reference) ----
 * hsc_assemble_bcoo for callers that cannot allocate device memory (code
 * under stream capture, serving loops and harnesses that forbid hipMalloc):
 * the same kernels and the same results, with the workspace in a 64 MiB
 * device array of the library (one per GPU, held until the process ends)
 * instead of memory the library allocates.  At most 1,048,576 triplets per
 * call (HSC_ERR_CAPACITY beyond, before any device work).  Every call on a
 * GPU shares that GPU's array: calls from one process are serialised by a
 * lock, and calls that overlap on one GPU from elsewhere race.  Nothing in
 * the array needs clearing between calls (every part is written before it
 * is read).  Synchronises `stream`. */
int32_t hsc_assemble_bcoo_noalloc(const int32_t* rows, const int32_t* cols, const double* values, int64_t count,
                                  int32_t fold_lower, int32_t* out_rows, int32_t* out_cols, double* out_values,
                                  int64_t* out_count, hsc_stream stream);

/* ---- measurement (no reference counterpart) ----
 * Times `repeats` back-to-back launches of one standalone operator kernel on
 * the set-up handle's D, between HIP events on `stream`: kernel 0 = FP64
 * y = D x, 1 = one FP64 block-Jacobi pass y = x + w Dinv (r - D x), 2 = the
 * same pass with FP32 values, inverse blocks and vectors (FP64 r), each with
 * `lanes_per_row` (4, 8, 16) lanes per block row (the persistent solve's
 * row mapping); kernels 3, 4, 5 = the same three operations with the
 * streaming kernels that hsc_apply_D, hsc_apply_schur, hsc_apply_precond,
 * hsc_apply_pass and the solve's acceptance D x run, where `lanes_per_row`
 * is instead the number of 64-block chunks per wavefront (1..64; 0 = the
 * library's choice).  The operands are the handle's own scratch vectors
 * (zeroed first; their contents are lost).  *ms_per_launch = elapsed time /
 * repeats. */
int32_t hsc_time_kernel(hsc_handle* h, int32_t kernel, int32_t lanes_per_row, int32_t repeats, hsc_stream stream,
                        double* ms_per_launch);
int32_t hsc_get_factors(hsc_handle* h, double* U, double* V, double* C, double* P,
                        int32_t* active, double* aca_history);

/* ---- captured batch (no reference counterpart) ----
 * The whole batched solve of `count` set-up handles for fixed b/x buffers
 * (right-hand-side reductions, one persistent PCG launch over all systems,
 * back-substitution and acceptance reductions) captured once into a
 * hipGraph.  hsc_batch_launch enqueues one replay on `stream` (no host
 * synchronisation); hsc_batch_stats waits for it and fills one hsc_stats per
 * system (pcg_ms is not measured inside a replay and reads 0).  The batch is
 * valid until one of its handles is set up again, reconfigured or
 * destroyed: after a setup or reconfiguration hsc_batch_launch and
 * hsc_batch_stats return HSC_ERR_STALE_BATCH before any device work (a
 * destroyed handle cannot be detected; destroy the batch first).  Handles must share device, precision, neumann_order, omega and
 * preconditioner (Neumann or Chebyshev; with Chebyshev the batch uses the
 * weights of the handle with the largest spectrum bound, valid for all). */
typedef struct hsc_batch hsc_batch;
int32_t hsc_batch_create(hsc_handle* const* handles, int32_t count, const double* const* b, double* const* x,
                         hsc_stream stream, hsc_batch** out);
int32_t hsc_batch_launch(hsc_batch* batch, hsc_stream stream);
int32_t hsc_batch_stats(hsc_batch* batch, hsc_stats* stats, hsc_stream stream);
int32_t hsc_batch_info(const hsc_batch* batch, int32_t* lanes_per_row, int32_t* wg_per_cu, int32_t* workgroups,
                       int32_t* cooperative_node, int32_t* schedule);
void hsc_batch_destroy(hsc_batch* batch);

/* ---- learned preconditioner (no reference counterpart) ----
 * preconditioner = HSC_PRECOND_LEARNED keeps the Chebyshev-weighted
 * block-Jacobi passes of HSC_PRECOND_CHEBYSHEV (neumann_order passes, one
 * fixed SPD polynomial of Dinv D per solve, so the persistent PCG runs
 * unchanged, in either precision) and lets a per-node graph MLP over D's
 * node graph (2 message-passing layers, hidden 64, BF16 weights and
 * activations, FP32 accumulation) choose the polynomial's interval: at
 * setup the network reads each node's features of D (hsc_get_learned_features),
 * and the node mean of its 3 outputs predicts log10(b/a) - 2.5 for degrees
 * 8, 12 and 16 (other degrees interpolate).  An output that is not finite
 * or outside log10(b/a) in [1, 4.5] is rejected: the plain Neumann series
 * is used (weight min(omega, 2/b), SPD at any order) and the stats report
 * fallback = 1, fallback_reason = 2.  hsc_set_learned_weights loads the
 * network's hsc_mlp_param_count() float32 parameters (host or device
 * pointer); setup without them fails with HSC_ERR_INVALID_ARGUMENT. */
int32_t hsc_set_learned_weights(hsc_handle* h, const void* theta_f32, int64_t bytes);

/* Node graph of D (neighbour CSR without self; device or host destination). */
int32_t hsc_get_node_graph(hsc_handle* h, int32_t* nbr_ptr, int32_t* nbr, int64_t* n_edges);

/* The network's input for h's set-up system: fem_nodes x 16 bfloat16 values
 * per node (device or host destination): D_ii / s_i (9, row-major, s_i =
 * tr(D_ii) / 3), log10(s_i / mean of the neighbours' s_j) / 2,
 * sum_{j != i} |D_ij|_F / |D_ii|_F / 4, neighbour count / 16, 1 when the
 * node is coupled to the ABD block (else 0), log10(fem_nodes) / 6, 0, 0.
 * Only D enters: the network runs at setup beside the power iteration. */
int32_t hsc_get_learned_features(hsc_handle* h, uint16_t* X);

/* The network object used for training on the GPU (hand-written BF16
 * forward/backward kernels + Adam) over one node graph (for training, the
 * disjoint union of the training systems' graphs).  X: [B x nb x 16] bf16
 * features; T: [B x nb x 3] fp32 targets; den: [B] per-sample normalisers
 * (loss = mean_b sum_i |o_bi - t_bi|^2 / den_b); O: [B x nb x 3] fp32
 * outputs.  Parameters are a flat float32 vector of hsc_mlp_param_count()
 * values. */
typedef struct hsc_mlp hsc_mlp;
int32_t hsc_mlp_param_count(void);
int32_t hsc_mlp_create(int32_t device, int32_t nb, const int32_t* nbr_ptr, const int32_t* nbr, int64_t n_edges,
                       int32_t max_batch, hsc_mlp** out);
void hsc_mlp_destroy(hsc_mlp* m);
int32_t hsc_mlp_set_params(hsc_mlp* m, const float* theta, int32_t device_ptr);
int32_t hsc_mlp_get_params(hsc_mlp* m, float* theta_host);
int32_t hsc_mlp_forward(hsc_mlp* m, const uint16_t* X, int32_t B, float* O, hsc_stream stream);
int32_t hsc_mlp_gradients(hsc_mlp* m, const uint16_t* X, const float* T, const float* den, int32_t B,
                          float* grad_host, float* loss, hsc_stream stream);
int32_t hsc_mlp_train_step(hsc_mlp* m, const uint16_t* X, const float* T, const float* den, int32_t B, float lr,
                           float* loss, hsc_stream stream);

#ifdef __cplusplus
}
#endif
#endif /* HSC_MI355X_H */
