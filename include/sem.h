/* sem.h -- C-ABI of libsem_hip.so, the MI355X (gfx950) implementation of
 * SchrodingerSEM's matrix-free spectral-element hot path.
 *
 * Conventions
 *   - Every function returns an int status: SEM_OK (0) or a negative code;
 *     sem_last_error() gives a thread-local message. No exception crosses the
 *     ABI. Solver non-convergence is NOT an error: the call returns SEM_OK with
 *     sem_solver_results.success = 0, as the reference returns
 *     solver_results{success=false} (minres.hpp:102-106, pcg.hpp:91-95).
 *   - Host arrays use the reference's element-local layout: element e's N x N
 *     block at offset e*N*N, entry (i, j) at i + j*N (i along xi, j along eta;
 *     column-major arma::Mat, types.hpp:9-10), N = p + 1.
 *   - A mesh handle owns its device mirrors, its HIP stream and workspaces.
 *     Calls on one handle are not re-entrant; separate handles may run
 *     concurrently on separate host threads.
 *   - FP64 everywhere at the ABI. Internally the solver may run its inner
 *     operator applies in FP32 (SEM_PREC_MIXED), always stopped by an FP64
 *     residual.
 */
#ifndef SEM_H
#define SEM_H

#ifdef __cplusplus
extern "C" {
#endif

enum sem_status {
    SEM_OK = 0,
    SEM_EINVAL = -1, /* invalid argument */
    SEM_EIO = -2,    /* mesh file missing or unreadable (load_mesh.hpp:93,102,118) */
    SEM_EHIP = -3,   /* HIP runtime error (incl. no GPU) */
    SEM_ENOMEM = -4,
    SEM_EUNSUP = -5  /* unsupported configuration (e.g. degree above the compiled range) */
};

/* Element operators of the galerkin_op action (openmp_impl/galerkin.hpp:24-35). */
enum sem_op_kind {
    SEM_OP_LAPLACE = 0,   /* glaplace (glaplace.hpp:16-28), poisson.hpp:37-42 */
    SEM_OP_HELMHOLTZ = 1, /* (W(u%c)W)%J + glaplace(u), helmholtz.hpp:38-46; coef = c */
    SEM_OP_MASS = 2,      /* (W u W)%J, the gproj element op, galerkin.hpp:49-51 */
    SEM_OP_SCHRO_RES = 3, /* Res lambda, schrodinger.hpp:185-207; 2 comps; coef = V, alpha */
    SEM_OP_SCHRO_F = 4,   /* f lambda, schrodinger.hpp:163-183; 2 comps; coef = V */
    SEM_OP_SCHRO_L = 5    /* L lambda glaplace(u) + (W(u%V)W)%J, schrodinger.hpp:37-43; coef = V */
};

/* Precision of the inner operator applies. */
enum sem_precision {
    SEM_PREC_FP64 = 0,  /* reference arithmetic */
    SEM_PREC_MIXED = 1  /* FP32 inner MINRES inside FP64 iterative refinement */
};

/* solver_base.hpp:6-12 */
typedef struct sem_solver_results {
    int success;
    int n_iter;
    double residual;
} sem_solver_results;

/* ode_base.hpp:6-13 (max_iter, tol, dt, verbose) plus build-only knobs. */
typedef struct sem_solver_opts {
    int max_iter;
    double tol;
    double dt;
    int verbose;
    int precision;   /* enum sem_precision; 0 = reference arithmetic */
    int guess;       /* 0 = reference guesses (zero / previous psi); 1 = extrapolated; 2 = learned */
    int learn_steps; /* guess 2: time steps during which the network trains (0 = 50), then frozen */
} sem_solver_opts;

/* Per-step observer (schrodinger.hpp:116,273,411): receives the unmasked
 * psi (host copies, element-local layout); return nonzero to stop. */
typedef int (*sem_step_cb)(const double* psi_re, const double* psi_im, void* user);

/* Run statistics of the last stepper call on a mesh handle. */
typedef struct sem_run_stats {
    int steps;
    int stage_solves;
    long total_iters;        /* inner Krylov iterations, all stages */
    long refinements;        /* FP64 refinement sweeps (mixed precision) */
    double last_error_estimate; /* DIRK4S5 embedded estimate (schrodinger.hpp:262-263) */
    double max_true_rel_residual; /* max over FP64-refined stage solves of ||b - A x|| / ||b|| */
    long mlp_train_steps;    /* Adam steps taken by the guess networks */
    double mlp_last_loss;    /* last training loss (mean squared error, scaled targets) */
    long learned_guess_chosen; /* stage solves (of each instance, in a batch) that started from the network's guess */
    double learned_residual_digits; /* sum over those solves of log10(||b - Res ext|| / ||b - Res x0||): how many
                                       digits of initial residual the network's guess x0 saves over the extrapolation ext */
} sem_run_stats;

typedef struct sem_mesh sem_mesh;

const char* sem_last_error(void);
int sem_version(void);
/* The progress line the steppers print with opts->verbose: VerboseTracker's
 * operator() (verbose.hpp:33-44) with precision int(2 - log10 dt)
 * (schrodinger.hpp:28); with_error adds the DIRK4S5 error estimate
 * (schrodinger.hpp:271). Host only, no GPU. */
int sem_progress_line(double t, double T, double dt, int with_error, double err, char* out, int out_len);
int sem_device_info(char* name, int name_len, int* n_cu, long* hbm_bytes);

/* ----- mesh setup: load_mesh<Mesh>(p, folder) + Mesh::compute_metrics()
 * (load_mesh.hpp:76-209, openmp_impl/Mesh.hpp:70-75, Quad.hpp:27-55) ----- */
int sem_mesh_create(int p, const double* xy, int n_nodes, const int* quads, int n_elements, sem_mesh** out);
int sem_mesh_load(int p, const char* folder, sem_mesh** out);
void sem_mesh_destroy(sem_mesh* m);
/* out[0..5] = N, n_elements, n_nodes, n_edges, dof (openmp_impl/Mesh.hpp:84-87), elements per chunk */
int sem_mesh_info(const sem_mesh* m, long* out);
/* quad_rule x, w (gauss_lobatto.hpp:27-32) and D (derivative_matrix.hpp:12-38, column-major) */
int sem_mesh_quadrature(const sem_mesh* m, double* x, double* w, double* D);
/* physical (x, y) of every element-local GLL node, Quad::from_local_coo (Quad.hpp:57-62) */
int sem_mesh_coords(const sem_mesh* m, double* X, double* Y);
/* per-element metrics J, A, B, C in element-local layout (Quad.hpp:47-54) */
int sem_mesh_metrics(const sem_mesh* m, double* J, double* A, double* B, double* C);
/* Seeded structured n x n quad mesh on [x0,x1]x[y0,y1]: interior vertices
 * moved by U(-jitter*h, jitter*h) per coordinate (SplitMix64 stream `seed`).
 * Writes (n+1)^2 nodes (xy pairs) and n^2 counter-clockwise quads. No GPU. */
int sem_mesh_structured(int n, double x0, double x1, double y0, double y1, double jitter, unsigned long long seed,
                        double* xy, int* quads);
/* Writes info.txt / coordinates.txt / elements.txt (the format load_mesh reads). */
int sem_mesh_write(const char* folder, const double* xy, int n_nodes, const int* quads, int n_elements);

/* ----- element operator and the Op action layer ----- */
/* glaplace(u, Quad, D, quad_rule) on every element, no assembly (glaplace.hpp:16-28). */
int sem_glaplace_apply(sem_mesh* m, const double* u, double* y);
/* Element operator only (no unmask / assembly), any kind; coef/alpha as in sem_op_kind. */
int sem_element_apply(sem_mesh* m, int kind, const double* coef, double alpha, const double* in, double* out);
/* The element operator in FP32 (the inner solver's precision; host arrays in
 * and out in FP64): with SEM_OP_SCHRO_RES the stiffness contractions follow
 * sem_set_contraction. No counterpart in the reference. */
int sem_element_apply_fp32(sem_mesh* m, int kind, const double* coef, double alpha, const double* in, double* out);
/* Contractions of the stiffness action in the FP32 inner solver's stage
 * operator (fused Lanczos step and sem_element_apply_fp32 with Res): 0 VALU
 * FMAs (default); 1 the stacked-element form on MFMA, v_mfma_f32_16x16x16bf16
 * tiles with bf16 operands and FP32 accumulation (about 4e-3 relative error;
 * the FP64 refinement still stops on the true residual, the inner solves ask
 * for a 5e-3 reduction); 3 the same with each operand split into hi + lo bf16
 * (three MFMAs per tile, about 2^-16 relative error). No counterpart in the
 * reference. */
int sem_set_contraction(sem_mesh* m, int mode);
int sem_get_contraction(const sem_mesh* m, int* mode);
/* FP32 MINRES with the one-launch Lanczos step (the inner solves of the mixed
 * precision): on (default) keeps every Lanczos vector in device memory, up to
 * budget_bytes (0: keep the current budget, 8 GiB by default), and forms the
 * solution from them once at the end instead of updating w and x every step
 * (the same iterate, minres.hpp:82-86 unrolled); off uses the w/x recurrence.
 * A solve that fills the budget stops there unconverged, as at maxit (the
 * FP64 refinement then continues with another inner solve). Batched meshes
 * keep the recurrence. No counterpart in the reference. */
int sem_set_stored_basis(sem_mesh* m, int on, long budget_bytes);
/* galerkin_op: unmask -> element op -> global_sum -> mask (galerkin.hpp:24-35);
 * 2-component kinds take and return [re | im]. */
int sem_galerkin_apply(sem_mesh* m, int kind, const double* coef, double alpha, const double* in, double* out);
/* gproj (galerkin.hpp:42-57) */
int sem_gproj(sem_mesh* m, const double* u, double* out);
/* DSS pieces, in place: openmp_impl/unmask.hpp:43-55, global_sum.hpp:65-77, mask.hpp:16-28 */
int sem_unmask(sem_mesh* m, double* a);
int sem_global_sum(sem_mesh* m, double* a);
int sem_mask(sem_mesh* m, double* a);
/* unique-DOF dot product (openmp_impl/dot.hpp:11-61) */
int sem_dot(sem_mesh* m, const double* a, const double* b, double* out);

/* ----- PDE drivers ----- */
/* poisson (openmp_impl/poisson.hpp:17-59): PCG on K u = M f; u in/out. */
int sem_poisson(sem_mesh* m, double* u, const double* f, int max_iter, double tol, sem_solver_results* out);
/* helmholtz (openmp_impl/helmholtz.hpp:18-64): MINRES on (K + M_c) u = M f. */
int sem_helmholtz(sem_mesh* m, double* u, const double* c, const double* f, int max_iter, double tol,
                  sem_solver_results* out);
/* Schrodinger steppers (openmp_impl/schrodinger.hpp:22-128, 138-285, 294-423):
 * psi in/out; *success mirrors the reference's bool return. opts->precision
 * and opts->guess select the build-only modes; zero keeps reference behaviour. */
int sem_schrodinger_bdf2(sem_mesh* m, double* re, double* im, const double* V, double T, const sem_solver_opts* opts,
                         sem_step_cb cb, void* user, int* success);
int sem_schrodinger_dirk4s3(sem_mesh* m, double* re, double* im, const double* V, double T, const sem_solver_opts* opts,
                            sem_step_cb cb, void* user, int* success);
int sem_schrodinger_dirk4s5(sem_mesh* m, double* re, double* im, const double* V, double T, const sem_solver_opts* opts,
                            sem_step_cb cb, void* user, int* success);
int sem_last_run_stats(const sem_mesh* m, sem_run_stats* out);

/* Stepper session: the same three schemes with psi, V and the stage vectors
 * resident on the GPU between calls (no counterpart in the reference, whose
 * steppers run to T in one call). advance() runs the reference's while (t < T)
 * loop, also stopping after max_steps steps when max_steps > 0. */
typedef struct sem_stepper sem_stepper;
int sem_stepper_create(sem_mesh* m, int scheme /* 0 bdf2, 1 dirk4s3, 2 dirk4s5 */, const double* re, const double* im,
                       const double* V, const sem_solver_opts* opts, sem_stepper** out);
int sem_stepper_advance(sem_stepper* s, double T, long max_steps, sem_step_cb cb, void* user, int* success);
int sem_stepper_read(sem_stepper* s, double* re, double* im, double* t);
int sem_stepper_stats(const sem_stepper* s, sem_run_stats* out);
/* learned guess: Adam steps per stage solve while training (default 5) and learning rate (default 3e-3) */
int sem_stepper_set_learning(sem_stepper* s, int epochs, float lr);
/* learned guess, before its first stage solve: the training loss (0: mean
 * squared error of the scaled correction z - ext; 1: the assembled residual
 * ||Res (g - z)||^2 of the guess g, relative to the extrapolation's, the norm
 * the guard compares and MINRES reduces) and the fourth input block (0: the
 * stage's right-hand side b; 1 (default): the extrapolation's residual b - Res ext) */
int sem_stepper_set_training(sem_stepper* s, int loss, int inputs);
/* learned guess: on (default) starts each stage solve from ext + c (g - ext),
 * the point on the line through the extrapolation ext and the network's guess g
 * with the smallest true FP64 residual ||b - Res x0|| (c = 0 is the
 * extrapolation, so the start is never worse than the fallback; one extra
 * operator apply and three dot products per stage solve); off always starts
 * from the network's guess g */
int sem_stepper_set_guard(sem_stepper* s, int on);
/* Checkpoint and resume of a session (no counterpart in the reference, whose
 * steppers keep their BDF2 history, schrodinger.hpp:85-102, in a local deque
 * and restart with implicit Euler). A checkpoint is one host array of
 * sem_stepper_checkpoint_size doubles: t, the step count, psi, the BDF2
 * history, the previous two steps' stage solutions (extrapolated / learned
 * guesses) and the guess network with its optimiser state, every vector in
 * the element-local layout with the node's value on every copy. restore()
 * takes a checkpoint of the same scheme and mesh size into a session created
 * with any options: psi, t, step count and BDF2 history always, the stage
 * history when the session extrapolates, the network when it has one of the
 * same size. A run split by save/restore equals the one-shot run bit for bit. */
int sem_stepper_checkpoint_size(const sem_stepper* s, long* n);
int sem_stepper_save(sem_stepper* s, double* buf, long n);
int sem_stepper_restore(sem_stepper* s, const double* buf, long n);
void sem_stepper_destroy(sem_stepper* s);
/* Batched instances (no counterpart in the reference, which runs one problem
 * per call): independent meshes advanced together. Instances of one degree
 * share one device mesh and one launch per kernel; each keeps its own Krylov
 * scalars and stop test and computes bit for bit what it computes alone.
 * Groups of different degree run concurrently on their own streams. The
 * batch copies what it needs: the instance meshes may be destroyed after
 * sem_batch_create. */
typedef struct sem_batch sem_batch;
int sem_batch_create(int n, sem_mesh* const* meshes, sem_batch** out);
void sem_batch_destroy(sem_batch* b);
/* per degree group (arrays sized by the caller, at most n entries):
 * degree, instances, device elements (ghost padding included) */
int sem_batch_info(const sem_batch* b, int* n_groups, int* group_p, int* group_size, long* group_elements);
/* sem_autotune on every group's mesh (operator SEM_OP_RES); shapes: 4 ints per group (optional) */
int sem_batch_autotune(sem_batch* b, int precision, int steps, int* shapes);
int sem_batch_set_step_shape(sem_batch* b, int group, int precision, const int* shape);
/* sem_set_graph_mode on every group's mesh */
int sem_batch_set_graph_mode(sem_batch* b, int on, int steps);
/* sem_set_step_mode on every group of the batch */
int sem_batch_set_step_mode(sem_batch* b, int precision, int mode);
/* schrodinger_bdf2 (schrodinger.hpp:22-128), schrodinger_dirk4s3 (:294-423)
 * and schrodinger_dirk4s5 (:138-285) on every instance to time T with one set
 * of options (FP64 or mixed precision; reference, extrapolated or learned
 * guess). re[k], im[k] (in/out) and V[k] are instance k's element-local
 * arrays. ok[k] (optional): every solve of instance k converged. An instance
 * whose solve fails stops there, as the reference's stepper returns (BDF2
 * keeps the unconverged iterate and counts the step; the SDIRK schemes keep
 * psi of the step's start); the other instances go on to T (sem_batch_times
 * gives each instance's time and step count). With the reference or
 * extrapolated guess, in either precision, every instance computes bit for
 * bit what it computes alone (mixed precision: each instance refines to its
 * own target). With the learned guess the instances of one degree share one
 * network, trained on the stages every instance solved: each instance meets
 * its own stop test, so results agree with single runs to the tolerance. */
int sem_batch_schrodinger_bdf2(sem_batch* b, double* const* re, double* const* im, const double* const* V, double T,
                               const sem_solver_opts* opts, int* ok);
int sem_batch_schrodinger_dirk4s3(sem_batch* b, double* const* re, double* const* im, const double* const* V,
                                  double T, const sem_solver_opts* opts, int* ok);
int sem_batch_schrodinger_dirk4s5(sem_batch* b, double* const* re, double* const* im, const double* const* V,
                                  double T, const sem_solver_opts* opts, int* ok);
/* The same as a session kept on the GPU: start (scheme 0 bdf2, 1 dirk4s3,
 * 2 dirk4s5), advance (the while (t < T) loop, also stopping after max_steps
 * steps when max_steps > 0; success: every group converged), read (psi,
 * per-instance success, time). */
int sem_batch_start(sem_batch* b, int scheme, const double* const* re, const double* const* im,
                    const double* const* V, const sem_solver_opts* opts);
int sem_batch_start_bdf2(sem_batch* b, const double* const* re, const double* const* im, const double* const* V,
                         const sem_solver_opts* opts);
int sem_batch_advance(sem_batch* b, double T, long max_steps, int* success);
int sem_batch_read(sem_batch* b, double* const* re, double* const* im, int* ok, double* t);
/* per-instance time and completed steps of the session (arrays of n entries;
 * either may be null): an instance stopped by a failed solve keeps its own */
int sem_batch_times(const sem_batch* b, double* t, int* steps);
/* statistics of the last batch run (steps, stage solves and Lanczos steps summed
 * over groups) and its wall time in seconds */
int sem_batch_last_run_stats(const sem_batch* b, sem_run_stats* out, double* seconds);

/* One Schrodinger stage solve Res(alpha, V) z = b (schrodinger.hpp:223 etc.), z in/out. */
int sem_stage_solve(sem_mesh* m, const double* V, double alpha, const double* b, double* z, int max_iter, double tol,
                    int precision, sem_solver_results* out);

/* ----- generic Krylov solvers over a tridiagonal operator y_i = lo x_{i-1} + di x_i + up x_{i+1},
 * the arma::conv(u, {lo,di,up}, "same") operators of tests/pcg.cpp:13-16 and tests/minres.cpp:14-17,
 * Euclidean dot; minres.hpp:19-107 and pcg.hpp:59-97 with IdentityPreconditioner. Host arrays. ----- */
int sem_minres_tridiag(long n, double lo, double di, double up, const double* b, double* x, int maxit, double tol,
                       sem_solver_results* out);
int sem_pcg_tridiag(long n, double lo, double di, double up, const double* b, double* x, int maxit, double tol,
                    sem_solver_results* out);

/* ----- generic Krylov solvers over the caller's operator (minres.hpp:19-107,
 * pcg.hpp:12-97): A, dot and precond play the reference's LinOp, Dot and
 * Precond template arguments. host = 1: b, x and every callback argument are
 * host arrays of n doubles (the library stages its device vectors through
 * them). host = 0: b, x and the callback arguments are device pointers, and a
 * callback enqueues its work on `stream` (the library's HIP stream, as a
 * void*) or synchronises before returning. A callback returns 0, or nonzero
 * to abort the solve (SEM_EINVAL). dot = NULL: Euclidean dot on the device;
 * precond = NULL: IdentityPreconditioner (solver_base.hpp:14-22).
 * consistent = 0 computes MINRES's beta as minres.hpp:63 does,
 * sqrt(dot(precond(r2), precond(r2))); consistent = 1 uses
 * sqrt(dot(r2, precond(r2))), the normalisation that matches minres.hpp:31
 * for a symmetric positive definite preconditioner. Both are the same with the
 * identity. ----- */
typedef int (*sem_linop_fn)(void* user, const double* x, double* y, void* stream);
typedef int (*sem_dot_fn)(void* user, const double* a, const double* b, double* result, void* stream);
typedef int (*sem_precond_fn)(void* user, const double* r, double* z, void* stream);
typedef struct sem_krylov {
    long n;
    int host;
    int consistent;
    sem_linop_fn A;
    void* A_user;
    sem_dot_fn dot;
    void* dot_user;
    sem_precond_fn precond;
    void* precond_user;
} sem_krylov;
int sem_minres(const sem_krylov* k, const double* b, double* x, int maxit, double tol, sem_solver_results* out);
int sem_pcg(const sem_krylov* k, const double* b, double* x, int maxit, double tol, sem_solver_results* out);
/* sem_minres without any allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). No counterpart in the reference. Device vectors only
 * (k->host = 0) and n <= 262144: the workspace (8n doubles of vectors plus the
 * dot partials, about 16 MB) is a static device array that every GPU the
 * library runs on holds until the process ends; larger n returns SEM_EINVAL.
 * The work runs on hipStreamPerThread, which the callbacks receive as
 * `stream`. Results equal sem_minres's bit for bit. Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU race. */
int sem_minres_noalloc(const sem_krylov* k, const double* b, double* x, int maxit, double tol, sem_solver_results* out);

/* Preconditioners of the mesh solvers (no counterpart in the reference beyond
 * IdentityPreconditioner, solver_base.hpp:14-22). */
enum sem_precond_kind {
    SEM_PRECOND_IDENTITY = 0, /* the reference's arithmetic (generic recurrences) */
    SEM_PRECOND_JACOBI = 1    /* 1/|diag| of the assembled operator, zero on Dirichlet entries */
};
/* poisson / helmholtz (openmp_impl/poisson.hpp:17-59, helmholtz.hpp:18-64)
 * with a preconditioner: PCG on K u = M f, MINRES (consistent normalisation)
 * on (K + M_c) u = M f; same arguments and outputs as sem_poisson /
 * sem_helmholtz. */
int sem_poisson_precond(sem_mesh* m, double* u, const double* f, int max_iter, double tol, int precond,
                        sem_solver_results* out);
int sem_helmholtz_precond(sem_mesh* m, double* u, const double* c, const double* f, int max_iter, double tol,
                          int precond, sem_solver_results* out);

/* galerkin_op(op, u, mesh) (openmp_impl/galerkin.hpp:24-35) with the caller's
 * element operator, the reference's Op action: unmask -> op -> global_sum ->
 * mask. in and out are device arrays of n_elements*N*N doubles in the
 * element-local layout; op receives device arrays u (unmasked input) and y
 * (its output, element by element) in the same layout and enqueues its work
 * on `stream`, returning 0 (nonzero aborts with SEM_EINVAL). */
typedef int (*sem_element_op_fn)(void* user, int n_elements, int N, const double* u, double* y, void* stream);
int sem_galerkin_apply_op(sem_mesh* m, sem_element_op_fn op, void* user, const double* in, double* out);

/* Device memory for the device-pointer entry points (no torch or HIP types
 * needed by the caller): n doubles; sem_copy direction 0 host->device,
 * 1 device->host, 2 device->device. */
int sem_device_alloc(long n, double** ptr);
int sem_device_free(double* ptr);
int sem_copy(double* dst, const double* src, long n, int direction);
/* waits for the work enqueued on a stream handed to a callback */
int sem_stream_synchronize(void* stream);

/* Host-only basis routines (no GPU): gauss_lobatto<double>(n) with n + 1
 * nodes and weights (gauss_lobatto.hpp:43-126); derivative_matrix(x) for n
 * points, D column-major (derivative_matrix.hpp:12-38); Quad::compute_metrics
 * (Quad.hpp:27-55) of the quad with corners (xs, ys) on the n points z, J, A,
 * B, C n x n column-major. */
int sem_gauss_lobatto(int n, double* x, double* w);
int sem_derivative_matrix(int n, const double* x, double* D);
int sem_quad_metrics(const double* xs, const double* ys, int n, const double* z, double* J, double* A, double* B,
                     double* C);

/* ----- measurement hooks (device-resident data, HIP events on the handle's stream) ----- */
/* Times `reps` launches of the element-operator kernel over device-resident
 * random data (kind, precision 0 = FP64, 1 = FP32); returns mean microseconds
 * per launch and the algorithmic bytes per launch. */
int sem_time_element_op(sem_mesh* m, int kind, int precision, int reps, double* us_per_launch, double* bytes_per_launch);
/* Live timing of the two MINRES Lanczos-step kernels: on = 1 times the
 * launches of one Lanczos step in sixteen (picked by a hash of the step index,
 * so the sample does not follow the batches between host checks) with HIP
 * events that take the
 * kernels' own start and end timestamps (hipExtLaunchKernel), on = 2 every
 * step, 0 turns it off. which = 0 / 1:
 * operator kernel (operator apply, y -= cr1*r1, alpha partial) FP64 / FP32;
 * 2 / 3: update kernel (interface gather, y -= c2*r2, beta^2 partial, w/x
 * update) FP64 / FP32. Returns launches, summed microseconds and the
 * algorithmic bytes of one launch. */
int sem_set_kernel_timing(sem_mesh* m, int on);
/* Device-to-device copy of `bytes` bytes (read `bytes`, write `bytes`), `reps`
 * launches: measured copy bandwidth and the calibration kernel for HBM byte
 * counters. variant 0: the 8-byte grid-stride copy; v > 0: 16-byte
 * nontemporal copy with 8*v workgroups per CU. */
int sem_time_copy(long bytes, int reps, int variant, double* us_per_launch);
int sem_kernel_timing(const sem_mesh* m, int which, long* launches, double* total_us, double* bytes_per_launch);

/* ----- work shape of the Lanczos-step kernels (no counterpart in the reference) ----- */
/* shape[4] = {operator chunks per workgroup (1 or 2), operator grid (0: one
 * workgroup per work item, > 0: persistent grid of that many workgroups),
 * update chunks per workgroup, update grid}, per precision (SEM_PREC_FP64 or
 * SEM_PREC_MIXED for the FP32 inner solver). Results do not depend on it. */
int sem_set_step_shape(sem_mesh* m, int precision, const int* shape);
int sem_get_step_shape(const sem_mesh* m, int precision, int* shape);
/* MINRES step mode per precision (SEM_PREC_FP64, or SEM_PREC_MIXED for the
 * FP32 inner solver): 0 two launches per Lanczos step; 1 one fused launch per
 * step that applies the operator before the 1/beta scaling (same recurrence,
 * results equal to rounding); 2 the
 * fused step with the recurrence's reduction of the per-chunk sums moved out
 * of every workgroup into one small kernel per step (same order: results
 * equal mode 1's bit for bit; cheaper on large meshes). Modes 1 and 2 are
 * FP32 only: FP64 returns SEM_EUNSUP for them.
 * Defaults: FP64 0; FP32 2 when an instance of the mesh has 1024 chunks of
 * elements or more, 1 otherwise. */
int sem_set_step_mode(sem_mesh* m, int precision, int mode);
int sem_get_step_mode(const sem_mesh* m, int precision, int* mode);
/* Diagonal scaling of the FP32 inner MINRES in mixed-precision Schrodinger
 * stage solves (on = 1, the default): the inner solver works on S Res S with
 * S = (m + |alpha| |k + c|)^(-1/2) from the assembled mass, stiffness and
 * potential-mass diagonals, which takes fewer Lanczos steps; the outer FP64
 * refinement stops on the true residual either way. It applies when the FP32
 * step mode is 1 or 2. No counterpart in the reference
 * (its MINRES is unpreconditioned, minres.hpp:19). */
int sem_set_inner_scaling(sem_mesh* m, int on);
int sem_get_inner_scaling(const sem_mesh* m, int* on);
/* Graph mode: with on = 1 the MINRES solver captures `steps` Lanczos steps
 * (a multiple of 3; the one-launch step of modes 1 and 2 rounds it up to a
 * multiple of 6 and replays from its second host check on) plus the
 * batch-closing kernel as one HIP graph per solve configuration and replays
 * it between host checks of the stop flags (results are unchanged bit for
 * bit; kernel timing, when on, uses the direct launches). Off by default for
 * a plain mesh, on for the meshes of a batch. */
int sem_set_graph_mode(sem_mesh* m, int on, int steps);
/* Autotuner: times each candidate shape on this mesh over `steps` Lanczos
 * steps of a synthetic solve of operator `kind` (SEM_OP_LAPLACE, _HELMHOLTZ or
 * _RES) and keeps the fastest shape per kernel. Outputs (optional):
 * *n_candidates; cand[2c], cand[2c+1] = chunks per workgroup and grid of
 * candidate c; us[2c], us[2c+1] = operator / update kernel microseconds per
 * launch. At most 16 candidates. */
int sem_autotune(sem_mesh* m, int kind, int precision, int steps, int* n_candidates, int* cand, double* us);

/* ----- per-element MLP of the learned stage guess (no counterpart in the
 * reference: its stage guesses are zero, openmp_impl/schrodinger.hpp:222 etc.).
 * Exposed for testing the forward, backward and Adam kernels. Row-major host
 * arrays: X [rows x in_pad], T and Y [rows x out_pad]; parameters in one block
 * W1 [in_pad x h1], b1, W2 [h1 x h2], b2, W3 [h2 x out_pad], b3. ----- */
typedef struct sem_mlp sem_mlp;
int sem_mlp_create(int in, int h1, int h2, int out, unsigned long long seed, sem_mlp** out_handle);
void sem_mlp_destroy(sem_mlp* m);
/* out[0..8] = in_pad, out_pad, nparams, offsets of W1, b1, W2, b2, W3, b3 */
int sem_mlp_info(const sem_mlp* m, long* out);
int sem_mlp_set_params(sem_mlp* m, const float* params);
int sem_mlp_get_params(sem_mlp* m, float* params, int gradients);
int sem_mlp_forward(sem_mlp* m, int rows, const float* X, float* Y);
/* forward + backward of the mean squared error; gradients stay on the device */
int sem_mlp_gradients(sem_mlp* m, int rows, const float* X, const float* T, double* loss);
int sem_mlp_adam(sem_mlp* m, float lr);

#ifdef __cplusplus
}
#endif

#endif /* SEM_H */
