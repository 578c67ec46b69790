/*
 * C ABI of the MI355X (gfx950) Newton-step library for the mixed variational
 * FEM of tytrusty/mfem. Plain C types only: host pointers are borrowed for the
 * duration of a call, device pointers are the caller's, every function
 * returns 0 on success or a nonzero code with a message in
 * mfem_last_error() (the reference prints and exits instead).
 *
 * The reference has no C ABI; its boundary is the C++ interface set of
 * src/variables/, src/energies/, src/linear_solvers/, src/optimizers/ and
 * src/simulation_state.cpp, selected by name from scene JSON. Each entry point
 * below names the reference interface it stands in for.
 */
#ifndef MFEM_HIP_H
#define MFEM_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mfem_sim mfem_sim;

/* Scene given as arrays (stands in for SimState<3>::load,
 * reference src/simulation_state.cpp:92-341, for one tetrahedral object with
 * per-element materials as with "material_ids", reference tet_mesh.h:11-15). */
typedef struct {
  int nV, nT;
  const double* V;        /* nV x dim, row per vertex */
  const int* T;           /* nT x (dim + 1) */
  const double* mu;       /* nT Lame mu (reference config.h:13-16) */
  const double* la;       /* nT Lame lambda */
  const int* mat;         /* nT material id: 0 SNH, 1 NH, 2 Corot, 3 FCR, 4 ARAP, 5 Fung */
  const int* fixed;       /* nV pin mask (1 = pinned), or NULL for bc_type/axis/ratio */
  const double* v0;       /* nV x dim initial velocity, or NULL */
  double density;
  double dt;              /* "dt" */
  double gravity[3];      /* "body_force" (stored as float, as the reference does) */
  int bdf_order;          /* "time_integrator": BDF1..BDF6 */
  int mixed;              /* 1: "mixed_variables": ["mixed-stretch"]; 0: "variables": ["stretch"] */
  const char* linear_solver; /* "linear_solver": "eigen-pcg-diag" (Schur complement, Jacobi
                              * CG), "eigen-pcg-IC" (Schur complement, CG with a multicolour
                              * block incomplete Cholesky), "eigen-llt", "eigen-ldlt",
                              * "cholmod", "eigen-lu" (dense factorization on the device) or
                              * "minres-indefinite-block" (KKT system, MINRES; mixed only) */
  double itr_tol;         /* "iterative_solver_tolerance" */
  int max_iters;          /* "max_iterative_solver_iters" */
  int max_newton;         /* "max_newton_iterations" */
  double newton_tol;      /* SimConfig::newton_tol (no JSON key in the reference) */
  int ls_iters;           /* "max_linesearch_iterations" */
  int precision;          /* 0: FP64, 1: mixed FP32/FP64 */
  int bc_type, bc_axis;   /* boundary condition when fixed == NULL (reference config.h:70-87) */
  double bc_ratio;
  int dim;                /* 3 (tetrahedra; 0 also means 3) or 2 (triangles, reference
                           * Tri2DMesh: V and v0 are nV x 2, T is nT x 3) */
  double bc_velocity;     /* "velocity" of the stretch / twist / bend boundary scripts, which
                           * move their pinned end groups after every timestep
                           * (reference boundary_conditions/*.h) */
  const double* x0;       /* nV x dim start positions, or NULL to start at V: the scene's
                           * "initial_state" x_path (reference simulation_state.cpp:263-285),
                           * with V the rest shape */
} mfem_sim_params;

typedef struct {
  int newton_iters;       /* Newton iterations of the last timestep */
  int pcg_iters;          /* CG iterations summed over them */
  int refinements;        /* defect-correction sweeps summed over them */
  int ls_iters;           /* line-search halvings summed over them */
  int ls_fail;            /* exhausted line searches */
  double decrement;       /* last "Newton dec" (max |dx|, |ds|) */
  double energy;          /* last E(0) */
  double alpha;           /* last accepted step */
  double res;             /* last energy residual */
  double rel_residual;    /* |b - K dx| / |b| of the last solve */
  double update_ms, solve_ms, ls_ms, step_ms; /* device time per phase (hipEvents; 0
                                                 unless phase timing is on, see
                                                 mfem_sim_set_profiling) */
  double spmv_ms;         /* summed device time of the CG kernels (when profiled) */
  int spmv_count;
} mfem_step_stats;

const char* mfem_last_error(void);
int mfem_version(void);

/* SimState<3>::load(json) + OptimizerFactory::create("newton") + reset()
 * (reference simulation_state.cpp:72-89, optimizer_factory.cpp:10-16,
 * newton_optimizer.cpp:127-140). */
int mfem_sim_create_json(const char* json_path, mfem_sim** out);
/* SimState<2>::load(json) for the 2D application (reference
 * simulation_state.cpp:46-60 and apps/sim2d.cpp): OBJ triangle meshes,
 * unreferenced vertices dropped, z truncated, no rescaling; positions come
 * back nV x 2. */
int mfem_sim_create_json_2d(const char* json_path, mfem_sim** out);
/* Same from arrays. */
int mfem_sim_create(const mfem_sim_params* p, mfem_sim** out);
int mfem_sim_destroy(mfem_sim* sim);

/* NewtonOptimizer<3>::step() (reference newton_optimizer.cpp:11-89). */
int mfem_sim_step(mfem_sim* sim);
/* Optimizer<3>::callback (reference optimizers/optimizer.h:16-42), which
 * NewtonOptimizer::step invokes once per Newton iteration, after the line
 * search (newton_optimizer.cpp:68). `cb` runs on the calling thread inside
 * mfem_sim_step with the simulation, the 1-based Newton iteration and `user`;
 * it may read the simulation (mfem_sim_get_positions gives the positions
 * after the iteration's accepted step, mfem_sim_get its arrays) but must not
 * step or destroy it. cb == NULL removes the callback. */
typedef void (*mfem_newton_callback)(mfem_sim* sim, int iteration, void* user);
int mfem_sim_set_callback(mfem_sim* sim, mfem_newton_callback cb, void* user);
/* Phases of step(): x_tilde + update_system() (:92-110), substep() (:113-124)
 * and the line-search energy E(alpha) (:34-49). */
int mfem_sim_update_system(mfem_sim* sim);
int mfem_sim_solve(mfem_sim* sim, double* decrement);
int mfem_sim_energy(mfem_sim* sim, double alpha, double* energy);
int mfem_sim_stats(mfem_sim* sim, mfem_step_stats* out);
/* Current vertex positions (nV x dim), i.e. Mesh::V_ after post_solve. */
int mfem_sim_get_positions(mfem_sim* sim, double* out);
/* Internal arrays by name, element-major, for parity checks: "R" (9/tet),
 * "S", "s", "lambda", "ds", "g", "gl" (6/tet), "dsdF" (54/tet, ds/dvecF
 * column-major), "H", "Hinv" (36/tet; H is Syminv Hpsi Syminv / vol as the
 * reference's H_), "blocks" (90/tet), "fe" (12/tet), "lhs" (scalar CSR
 * values in the reference's row order), "rhs", "dx" (3/free vertex),
 * "x" (3/vertex), "x_tilde", "mass" (per nonzero block), "cg_stamps"
 * (profiling: per-workgroup start/end of the last two CG launches, in us
 * from the earliest stamp, [2][grid][2]), "ic_shift" ("eigen-pcg-IC":
 * diagonal shift of the last factorization). */
long mfem_sim_size(mfem_sim* sim, const char* name);
int mfem_sim_get(mfem_sim* sim, const char* name, double* out);
/* "rowptr", "colind" (scalar CSR pattern), "free_map", "fixed",
 * "bsr_rowptr", "bsr_col", "slot_ptr", "slots", "pcg_iters" (CG iterations
 * of the last solve); "eigen-pcg-IC": "ic_color" (colour of every free
 * vertex), "ic_colors", "ic_attempts" (factorizations of the last solve) */
int mfem_sim_get_int(mfem_sim* sim, const char* name, int* out);
/* The last step's Newton iterations, one row of 10 doubles each: iteration,
 * energy, energy res, Newton dec, alpha (the columns the reference records,
 * newton_optimizer.cpp:60-66, and prints with "print_stats"), then the
 * device time in ms of the update, solve and line search ("print_timing"'s
 * update / global / LS), the linear solver's iterations and the line-search
 * halvings. out may be NULL to query the count; *rows is its capacity. */
int mfem_sim_newton_history(mfem_sim* sim, double* out, int* rows);
/* Per-step output of the reference's application (apps/polyscope_app.h:
 * 131-149 write_obj, 335-370 the .mesh export and the sim_x0 / sim_v DMAT
 * pair): the current positions as an OBJ surface (3D: the boundary
 * triangles, normals out of the mesh; 2D: the triangles, z = 0) or as a
 * Medit .mesh (3D), and the integrator's last positions and velocities
 * (x_prev, v_prev, nV x dim) as DMAT files that a scene's "initial_state"
 * x_path / v_path restarts from. Either path of write_state may be NULL. */
int mfem_sim_write_obj(mfem_sim* sim, const char* path);
int mfem_sim_write_mesh(mfem_sim* sim, const char* path);
int mfem_sim_write_state(mfem_sim* sim, const char* x_path, const char* v_path);
/* Boundary triangles of a tetrahedral mesh (V nV x 3, T nT x 4): the faces
 * of exactly one tet, oriented out of it, in (tet, face) order. Host only.
 * F may be NULL to query the count; *nF is its capacity in triangles. */
int mfem_boundary_faces(const double* V, int nV, const int* T, int nT, int* F, int* nF);

/* Kernel of the mixed stiffness blocks vol B^T Hpsi B (no counterpart in the
 * reference; the blocks are the Aloc of mixed_stretch.cpp:168-175): 0 (default)
 * forms the products on the vector units, 1 on the matrix cores
 * (v_mfma_f32_4x4x1f32) when the blocks are stored in FP32 (mixed
 * precision, 3D); FP64 storage and 2D keep the vector-unit kernel. */
int mfem_sim_set_element_kernel(mfem_sim* sim, int variant);
/* bit 0: time every CG kernel launch with per-workgroup device clock
 * stamps (span of the launch, summed in the stats as spmv_ms/spmv_count);
 * bit 1: time each Newton iteration's phases with device events (update_ms,
 * solve_ms, ls_ms; also on with "print_timing" or a callback installed) */
int mfem_sim_set_profiling(mfem_sim* sim, int flags);
/* Sizes used by the roofline model of the CG SpMV: free vertices, nonzero
 * 3x3 blocks, bytes per stored value. */
int mfem_sim_matrix_info(mfem_sim* sim, int* nfree, int* nnzb, int* value_bytes);
/* Time `reps` back-to-back launches of the SpMV kernel (the autotuned shape)
 * on the current K and a device vector, returning the mean device time per
 * launch (ms). */
int mfem_sim_time_spmv(mfem_sim* sim, int reps, double* ms_per_launch);
/* SpMV autotuner result (no counterpart in the reference): mean device time
 * per launch of every shape (times[count], ms; pass NULL to query count),
 * the chosen shape, and its lanes-per-row / unroll / non-temporal settings. */
int mfem_sim_spmv_tuning(mfem_sim* sim, double* times, int* count, int* chosen, int* lanes,
                         int* unroll, int* nontemporal);
/* CG kernel shape (no counterpart in the reference; the iteration itself
 * follows Eigen's ConjugateGradient as linear_solver_factory.cpp:144-154
 * configures it). Force shape `variant` (0-7: lanes per block row, blocks
 * in flight per lane and workgroup size as listed in k_solver.hip) on a grid
 * of `blocks` workgroups (a multiple of 8 in [8, 4096], or 0 for the
 * default); variant -1 restores autotuning on the next solve. */
int mfem_sim_set_cg_shape(mfem_sim* sim, int variant, int blocks);
/* CG kernel autotuner result: every (shape, grid) pair timed on the first
 * solve (times[count] ms per iteration, variants[count], grids[count]; pass
 * NULL arrays to query count), the index of the chosen pair (-1 when untuned
 * or forced), and the chosen lanes-per-row / blocks in flight per lane /
 * workgroup size / grid. */
int mfem_sim_cg_tuning(mfem_sim* sim, double* times, int* variants, int* grids, int* count,
                       int* chosen, int* lanes, int* unroll, int* threads, int* blocks);

/* Indefinite (KKT) system of "minres-indefinite-block" (reference
 * SystemMatrixIndefinite, linear_system.h:60-99, and BlockMatrix::
 * scaleAndAddTo, block_matrix.h:74-125) at the current state (after
 * mfem_sim_update_system). Vectors are [dx (D nfree); ds (ns nT); lambda
 * (ns nT)] with element-major s-vectors (D = 3, ns = 6 on tetrahedra; D = 2,
 * ns = 3 on triangles), size from mfem_sim_kkt_size.
 * which 0: out = K in; 1: out = P in (the solver's block preconditioner);
 * 2: out = the right-hand side (in unused). Errors unless the simulation's
 * linear solver is "minres-indefinite-block". */
int mfem_sim_kkt_size(mfem_sim* sim, long* size);
int mfem_sim_kkt_apply(mfem_sim* sim, int which, const double* in, double* out);

/* Batch of independent mixed-stretch simulations (no counterpart in the
 * reference, where each would be its own SimState + NewtonOptimizer,
 * reference newton_optimizer.cpp:11-89). The meshes are advanced together
 * as one block-diagonal system, one launch per Newton phase for all of
 * them, while each keeps its own CG scalars, line search and Newton
 * termination. Time step, integrator, gravity, solver settings and
 * precision must agree across `sims`; "mixed" must be 1. */
typedef struct mfem_batch mfem_batch;
int mfem_batch_create(const mfem_sim_params* sims, int count, mfem_batch** out);
int mfem_batch_destroy(mfem_batch* batch);
/* One time step of every simulation. */
int mfem_batch_step(mfem_batch* batch);
/* count, then per simulation vertex and tet counts (arrays of count; NULL
 * allowed). */
int mfem_batch_sizes(mfem_batch* batch, int* count, int* nV, int* nT);
int mfem_batch_get_positions(mfem_batch* batch, int sim, double* out);
/* Per-simulation stats of the last step (arrays of count; NULL allowed) and
 * the step's wall time. */
int mfem_batch_stats(mfem_batch* batch, int* newton_iters, int* pcg_iters, int* refinements,
                     int* ls_fail, double* decrement, double* step_ms);
/* Batched CG shape (as mfem_sim_set_cg_shape) and rows per workgroup;
 * variant -1 restores autotuning on the next solve. */
int mfem_batch_set_cg_shape(mfem_batch* batch, int variant, int rows_per_workgroup);
int mfem_batch_cg_shape(mfem_batch* batch, int* variant, int* rows_per_workgroup);
/* hipGraph replay of the batch's Newton phases (no counterpart in the
 * reference; off by default): with on = 1 the element update + assembly phase of every
 * Newton iteration is one captured graph from the second step on, and CG
 * iterations are replayed in captured groups of 16; on = 0 launches every
 * kernel separately. Results are identical bit for bit. */
int mfem_batch_set_graphs(mfem_batch* batch, int on);
/* Host submissions since creation: of the update + assembly phase, of the CG
 * iterations, and the number of Newton iterations (per-simulation Newton
 * loops advance together, so one per pass). Any pointer may be NULL. */
int mfem_batch_launch_counts(mfem_batch* batch, long* update_submissions, long* cg_submissions,
                             long* newton_iterations);

/* LinearSolver::solve() of "eigen-pcg-diag" on a caller's system (reference
 * linear_solvers/linear_solver.h:8-21; EigenIterativeSolver<ConjugateGradient<
 * SpMat, Lower|Upper>> with Eigen's default DiagonalPreconditioner,
 * eigen_iterative_solver.h:14-29, as linear_solver_factory.cpp:75-85
 * configures it): x = K^-1 b by Jacobi-preconditioned CG in FP64 with Eigen's
 * semantics -- x0 = 0; preconditioner 1 / K_jj (1 where the diagonal entry is
 * zero or absent); stop when |b - K x|^2 < max(tol^2 |b|^2, DBL_MIN) or after
 * max_iters iterations. K is symmetric positive definite, stored as 3x3-block
 * CSR (the reference's assembled lhs holds full 3x3 blocks,
 * utils/sparse_utils.cpp:19-138): nb block rows, rowptr[nb + 1] (rowptr[0] =
 * 0, non-decreasing, at most nnzb), col[nnzb] block columns in [0, nb),
 * vals[9 nnzb] with each block row-major. All arrays are device pointers; b
 * and x (out) hold 3 nb values. The pattern is checked on the device first
 * (an error for a malformed rowptr or an out-of-range column, before any CG
 * work). iterations receives Eigen's iterations(), rel_residual its error()
 * (|r| / |b|); either may be NULL. The work arrays (72 bytes per scalar row
 * plus 320 KB) are allocated for the call. Runs on `stream` (NULL: the
 * default stream) and returns after the solve. */
int mfem_pcg_bsr3(int nb, int nnzb, const int* rowptr, const int* col, const double* vals,
                  const double* b, double* x, double tol, int max_iters, int* iterations,
                  double* rel_residual, void* stream);

/* mfem_pcg_bsr3 without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code run
 * while a stream is being captured, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, semantics and results as mfem_pcg_bsr3 (the
 * same kernels on the same launch shape). The work arrays live in one fixed
 * device array of the library per GPU (about 38 MB, held until the process
 * ends), which limits the system to 3 nb <= MFEM_PCG_NOALLOC_MAX_ROWS scalar
 * rows; larger systems are rejected with an error. All calls on one GPU share
 * that array and its contents from the last call: calls that overlap on one
 * GPU would race, so the library serializes them within a process. */
#define MFEM_PCG_NOALLOC_MAX_ROWS 524288
int mfem_pcg_bsr3_noalloc(int nb, int nnzb, const int* rowptr, const int* col,
                          const double* vals, const double* b, double* x, double tol,
                          int max_iters, int* iterations, double* rel_residual, void* stream);

/* Batched MaterialModel evaluation on the 6-vector s (reference
 * energies/material_model.h:29-38): device pointers, n points, s n x 6,
 * E n, g n x 6, H n x 36 column-major; psd_fix as hessian(s, psd_fix). */
int mfem_material_s6_batch(int n, const int* type, const double* mu, const double* la,
                           const double* s, int psd_fix, double* E, double* g, double* H,
                           void* stream);
/* Batched plain SNH on vec(F) (reference stable_neohookean.cpp:97-234). */
int mfem_snh_f9_batch(int n, const double* mu, const double* la, const double* F, double* E,
                      double* g, double* H, void* stream);
/* Batched polar fit: newton_procrustes(R, I, F, tol 1e-6, 100 iterations)
 * warm-started from R (in/out), s = sym(R^T F), ds/dvecF (n x 54), iteration
 * counts (reference svd/newton_procrustes.h:108-189, mixed_stretch.cpp:14-42).
 * dsdF and iters may be NULL. */
int mfem_polar_batch(int n, const double* F, double* R, double* s, double* dsdF, int* iters,
                     void* stream);

#ifdef __cplusplus
}
#endif
#endif /* MFEM_HIP_H */
