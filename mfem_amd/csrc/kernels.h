// Launchers for the gfx950 kernels of the Newton-step hot path. Host code
// (the variable / solver / optimizer classes and the C ABI) calls these; all
// pointers are device pointers unless stated, all launches are asynchronous
// on `stream`.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace mfemhip {

// Mesh data on the device, structure-of-arrays with stride nT.
struct TetMeshDev {
  int nV = 0, nT = 0, nfree = 0;
  int dim = 3;                     // 2: triangles in the z = 0 plane (corners x, y, z of T)
  const int4* T = nullptr;         // [nT] vertex ids
  const double* dX = nullptr;      // [12][nT]: dphi_k/dX_c at (3k+c)*nT + e
  const double* vol = nullptr;     // [nT]
  const double* mu = nullptr;      // [nT]
  const double* la = nullptr;      // [nT]
  const int* mat = nullptr;        // [nT]
  const int* free_map = nullptr;   // [nV] free index or -1
  double density = 1000.0;
};

// Deterministic reduction target: per-block partials, summed in index order
// by a one-block finalize kernel into `result`.
struct ReduceSlot {
  double* partials = nullptr;  // >= grid size of the reducing kernel
  double* result = nullptr;    // 1 value
};

struct Inertia {
  const double* x_tilde = nullptr;  // [3nV]
  double h2g[3] = {0, 0, 0};        // h^2 * gravity
};

// --- mixed stretch (reference src/variables/mixed_stretch.cpp) ------------
// Outputs of the element Hessian phase, [21][nT] packed symmetric (6 packed
// per triangle in 2D). The projection and the inverse are computed in FP64;
// Hpsi (= h^2 Hessian) and Hinv are stored at the storage precision Real
// (stiffness blocks, KKT product and preconditioner); Hloc / Hlinv, when not
// null, receive the same two matrices in FP64 for the local solve of the
// FP32 storage mode. Hinv may be null when nothing reads it.
template <typename Real>
struct HessOut {
  Real* Hpsi = nullptr;
  Real* Hinv = nullptr;
  double* Hloc = nullptr;
  double* Hlinv = nullptr;
};
// Phase 1: polar fit (warm start R), S, ds/dF, material gradient g, PSD-
// projected Hessian Hpsi, its inverse, and gl.
template <typename Real>
void launch_mixed_polar_material(const TetMeshDev& m, const double* x_full, double* R,
                                 double* S, const double* s, Real* dsdF, const HessOut<Real>& ho,
                                 double* g, double* gl, double h2, hipStream_t st);
// Phase 2: element stiffness blocks vol * B^T Hpsi B (10 unique 3x3 blocks
// per tet) and the element residual f_e (12 per tet). mfma (FP32 only):
// the B^T C products on the matrix cores instead of the vector units.
template <typename Real>
void launch_mixed_blocks_rhs(const TetMeshDev& m, const double* x_full, const Inertia& in,
                             const Real* dsdF, const Real* Hpsi, const double* gl,
                             Real* blocks, double* fe, hipStream_t st, bool mfma = false);
// Local solve (reference mixed_stretch.cpp:244-259): lambda and ds from dx,
// with the FP64 Hessian and inverse (Hpsi, Hinv as in HessOut: the Real
// arrays in FP64 mode, Hloc in FP32 storage mode, where Hinv is null and the
// kernel solves with Hloc by LU with partial pivoting).
template <typename Real>
void launch_mixed_local_solve(const TetMeshDev& m, const double* dx, const Real* dsdF,
                              const double* Hpsi, const double* Hinv, const double* g,
                              const double* gl, double* lam, double* ds, ReduceSlot max_ds,
                              hipStream_t st);
// Line-search energy E(alpha) of the mixed formulation (reference
// newton_optimizer.cpp:34-49 with mixed_stretch.cpp:83-123 and
// displacement.cpp:21-31), summed per element.
void launch_mixed_energy(const TetMeshDev& m, const double* x_full, const double* dx,
                         double alpha, const Inertia& in, const double* R, const double* s,
                         const double* ds, const double* lam, double h2, ReduceSlot out,
                         hipStream_t st);
// E(alpha0) into out0 and E(alpha1) into out1 in one pass, each bit-identical
// to launch_mixed_energy's
void launch_mixed_energy2(const TetMeshDev& m, const double* x_full, const double* dx,
                          double alpha0, double alpha1, const Inertia& in, const double* R,
                          const double* s, const double* ds, const double* lam, double h2,
                          ReduceSlot out0, ReduceSlot out1, hipStream_t st);

// --- plain stretch (reference src/variables/stretch.cpp) -------------------
template <typename Real>
void launch_plain_update(const TetMeshDev& m, const double* x_full, const Inertia& in,
                         double h2, Real* blocks, double* fe, hipStream_t st);
void launch_plain_energy(const TetMeshDev& m, const double* x_full, const double* dx,
                         double alpha, const Inertia& in, double h2, ReduceSlot out,
                         hipStream_t st);
// 2D mixed stretch on triangles (same phases as the 3D launchers above;
// per triangle R 4, S/s/g/gl 3, dsdF 12, Hpsi/Hinv 6 packed, in the leading
// rows of the 3D arrays).
template <typename Real>
void launch_mixed2d_polar_material(const TetMeshDev& m, const double* x_full, double* R, double* S,
                                   const double* s, Real* dsdF, const HessOut<Real>& ho,
                                   double* g, double* gl, double h2, hipStream_t st);
template <typename Real>
void launch_mixed2d_blocks_rhs(const TetMeshDev& m, const double* x_full, const Inertia& in,
                               const Real* dsdF, const Real* Hpsi, const double* gl, Real* blocks,
                               double* fe, hipStream_t st);
template <typename Real>
void launch_mixed2d_local_solve(const TetMeshDev& m, const double* dx, const Real* dsdF,
                                const double* Hpsi, const double* Hinv, const double* g,
                                const double* gl, double* lam, double* ds, ReduceSlot max_ds,
                                hipStream_t st);
void launch_mixed2d_energy(const TetMeshDev& m, const double* x_full, const double* dx,
                           double alpha, const Inertia& in, const double* s, const double* ds,
                           const double* lam, double h2, ReduceSlot out, hipStream_t st);
// 2D (triangles, reference tri2d_mesh.cpp and the 4-vector SNH API,
// stable_neohookean.cpp:279-328, no PSD projection): the same outputs with
// the z components zero (z rows carry only the mass).
template <typename Real>
void launch_plain2d_update(const TetMeshDev& m, const double* x_full, const Inertia& in,
                           double h2, Real* blocks, double* fe, hipStream_t st);
void launch_plain2d_energy(const TetMeshDev& m, const double* x_full, const double* dx,
                           double alpha, const Inertia& in, double h2, ReduceSlot out,
                           hipStream_t st);

// --- assembly (reference src/utils/sparse_utils.cpp:141-203) ---------------
// invdiag[3i+r] = 1 / K_(3i+r),(3i+r) from the diagonal blocks (Jacobi).
template <typename Real, typename VT>
void launch_jacobi_diag(int nfree, const int* diag_block, const Real* K, VT* invdiag,
                        hipStream_t st);
// Dense column-major FP64 copy of a BSR 3x3 matrix (zero A first).
template <typename Real>
void launch_bsr_to_dense(int nrows, const int* rowptr, const int* col, const Real* K, double* A,
                         long lda, hipStream_t st);
// The K and rhs gathers in one launch.
// K: block b = mass[b] * I + sum over its slot list of element blocks, with
// slot = (e << 4) | (a << 2) | bb (local vertex pair a, bb of tet e); runs
// over the nub blocks j >= i (pattern positions `upper`) and writes each also
// transposed at `tpos` (the block (j, i)).
// rhs: at free vertex i, the sum over its (tet, corner) incidences of
// f_e[corner]; rows lists the free rows in visiting order (a permutation of
// 0 .. nfree-1).
template <typename Real>
void launch_assemble_bsr_rhs(int nub, const int* upper, const int* tpos, const int* slot_ptr,
                             const int* slots, const double* mass, const Real* blocks, Real* K,
                             int nfree, const int* rows, const int* inc_ptr, const int* inc,
                             const double* fe, double* rhs, hipStream_t st);

// --- BSR 3x3 SpMV / Jacobi PCG (reference Eigen ConjugateGradient semantics)
// Control block of one CG solve (device memory, read by the host at polls).
struct PcgCtrl {
  double rr, b2, threshold;
  double kern_ticks;  // profiling: summed wall-clock ticks of the CG kernels
  int iter, done, max_iters, stop_it;
  int kern_launches;  // profiling: CG kernels accounted in kern_ticks
  int started;        // index of the last CG kernel that did not exit at once
};
// SpMV shapes (lanes per block row G, blocks per lane per trip U,
// non-temporal value loads NT), chosen per matrix by the host autotuner.
int spmv_variant_count();
void spmv_variant_shape(int variant, int* G, int* U, int* NT);
template <typename MT, typename VT>
void launch_spmv(int nrows, const int* rowptr, const int* col, const MT* val, const VT* x,
                 VT* y, hipStream_t st, int variant = 0);
// y = b - K x in FP64 (defect correction residual)
template <typename MT>
void launch_residual(int nrows, const int* rowptr, const int* col, const MT* val,
                     const double* x, const double* b, double* r, ReduceSlot r2,
                     hipStream_t st);

// One CG iteration is one kernel (see k_solver.hip). Per scalar row the
// iteration state is the 4-tuple {r, q, p, D^-1} (16 B in FP32, 32 B in
// FP64), double-buffered; every kernel writes 5 partial sums per block,
// double-buffered as well.
constexpr int kCgParts = 5;
// (lanes per block row, blocks per lane in flight, workgroup size) of the CG
// kernel shapes
int cg_variant_count();
void cg_variant_shape(int variant, int* G, int* U, int* BS);
// Grid for shape `variant` on n scalar rows with `per_cu` workgroups on
// every CU (fewer when the rows do not fill that many workgroups).
int cg_blocks_per_cu(int variant, int n, int per_cu);
template <typename T>
struct PcgBuffers {
  int n = 0;  // scalar rows (3 * nfree)
  const int* rowptr = nullptr;
  const int* col = nullptr;
  const T* K = nullptr;
  const T* invdiag = nullptr;
  const double* b = nullptr;   // right-hand side
  double* x = nullptr;         // solution (FP64 accumulation)
  T* state[2] = {nullptr, nullptr};            // 4 n values each
  PcgCtrl* ctrl = nullptr;
  // optional: mapped pinned host memory that receives the control block when
  // the solve stops (the host polls it instead of copying ctrl)
  PcgCtrl* mirror = nullptr;
  double* parts[2] = {nullptr, nullptr};       // kCgParts * blocks each
  unsigned long long* stamps[2] = {nullptr, nullptr};  // 2 * blocks each (profiling) or null
  int variant = 0;  // CG kernel shape
  int blocks = 0;   // grid of the CG kernel, a multiple of 8, <= kCgMaxBlocks
};
constexpr int kCgMaxBlocks = 4096;
// Default grid of the CG kernel for n scalar rows.
int pcg_default_blocks(int n);
template <typename T>
void launch_pcg_init(const PcgBuffers<T>& B, double tol, int max_iters, hipStream_t st);
// CG kernel `it` (it = 0 forms p_0 = D^-1 b and q_0 = K p_0; kernel it >= 1
// also applies the update of iteration it-1 and tests its residual).
template <typename T>
void launch_pcg_iteration(const PcgBuffers<T>& B, int it, hipStream_t st);

// Batched CG over the diagonal blocks of a block-diagonal K (independent
// simulations): workgroups [wg0[b], wg0[b+1]) own the block rows
// [row0[b], row0[b+1]) of system b; ctrls[b] is that system's control block.
template <typename T>
struct BatchCgBuffers {
  int n = 0;  // scalar rows of all systems
  const int* rowptr = nullptr;
  const int* col = nullptr;
  const T* K = nullptr;
  const T* invdiag = nullptr;
  const double* b = nullptr;
  double* x = nullptr;
  T* state[2] = {nullptr, nullptr};
  PcgCtrl* ctrls = nullptr;
  double* parts[2] = {nullptr, nullptr};  // kCgParts * nwg each
  int nwg = 0;
  const int* wg_sys = nullptr;  // [nwg] system of each workgroup
  const int* wg0 = nullptr;     // [nsys + 1] first workgroup of each system
  const int* row0 = nullptr;    // [nsys + 1] first block row of each system
  // optional: the workgroups to run ([nactive], every workgroup of each
  // system still iterating); null runs all nwg
  const int* active = nullptr;
  int nactive = 0;
  int variant = 0;
  // optional, in mapped host memory: a copy of every control block, written
  // when the system's solve starts (its threshold, or done) and when it
  // stops (every field, then done), and progress->started, the index of the
  // last kernel that began
  PcgCtrl* mirror = nullptr;
  PcgCtrl* progress = nullptr;
};
// x = 0 and the initial state for all rows (scratch: >= 256 doubles)
template <typename T>
void launch_pcg_batch_init(const BatchCgBuffers<T>& B, double* scratch, hipStream_t st);
// control blocks from per-system |b|^2, tolerance, iteration budget, on flag
void launch_pcg_batch_setup(int nsys, PcgCtrl* ctrls, const double* b2, const double* tol,
                            const int* max_iters, const int* on, hipStream_t st,
                            PcgCtrl* mirror = nullptr);
// Batched CG kernel `it`. For graph replay pass it_in / it_out (two distinct
// device words): the kernel then takes its iteration index from *it_in and
// stores it + 1 in *it_out; `it` only selects the buffer parity.
template <typename T>
void launch_pcg_batch_iteration(const BatchCgBuffers<T>& B, int it, hipStream_t st,
                                const int* it_in = nullptr, int* it_out = nullptr);
// *p = v on the device (a kernel, so that it may sit in a captured graph)
void launch_set_int(int* p, int v, hipStream_t st);

// Per-segment reductions in a fixed order: out[b] = sum v (mode 0), sum v^2
// (mode 1) or max |v| (mode 2) over indices [seg[b], seg[b+1]) of each of
// ncomp components (component c at v + c * comp_stride). Each segment is cut
// into kSegSlices equal slices, one workgroup each; `part` holds
// nseg * kSegSlices partial results, combined per segment in slice order.
constexpr int kSegSlices = 16;
void launch_seg_reduce(int nseg, const double* v, int ncomp, long comp_stride, const int* seg,
                       int mode, double* out, double* part, hipStream_t st);
// Batched line-search acceptance: x += alpha[sim of vertex] dx on free
// vertices, s += alpha[sim of tet] ds.
void launch_accept_batch(int nV, const int* free_map, const int* vert_sim, int nT,
                         const int* tet_sim, const double* alpha_sim, double* x_full,
                         const double* dx, double* s, const double* ds, hipStream_t st);
// Batched line-search energies: e_tet[e] = energy of tet e at the step length
// of its simulation (0 for simulations that are off).
void launch_mixed_energy_batch(const TetMeshDev& m, const double* x_full, const double* dx,
                               const int* tet_sim, const double* alpha_sim, const int* sim_on,
                               const Inertia& in, const double* R, const double* s,
                               const double* ds, const double* lam, double h2, double* e_tet,
                               hipStream_t st);

// --- indefinite (KKT) system and MINRES ("minres-indefinite-block";
// reference linear_system.h:60-127, block_matrix.h:74-125) ----------------
// Scalars of one MINRES solve (Eigen 3.4's recurrence), device resident.
struct MinresCtrl {
  double beta, beta_prev, alpha;     // beta_k, beta_{k-1}, alpha_k
  double c, c_old, s, s_old, eta;    // Givens state
  double beta_one, res2, rhs2, threshold2;
  double r1, r2, r3, coef;           // update of the last iteration: p and x
  int iter, done, max_iters, pending, fail, pad_;
};
// A KKT vector: x part (3 nfree) and the element parts (ns x nT, stride nT;
// ns = 6 on tetrahedra, 3 on triangles, storage 6 x nT either way).
struct KktVec {
  double* x = nullptr;
  double* s = nullptr;
  double* l = nullptr;
};
// Chebyshev steps of the mass block of the preconditioner
constexpr int kMassChebSteps = 4;
template <typename Real>
struct KktSystem {
  TetMeshDev m;
  const int* rowptr = nullptr;   // BSR pattern over free vertices
  const int* col = nullptr;
  const double* mass = nullptr;  // PMP coefficient per block
  const int* inc_ptr = nullptr;  // (tet, corner) incidences per free vertex
  const int* inc = nullptr;
  const Real* dsdF = nullptr;    // [54][nT]
  const Real* Hpsi = nullptr;    // [21][nT], h^2 Hessian (PSD-projected)
  const Real* Hinv = nullptr;    // [21][nT]
  double* Sinv = nullptr;        // [21][nT] inverted element Schur blocks
  const double* mlump_inv = nullptr;  // [nfree] 1 / row sum of PMP
  const double* mdinv = nullptr;      // [nfree] 1 / diagonal of PMP
  double* fe = nullptr;               // [12 nT] scratch
  double* cheb_r = nullptr;           // [3 nfree] scratch
  double* cheb_d[2] = {nullptr, nullptr};
  double* parts[2] = {nullptr, nullptr};  // kkt_tet_blocks + kkt_vertex_blocks each
  MinresCtrl* ctrl = nullptr;
};
struct MinresBuffers {
  KktVec v[2], w[2], p[3];  // Lanczos vectors (unnormalised) and search directions
  KktVec x;                 // solution [dx; ds; lambda]
};
int kkt_tet_blocks(int nT);
int kkt_vertex_blocks(int nfree);
// Inverted element Schur blocks of the preconditioner from the current state.
template <typename Real>
void launch_kkt_setup(const KktSystem<Real>& S, hipStream_t st);
// Right-hand side into b (= v_0 before scaling), w = P b, and the scalars.
template <typename Real>
void launch_minres_init(const KktSystem<Real>& S, const double* x_full, const Inertia& in,
                        const double* g, const double* Smat, const double* s, KktVec b, KktVec w,
                        double tol, int max_iters, hipStream_t st);
// MINRES iteration k (k = 0, 1, ...): applies the solution update of
// iteration k-1 first, so one more launch after the stop completes x.
// tet_ev: optional pair of events recorded around the element kernel
// (k_mr_tet), for profiling.
template <typename Real>
void launch_minres_iteration(const KktSystem<Real>& S, const MinresBuffers& B, long k,
                             hipStream_t st, hipEvent_t* tet_ev = nullptr);
// Diagnostics: out = K in (which 0) or out = P in (which 1) at the current
// state; scratch is a KKT vector the product may write.
template <typename Real>
void launch_kkt_apply(const KktSystem<Real>& S, int which, KktVec in, KktVec out, KktVec scratch,
                      hipStream_t st);

// --- vectors ---------------------------------------------------------------
// x_full[free] += alpha * dx ; s += alpha * ds (line-search acceptance,
// reference linesearch.h:120-126)
void launch_accept_step(const TetMeshDev& m, double* x_full, const double* dx, double* s,
                        const double* ds, double alpha, int with_s, hipStream_t st);
void launch_bdf_xtilde(int n, int order, const double* const* xh, const double* const* vh,
                       const double* a, double dt, double* x_tilde, hipStream_t st);
void launch_bdf_update(int n, int order, const double* const* xh, const double* a, double dt,
                       const double* x_full, double* x_new, double* v_new, hipStream_t st);
void launch_reduce_maxabs(int n, const double* v, ReduceSlot out, hipStream_t st);
void launch_fill(double* p, double v, long n, hipStream_t st);
// x[3 ids[k] + r] = pos[3k + r] for k < n (pinned vertices' prescribed positions)
void launch_scatter3(int n, const int* ids, const double* pos, double* x, hipStream_t st);
void launch_axpy(long n, double alpha, const double* x, double* y, hipStream_t st);  // y += a x

// --- batched element entry points (MaterialModel / polar on many points) ---
void launch_batch_polar(int n, const double* F, double* R, double* s, double* dsdF,
                        int* iters, hipStream_t st);
void launch_batch_material(int n, const int* type, const double* mu, const double* la,
                           const double* s, int psd, double* E, double* g, double* H,
                           hipStream_t st);
void launch_batch_snh_f9(int n, const double* mu, const double* la, const double* F,
                         double* E, double* g, double* H, hipStream_t st);

int reduce_grid(long n);  // grid size used by the reductions for n items

// --- "eigen-pcg-IC" (k_ic.hip; reference linear_solver_factory.cpp:87-97) --
// CG scalars of one solve, device resident.
struct IcCtrl {
  double b2, threshold, rr, abs_new, alpha, beta, tol2, mindiag;
  int iter, done, max_iters, first;
};
// Multicolour block IC(0) factor of K over the free vertices. Rows are
// numbered by position in the colour order; the lists, L and Dinv are
// stored by position, vectors and `scale` by row.
struct IcFactorDev {
  int nfree = 0, dim = 3;
  const int* rowptr = nullptr;  // BSR pattern of K (by row)
  const int* col = nullptr;
  const int* diag = nullptr;    // diagonal block of each row
  const int* color_rows = nullptr;  // row at each position
  // strictly lower blocks (earlier colours) of each position, sorted by
  // position: column row, column position, K block
  const int* lo_ptr = nullptr;
  const int* lo_col = nullptr;
  const int* lo_pos = nullptr;
  const int* lo_k = nullptr;
  // strictly upper blocks: row j and the index of L_ji in the lower lists
  const int* up_ptr = nullptr;
  const int* up_col = nullptr;
  const int* up_l = nullptr;
  double* L = nullptr;      // 9 per lower block, row-major
  double* Dinv = nullptr;   // 9 per position: inverse of the diagonal factor
  double* scale = nullptr;  // 3 per row
  int* fail = nullptr;      // set when a pivot of the last factorization failed
};
struct IcVectors {
  const double* b = nullptr;
  double *x = nullptr, *r = nullptr, *q = nullptr, *y = nullptr, *zt = nullptr, *z = nullptr;
  double* p[2] = {nullptr, nullptr};  // search direction, ping-pong between iterations
  double* parts = nullptr;            // ic_partials(...) values
  IcCtrl* ctrl = nullptr;
};
// partial-sum slots the solver kernels need (color_ptr: positions per colour)
int ic_partials(int nfree, const int* color_ptr, int ncolors);
// Row scaling and min scaled diagonal (into ctrl->mindiag)
template <typename MT>
void launch_ic_scale(const IcFactorDev& f, const MT* K, double* parts, IcCtrl* ctrl,
                     hipStream_t st);
// The factorization, colour by colour (color_ptr on the host)
template <typename MT>
void launch_ic_factor(const IcFactorDev& f, const MT* K, double shift, const int* color_ptr,
                      int ncolors, hipStream_t st);
// CG set-up (x = 0, r = b, first preconditioned residual) and iteration k
// (counted from 0; selects the ping-pong direction buffers)
void launch_ic_cg_init(const IcFactorDev& f, const IcVectors& v, double tol, int max_iters,
                       const int* color_ptr, int ncolors, hipStream_t st);
template <typename MT>
void launch_ic_cg_iteration(const IcFactorDev& f, const MT* K, const IcVectors& v, int k,
                            const int* color_ptr, int ncolors, hipStream_t st);

}  // namespace mfemhip
