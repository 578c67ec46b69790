// Host framework of the device Newton step, mirroring the reference's
// class structure and names:
//   ImplicitIntegrator / BDF<I>           reference src/time_integrators/
//   Variable / MixedVariable              reference src/variables/variable.h,
//   Displacement / Stretch / MixedStretch              mixed_variable.h, ...
//   LinearSolver / SystemMatrixPD / CG    reference src/linear_solvers/
//   Optimizer / NewtonOptimizer           reference src/optimizers/
//   SimState                              reference src/simulation_state.h
// Vectors and per-element state live in device memory; the host keeps the
// control flow (Newton loop, line search, termination tests), exactly as the
// reference's NewtonOptimizer does.
#pragma once
#include <array>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "../kernels.h"
#include "config.h"
#include "device.h"
#include "factory.h"
#include "json.h"
#include "material_model.h"
#include "mesh.h"

namespace mfem {

// Pool of deterministic reduction slots (see reduce.h) with pinned read-back.
class Reducer {
 public:
  static constexpr int kSlots = 12;
  void init(long max_items);
  mfemhip::ReduceSlot slot(int k);
  double read(int k, hipStream_t st);
  // results of slots k0 .. k0 + n - 1 with one copy and one synchronisation
  void read_range(int k0, int n, double* out, hipStream_t st);

 private:
  long stride_ = 0;
  DeviceArray<double> partials_;
  std::unique_ptr<PinnedBuffer> results_;
};

// Device resources shared by the variables and solvers of one simulation.
struct DeviceContext {
  hipStream_t stream = nullptr;
  bool own_stream = false;
  Precision precision = Precision::FP64;
  bool fp32() const { return precision == Precision::MIXED; }
  Reducer reducer;
  RealArray blocks;           // element stiffness blocks, 90 per tet
  DeviceArray<double> fe;     // element residual vectors, 12 per tet
  DeviceArray<double> x_tilde;
  double h2g[3] = {0, 0, 0};
  mfemhip::Inertia inertia() const;
  ~DeviceContext();
};

// ---------------------------------------------------------------------------
class ImplicitIntegrator {
 public:
  ImplicitIntegrator(double h) : h_(h) {}
  virtual ~ImplicitIntegrator() = default;
  virtual void x_tilde(double* out, hipStream_t st) const = 0;  // device out
  virtual double dt() const = 0;
  virtual void update(const double* x, hipStream_t st) = 0;      // device x
  virtual const double* x_prev() const = 0;
  virtual const double* v_prev() const = 0;

 protected:
  double h_;
};

template <int I>
class BDF : public ImplicitIntegrator {
 public:
  static std::string name() { return "BDF" + std::to_string(I); }
  BDF(const std::vector<double>& x0, const std::vector<double>& v0, double h, hipStream_t st);
  void x_tilde(double* out, hipStream_t st) const override;
  double dt() const override { return beta() * h_; }
  void update(const double* x, hipStream_t st) override;
  const double* x_prev() const override { return xs_[head_].data(); }
  const double* v_prev() const override { return vs_[head_].data(); }

 private:
  static std::array<double, I> alphas();
  static double beta();
  std::array<DeviceArray<double>, I> xs_, vs_;
  int head_ = 0;
  int n_ = 0;
};

class IntegratorFactory
    : public Factory<TimeIntegratorType, ImplicitIntegrator, const std::vector<double>&,
                     const std::vector<double>&, double, hipStream_t> {
 public:
  IntegratorFactory();
};

// ---------------------------------------------------------------------------
template <int DIM>
class Variable {
 public:
  Variable(std::shared_ptr<Mesh> mesh, DeviceContext* ctx) : mesh_(std::move(mesh)), ctx_(ctx) {}
  virtual ~Variable() = default;
  // element phase at the full configuration x (device, 3 nV) for time step dt
  virtual void update(const double* x_full, double dt) = 0;
  virtual void reset() = 0;
  virtual void pre_solve() {}
  virtual void post_solve() {}
  virtual int size() const = 0;
  // energy of this variable at (x + alpha dx, s + alpha ds), inertia included
  // when the variable owns the element residual (device backend detail),
  // launched into `out` on the context's stream; false if it has none
  virtual bool energy_launch(const double* x_full, const double* dx, double alpha,
                             mfemhip::ReduceSlot out) {
    return false;
  }
  // energies at alpha0 and alpha1 into out0 and out1 (the same values as two
  // energy_launch calls)
  virtual bool energy_launch2(const double* x_full, const double* dx, double alpha0,
                              mfemhip::ReduceSlot out0, double alpha1, mfemhip::ReduceSlot out1) {
    if (!energy_launch(x_full, dx, alpha0, out0)) return false;
    energy_launch(x_full, dx, alpha1, out1);
    return true;
  }
  std::shared_ptr<Mesh> mesh() const { return mesh_; }

 protected:
  std::shared_ptr<Mesh> mesh_;
  DeviceContext* ctx_;
};

template <int DIM>
class MixedVariable : public Variable<DIM> {
 public:
  using Variable<DIM>::Variable;
  // local solve for (lambda, ds) given the displacement step dx (device)
  virtual void solve(const double* dx) = 0;
  virtual DeviceArray<double>& value() = 0;   // s
  virtual DeviceArray<double>& delta() = 0;   // ds
  virtual DeviceArray<double>& lambda() = 0;
  virtual int size_dual() const = 0;
  // max |ds| of the last solve, in the reducer slot the local solve wrote
  virtual int max_abs_delta_slot() const = 0;
};

template <int DIM>
class Displacement : public Variable<DIM> {
 public:
  Displacement(std::shared_ptr<Mesh> mesh, std::shared_ptr<SimConfig> config, DeviceContext* ctx);
  void update(const double*, double) override {}
  void reset() override;
  void pre_solve() override;   // x_tilde for this timestep
  void post_solve() override;  // BDF update, mesh positions
  int size() const override { return 3 * this->mesh_->nfree(); }
  DeviceArray<double>& value() { return x_full_; }  // full configuration (3 nV)
  DeviceArray<double>& delta() { return dx_; }      // free-DOF step (3 nfree)
  const std::shared_ptr<ImplicitIntegrator>& integrator() const { return integrator_; }
  // launch max |dx| into the reducer; returns its slot
  int launch_max_abs_delta();

 private:
  std::shared_ptr<SimConfig> config_;
  std::shared_ptr<ImplicitIntegrator> integrator_;
  DeviceArray<double> x_full_, dx_;
  DeviceArray<int> pinned_ids_;      // scripted boundary motion
  DeviceArray<double> pinned_pos_;
};

template <int DIM>
class Stretch : public Variable<DIM> {
 public:
  static std::string name() { return "stretch"; }
  // Needs the 9-vector (F) API of every element's material: SNH, NH or Fung
  // (Corot, FCR and ARAP leave it unimplemented, reference
  // material_model.h:43-62).
  Stretch(std::shared_ptr<Mesh> mesh, std::shared_ptr<SimConfig> config, DeviceContext* ctx);
  void update(const double* x_full, double dt) override;
  void reset() override {}
  int size() const override { return 0; }
  bool energy_launch(const double* x_full, const double* dx, double alpha,
                     mfemhip::ReduceSlot out) override;

 private:
  double h2_ = 0;
};

template <int DIM>
class MixedStretch : public MixedVariable<DIM> {
 public:
  static std::string name() { return "mixed-stretch"; }
  MixedStretch(std::shared_ptr<Mesh> mesh, std::shared_ptr<SimConfig> config, DeviceContext* ctx)
      : MixedVariable<DIM>(std::move(mesh), ctx) {}
  void update(const double* x_full, double dt) override;
  void reset() override;
  void post_solve() override;  // lambda <- 0 (reference mixed_stretch.cpp:307-310)
  int size() const override { return 6 * this->mesh_->nT(); }
  int size_dual() const override { return 6 * this->mesh_->nT(); }
  void solve(const double* dx) override;
  bool energy_launch(const double* x_full, const double* dx, double alpha,
                     mfemhip::ReduceSlot out) override;
  bool energy_launch2(const double* x_full, const double* dx, double alpha0,
                      mfemhip::ReduceSlot out0, double alpha1, mfemhip::ReduceSlot out1) override;
  DeviceArray<double>& value() override { return s_; }
  DeviceArray<double>& delta() override { return ds_; }
  DeviceArray<double>& lambda() override { return la_; }
  int max_abs_delta_slot() const override;

  // element state (structure of arrays, stride nT)
  DeviceArray<double> R_, S_, s_, la_, ds_, g_, gl_;
  RealArray dsdF_, Hpsi_, Hinv_;
  // FP64 copies of Hpsi (and on triangles its inverse) for the local solve
  // when the element arrays above are stored in FP32 (empty in FP64 mode,
  // where Hpsi_ and Hinv_ are the FP64 matrices); the tet local solve
  // solves with Hloc_ by LU
  DeviceArray<double> Hloc_, Hlinv_;
  // the FP64 Hessian and inverse the local solve reads (inverse null: the
  // kernel forms it)
  const double* local_hpsi() const { return Hpsi_.fp32() ? Hloc_.data() : Hpsi_.f64(); }
  const double* local_hinv() const {
    if (!Hpsi_.fp32()) return Hinv_.f64();
    return Hlinv_.size() ? Hlinv_.data() : nullptr;
  }
  // false when the linear solver does not use the Schur-complement system
  // (its element stiffness blocks and residual are then not formed)
  bool schur_blocks_ = true;
  // 3D FP32 storage: form the stiffness blocks' products on the matrix
  // cores (MFMA) instead of the vector units
  bool mfma_blocks_ = false;

 private:
  template <typename Real>
  mfemhip::HessOut<Real> hess_out();
  double h2_ = 0;
};

template <int DIM>
class VariableFactory
    : public Factory<VariableType, Variable<DIM>, std::shared_ptr<Mesh>,
                     std::shared_ptr<SimConfig>, DeviceContext*> {
 public:
  VariableFactory();
};

template <int DIM>
class MixedVariableFactory
    : public Factory<VariableType, MixedVariable<DIM>, std::shared_ptr<Mesh>,
                     std::shared_ptr<SimConfig>, DeviceContext*> {
 public:
  MixedVariableFactory();
};

// ---------------------------------------------------------------------------
template <int DIM>
struct SimState;

struct SolveStats {
  int iterations = 0;       // CG iterations (all refinement sweeps)
  int refinements = 0;      // defect-correction sweeps (MIXED)
  double rel_residual = 0;  // |b - K x| / |b| at exit (FP64)
  double spmv_ms = 0;       // summed device time of the CG SpMV kernel, when profiled
  int spmv_count = 0;
};

// K = PMP + sum of element blocks on the BSR pattern, rhs from element
// residuals (reference linear_solvers/linear_system.h:9-58).
template <int DIM>
class SystemMatrixPD {
 public:
  void pre_solve(SimState<DIM>* state);
  void post_solve(SimState<DIM>* state, const double* dx);
  RealArray K;                // nnzb * 9
  RealArray invdiag;          // 3 nfree
  DeviceArray<double> rhs;    // 3 nfree
};

template <int DIM>
class LinearSolver {
 public:
  explicit LinearSolver(SimState<DIM>* state) : state_(state) {}
  virtual ~LinearSolver() = default;
  virtual void solve() = 0;
  // solve() with the solver's last acceptance test (if it reads a result
  // back) left for settle(): work that depends only on the solution can be
  // queued behind that test meanwhile, with no host round trip before it.
  void solve_deferred() {
    struct Reset {
      bool& f;
      ~Reset() { f = false; }
    } reset{defer_};
    defer_ = true;
    solve();
  }
  // Completes a deferred solve. False: the solution changed after solve()
  // returned, and work queued behind it has to be queued again.
  virtual bool settle() { return true; }
  const SolveStats& stats() const { return stats_; }
  bool profile_spmv = false;

 protected:
  SimState<DIM>* state_;
  SolveStats stats_;
  bool defer_ = false;
};

// "eigen-pcg-diag": Jacobi-preconditioned CG with Eigen's stopping rule.
// FP64: one CG solve on the FP64 system. MIXED: FP32 CG sweeps inside an FP64
// defect-correction loop that stops on the true FP64 residual.
template <int DIM>
class PcgSolver : public LinearSolver<DIM> {
 public:
  explicit PcgSolver(SimState<DIM>* state);
  ~PcgSolver() override;
  void solve() override;
  bool settle() override;
  SystemMatrixPD<DIM>& system() { return system_; }
  // Plain SpMV shape chosen by its autotuner (-1 before the first solve) and
  // the mean device time of every shape on this matrix.
  int spmv_variant() const { return variant_; }
  const std::vector<double>& spmv_tuning_ms() const { return tune_ms_; }
  // CG kernel shape and grid: picked by timing every shape on a few balanced
  // grids on the first solve of a large system, or forced with
  // force_cg_shape (variant -1 re-enables the tuner; blocks 0 takes the
  // default grid).
  struct CgTrial {
    int variant, blocks;
    double ms;  // mean device time per CG iteration
  };
  int cg_variant() const { return cg_variant_; }
  int cg_blocks() const { return cg_blocks_; }
  const std::vector<CgTrial>& cg_trials() const { return cg_trials_; }
  void force_cg_shape(int variant, int blocks);
  // Profiling stamps of the last two CG launches: [2][blocks][start, end]
  // device wall-clock ticks, and the clock rate in kHz.
  std::vector<unsigned long long> cg_stamps(hipStream_t st) const;
  double wall_clock_khz() const { return wall_khz_; }

 private:
  template <typename T>
  mfemhip::PcgBuffers<T> buffers(const double* b, double* x);
  template <typename T>
  int run_cg(const double* b, double* x, double tol, int max_iters);
  template <typename T>
  void tune_spmv();
  template <typename T>
  void tune_cg();
  void refine(double* xo);
  // the FP32 defect-correction loop's state, kept across a deferred check
  struct Refinement {
    int budget = 0, sweeps = 0;
    double b2 = 0, r2 = 0, thr = 0;
    bool pending = false;
  } rf_;
  int variant_ = -1;
  std::vector<double> tune_ms_;
  int cg_variant_ = -1, cg_blocks_ = 0;
  std::vector<CgTrial> cg_trials_;
  SystemMatrixPD<DIM> system_;
  DeviceArray<double> x_, r64_, dx64_;
  RealArray q_;                  // plain SpMV output (autotuner)
  RealArray state0_, state1_;    // CG state {r, q, p, D^-1}, 4 n values each
  DeviceArray<double> parts_;    // 2 x kCgParts x kCgMaxBlocks partial sums
  DeviceArray<unsigned long long> stamps_;  // 2 x 2 x kCgMaxBlocks clock stamps
  DeviceArray<mfemhip::PcgCtrl> ctrl_;
  mfemhip::PcgCtrl last_{};      // control block at the end of the last CG solve
  std::unique_ptr<PinnedBuffer> pinned_;  // mapped host copy of the control block
  double wall_khz_ = 100000.0;   // device wall-clock rate
};

// "eigen-llt", "eigen-ldlt", "cholmod" and "eigen-lu": a direct solve of the
// Schur-complement system K dx = rhs, refactored on every call as the
// reference's EigenSolver does (eigen_solver.h:17-31), here as a dense FP64
// factorization on the device (rocSOLVER potrf/potrs; getrf/getrs for LU).
// K is SPD, so the LLT, LDLT and supernodal Cholesky variants give the same
// solution. Systems above kMaxRows scalar rows are rejected.
template <int DIM>
class DirectSolver : public LinearSolver<DIM> {
 public:
  static constexpr long kMaxRows = 50000;
  DirectSolver(SimState<DIM>* state, bool lu);
  ~DirectSolver() override;
  void solve() override;
  SystemMatrixPD<DIM>& system() { return system_; }

 private:
  bool lu_ = false;
  void* handle_ = nullptr;  // rocblas_handle
  SystemMatrixPD<DIM> system_;
  DeviceArray<double> dense_, x_;
  DeviceArray<int> ipiv_, info_;
};

// "eigen-pcg-IC": CG with an incomplete-Cholesky preconditioner (reference
// linear_solver_factory.cpp:87-97, Eigen's ConjugateGradient with
// IncompleteCholesky). The factor is IC(0) on the 3x3 vertex blocks of K in
// smallest-last multicolour order, with Eigen's row scaling and diagonal-shift
// restarts (k_ic.hip), refactored on every call as Eigen's compute() does.
// Vectors and the factor are FP64; K is read in the simulation's precision.
template <int DIM>
class IcPcgSolver : public LinearSolver<DIM> {
 public:
  explicit IcPcgSolver(SimState<DIM>* state);
  ~IcPcgSolver() override;
  void solve() override;
  SystemMatrixPD<DIM>& system() { return system_; }
  int colors() const { return color_ptr_.empty() ? 0 : (int)color_ptr_.size() - 1; }
  const std::vector<int>& vertex_colors() const { return color_; }  // per free vertex
  double shift() const { return shift_; }     // diagonal shift of the last factor
  int attempts() const { return attempts_; }  // factorizations of the last solve

 private:
  void setup();
  mfemhip::IcFactorDev factor_dev();
  template <typename MT>
  void run(const MT* K);
  SystemMatrixPD<DIM> system_;
  bool ready_ = false;
  std::vector<int> color_, color_ptr_;
  DeviceArray<int> color_rows_, lo_ptr_, lo_col_, lo_pos_, lo_k_, up_ptr_, up_col_, up_l_, fail_;
  DeviceArray<double> L_, Dinv_, scale_, vecs_, parts_;
  DeviceArray<mfemhip::IcCtrl> ctrl_;
  mfemhip::IcCtrl last_{};
  std::unique_ptr<PinnedBuffer> pinned_;  // two control-block snapshots + a flag
  hipEvent_t poll_ev_[2] = {nullptr, nullptr};
  // two CG iterations (both parities of the direction buffers) captured as
  // one graph per matrix precision, replayed for every pair
  hipGraphExec_t pair_graph_[2] = {nullptr, nullptr};
  const void* pair_graph_K_[2] = {nullptr, nullptr};
  double shift_ = 0;
  int attempts_ = 0;
};

// "minres-indefinite-block": MINRES on the KKT system of the mixed
// formulation (reference linear_solver_factory.cpp:117-134 with
// SystemMatrixIndefinite, linear_system.h:60-99, and the matrix-free
// BlockMatrix product, block_matrix.h:74-125). dx, ds and lambda come from the
// solution directly; there is no local solve. The block preconditioner is
// the GPU form described in k_minres.hip. Needs one mixed-stretch variable.
template <int DIM>
class MinresSolver : public LinearSolver<DIM> {
 public:
  explicit MinresSolver(SimState<DIM>* state);
  ~MinresSolver() override;
  void solve() override;
  // |r|_P / |b| as estimated by the recurrence at exit
  double estimated_error() const { return last_.rhs2 > 0 ? std::sqrt(last_.res2 / last_.rhs2) : 0; }
  // Diagnostics at the current state, KKT vectors in the reference's layout
  // [dx (D per free vertex); ds; lambda] with element-major s-vectors (6 in
  // 3D, 3 in 2D): which 0: K in, 1: P in,
  // 2: the right-hand side (in unused).
  std::vector<double> apply(int which, const std::vector<double>& in);
  long kkt_size() const { return Nref_; }  // reference layout: D nfree + 2 ns nT

 private:
  template <typename Real>
  mfemhip::KktSystem<Real> system();
  template <typename Real>
  void run();
  template <typename Real>
  void apply_device(int which);
  MixedStretch<DIM>* mixed_ = nullptr;
  long n_ = 0, N_ = 0, Nref_ = 0;  // device x part; device and reference vector sizes
  int ns_ = 6;                     // entries of s per element (3 in 2D)
  DeviceArray<double> vecs_;   // v[2], w[2], p[3]: 7 KKT vectors
  DeviceArray<double> sinv_, mlump_inv_, mdinv_, cheb_, parts_;
  DeviceArray<mfemhip::MinresCtrl> ctrl_;
  mfemhip::MinresCtrl last_{};
  std::unique_ptr<PinnedBuffer> pinned_;  // two control-block snapshots
  hipEvent_t poll_ev_[2] = {nullptr, nullptr};
  // profiling: event pairs around the element kernel of the first
  // kProfiledIters iterations of a solve
  static constexpr int kProfiledIters = 32;
  std::vector<hipEvent_t> prof_ev_;
};

template <int DIM>
class LinearSolverFactory
    : public Factory<LinearSolverType, LinearSolver<DIM>, SimState<DIM>*> {
 public:
  LinearSolverFactory();
};

// ---------------------------------------------------------------------------
struct StepStats {
  int newton_iters = 0;
  int pcg_iters = 0;
  int refinements = 0;
  int ls_iters = 0;
  int ls_fail = 0;
  double decrement = 0;
  double energy = 0;
  double alpha = 0;
  double res = 0;
  double rel_residual = 0;
  double update_ms = 0, solve_ms = 0, ls_ms = 0, step_ms = 0;
  double spmv_ms = 0;
  int spmv_count = 0;
};

template <int DIM>
class Optimizer {
 public:
  // Takes ownership of the state (the reference moves it, optimizer.h:19).
  explicit Optimizer(std::unique_ptr<SimState<DIM>> state) : state_(std::move(state)) {}
  virtual ~Optimizer() = default;
  static std::string name() { return "base"; }
  virtual void reset();
  virtual void step() = 0;
  SimState<DIM>& state() { return *state_; }
  std::function<void(const SimState<DIM>&)> callback = [](const SimState<DIM>&) {};
  // set when `callback` is the caller's: each Newton iteration then ends with
  // a synchronisation, so the callback sees that iteration's timings
  bool sync_each_iteration = false;
  // device-event timings of each Newton iteration's phases (update, solve,
  // line search); also on with show_timing or a caller's callback. Each
  // event is a marker in the stream that costs the device a few
  // microseconds, so they are off by default.
  bool phase_timing = false;

 protected:
  std::unique_ptr<SimState<DIM>> state_;
};

template <int DIM>
class NewtonOptimizer : public Optimizer<DIM> {
 public:
  static std::string name() { return "newton"; }
  using Optimizer<DIM>::Optimizer;
  ~NewtonOptimizer() override;
  void reset() override;
  void step() override;
  // pieces of step(), exposed for phase-level checks
  void update_system();
  void substep(double& decrement);
  double energy(double alpha);
  const StepStats& last_stats() const { return stats_; }
  LinearSolver<DIM>* solver() { return solver_.get(); }
  // One row per Newton iteration of the last step: the reference's
  // OptimizerData columns (newton_optimizer.cpp:60-66) and the device time
  // of its phases.
  struct Record {
    int iteration;
    double energy, energy_res, newton_dec, alpha;
    double update_ms, solve_ms, ls_ms;
    int solver_iters, ls_halvings;
  };
  const std::vector<Record>& history() const { return history_; }

 private:
  // the per-step table and timing of the reference's print_stats /
  // print_timing (optimizer_data.cpp:21-183)
  void print_history() const;
  std::vector<Record> history_;
  // E(alpha) as one partial per elastic variable in reducer slots base, ...;
  // returns the number of slots written (at most 2)
  int launch_energy(double alpha, int base);
  // E(alpha0) and E(alpha1) into slots base0.. and base1.. in one pass where
  // a variable supports it (same values as two launch_energy calls)
  int launch_energy2(double alpha0, int base0, double alpha1, int base1);
  // the decrement's slots after a solve: max |dx| and, with a mixed
  // variable, max |ds| (returns the number of leading slots to read)
  int launch_decrement();
  std::unique_ptr<LinearSolver<DIM>> solver_;
  StepStats stats_;
  // four timing events per Newton iteration of the current step
  std::vector<hipEvent_t> ev_;
};

template <int DIM>
class OptimizerFactory
    : public Factory<OptimizerType, Optimizer<DIM>, std::unique_ptr<SimState<DIM>>> {
 public:
  OptimizerFactory();
};

// ---------------------------------------------------------------------------
// Scene description accepted from the C ABI (arrays) or built by load(json).
struct SceneArrays {
  int dim = 3;                // 2: triangles in the plane
  std::vector<double> V;      // nV x 3 (z = 0 for dim 2)
  std::vector<int> T;         // nT x (dim + 1)
  std::vector<double> mu, la;
  std::vector<int> mat;
  std::vector<int> fixed;     // optional pin mask
  std::vector<double> v0;     // optional initial velocity
  std::vector<double> x0;     // optional start positions (V is then the rest shape)
  double density = 1000.0;
  BoundaryConditionConfig bc;  // used when `fixed` is empty
};

template <int DIM>
struct SimState {
  std::shared_ptr<Mesh> mesh_;
  std::shared_ptr<SimConfig> config_;
  std::vector<std::shared_ptr<MaterialModel>> material_models_;
  std::unique_ptr<Displacement<DIM>> x_;
  std::vector<std::unique_ptr<MixedVariable<DIM>>> mixed_vars_;
  std::vector<std::shared_ptr<Variable<DIM>>> vars_;
  std::unique_ptr<DeviceContext> ctx_;

  int size() const {
    int n = x_->size();
    for (const auto& v : mixed_vars_) n += v->size() + v->size_dual();
    return n;
  }
  // Scene JSON (reference simulation_state.cpp:72-366 key set). dim 3:
  // Medit .mesh objects scaled by 1/max coefficient (SimState<3>); dim 2:
  // OBJ triangle meshes, unreferenced vertices dropped and z truncated
  // (SimState<2>, simulation_state.cpp:46-60).
  bool load(const std::string& json_file, int dim = 3);
  bool load(const Json& args, const std::string& base_dir = ".", int dim = 3);
  // Arrays (per-element materials as in the material_ids constructor,
  // reference tet_mesh.h:11-15); names of variables / solver as in JSON.
  void load_arrays(const SceneArrays& sc, std::shared_ptr<SimConfig> config,
                   const std::vector<std::string>& mixed_variables,
                   const std::vector<std::string>& variables);
  std::vector<int> fixed_override_;

 private:
  void finish_setup();
  void load_params(const Json& args);
};

// Medit .mesh and libigl .dmat readers used by the scene loader.
void read_mesh_file(const std::string& path, std::vector<double>& V, std::vector<int>& T);
std::vector<double> read_dmat(const std::string& path, int& rows, int& cols);
// OBJ as a triangle mesh (igl::read_triangle_mesh: "v" rows, "f" polygons
// fan-triangulated from their first corner; v/vt/vn and negative indices
// accepted). V is nV x 3.
void read_obj_file(const std::string& path, std::vector<double>& V, std::vector<int>& T);
// Drops the vertices no element references and renumbers T, keeping the
// order of the others (igl::remove_unreferenced). `nodes` corners per element.
void remove_unreferenced(std::vector<double>& V, std::vector<int>& T, int nodes);

}  // namespace mfem
