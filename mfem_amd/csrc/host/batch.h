// Many independent mixed-stretch simulations advanced together on one GPU.
//
// The reference runs each simulation as its own SimState with its own
// NewtonOptimizer (newton_optimizer.cpp:11-89); a multi-object scene would
// instead share one Newton loop and line search (simulation_state.cpp:260),
// which is not what independent simulations need. Here the meshes are
// concatenated into one block-diagonal system, so every Newton phase is one
// launch for all simulations (element update, assembly, local solve, line-
// search energies), while each simulation keeps its own CG scalars and
// stopping rule, its own line search, acceptance and Newton termination:
// every simulation follows the single-simulation algorithm exactly.
#pragma once

#include <memory>
#include <vector>

#include "sim.h"

namespace mfem {

class BatchSim {
 public:
  // One entry per simulation; time step, gravity, integrator, solver
  // settings and precision come from `config` and are shared.
  BatchSim(const std::vector<SceneArrays>& sims, std::shared_ptr<SimConfig> config);
  ~BatchSim();

  // One time step of every simulation (NewtonOptimizer::step semantics each).
  void step();

  int count() const { return nsim_; }
  int vertices(int sim) const { return v0_[sim + 1] - v0_[sim]; }
  int tets(int sim) const { return t0_[sim + 1] - t0_[sim]; }
  // nV x 3 positions of simulation `sim`
  std::vector<double> positions(int sim) const;

  struct SimStats {
    int newton_iters = 0, pcg_iters = 0, refinements = 0, ls_iters = 0, ls_fail = 0;
    double decrement = 0, energy = 0;
  };
  const std::vector<SimStats>& stats() const { return stats_; }
  double last_step_ms() const { return step_ms_; }
  // CG kernel shape / rows per workgroup chosen by the batch autotuner
  int cg_variant() const { return variant_; }
  int rows_per_workgroup() const { return rows_per_wg_; }
  // Force the CG shape and rows per workgroup (variant -1: autotune again).
  void force_cg_shape(int variant, int rows_per_wg);
  // Replay the element-update + assembly phase and groups of 16 CG
  // iterations as captured hipGraphs; results are identical either way.
  void set_graphs(bool on);
  bool graphs() const { return graphs_; }
  // Host submissions since creation: of the update + assembly phase (kernel
  // launches, or one graph launch), of the CG iterations (one launch per
  // iteration, or one graph launch per 16), and Newton iterations run.
  long update_submissions() const { return subs_update_; }
  long cg_submissions() const { return subs_cg_; }
  long newton_iterations() const { return newton_total_; }

 private:
  struct CgGraph {
    const void* b = nullptr;
    const void* x = nullptr;
    int variant = -1;
    bool f32 = false;
    hipGraphExec_t exec = nullptr;
  };
  void clear_graphs();
  template <typename T>
  hipGraphExec_t cg_pair_graph(const mfemhip::BatchCgBuffers<T>& B);
  // update_system + SystemMatrixPD::pre_solve, eager or replayed
  void update_and_assemble(double dt);
  template <typename T>
  mfemhip::BatchCgBuffers<T> buffers(const double* b, double* x);
  template <typename T>
  void run_cg(const double* b, double* x, const std::vector<double>& tol,
              const std::vector<int>& budget, const std::vector<int>& on,
              std::vector<int>& iters);
  template <typename T>
  void tune();
  void layout_workgroups(int rows_per_wg);
  void solve(const std::vector<int>& on);
  // per-simulation results into res_ (host), one synchronisation
  void read_results(int n);
  // Copy a small per-simulation host array to the device through pinned
  // staging region `region`; a region is rewritten only after the stream has
  // synchronised since its previous use.
  template <typename T>
  void stage(DeviceArray<T>& dst, const std::vector<T>& src, int region);

  std::unique_ptr<SimState<3>> state_;
  MixedStretch<3>* mixed_ = nullptr;
  SystemMatrixPD<3> system_;
  int nsim_ = 0;
  std::vector<int> v0_, t0_, f0_;  // prefix offsets: vertices, tets, free vertices
  DeviceArray<int> d_vert_sim_, d_tet_sim_, d_seg_rows_, d_seg_tets_, d_row0_;
  // CG workgroup layout
  int variant_ = -1, rows_per_wg_ = 0, nwg_ = 0;
  std::vector<int> wg0_;
  DeviceArray<int> d_wg_sys_, d_wg0_;
  // CG buffers
  DeviceArray<double> x_, r64_, dx64_, parts_, scratch_, d_b2_, d_tol_;
  RealArray state0_, state1_;
  DeviceArray<mfemhip::PcgCtrl> ctrls_;
  DeviceArray<int> d_budget_, d_on_;
  std::vector<mfemhip::PcgCtrl> ctrl_host_;
  // mapped copy of every system's control block, then the progress word
  std::unique_ptr<PinnedBuffer> pinned_ctrl_;
  // Compacted lists of the workgroups of the systems still iterating (eager
  // CG only), a ring of three (device lists and their pinned sources) so
  // that a list is rewritten only after the launches that read it have run.
  DeviceArray<int> d_active_[3];
  std::unique_ptr<PinnedBuffer> pinned_active_;
  int pinned_active_n_ = 0;
  // line search / decrement
  DeviceArray<double> e_tet_, d_alpha_, d_alpha1_, d_res_;
  DeviceArray<double> seg_part_;  // slice partials of the per-simulation reductions
  DeviceArray<int> d_on1_;
  std::unique_ptr<PinnedBuffer> pinned_res_;
  std::unique_ptr<PinnedBuffer> staging_;  // kStageRegions x nsim doubles
  std::vector<double> res_;
  std::vector<SimStats> stats_;
  std::vector<int> solve_iters_, solve_sweeps_;
  double step_ms_ = 0;
  // graph replay (off by default: it gives the same results and cuts host
  // submissions ~16x, but the device is the bottleneck and it measured ~1%
  // slower on the 256-beam batch)
  bool graphs_ = false;
  std::vector<CgGraph> cg_graphs_;
  hipGraphExec_t update_graph_ = nullptr;
  std::vector<const void*> update_sig_;
  double update_dt_ = 0;
  DeviceArray<int> it_ctr_;  // the CG iteration index of graph replays (two words)
  long subs_update_ = 0, subs_cg_ = 0, newton_total_ = 0;
  int steps_done_ = 0;
};

}  // namespace mfem
