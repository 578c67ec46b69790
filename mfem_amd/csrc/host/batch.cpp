// Batched simulations: see batch.h.
#include "batch.h"

#include <immintrin.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <stdexcept>

namespace mfem {

using namespace mfemhip;

namespace {
// per-simulation result rows in res_: max |dx|, max |ds|, E(0), E(1)
enum { RES_DX = 0, RES_DS = 1, RES_E0 = 2, RES_E1 = 3, RES_ROWS = 4 };
// pinned staging regions of the small uploads
enum { ST_TOL, ST_BUDGET, ST_ON_CG, ST_ALPHA0, ST_ON0, ST_ALPHA1, ST_ON1, ST_ACCEPT, kStageRegions };
constexpr int kTuneMinRows = 20000;
// CG iteration pairs per replayed graph (16 iterations: one host poll's worth)
constexpr int kCgGraphPairs = 8;
}  // namespace

BatchSim::BatchSim(const std::vector<SceneArrays>& sims, std::shared_ptr<SimConfig> config) {
  nsim_ = (int)sims.size();
  if (nsim_ == 0) throw std::runtime_error("batch: no simulations");
  if (!config) throw std::runtime_error("batch: null config");
  SceneArrays all;
  all.density = sims[0].density;
  v0_.assign(1, 0);
  t0_.assign(1, 0);
  std::vector<int> vert_sim, tet_sim;
  for (int b = 0; b < nsim_; ++b) {
    const SceneArrays& sc = sims[b];
    const int nV = (int)(sc.V.size() / 3), nT = (int)(sc.T.size() / 4);
    if (nV <= 0 || nT <= 0) throw std::runtime_error("batch: empty mesh");
    if (sc.density != all.density)
      throw std::runtime_error("batch: all simulations must share one density");
    if ((int)sc.mu.size() != nT || (int)sc.la.size() != nT || (int)sc.mat.size() != nT)
      throw std::runtime_error("batch: per-element arrays must have nT entries");
    for (int v : sc.T)
      if (v < 0 || v >= nV) throw std::runtime_error("batch: tet vertex index out of range");
    std::vector<int> fixed = sc.fixed.empty() ? boundary_fixed(sc.bc, sc.V, 3) : sc.fixed;
    if ((int)fixed.size() != nV) throw std::runtime_error("batch: pin mask must have nV entries");
    const int vo = v0_.back();
    all.V.insert(all.V.end(), sc.V.begin(), sc.V.end());
    for (int v : sc.T) all.T.push_back(v + vo);
    all.mu.insert(all.mu.end(), sc.mu.begin(), sc.mu.end());
    all.la.insert(all.la.end(), sc.la.begin(), sc.la.end());
    all.mat.insert(all.mat.end(), sc.mat.begin(), sc.mat.end());
    all.fixed.insert(all.fixed.end(), fixed.begin(), fixed.end());
    if (sc.v0.empty()) all.v0.insert(all.v0.end(), 3 * (size_t)nV, 0.0);
    else if ((int)sc.v0.size() == 3 * nV) all.v0.insert(all.v0.end(), sc.v0.begin(), sc.v0.end());
    else throw std::runtime_error("batch: initial velocity must have 3 nV entries");
    vert_sim.insert(vert_sim.end(), nV, b);
    tet_sim.insert(tet_sim.end(), nT, b);
    v0_.push_back(vo + nV);
    t0_.push_back(t0_.back() + nT);
  }
  state_ = std::make_unique<SimState<3>>();
  state_->load_arrays(all, std::move(config), {"mixed-stretch"}, {});
  auto& S = *state_;
  mixed_ = dynamic_cast<MixedStretch<3>*>(S.mixed_vars_.at(0).get());
  if (!mixed_) throw std::runtime_error("batch: mixed-stretch variable missing");
  S.x_->reset();
  mixed_->reset();
  // free vertices are numbered in vertex order, so each simulation's block
  // rows are contiguous
  const auto& fm = S.mesh_->free_map_;
  f0_.assign(nsim_ + 1, 0);
  for (int b = 0; b < nsim_; ++b) {
    int c = 0;
    for (int v = v0_[b]; v < v0_[b + 1]; ++v) c += fm[v] >= 0;
    f0_[b + 1] = f0_[b] + c;
  }
  hipStream_t st = S.ctx_->stream;
  std::vector<int> seg_rows(nsim_ + 1);
  for (int b = 0; b <= nsim_; ++b) seg_rows[b] = 3 * f0_[b];
  d_vert_sim_.upload(vert_sim, st);
  d_tet_sim_.upload(tet_sim, st);
  d_seg_rows_.upload(seg_rows, st);
  d_seg_tets_.upload(t0_, st);
  d_row0_.upload(f0_, st);
  ctrls_.resize(nsim_);
  ctrl_host_.resize(nsim_);
  pinned_ctrl_ = std::make_unique<PinnedBuffer>(sizeof(PcgCtrl) * (nsim_ + 1) + 64,
                                                hipHostMallocMapped | hipHostMallocCoherent);
  pinned_res_ = std::make_unique<PinnedBuffer>(sizeof(double) * RES_ROWS * nsim_ + 64);
  staging_ = std::make_unique<PinnedBuffer>(sizeof(double) * kStageRegions * nsim_ + 64);
  d_res_.resize((size_t)RES_ROWS * nsim_);
  d_alpha_.resize(nsim_);
  d_b2_.resize(nsim_);
  seg_part_.resize((size_t)kSegSlices * nsim_);
  d_tol_.resize(nsim_);
  d_budget_.resize(nsim_);
  d_on_.resize(nsim_);
  e_tet_.resize((size_t)S.mesh_->nT());
  scratch_.resize(256);
  res_.assign((size_t)RES_ROWS * nsim_, 0.0);
  stats_.resize(nsim_);
  MFEM_HIP_CHECK(hipStreamSynchronize(st));
}

BatchSim::~BatchSim() { clear_graphs(); }

void BatchSim::clear_graphs() {
  for (auto& g : cg_graphs_)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  cg_graphs_.clear();
  if (update_graph_) (void)hipGraphExecDestroy(update_graph_);
  update_graph_ = nullptr;
  update_sig_.clear();
}

void BatchSim::set_graphs(bool on) {
  if (!on) clear_graphs();
  graphs_ = on;
}

// kCgGraphPairs pairs of CG iterations (even, then odd buffer parity)
// captured once per (right-hand side, solution, shape, precision) and
// replayed; the kernels take their iteration index from it_ctr_ (reset to 0
// before the first replay of a solve), so one graph serves a whole solve.
template <typename T>
hipGraphExec_t BatchSim::cg_pair_graph(const BatchCgBuffers<T>& B) {
  const bool f32 = sizeof(T) == 4;
  for (auto& g : cg_graphs_)
    if (g.b == B.b && g.x == B.x && g.variant == B.variant && g.f32 == f32) return g.exec;
  hipStream_t st = state_->ctx_->stream;
  if (it_ctr_.size() != 2) it_ctr_.resize(2);
  hipGraph_t graph = nullptr;
  MFEM_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  for (int p = 0; p < kCgGraphPairs; ++p) {
    launch_pcg_batch_iteration(B, 0, st, it_ctr_.data(), it_ctr_.data() + 1);
    launch_pcg_batch_iteration(B, 1, st, it_ctr_.data() + 1, it_ctr_.data());
  }
  MFEM_HIP_CHECK(hipStreamEndCapture(st, &graph));
  CgGraph g;
  g.b = B.b;
  g.x = B.x;
  g.variant = B.variant;
  g.f32 = f32;
  MFEM_HIP_CHECK(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
  MFEM_HIP_CHECK(hipGraphDestroy(graph));
  cg_graphs_.push_back(g);
  return g.exec;
}

// update_system (reference newton_optimizer.cpp:92-110) and the assembly of
// SystemMatrixPD::pre_solve (linear_system.h:18-31). From the second step on,
// with graphs enabled, the four launches are one captured graph, re-captured
// when a buffer or the time step changed.
void BatchSim::update_and_assemble(double dt) {
  auto& S = *state_;
  hipStream_t st = S.ctx_->stream;
  const bool f32 = S.ctx_->fp32();
  if (!graphs_ || steps_done_ == 0) {
    mixed_->update(S.x_->value().data(), dt);
    system_.pre_solve(&S);
    subs_update_ += 4;
    return;
  }
  const std::vector<const void*> sig = {
      S.x_->value().data(), S.ctx_->blocks.as<float>(), S.ctx_->blocks.as<double>(),
      S.ctx_->fe.data(), system_.K.as<float>(), system_.K.as<double>(),
      system_.invdiag.as<float>(), system_.invdiag.as<double>(), system_.rhs.data(),
      mixed_->Hloc_.data(), mixed_->dsdF_.as<float>(), mixed_->dsdF_.as<double>()};
  if (!update_graph_ || sig != update_sig_ || dt != update_dt_) {
    if (update_graph_) MFEM_HIP_CHECK(hipGraphExecDestroy(update_graph_));
    update_graph_ = nullptr;
    hipGraph_t graph = nullptr;
    MFEM_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    mixed_->update(S.x_->value().data(), dt);
    system_.pre_solve(&S);
    MFEM_HIP_CHECK(hipStreamEndCapture(st, &graph));
    MFEM_HIP_CHECK(hipGraphInstantiate(&update_graph_, graph, nullptr, nullptr, 0));
    MFEM_HIP_CHECK(hipGraphDestroy(graph));
    update_sig_ = sig;
    update_dt_ = dt;
  }
  (void)f32;
  MFEM_HIP_CHECK(hipGraphLaunch(update_graph_, st));
  subs_update_ += 1;
}

std::vector<double> BatchSim::positions(int sim) const {
  if (sim < 0 || sim >= nsim_) throw std::runtime_error("batch: simulation index out of range");
  auto& S = *state_;
  const auto x = S.x_->value().to_host(S.ctx_->stream);
  return std::vector<double>(x.begin() + 3 * (size_t)v0_[sim], x.begin() + 3 * (size_t)v0_[sim + 1]);
}

template <typename T>
void BatchSim::stage(DeviceArray<T>& dst, const std::vector<T>& src, int region) {
  static_assert(sizeof(T) <= sizeof(double), "staging holds at most 8-byte values");
  if ((int)src.size() != nsim_) throw std::logic_error("batch: staged array has the wrong size");
  T* h = reinterpret_cast<T*>(staging_->as<double>() + (size_t)region * nsim_);
  std::copy(src.begin(), src.end(), h);
  dst.resize(src.size());
  MFEM_HIP_CHECK(hipMemcpyAsync(dst.data(), h, sizeof(T) * src.size(), hipMemcpyHostToDevice,
                                state_->ctx_->stream));
}

void BatchSim::force_cg_shape(int variant, int rows_per_wg) {
  if (variant < -1 || variant >= cg_variant_count())
    throw std::invalid_argument("CG kernel shape out of range");
  if (variant >= 0 && rows_per_wg <= 0) throw std::invalid_argument("rows per workgroup must be > 0");
  variant_ = variant;
  rows_per_wg_ = variant < 0 ? 0 : rows_per_wg;
  if (variant >= 0) layout_workgroups(rows_per_wg_);
}

// Workgroups per simulation: ceil(rows / rows_per_wg), at least one.
void BatchSim::layout_workgroups(int rows_per_wg) {
  hipStream_t st = state_->ctx_->stream;
  wg0_.assign(nsim_ + 1, 0);
  std::vector<int> wg_sys;
  for (int b = 0; b < nsim_; ++b) {
    const int rows = f0_[b + 1] - f0_[b];
    const int nw = std::max(1, (rows + rows_per_wg - 1) / rows_per_wg);
    wg0_[b + 1] = wg0_[b] + nw;
    wg_sys.insert(wg_sys.end(), nw, b);
  }
  nwg_ = wg0_.back();
  for (auto& g : cg_graphs_)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  cg_graphs_.clear();
  d_wg_sys_.upload(wg_sys, st);
  d_wg0_.upload(wg0_, st);
  parts_.resize(2 * (size_t)kCgParts * nwg_);
  MFEM_HIP_CHECK(hipStreamSynchronize(st));  // wg_sys is a local
}

template <typename T>
BatchCgBuffers<T> BatchSim::buffers(const double* b, double* x) {
  auto& m = *state_->mesh_;
  BatchCgBuffers<T> B;
  B.n = 3 * m.nfree();
  B.rowptr = m.d_rowptr.data();
  B.col = m.d_col.data();
  B.K = system_.K.template as<T>();
  B.invdiag = system_.invdiag.template as<T>();
  B.b = b;
  B.x = x;
  B.state[0] = state0_.template as<T>();
  B.state[1] = state1_.template as<T>();
  B.ctrls = ctrls_.data();
  B.parts[0] = parts_.data();
  B.parts[1] = parts_.data() + (size_t)kCgParts * nwg_;
  B.nwg = nwg_;
  B.wg_sys = d_wg_sys_.data();
  B.wg0 = d_wg0_.data();
  B.row0 = d_row0_.data();
  B.variant = std::max(variant_, 0);
  B.mirror = pinned_ctrl_->device<PcgCtrl>();
  B.progress = B.mirror + nsim_;
  return B;
}

// One batched CG solve (Eigen semantics per system, reference
// eigen_iterative_solver.h:20-29): x = 0, per-system tolerance and
// iteration budget; systems with on[b] = 0 are skipped (x stays 0).
template <typename T>
void BatchSim::run_cg(const double* b, double* x, const std::vector<double>& tol,
                      const std::vector<int>& budget, const std::vector<int>& on,
                      std::vector<int>& iters) {
  auto& S = *state_;
  hipStream_t st = S.ctx_->stream;
  const BatchCgBuffers<T> B = buffers<T>(b, x);
  // Every system's control block is mirrored in mapped host memory (the
  // setup writes each one, the kernel that stops a system writes it again),
  // and the first workgroup of every CG kernel writes the kernel's index
  // there. The host keeps a few kernels queued beyond the last one that
  // began and puts no event or copy into the stream.
  volatile PcgCtrl* mirror = pinned_ctrl_->as<PcgCtrl>();
  volatile PcgCtrl* progress = mirror + nsim_;
  for (int k = 0; k < nsim_; ++k) mirror[k].done = 0;
  progress->started = -1;
  launch_pcg_batch_init(B, scratch_.data(), st);
  launch_seg_reduce(nsim_, b, 1, 0, d_seg_rows_.data(), 1, d_b2_.data(), seg_part_.data(),
                    st);
  stage(d_tol_, tol, ST_TOL);
  stage(d_budget_, budget, ST_BUDGET);
  stage(d_on_, on, ST_ON_CG);
  launch_pcg_batch_setup(nsim_, ctrls_.data(), d_b2_.data(), d_tol_.data(), d_budget_.data(),
                         d_on_.data(), st, B.mirror);
  int max_budget = 0;
  for (int k = 0; k < nsim_; ++k)
    if (on[k]) max_budget = std::max(max_budget, budget[k]);
  const long limit = (long)max_budget + 2;
  const int poll = std::max(1, S.config_->pcg_poll);
  long launched = 0;
  // Eager launches run only the workgroups of systems that had stopped at
  // no earlier look at the mirror: a launch late in a solve, when a few slow
  // systems are left, dispatches their workgroups instead of the whole
  // batch's. Every workgroup of a listed system runs, so each system's
  // iteration is the same kernel work as with the full grid.
  BatchCgBuffers<T> Bc = B;
  if (!graphs_) {
    for (auto& d : d_active_) d.resize(nwg_);
    if (pinned_active_n_ != nwg_) {
      pinned_active_ = std::make_unique<PinnedBuffer>(3 * sizeof(int) * (size_t)nwg_);
      pinned_active_n_ = nwg_;
    }
  }
  int list_slot = -1;
  // a list slot is rewritten only once the kernel with this index has begun:
  // every kernel that read the slot's previous list has then completed, and
  // so has the copy of that list from its pinned source
  long free_from[3] = {0, 0, 0};
  auto compact = [&](long started) {
    if (graphs_) return;
    const int cur = Bc.active ? Bc.nactive : nwg_;
    int count = 0;
    for (int k = 0; k < nsim_; ++k)
      if (!mirror[k].done) count += wg0_[k + 1] - wg0_[k];
    if (count == 0 || count > cur * 3 / 4) return;
    const int slot = (list_slot + 1) % 3;
    if (started < free_from[slot]) return;
    int* h = pinned_active_->as<int>() + (size_t)slot * nwg_;
    int n = 0;
    for (int k = 0; k < nsim_; ++k)
      if (!mirror[k].done)
        for (int w = wg0_[k]; w < wg0_[k + 1]; ++w) h[n++] = w;
    MFEM_HIP_CHECK(hipMemcpyAsync(d_active_[slot].data(), h, sizeof(int) * (size_t)n,
                                  hipMemcpyHostToDevice, st));
    if (list_slot >= 0) free_from[list_slot] = launched;
    list_slot = slot;
    Bc.active = d_active_[slot].data();
    Bc.nactive = n;
  };
  // graph replay: groups of kCgGraphPairs pairs of iterations; a group past
  // the stop exits at once, like any launch queued after it
  hipGraphExec_t pair = nullptr;
  if (graphs_) {
    pair = cg_pair_graph(B);
    launch_set_int(it_ctr_.data(), 0, st);
  }
  const long per = pair ? 2L * kCgGraphPairs : 1;
  const long ahead = std::max<long>(std::max(2, poll / 4), per);
  auto launch_to = [&](long n) {
    n = std::min(n, limit);
    if (launched >= n) return;
    for (; launched < n; launched += per) {
      if (pair) MFEM_HIP_CHECK(hipGraphLaunch(pair, st));
      else launch_pcg_batch_iteration(Bc, (int)launched, st);
      ++subs_cg_;
    }
    MFEM_HIP_CHECK(hipGetLastError());
  };
  auto all_done = [&] {
    for (int k = 0; k < nsim_; ++k)
      if (!mirror[k].done) return false;
    return true;
  };
  launch_to(ahead);
  long seen = -2;
  while (true) {
    const long started = progress->started;
    if (started != seen) {  // a kernel began: look at the systems' flags
      seen = started;
      if (all_done()) break;
      compact(started);
    }
    if (launched < limit) {
      launch_to(started + 1 + ahead);
      _mm_pause();
      continue;
    }
    // all kernels queued: the last one stops every system at the latest
    MFEM_HIP_CHECK(hipStreamSynchronize(st));
    if (!all_done()) throw std::runtime_error("batched CG control blocks did not terminate");
    break;
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  for (int k = 0; k < nsim_; ++k) {
    PcgCtrl& c = ctrl_host_[k];
    c.rr = mirror[k].rr;
    c.b2 = mirror[k].b2;
    c.threshold = mirror[k].threshold;
    c.kern_ticks = mirror[k].kern_ticks;
    c.iter = mirror[k].iter;
    c.done = 1;
    c.max_iters = mirror[k].max_iters;
    c.stop_it = mirror[k].stop_it;
    c.kern_launches = mirror[k].kern_launches;
  }
  // the kernels queued past the last stop exit at once; they finish before
  // the mirror is reset for the next solve
  MFEM_HIP_CHECK(hipStreamSynchronize(st));
  iters.assign(nsim_, 0);
  for (int k = 0; k < nsim_; ++k) iters[k] = on[k] ? ctrl_host_[k].iter : 0;
}

// Batch autotuner: every CG shape with 1, 2, 4 and 8 passes of rows per
// workgroup, timed over 10 iterations of a real solve at tolerance 0.
template <typename T>
void BatchSim::tune() {
  auto& S = *state_;
  hipStream_t st = S.ctx_->stream;
  if (S.mesh_->nfree() < kTuneMinRows) {
    variant_ = 0;
    rows_per_wg_ = 256;
    layout_workgroups(rows_per_wg_);
    return;
  }
  hipEvent_t a, b;
  MFEM_HIP_CHECK(hipEventCreate(&a));
  MFEM_HIP_CHECK(hipEventCreate(&b));
  const std::vector<double> tol(nsim_, 0.0);
  const std::vector<int> budget(nsim_, 1 << 30), on(nsim_, 1);
  double best = 1e300;
  int best_v = 0, best_r = 256;
  for (int v = 0; v < cg_variant_count(); ++v) {
    int G, U, BS;
    cg_variant_shape(v, &G, &U, &BS);
    for (int passes : {1, 2, 4, 8}) {
      const int rpw = passes * BS / G;
      variant_ = v;
      layout_workgroups(rpw);
      BatchCgBuffers<T> B = buffers<T>(system_.rhs.data(), x_.data());
      launch_pcg_batch_init(B, scratch_.data(), st);
      launch_seg_reduce(nsim_, B.b, 1, 0, d_seg_rows_.data(), 1, d_b2_.data(),
                        seg_part_.data(), st);
      stage(d_tol_, tol, ST_TOL);
      stage(d_budget_, budget, ST_BUDGET);
      stage(d_on_, on, ST_ON_CG);
      launch_pcg_batch_setup(nsim_, ctrls_.data(), d_b2_.data(), d_tol_.data(), d_budget_.data(),
                             d_on_.data(), st);
      launch_pcg_batch_iteration(B, 0, st);
      launch_pcg_batch_iteration(B, 1, st);
      MFEM_HIP_CHECK(hipEventRecord(a, st));
      for (int it = 2; it < 12; ++it) launch_pcg_batch_iteration(B, it, st);
      MFEM_HIP_CHECK(hipEventRecord(b, st));
      MFEM_HIP_CHECK(hipEventSynchronize(b));
      MFEM_HIP_CHECK(hipStreamSynchronize(st));
      float ms = 0;
      MFEM_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
      if (ms < best) {
        best = ms;
        best_v = v;
        best_r = rpw;
      }
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  variant_ = best_v;
  rows_per_wg_ = best_r;
  layout_workgroups(rows_per_wg_);
}

// Linear solve of every simulation that is on (PcgSolver::solve per system:
// FP64 CG, or FP32 CG sweeps in FP64 defect correction); dx in x_->delta()
// and the local solve for (lambda, ds) afterwards.
void BatchSim::solve(const std::vector<int>& on) {
  auto& S = *state_;
  auto& m = *S.mesh_;
  auto& cfg = *S.config_;
  hipStream_t st = S.ctx_->stream;
  const int n = 3 * m.nfree();
  const bool f32 = S.ctx_->fp32();
  x_.resize(n);
  state0_.resize(4 * (size_t)n, f32);
  state1_.resize(4 * (size_t)n, f32);
  if (variant_ < 0) {
    if (f32) tune<float>();
    else tune<double>();
  } else if (nwg_ == 0) {
    layout_workgroups(rows_per_wg_);
  }
  solve_iters_.assign(nsim_, 0);
  solve_sweeps_.assign(nsim_, 0);
  std::vector<int> it;
  std::vector<int> budget(nsim_, cfg.max_iterative_solver_iters);
  if (!f32) {
    run_cg<double>(system_.rhs.data(), x_.data(), std::vector<double>(nsim_, cfg.itr_tol), budget,
                   on, it);
    for (int k = 0; k < nsim_; ++k) {
      solve_iters_[k] = it[k];
      solve_sweeps_[k] = on[k];
    }
  } else {
    r64_.resize(n);
    dx64_.resize(n);
    const double inner0 = std::min(std::max(cfg.inner_tol, 0.5 * cfg.itr_tol), 0.5);
    std::vector<double> b2(nsim_, 0.0), thr(nsim_, 0.0), r2(nsim_, 0.0);
    if (cfg.max_refinements > 0) {
      run_cg<float>(system_.rhs.data(), x_.data(), std::vector<double>(nsim_, inner0), budget, on,
                    it);
      for (int k = 0; k < nsim_; ++k) {
        if (!on[k]) continue;
        solve_iters_[k] += it[k];
        budget[k] -= std::max(it[k], 1);
        solve_sweeps_[k] = 1;
        b2[k] = ctrl_host_[k].b2;
        thr[k] = std::max(cfg.itr_tol * cfg.itr_tol * b2[k], DBL_MIN);
      }
    } else {
      x_.zero(st);
    }
    auto& R = S.ctx_->reducer;
    while (true) {
      launch_residual<float>(m.nfree(), m.d_rowptr.data(), m.d_col.data(), system_.K.f32(),
                             x_.data(), system_.rhs.data(), r64_.data(), R.slot(8), st);
      launch_seg_reduce(nsim_, r64_.data(), 1, 0, d_seg_rows_.data(), 1, d_res_.data(),
                        seg_part_.data(), st);
      MFEM_HIP_CHECK(hipMemcpyAsync(pinned_res_->as<double>(), d_res_.data(),
                                    sizeof(double) * nsim_, hipMemcpyDeviceToHost, st));
      MFEM_HIP_CHECK(hipStreamSynchronize(st));
      std::vector<int> need(nsim_, 0);
      std::vector<double> inner(nsim_, 0.5);
      bool any = false;
      for (int k = 0; k < nsim_; ++k) {
        r2[k] = pinned_res_->as<double>()[k];
        need[k] = on[k] && b2[k] > 0 && r2[k] >= thr[k] && budget[k] > 0 &&
                  solve_sweeps_[k] < cfg.max_refinements;
        if (need[k]) inner[k] = std::min(std::max(cfg.inner_tol, std::sqrt(thr[k] / r2[k]) * 0.5), 0.5);
        any = any || need[k];
      }
      if (!any) break;
      run_cg<float>(r64_.data(), dx64_.data(), inner, budget, need, it);
      for (int k = 0; k < nsim_; ++k)
        if (need[k]) {
          solve_iters_[k] += it[k];
          budget[k] -= std::max(it[k], 1);
          ++solve_sweeps_[k];
        }
      launch_axpy(n, 1.0, dx64_.data(), x_.data(), st);
    }
  }
  system_.post_solve(&S, x_.data());
}

void BatchSim::read_results(int n) {
  hipStream_t st = state_->ctx_->stream;
  MFEM_HIP_CHECK(hipMemcpyAsync(pinned_res_->as<double>(), d_res_.data(), sizeof(double) * n,
                                hipMemcpyDeviceToHost, st));
  MFEM_HIP_CHECK(hipStreamSynchronize(st));
  std::copy(pinned_res_->as<double>(), pinned_res_->as<double>() + n, res_.begin());
}

// Every simulation follows NewtonOptimizer::step (reference
// newton_optimizer.cpp:11-89, line search linesearch.h:78-131 with c = 0 and
// p = 0.5); a simulation whose Newton loop has ended is carried along with
// step length 0 until all have ended.
void BatchSim::step() {
  auto& S = *state_;
  auto* c = S.ctx_.get();
  auto& cfg = *S.config_;
  hipStream_t st = c->stream;
  const auto t_start = std::chrono::steady_clock::now();
  const double dt = S.x_->integrator()->dt();
  const double h2 = dt * dt;
  auto m = S.mesh_->dev();
  const int nT = S.mesh_->nT();
  for (auto& s : stats_) s = SimStats();
  S.x_->pre_solve();
  mixed_->pre_solve();
  std::vector<int> on(nsim_, 1);
  std::vector<double> E_prev(nsim_, 0.0), alpha(nsim_, 1.0);
  // E(alpha) per simulation into result row `row`; E(0) and E(1) are staged
  // through different regions because they are issued back to back
  auto energies = [&](const std::vector<double>& a, const std::vector<int>& which, int row,
                      bool second) {
    DeviceArray<double>& da = second ? d_alpha1_ : d_alpha_;
    DeviceArray<int>& dn = second ? d_on1_ : d_on_;
    stage(da, a, second ? ST_ALPHA1 : ST_ALPHA0);
    stage(dn, which, second ? ST_ON1 : ST_ON0);
    launch_mixed_energy_batch(m, S.x_->value().data(), S.x_->delta().data(), d_tet_sim_.data(),
                              da.data(), dn.data(), c->inertia(), mixed_->R_.data(),
                              mixed_->s_.data(), mixed_->ds_.data(), mixed_->la_.data(), h2,
                              e_tet_.data(), st);
    launch_seg_reduce(nsim_, e_tet_.data(), 1, 0, d_seg_tets_.data(), 0,
                      d_res_.data() + (size_t)row * nsim_, seg_part_.data(), st);
  };
  const int ls_iters = cfg.ls_iters;
  for (int i = 0; i < cfg.outer_steps; ++i) {
    if (std::find(on.begin(), on.end(), 1) == on.end()) break;
    // update_system (reference newton_optimizer.cpp:92-110) and the assembly
    update_and_assemble(dt);
    ++newton_total_;
    // substep (:113-124) for the simulations that are on
    solve(on);
    // decrement, E(0) and E(1), one read
    launch_seg_reduce(nsim_, S.x_->delta().data(), 1, 0, d_seg_rows_.data(), 2,
                      d_res_.data() + (size_t)RES_DX * nsim_, seg_part_.data(), st);
    launch_seg_reduce(nsim_, mixed_->ds_.data(), 6, nT, d_seg_tets_.data(), 2,
                      d_res_.data() + (size_t)RES_DS * nsim_, seg_part_.data(), st);
    energies(std::vector<double>(nsim_, 0.0), on, RES_E0, false);
    if (ls_iters > 0) energies(std::vector<double>(nsim_, 1.0), on, RES_E1, true);
    read_results(RES_ROWS * nsim_);
    std::vector<double> E0(nsim_), dec(nsim_), res(nsim_), fx(nsim_);
    std::vector<int> lsit(nsim_, 0), pending(nsim_, 0);
    for (int k = 0; k < nsim_; ++k) {
      if (!on[k]) continue;
      dec[k] = std::max(res_[RES_DX * nsim_ + k], res_[RES_DS * nsim_ + k]);
      E0[k] = res_[RES_E0 * nsim_ + k];
      res[k] = std::fabs((E0[k] - E_prev[k]) / (E0[k] + 1e-6));
      E_prev[k] = E0[k];
      alpha[k] = 1.0;
      if (ls_iters > 0) {
        fx[k] = res_[RES_E1 * nsim_ + k];
        pending[k] = !(fx[k] < E0[k]);
      }
    }
    // backtracking for the simulations whose full step did not decrease E
    while (true) {
      bool any = false;
      for (int k = 0; k < nsim_; ++k)
        if (pending[k]) {
          alpha[k] *= 0.5;
          if (++lsit[k] >= ls_iters) pending[k] = 0;  // exhausted
          any = any || pending[k];
        }
      if (!any) break;
      energies(alpha, pending, RES_E1, false);
      read_results(RES_ROWS * nsim_);
      for (int k = 0; k < nsim_; ++k)
        if (pending[k] && res_[RES_E1 * nsim_ + k] < E0[k]) pending[k] = 0;
    }
    std::vector<double> accept(nsim_, 0.0);
    for (int k = 0; k < nsim_; ++k) {
      if (!on[k]) continue;
      auto& s = stats_[k];
      const bool ok = ls_iters > 0 && lsit[k] < ls_iters;
      if (ok) accept[k] = alpha[k];
      else s.ls_fail++;
      s.ls_iters += lsit[k];
      s.pcg_iters += solve_iters_[k];
      s.refinements += solve_sweeps_[k];
      s.newton_iters++;
      s.decrement = dec[k];
      s.energy = E0[k];
      on[k] = s.newton_iters < cfg.outer_steps && dec[k] > cfg.newton_tol && res[k] > 1e-12;
    }
    stage(d_alpha_, accept, ST_ACCEPT);
    launch_accept_batch(S.mesh_->nV(), S.mesh_->d_free_map.data(), d_vert_sim_.data(), nT,
                        d_tet_sim_.data(), d_alpha_.data(), S.x_->value().data(),
                        S.x_->delta().data(), mixed_->s_.data(), mixed_->ds_.data(), st);
  }
  S.x_->post_solve();
  mixed_->post_solve();
  MFEM_HIP_CHECK(hipStreamSynchronize(st));
  MFEM_HIP_CHECK(hipGetLastError());
  ++steps_done_;
  step_ms_ = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start)
                 .count();
}

}  // namespace mfem
