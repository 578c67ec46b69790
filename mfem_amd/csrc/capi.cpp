// extern "C" boundary of the device Newton step (declared in
// include/mfem_hip.h). Exceptions never cross it: every entry point returns 0
// or a nonzero code and leaves a message for mfem_last_error().
#include "../../include/mfem_hip.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "host/abi_error.h"
#include "host/batch.h"
#include "host/output.h"
#include "host/sim.h"
#include "kernels.h"

using namespace mfem;

struct mfem_sim {
  std::unique_ptr<NewtonOptimizer<3>> opt;
  SimState<3>& state() { return opt->state(); }
  std::vector<int> surface;  // output triangles (boundary faces in 3D), built on first use
};

struct mfem_batch {
  std::unique_ptr<BatchSim> sim;
};

namespace {
thread_local std::string g_last_error;
}  // namespace

void mfem::set_last_error(const std::string& msg) { g_last_error = msg; }

namespace {

template <typename F>
int guarded(F&& f) {
  return mfem::abi_guarded(std::forward<F>(f));
}

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// structure-of-arrays (stride nT) -> element-major
std::vector<double> soa_to_elem(const std::vector<double>& a, int nT, int k) {
  std::vector<double> o((size_t)nT * k);
  for (int e = 0; e < nT; ++e)
    for (int i = 0; i < k; ++i) o[(size_t)e * k + i] = a[(size_t)i * nT + e];
  return o;
}

int sym6(int i, int j) {
  if (i > j) std::swap(i, j);
  return i * 6 - (i * (i - 1)) / 2 + (j - i);
}

// Inverses of packed symmetric 6x6 matrices (21 values, stride nT) by
// Gauss-Jordan elimination with partial pivoting; the symmetric part of each
// inverse, same layout.
std::vector<double> packed_sym_inverse6(const std::vector<double>& P, int nT) {
  std::vector<double> out(P.size());
  for (int e = 0; e < nT; ++e) {
    double A[6][12];
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) {
        A[i][j] = P[(size_t)sym6(i, j) * nT + e];
        A[i][6 + j] = i == j ? 1.0 : 0.0;
      }
    for (int c = 0; c < 6; ++c) {
      int p = c;
      for (int r = c + 1; r < 6; ++r)
        if (std::fabs(A[r][c]) > std::fabs(A[p][c])) p = r;
      if (p != c)
        for (int j = 0; j < 12; ++j) std::swap(A[c][j], A[p][j]);
      const double d = 1.0 / A[c][c];
      for (int j = 0; j < 12; ++j) A[c][j] *= d;
      for (int r = 0; r < 6; ++r)
        if (r != c && A[r][c] != 0.0) {
          const double f = A[r][c];
          for (int j = 0; j < 12; ++j) A[r][j] -= f * A[c][j];
        }
    }
    for (int i = 0; i < 6; ++i)
      for (int j = i; j < 6; ++j)
        out[(size_t)sym6(i, j) * nT + e] = 0.5 * (A[i][6 + j] + A[j][6 + i]);
  }
  return out;
}

std::vector<double> fetch(mfem_sim* sim, const std::string& k) {
  auto& S = sim->state();
  auto& m = *S.mesh_;
  hipStream_t st = S.ctx_->stream;
  const int nT = m.nT();
  MFEM_HIP_CHECK(hipStreamSynchronize(st));
  MixedStretch<3>* ms = S.mixed_vars_.empty() ? nullptr
                                               : dynamic_cast<MixedStretch<3>*>(S.mixed_vars_[0].get());
  auto need_ms = [&]() {
    if (!ms) throw std::runtime_error("'" + k + "' needs a mixed-stretch variable");
    return ms;
  };
  // element widths: 3D s is a 6-vector, 2D a 3-vector (leading rows)
  const bool two = m.dim() == 2;
  const int ws = two ? 3 : 6;
  if (k == "R") return soa_to_elem(need_ms()->R_.to_host(st), nT, two ? 4 : 9);
  if (k == "S") return soa_to_elem(need_ms()->S_.to_host(st), nT, ws);
  if (k == "s") return soa_to_elem(need_ms()->s_.to_host(st), nT, ws);
  if (k == "lambda") return soa_to_elem(need_ms()->la_.to_host(st), nT, ws);
  if (k == "ds") return soa_to_elem(need_ms()->ds_.to_host(st), nT, ws);
  if (k == "g") return soa_to_elem(need_ms()->g_.to_host(st), nT, ws);
  if (k == "gl") return soa_to_elem(need_ms()->gl_.to_host(st), nT, ws);
  if (k == "dsdF") return soa_to_elem(need_ms()->dsdF_.to_host_double(st), nT, two ? 12 : 54);
  // H and Hinv: the FP64 matrices when the variable keeps them (the local
  // solve's copies in FP32 storage mode), else the stored ones
  auto hess_host = [&](bool inv) {
    auto* ms = need_ms();
    const auto& d = inv ? ms->Hlinv_ : ms->Hloc_;
    if (d.size()) return d.to_host(st);
    // FP32 storage on tets keeps no inverse: that of the FP64 Hessian
    if (inv && ms->Hloc_.size()) return packed_sym_inverse6(ms->Hloc_.to_host(st), nT);
    return inv ? ms->Hinv_.to_host_double(st) : ms->Hpsi_.to_host_double(st);
  };
  if (two && (k == "H" || k == "Hinv")) {
    auto P = soa_to_elem(hess_host(k == "Hinv"), nT, 6);
    std::vector<double> o((size_t)nT * 9);
    const double syminv[3] = {1, 1, 0.5};
    for (int e = 0; e < nT; ++e)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          const int a = std::min(i, j), b = std::max(i, j);
          double v = P[(size_t)e * 6 + (a * 3 - (a * (a - 1)) / 2 + (b - a))];
          if (k == "H") v = syminv[i] * v * syminv[j] / m.vols_[e];
          o[(size_t)e * 9 + i + 3 * j] = v;
        }
    return o;
  }
  if (k == "H" || k == "Hinv") {
    auto P = soa_to_elem(hess_host(k == "Hinv"), nT, 21);
    std::vector<double> o((size_t)nT * 36);
    const double syminv[6] = {1, 1, 1, 0.5, 0.5, 0.5};
    for (int e = 0; e < nT; ++e)
      for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
          double v = P[(size_t)e * 21 + sym6(i, j)];
          if (k == "H") v = syminv[i] * v * syminv[j] / m.vols_[e];
          o[(size_t)e * 36 + i + 6 * j] = v;
        }
    return o;
  }
  if (k == "blocks") return S.ctx_->blocks.to_host_double(st);
  if (k == "fe") return S.ctx_->fe.to_host(st);
  if (k == "x") return S.x_->value().to_host(st);
  if (k == "x_tilde") return S.ctx_->x_tilde.to_host(st);
  if (k == "dx") return S.x_->delta().to_host(st);
  auto* pcg = dynamic_cast<PcgSolver<3>*>(sim->opt->solver());
  auto* ic = dynamic_cast<IcPcgSolver<3>*>(sim->opt->solver());
  if (k == "ic_shift") {
    if (!ic) throw std::runtime_error("the linear solver is not \"eigen-pcg-IC\"");
    return {ic->shift()};
  }
  if (k == "rhs" || k == "lhs") {
    SystemMatrixPD<3>* sys = pcg ? &pcg->system() : ic ? &ic->system() : nullptr;
    if (auto* d = dynamic_cast<DirectSolver<3>*>(sim->opt->solver())) sys = &d->system();
    if (!sys) throw std::runtime_error("no assembled system");
    if (k == "rhs") return sys->rhs.to_host(st);
    auto Kb = sys->K.to_host_double(st);
    std::vector<double> o;
    o.reserve(Kb.size());
    for (int i = 0; i < m.nfree(); ++i)
      for (int r = 0; r < 3; ++r)
        for (int b = m.rowptr_[i]; b < m.rowptr_[i + 1]; ++b)
          for (int c = 0; c < 3; ++c) o.push_back(Kb[(size_t)9 * b + 3 * r + c]);
    return o;
  }
  if (k == "mass") return m.d_mass.to_host(st);
  if (k == "cg_stamps") {
    if (!pcg) throw std::runtime_error("no CG solver");
    auto t = pcg->cg_stamps(st);
    std::vector<double> o(t.size());
    const double us_per_tick = 1e3 / pcg->wall_clock_khz();
    const unsigned long long t0 = t.empty() ? 0 : *std::min_element(t.begin(), t.end());
    for (size_t i = 0; i < t.size(); ++i) o[i] = (double)(t[i] - t0) * us_per_tick;
    return o;
  }
  throw std::runtime_error("unknown array '" + k + "'");
}

std::vector<int> fetch_int(mfem_sim* sim, const std::string& k) {
  auto& m = *sim->state().mesh_;
  hipStream_t st = sim->state().ctx_->stream;
  if (k == "free_map") return m.free_map_;
  if (k == "fixed") return m.is_fixed_;
  if (k == "bsr_rowptr") return m.rowptr_;
  if (k == "bsr_col") return m.col_;
  if (k == "slot_ptr") return m.d_slot_ptr.to_host(st);
  if (k == "slots") return m.d_slots.to_host(st);
  if (k == "pcg_iters") return {sim->opt->solver()->stats().iterations};
  if (k == "ic_color" || k == "ic_colors" || k == "ic_attempts") {
    auto* ic = dynamic_cast<IcPcgSolver<3>*>(sim->opt->solver());
    if (!ic) throw std::runtime_error("the linear solver is not \"eigen-pcg-IC\"");
    if (k == "ic_color") return ic->vertex_colors();
    return {k == "ic_colors" ? ic->colors() : ic->attempts()};
  }
  if (k == "rowptr" || k == "colind") {
    std::vector<int> rp(1, 0), ci;
    for (int i = 0; i < m.nfree(); ++i)
      for (int r = 0; r < 3; ++r) {
        for (int b = m.rowptr_[i]; b < m.rowptr_[i + 1]; ++b)
          for (int c = 0; c < 3; ++c) ci.push_back(3 * m.col_[b] + c);
        rp.push_back((int)ci.size());
      }
    return k == "rowptr" ? rp : ci;
  }
  throw std::runtime_error("unknown array '" + k + "'");
}

void finish_create(std::unique_ptr<SimState<3>> state, mfem_sim** out) {
  OptimizerFactory<3> of;
  if (state->config_->optimizer != OPTIMIZER_NEWTON)
    throw std::runtime_error("only the \"newton\" optimizer is implemented");
  auto base = of.create(OPTIMIZER_NEWTON, std::move(state));
  auto* raw = dynamic_cast<NewtonOptimizer<3>*>(base.get());
  auto sim = std::make_unique<mfem_sim>();
  sim->opt.reset(raw);
  base.release();
  sim->opt->reset();
  *out = sim.release();
}

SceneArrays scene_of(const mfem_sim_params& p) {
  if (p.nV <= 0 || p.nT <= 0) throw std::runtime_error("empty mesh");
  if (!p.V || !p.T || !p.mu || !p.la || !p.mat) throw std::runtime_error("null mesh array");
  SceneArrays sc;
  sc.dim = p.dim == 0 ? 3 : p.dim;
  if (sc.dim != 2 && sc.dim != 3) throw std::runtime_error("dim must be 2 or 3");
  const int D = sc.dim;
  // vertices stored as nV x 3 (z = 0 in 2D)
  sc.V.assign(3L * p.nV, 0.0);
  for (long i = 0; i < p.nV; ++i)
    for (int d = 0; d < D; ++d) sc.V[3 * i + d] = p.V[D * i + d];
  sc.T.assign(p.T, p.T + (long)(D + 1) * p.nT);
  for (int v : sc.T)
    if (v < 0 || v >= p.nV) throw std::runtime_error("element vertex index out of range");
  sc.mu.assign(p.mu, p.mu + p.nT);
  sc.la.assign(p.la, p.la + p.nT);
  sc.mat.assign(p.mat, p.mat + p.nT);
  if (p.fixed) sc.fixed.assign(p.fixed, p.fixed + p.nV);
  if (p.v0) {
    sc.v0.assign(3L * p.nV, 0.0);
    for (long i = 0; i < p.nV; ++i)
      for (int d = 0; d < D; ++d) sc.v0[3 * i + d] = p.v0[D * i + d];
  }
  if (p.x0) {
    sc.x0.assign(3L * p.nV, 0.0);
    for (long i = 0; i < p.nV; ++i)
      for (int d = 0; d < D; ++d) sc.x0[3 * i + d] = p.x0[D * i + d];
  }
  sc.density = p.density;
  sc.bc.type = (BCScriptType)p.bc_type;
  sc.bc.axis = p.bc_axis;
  sc.bc.ratio = p.bc_ratio;
  sc.bc.velocity = p.bc_velocity;
  return sc;
}

std::shared_ptr<SimConfig> config_of(const mfem_sim_params& p) {
  auto cfg = std::make_shared<SimConfig>();
  // no print_stats / print_timing keys on this path: quiet
  cfg->show_data = cfg->show_timing = false;
  cfg->h = p.dt;
  for (int i = 0; i < 3; ++i) cfg->ext[i] = (float)p.gravity[i];
  if (p.bdf_order < 1 || p.bdf_order > 6) throw std::runtime_error("bdf_order must be 1..6");
  cfg->ti_type = (TimeIntegratorType)(p.bdf_order - 1);
  LinearSolverFactory<3> lf;
  const std::string sname = p.linear_solver ? p.linear_solver : "eigen-pcg-diag";
  if (!lf.has_name(sname)) throw std::runtime_error("unknown linear solver '" + sname + "'");
  cfg->solver_type = lf.type_by_name(sname);
  cfg->itr_tol = p.itr_tol;
  cfg->max_iterative_solver_iters = p.max_iters;
  cfg->outer_steps = p.max_newton;
  cfg->newton_tol = p.newton_tol;
  cfg->ls_iters = p.ls_iters;
  cfg->precision = p.precision ? Precision::MIXED : Precision::FP64;
  cfg->mixed_variables.clear();
  cfg->variables.clear();
  return cfg;
}

}  // namespace

extern "C" {

const char* mfem_last_error(void) { return g_last_error.c_str(); }

int mfem_version(void) { return 1; }

int mfem_sim_create_json(const char* json_path, mfem_sim** out) {
  return guarded([&] {
    if (!out) throw std::runtime_error("null output handle");
    auto state = std::make_unique<SimState<3>>();
    if (!state->load(std::string(json_path))) throw std::runtime_error("scene load failed");
    finish_create(std::move(state), out);
  });
}

int mfem_sim_create_json_2d(const char* json_path, mfem_sim** out) {
  return guarded([&] {
    if (!out) throw std::runtime_error("null output handle");
    auto state = std::make_unique<SimState<3>>();
    if (!state->load(std::string(json_path), 2)) throw std::runtime_error("scene load failed");
    finish_create(std::move(state), out);
  });
}

int mfem_sim_create(const mfem_sim_params* p, mfem_sim** out) {
  return guarded([&] {
    if (!p || !out) throw std::runtime_error("null argument");
    SceneArrays sc = scene_of(*p);
    auto cfg = config_of(*p);
    auto state = std::make_unique<SimState<3>>();
    std::vector<std::string> mixed, plain;
    if (p->mixed) mixed.push_back("mixed-stretch");
    else plain.push_back("stretch");
    state->load_arrays(sc, cfg, mixed, plain);
    finish_create(std::move(state), out);
  });
}

int mfem_sim_destroy(mfem_sim* sim) {
  return guarded([&] { delete sim; });
}

int mfem_sim_step(mfem_sim* sim) {
  return guarded([&] { sim->opt->step(); });
}

int mfem_sim_set_callback(mfem_sim* sim, mfem_newton_callback cb, void* user) {
  return guarded([&] {
    if (!sim) throw std::invalid_argument("null simulation");
    if (!cb) {
      sim->opt->callback = [](const SimState<3>&) {};
      sim->opt->sync_each_iteration = false;
      return;
    }
    NewtonOptimizer<3>* opt = sim->opt.get();
    sim->opt->callback = [sim, opt, cb, user](const SimState<3>&) {
      cb(sim, (int)opt->history().size(), user);
    };
    sim->opt->sync_each_iteration = true;
  });
}

int mfem_sim_update_system(mfem_sim* sim) {
  return guarded([&] {
    sim->state().x_->pre_solve();
    sim->opt->update_system();
    MFEM_HIP_CHECK(hipStreamSynchronize(sim->state().ctx_->stream));
  });
}

int mfem_sim_solve(mfem_sim* sim, double* decrement) {
  return guarded([&] {
    double d = 0;
    sim->opt->substep(d);
    if (decrement) *decrement = d;
  });
}

int mfem_sim_energy(mfem_sim* sim, double alpha, double* energy) {
  return guarded([&] { *energy = sim->opt->energy(alpha); });
}

int mfem_sim_stats(mfem_sim* sim, mfem_step_stats* o) {
  return guarded([&] {
    const auto& s = sim->opt->last_stats();
    o->newton_iters = s.newton_iters;
    o->pcg_iters = s.pcg_iters;
    o->refinements = s.refinements;
    o->ls_iters = s.ls_iters;
    o->ls_fail = s.ls_fail;
    o->decrement = s.decrement;
    o->energy = s.energy;
    o->alpha = s.alpha;
    o->res = s.res;
    o->rel_residual = s.rel_residual;
    o->update_ms = s.update_ms;
    o->solve_ms = s.solve_ms;
    o->ls_ms = s.ls_ms;
    o->step_ms = s.step_ms;
    o->spmv_ms = s.spmv_ms;
    o->spmv_count = s.spmv_count;
  });
}

int mfem_sim_get_positions(mfem_sim* sim, double* out) {
  return guarded([&] {
    auto v = sim->state().x_->value().to_host(sim->state().ctx_->stream);
    const int D = sim->state().mesh_->dim();
    const size_t nV = v.size() / 3;
    for (size_t i = 0; i < nV; ++i)
      for (int d = 0; d < D; ++d) out[D * i + d] = v[3 * i + d];
  });
}

long mfem_sim_size(mfem_sim* sim, const char* name) {
  long n = -1;
  int rc = guarded([&] {
    const std::string k(name);
    if (k == "rowptr" || k == "colind" || k == "free_map" || k == "fixed" || k == "bsr_rowptr" ||
        k == "bsr_col" || k == "slot_ptr" || k == "slots" || k == "pcg_iters" ||
        k == "ic_color" || k == "ic_colors" || k == "ic_attempts")
      n = (long)fetch_int(sim, k).size();
    else
      n = (long)fetch(sim, k).size();
  });
  return rc == 0 ? n : -1;
}

int mfem_sim_get(mfem_sim* sim, const char* name, double* out) {
  return guarded([&] {
    auto v = fetch(sim, name);
    std::memcpy(out, v.data(), v.size() * sizeof(double));
  });
}

int mfem_sim_get_int(mfem_sim* sim, const char* name, int* out) {
  return guarded([&] {
    auto v = fetch_int(sim, name);
    std::memcpy(out, v.data(), v.size() * sizeof(int));
  });
}

int mfem_boundary_faces(const double* V, int nV, const int* T, int nT, int* F, int* nF) {
  return guarded([&] {
    if (!V || !T || !nF || nV < 0 || nT < 0) throw std::runtime_error("null argument");
    std::vector<int> t(T, T + 4L * nT);
    for (int v : t)
      if (v < 0 || v >= nV) throw std::runtime_error("tet vertex index out of range");
    auto f = boundary_faces(std::vector<double>(V, V + 3L * nV), t);
    if (F) {
      if (*nF < (int)(f.size() / 3)) throw std::runtime_error("face buffer too small");
      std::copy(f.begin(), f.end(), F);
    }
    *nF = (int)(f.size() / 3);
  });
}

namespace {
const std::vector<int>& output_surface(mfem_sim* sim) {
  auto& m = *sim->state().mesh_;
  if (sim->surface.empty()) sim->surface = m.dim() == 2 ? m.T_ : boundary_faces(m.Vref_, m.T_);
  return sim->surface;
}

std::vector<double> current_positions(mfem_sim* sim) {
  return sim->state().x_->value().to_host(sim->state().ctx_->stream);
}
}  // namespace

int mfem_sim_write_obj(mfem_sim* sim, const char* path) {
  return guarded([&] {
    if (!path) throw std::runtime_error("null path");
    write_obj(path, current_positions(sim), output_surface(sim));
  });
}

int mfem_sim_write_mesh(mfem_sim* sim, const char* path) {
  return guarded([&] {
    if (!path) throw std::runtime_error("null path");
    auto& m = *sim->state().mesh_;
    if (m.dim() != 3) throw std::runtime_error(".mesh output is for tetrahedral meshes");
    write_medit(path, current_positions(sim), m.T_, output_surface(sim));
  });
}

int mfem_sim_write_state(mfem_sim* sim, const char* x_path, const char* v_path) {
  return guarded([&] {
    auto& S = sim->state();
    const int D = S.mesh_->dim(), nV = S.mesh_->nV();
    hipStream_t st = S.ctx_->stream;
    const auto& I = *S.x_->integrator();
    std::vector<double> x(3L * nV), v(3L * nV);
    MFEM_HIP_CHECK(hipMemcpyAsync(x.data(), I.x_prev(), x.size() * sizeof(double),
                                  hipMemcpyDeviceToHost, st));
    MFEM_HIP_CHECK(hipMemcpyAsync(v.data(), I.v_prev(), v.size() * sizeof(double),
                                  hipMemcpyDeviceToHost, st));
    MFEM_HIP_CHECK(hipStreamSynchronize(st));
    auto take = [&](const std::vector<double>& a) {
      std::vector<double> o((size_t)nV * D);
      for (int i = 0; i < nV; ++i)
        for (int d = 0; d < D; ++d) o[(size_t)i * D + d] = a[3L * i + d];
      return o;
    };
    if (x_path) write_dmat(x_path, take(x), nV, D);
    if (v_path) write_dmat(v_path, take(v), nV, D);
  });
}

int mfem_sim_newton_history(mfem_sim* sim, double* out, int* rows) {
  return guarded([&] {
    if (!rows) throw std::runtime_error("null argument");
    const auto& h = sim->opt->history();
    if (out) {
      if (*rows < (int)h.size()) throw std::runtime_error("history buffer too small");
      for (size_t k = 0; k < h.size(); ++k) {
        const double row[10] = {(double)h[k].iteration, h[k].energy, h[k].energy_res,
                                h[k].newton_dec, h[k].alpha, h[k].update_ms, h[k].solve_ms,
                                h[k].ls_ms, (double)h[k].solver_iters, (double)h[k].ls_halvings};
        std::copy(row, row + 10, out + 10 * k);
      }
    }
    *rows = (int)h.size();
  });
}

int mfem_sim_set_element_kernel(mfem_sim* sim, int variant) {
  return guarded([&] {
    if (variant != 0 && variant != 1) throw std::runtime_error("element kernel variant must be 0 or 1");
    auto& S = sim->state();
    MixedStretch<3>* ms =
        S.mixed_vars_.empty() ? nullptr : dynamic_cast<MixedStretch<3>*>(S.mixed_vars_[0].get());
    if (!ms) throw std::runtime_error("the element kernel variant needs a mixed-stretch variable");
    ms->mfma_blocks_ = variant == 1;
  });
}

int mfem_sim_set_profiling(mfem_sim* sim, int flags) {
  return guarded([&] {
    sim->opt->solver()->profile_spmv = (flags & 1) != 0;
    sim->opt->phase_timing = (flags & 2) != 0;
  });
}

int mfem_sim_matrix_info(mfem_sim* sim, int* nfree, int* nnzb, int* value_bytes) {
  return guarded([&] {
    auto& m = *sim->state().mesh_;
    *nfree = m.nfree();
    *nnzb = m.nnzb();
    *value_bytes = sim->state().ctx_->fp32() ? 4 : 8;
  });
}

int mfem_sim_time_spmv(mfem_sim* sim, int reps, double* ms_per_launch) {
  return guarded([&] {
    auto& S = sim->state();
    auto* pcg = dynamic_cast<PcgSolver<3>*>(sim->opt->solver());
    if (!pcg || pcg->system().K.bytes() == 0) throw std::runtime_error("no assembled system");
    auto& m = *S.mesh_;
    hipStream_t st = S.ctx_->stream;
    const int n = 3 * m.nfree();
    const int var = std::max(pcg->spmv_variant(), 0);
    hipEvent_t a, b;
    MFEM_HIP_CHECK(hipEventCreate(&a));
    MFEM_HIP_CHECK(hipEventCreate(&b));
    if (S.ctx_->fp32()) {
      DeviceArray<float> x(n), y(n);
      std::vector<float> ones(n, 1.0f);
      x.upload(ones, st);
      mfemhip::launch_spmv<float, float>(m.nfree(), m.d_rowptr.data(), m.d_col.data(),
                                         pcg->system().K.f32(), x.data(), y.data(), st, var);
      MFEM_HIP_CHECK(hipEventRecord(a, st));
      for (int r = 0; r < reps; ++r)
        mfemhip::launch_spmv<float, float>(m.nfree(), m.d_rowptr.data(), m.d_col.data(),
                                           pcg->system().K.f32(), x.data(), y.data(), st, var);
      MFEM_HIP_CHECK(hipEventRecord(b, st));
      MFEM_HIP_CHECK(hipEventSynchronize(b));
    } else {
      DeviceArray<double> x(n), y(n);
      std::vector<double> ones(n, 1.0);
      x.upload(ones, st);
      mfemhip::launch_spmv<double, double>(m.nfree(), m.d_rowptr.data(), m.d_col.data(),
                                           pcg->system().K.f64(), x.data(), y.data(), st, var);
      MFEM_HIP_CHECK(hipEventRecord(a, st));
      for (int r = 0; r < reps; ++r)
        mfemhip::launch_spmv<double, double>(m.nfree(), m.d_rowptr.data(), m.d_col.data(),
                                             pcg->system().K.f64(), x.data(), y.data(), st, var);
      MFEM_HIP_CHECK(hipEventRecord(b, st));
      MFEM_HIP_CHECK(hipEventSynchronize(b));
    }
    float ms = 0;
    MFEM_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    *ms_per_launch = reps > 0 ? ms / reps : 0.0;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
  });
}

int mfem_sim_spmv_tuning(mfem_sim* sim, double* times, int* count, int* chosen, int* lanes,
                         int* unroll, int* nontemporal) {
  return guarded([&] {
    auto* pcg = dynamic_cast<PcgSolver<3>*>(sim->opt->solver());
    if (!pcg) throw std::runtime_error("no CG solver");
    const auto& t = pcg->spmv_tuning_ms();
    if (count) *count = mfemhip::spmv_variant_count();
    if (times)
      for (int i = 0; i < mfemhip::spmv_variant_count(); ++i)
        times[i] = i < (int)t.size() ? t[i] : 0.0;
    const int v = pcg->spmv_variant();
    if (chosen) *chosen = v;
    int G = 0, U = 0, NT = 0;
    mfemhip::spmv_variant_shape(std::max(v, 0), &G, &U, &NT);
    if (lanes) *lanes = G;
    if (unroll) *unroll = U;
    if (nontemporal) *nontemporal = NT;
  });
}

int mfem_sim_kkt_size(mfem_sim* sim, long* size) {
  return guarded([&] {
    auto* mr = dynamic_cast<MinresSolver<3>*>(sim->opt->solver());
    if (!mr) throw std::runtime_error("the linear solver is not \"minres-indefinite-block\"");
    *size = mr->kkt_size();
  });
}

int mfem_sim_kkt_apply(mfem_sim* sim, int which, const double* in, double* out) {
  return guarded([&] {
    auto* mr = dynamic_cast<MinresSolver<3>*>(sim->opt->solver());
    if (!mr) throw std::runtime_error("the linear solver is not \"minres-indefinite-block\"");
    std::vector<double> v;
    if (which != 2) v.assign(in, in + mr->kkt_size());
    auto y = mr->apply(which, v);
    std::copy(y.begin(), y.end(), out);
  });
}

int mfem_sim_set_cg_shape(mfem_sim* sim, int variant, int blocks) {
  return guarded([&] {
    auto* pcg = dynamic_cast<PcgSolver<3>*>(sim->opt->solver());
    if (!pcg) throw std::runtime_error("no CG solver");
    pcg->force_cg_shape(variant, blocks);
  });
}

int mfem_sim_cg_tuning(mfem_sim* sim, double* times, int* variants, int* grids, int* count,
                       int* chosen, int* lanes, int* unroll, int* threads, int* blocks) {
  return guarded([&] {
    auto* pcg = dynamic_cast<PcgSolver<3>*>(sim->opt->solver());
    if (!pcg) throw std::runtime_error("no CG solver");
    const auto& t = pcg->cg_trials();
    if (count) *count = (int)t.size();
    int ch = -1;
    for (size_t i = 0; i < t.size(); ++i) {
      if (times) times[i] = t[i].ms;
      if (variants) variants[i] = t[i].variant;
      if (grids) grids[i] = t[i].blocks;
      if (t[i].variant == pcg->cg_variant() && t[i].blocks == pcg->cg_blocks()) ch = (int)i;
    }
    if (chosen) *chosen = ch;
    int G = 0, U = 0, BS = 0;
    mfemhip::cg_variant_shape(std::max(pcg->cg_variant(), 0), &G, &U, &BS);
    if (lanes) *lanes = G;
    if (unroll) *unroll = U;
    if (threads) *threads = BS;
    if (blocks) *blocks = pcg->cg_blocks();
  });
}

int mfem_material_s6_batch(int n, const int* type, const double* mu, const double* la,
                           const double* s, int psd_fix, double* E, double* g, double* H,
                           void* stream) {
  return guarded([&] {
    mfemhip::launch_batch_material(n, type, mu, la, s, psd_fix, E, g, H, as_stream(stream));
    MFEM_HIP_CHECK(hipGetLastError());
  });
}

int mfem_snh_f9_batch(int n, const double* mu, const double* la, const double* F, double* E,
                      double* g, double* H, void* stream) {
  return guarded([&] {
    mfemhip::launch_batch_snh_f9(n, mu, la, F, E, g, H, as_stream(stream));
    MFEM_HIP_CHECK(hipGetLastError());
  });
}

int mfem_polar_batch(int n, const double* F, double* R, double* s, double* dsdF, int* iters,
                     void* stream) {
  return guarded([&] {
    mfemhip::launch_batch_polar(n, F, R, s, dsdF, iters, as_stream(stream));
    MFEM_HIP_CHECK(hipGetLastError());
  });
}

int mfem_batch_create(const mfem_sim_params* sims, int count, mfem_batch** out) {
  return guarded([&] {
    if (!sims || !out || count <= 0) throw std::runtime_error("null argument or empty batch");
    const mfem_sim_params& p0 = sims[0];
    std::vector<SceneArrays> scenes;
    for (int k = 0; k < count; ++k) {
      const mfem_sim_params& p = sims[k];
      if (p.dt != p0.dt || p.bdf_order != p0.bdf_order || p.precision != p0.precision ||
          p.itr_tol != p0.itr_tol || p.max_iters != p0.max_iters || p.max_newton != p0.max_newton ||
          p.newton_tol != p0.newton_tol || p.ls_iters != p0.ls_iters ||
          p.gravity[0] != p0.gravity[0] || p.gravity[1] != p0.gravity[1] ||
          p.gravity[2] != p0.gravity[2])
        throw std::runtime_error("batch: time step, integrator, gravity, solver settings and "
                                 "precision must be the same for every simulation");
      if (!p.mixed) throw std::runtime_error("batch: only mixed-stretch simulations are batched");
      if (p.dim != 0 && p.dim != 3) throw std::runtime_error("batch: tetrahedral meshes only");
      scenes.push_back(scene_of(p));
    }
    auto cfg = config_of(p0);
    if (cfg->solver_type != SOLVER_EIGEN_CG_DIAG)
      throw std::runtime_error("batch: the linear solver must be \"eigen-pcg-diag\"");
    auto b = std::make_unique<mfem_batch>();
    b->sim = std::make_unique<BatchSim>(scenes, cfg);
    *out = b.release();
  });
}

int mfem_batch_destroy(mfem_batch* b) {
  return guarded([&] { delete b; });
}

int mfem_batch_step(mfem_batch* b) {
  return guarded([&] { b->sim->step(); });
}

int mfem_batch_sizes(mfem_batch* b, int* count, int* nV, int* nT) {
  return guarded([&] {
    const int n = b->sim->count();
    if (count) *count = n;
    for (int k = 0; k < n; ++k) {
      if (nV) nV[k] = b->sim->vertices(k);
      if (nT) nT[k] = b->sim->tets(k);
    }
  });
}

int mfem_batch_get_positions(mfem_batch* b, int sim, double* out) {
  return guarded([&] {
    auto x = b->sim->positions(sim);
    std::memcpy(out, x.data(), x.size() * sizeof(double));
  });
}

int mfem_batch_stats(mfem_batch* b, int* newton_iters, int* pcg_iters, int* refinements,
                     int* ls_fail, double* decrement, double* step_ms) {
  return guarded([&] {
    const auto& st = b->sim->stats();
    for (size_t k = 0; k < st.size(); ++k) {
      if (newton_iters) newton_iters[k] = st[k].newton_iters;
      if (pcg_iters) pcg_iters[k] = st[k].pcg_iters;
      if (refinements) refinements[k] = st[k].refinements;
      if (ls_fail) ls_fail[k] = st[k].ls_fail;
      if (decrement) decrement[k] = st[k].decrement;
    }
    if (step_ms) *step_ms = b->sim->last_step_ms();
  });
}

int mfem_batch_set_cg_shape(mfem_batch* b, int variant, int rows_per_workgroup) {
  return guarded([&] { b->sim->force_cg_shape(variant, rows_per_workgroup); });
}

int mfem_batch_set_graphs(mfem_batch* b, int on) {
  return guarded([&] { b->sim->set_graphs(on != 0); });
}

int mfem_batch_launch_counts(mfem_batch* b, long* update_submissions, long* cg_submissions,
                             long* newton_iterations) {
  return guarded([&] {
    if (update_submissions) *update_submissions = b->sim->update_submissions();
    if (cg_submissions) *cg_submissions = b->sim->cg_submissions();
    if (newton_iterations) *newton_iterations = b->sim->newton_iterations();
  });
}

int mfem_batch_cg_shape(mfem_batch* b, int* variant, int* rows_per_workgroup) {
  return guarded([&] {
    if (variant) *variant = b->sim->cg_variant();
    if (rows_per_workgroup) *rows_per_workgroup = b->sim->rows_per_workgroup();
  });
}

}  // extern "C"
