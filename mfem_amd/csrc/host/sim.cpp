#include "sim.h"

#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <atomic>
#include <immintrin.h>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <queue>
#include <sstream>
#include <type_traits>

namespace mfem {

using namespace mfemhip;

// ============================================================ Reducer
void Reducer::init(long max_items) {
  // Largest grid of any reducing kernel: element kernels (n/128), BSR SpMV
  // with 8 lanes per block row (n/32 rows), vector kernels (n/256); the
  // two-value CG reduction stores 2 partials per block.
  long g = std::max<long>(reduce_grid(max_items), max_items / 32 + 1);
  stride_ = std::max<long>(2 * g + 64, 4096);
  partials_.resize((size_t)kSlots * stride_);
  // the finalize kernels write the results straight into mapped pinned
  // memory: a read is a stream synchronisation, with no copy on the stream
  results_ = std::make_unique<PinnedBuffer>(kSlots * sizeof(double),
                                            hipHostMallocMapped | hipHostMallocCoherent);
  for (int k = 0; k < kSlots; ++k) results_->as<double>()[k] = 0.0;
  MFEM_HIP_CHECK(hipDeviceSynchronize());
}

ReduceSlot Reducer::slot(int k) {
  ReduceSlot r;
  r.partials = partials_.data() + (size_t)k * stride_;
  r.result = results_->device<double>() + k;
  return r;
}

double Reducer::read(int k, hipStream_t st) {
  double v = 0;
  read_range(k, 1, &v, st);
  return v;
}

void Reducer::read_range(int k0, int n, double* out, hipStream_t st) {
  if (k0 < 0 || n < 0 || k0 + n > kSlots) throw std::logic_error("reducer slot out of range");
  MFEM_HIP_CHECK(hipStreamSynchronize(st));
  const volatile double* r = results_->as<double>() + k0;
  for (int i = 0; i < n; ++i) out[i] = r[i];
}

Inertia DeviceContext::inertia() const {
  Inertia in;
  in.x_tilde = x_tilde.data();
  for (int i = 0; i < 3; ++i) in.h2g[i] = h2g[i];
  return in;
}

DeviceContext::~DeviceContext() {
  if (own_stream && stream) (void)hipStreamDestroy(stream);
}

// reduction slot ids: the decrement pair and the two line-search energies of
// a Newton iteration are adjacent so that one read returns all of them
enum {
  RS_MAXDX = 0,
  RS_MAXDS = 1,
  RS_E0 = 2,      // E(0), 2 slots (one per elastic variable)
  RS_E1 = 4,      // E(1), 2 slots
  RS_ENERGY = 6,  // other E(alpha), 2 slots
  RS_RES = 8
};

// ============================================================ BDF
template <int I>
std::array<double, I> BDF<I>::alphas() {
  // reference src/time_integrators/BDF.cpp:45-75
  if constexpr (I == 1) return {1.0};
  if constexpr (I == 2) return {4.0 / 3.0, -1.0 / 3.0};
  if constexpr (I == 3) return {18.0 / 11.0, -9.0 / 11.0, 2.0 / 11.0};
  if constexpr (I == 4) return {48.0 / 25.0, -36.0 / 25.0, 16.0 / 25.0, -3.0 / 25.0};
  if constexpr (I == 5)
    return {300.0 / 137.0, -300.0 / 137.0, 200.0 / 137.0, -75.0 / 137.0, 12.0 / 137.0};
  if constexpr (I == 6)
    return {360.0 / 147.0, -450.0 / 147.0, 400.0 / 147.0, -225.0 / 147.0, 72.0 / 147.0,
            -10.0 / 147.0};
}

template <int I>
double BDF<I>::beta() {
  // reference src/time_integrators/BDF.cpp:77-93
  switch (I) {
    case 1: return 1.0;
    case 2: return 2.0 / 3.0;
    case 3: return 6.0 / 11.0;
    case 4: return 12.0 / 25.0;
    case 5: return 60.0 / 137.0;
    default: return 60.0 / 147.0;
  }
}

template <int I>
BDF<I>::BDF(const std::vector<double>& x0, const std::vector<double>& v0, double h, hipStream_t st)
    : ImplicitIntegrator(h) {
  n_ = (int)x0.size();
  for (int j = 0; j < I; ++j) {  // I copies of (x0, v0) (reference BDF.h:19-26)
    xs_[j].upload(x0, st);
    vs_[j].upload(v0, st);
  }
  head_ = 0;
}

template <int I>
void BDF<I>::x_tilde(double* out, hipStream_t st) const {
  const double* xh[6];
  const double* vh[6];
  for (int j = 0; j < I; ++j) {
    xh[j] = xs_[(head_ + j) % I].data();
    vh[j] = vs_[(head_ + j) % I].data();
  }
  auto a = alphas();
  launch_bdf_xtilde(n_, I, xh, vh, a.data(), dt(), out, st);
}

template <int I>
void BDF<I>::update(const double* x, hipStream_t st) {
  const double* xh[6];
  for (int j = 0; j < I; ++j) xh[j] = xs_[(head_ + j) % I].data();
  const int oldest = (head_ + I - 1) % I;
  auto a = alphas();
  launch_bdf_update(n_, I, xh, a.data(), dt(), x, xs_[oldest].data(), vs_[oldest].data(), st);
  head_ = oldest;
}

IntegratorFactory::IntegratorFactory() {
#define MFEM_REG_BDF(k)                                                                    \
  register_type(TI_BDF##k, BDF<k>::name(),                                                 \
                [](const std::vector<double>& x0, const std::vector<double>& v0, double h, \
                   hipStream_t st) -> std::unique_ptr<ImplicitIntegrator> {                \
                  return std::make_unique<BDF<k>>(x0, v0, h, st);                          \
                });
  MFEM_REG_BDF(1) MFEM_REG_BDF(2) MFEM_REG_BDF(3) MFEM_REG_BDF(4) MFEM_REG_BDF(5) MFEM_REG_BDF(6)
#undef MFEM_REG_BDF
}

// ============================================================ Displacement
template <int DIM>
Displacement<DIM>::Displacement(std::shared_ptr<Mesh> mesh, std::shared_ptr<SimConfig> config,
                                DeviceContext* ctx)
    : Variable<DIM>(std::move(mesh), ctx), config_(std::move(config)) {}

// reference displacement.cpp:76-98
template <int DIM>
void Displacement<DIM>::reset() {
  auto& m = *this->mesh_;
  hipStream_t st = this->ctx_->stream;
  x_full_.upload(m.V_, st);
  IntegratorFactory f;
  integrator_ = f.create(config_->ti_type, m.V_, m.initial_velocity_, config_->h, st);
  dx_.resize(3 * m.nfree());
  dx_.zero(st);
  this->ctx_->x_tilde.resize(m.V_.size());
  const double h = integrator_->dt();
  for (int i = 0; i < 3; ++i)  // body force is stored as float (reference config.h:93)
    this->ctx_->h2g[i] = h * h * (double)config_->ext[i];
}

template <int DIM>
void Displacement<DIM>::pre_solve() {
  integrator_->x_tilde(this->ctx_->x_tilde.data(), this->ctx_->stream);
}

// reference displacement.cpp:33-50: the pinned entries of x_full take the
// mesh's (possibly scripted) pinned positions, then the integrator update.
template <int DIM>
void Displacement<DIM>::post_solve() {
  auto& m = *this->mesh_;
  hipStream_t st = this->ctx_->stream;
  if (m.bc_moves() && !m.pinned().empty()) {
    const auto pos = m.pinned_positions();
    if (pinned_ids_.size() != m.pinned().size()) pinned_ids_.upload(m.pinned(), st);
    pinned_pos_.upload(pos, st);
    launch_scatter3((int)m.pinned().size(), pinned_ids_.data(), pinned_pos_.data(),
                    x_full_.data(), st);
    MFEM_HIP_CHECK(hipStreamSynchronize(st));  // the host copies above are pageable
  }
  integrator_->update(x_full_.data(), st);
}

template <int DIM>
int Displacement<DIM>::launch_max_abs_delta() {
  auto& R = this->ctx_->reducer;
  launch_reduce_maxabs((int)dx_.size(), dx_.data(), R.slot(RS_MAXDX), this->ctx_->stream);
  return RS_MAXDX;
}

// ============================================================ Stretch
template <int DIM>
Stretch<DIM>::Stretch(std::shared_ptr<Mesh> mesh, std::shared_ptr<SimConfig> config,
                      DeviceContext* ctx)
    : Variable<DIM>(std::move(mesh), ctx) {
  if (this->mesh_->dim() == 2) {
    // the 4-vector (2D) F API exists only for SNH (stable_neohookean.cpp:279-328)
    for (int t : this->mesh_->element_types())
      if (t != MATERIAL_SNH)
        throw std::runtime_error("material '" +
                                 MaterialModelFactory().name_by_type((MaterialModelType)t) +
                                 "' has no 2D displacement-only (4-vector) energy; 2D \"stretch\" "
                                 "needs Stable-Neohookean");
    return;
  }
  for (int t : this->mesh_->element_types())
    if (t != MATERIAL_SNH && t != MATERIAL_NH && t != MATERIAL_FUNG)
      throw std::runtime_error("material '" +
                               MaterialModelFactory().name_by_type((MaterialModelType)t) +
                               "' has no displacement-only (9-vector) energy; \"stretch\" needs "
                               "Stable-Neohookean, Neohookean or Fung");
}

template <int DIM>
void Stretch<DIM>::update(const double* x_full, double dt) {
  h2_ = dt * dt;
  auto* c = this->ctx_;
  auto m = this->mesh_->dev();
  if (m.dim == 2) {
    if (c->fp32())
      launch_plain2d_update<float>(m, x_full, c->inertia(), h2_, c->blocks.f32(), c->fe.data(),
                                   c->stream);
    else
      launch_plain2d_update<double>(m, x_full, c->inertia(), h2_, c->blocks.f64(),
                                    c->fe.data(), c->stream);
    return;
  }
  if (c->fp32())
    launch_plain_update<float>(m, x_full, c->inertia(), h2_, c->blocks.f32(), c->fe.data(),
                               c->stream);
  else
    launch_plain_update<double>(m, x_full, c->inertia(), h2_, c->blocks.f64(),
                                c->fe.data(), c->stream);
}

template <int DIM>
bool Stretch<DIM>::energy_launch(const double* x_full, const double* dx, double alpha,
                                 ReduceSlot out) {
  auto* c = this->ctx_;
  const auto m = this->mesh_->dev();
  if (m.dim == 2)
    launch_plain2d_energy(m, x_full, dx, alpha, c->inertia(), h2_, out, c->stream);
  else
    launch_plain_energy(m, x_full, dx, alpha, c->inertia(), h2_, out, c->stream);
  return true;
}

// ============================================================ MixedStretch
// reference mixed_stretch.cpp:262-304
template <int DIM>
void MixedStretch<DIM>::reset() {
  const bool two = this->mesh_->dim() == 2;
  if (two)
    for (int t : this->mesh_->element_types())
      if (t != MATERIAL_SNH && t != MATERIAL_ARAP && t != MATERIAL_COROT && t != MATERIAL_FCR)
        throw std::runtime_error("material '" +
                                 MaterialModelFactory().name_by_type((MaterialModelType)t) +
                                 "' has no 2D mixed (3-vector) energy; 2D \"mixed-stretch\" needs "
                                 "Stable-Neohookean, ARAP, Corot or FixedCorot");
  auto* c = this->ctx_;
  const int nT = this->mesh_->nT();
  hipStream_t st = c->stream;
  R_.resize(9 * (size_t)nT);
  S_.resize(6 * (size_t)nT);
  s_.resize(6 * (size_t)nT);
  la_.resize(6 * (size_t)nT);
  ds_.resize(6 * (size_t)nT);
  g_.resize(6 * (size_t)nT);
  gl_.resize(6 * (size_t)nT);
  dsdF_.resize(54 * (size_t)nT, c->fp32());
  Hpsi_.resize(21 * (size_t)nT, c->fp32());
  Hinv_.resize(21 * (size_t)nT, c->fp32());
  la_.zero(st);
  ds_.zero(st);
  s_.zero(st);
  // R = I, then fit at the rest configuration and set s = S
  std::vector<double> Rh(9 * (size_t)nT, 0.0);
  for (int e = 0; e < nT; ++e) {
    if (two) {
      Rh[0 * nT + e] = 1; Rh[3 * nT + e] = 1;
    } else {
      Rh[0 * nT + e] = 1; Rh[4 * nT + e] = 1; Rh[8 * nT + e] = 1;
    }
  }
  R_.upload(Rh, st);
  DeviceArray<double> x0;
  x0.upload(this->mesh_->V_, st);
  auto m = this->mesh_->dev();
  S_.zero(st);
  if (two && c->fp32())
    launch_mixed2d_polar_material<float>(m, x0.data(), R_.data(), S_.data(), s_.data(),
                                         dsdF_.f32(), hess_out<float>(), g_.data(), gl_.data(),
                                         0.0, st);
  else if (two)
    launch_mixed2d_polar_material<double>(m, x0.data(), R_.data(), S_.data(), s_.data(),
                                          dsdF_.f64(), hess_out<double>(), g_.data(), gl_.data(),
                                          0.0, st);
  else if (c->fp32())
    launch_mixed_polar_material<float>(m, x0.data(), R_.data(), S_.data(), s_.data(),
                                       dsdF_.f32(), hess_out<float>(), g_.data(), gl_.data(), 0.0,
                                       st);
  else
    launch_mixed_polar_material<double>(m, x0.data(), R_.data(), S_.data(), s_.data(),
                                        dsdF_.f64(), hess_out<double>(), g_.data(), gl_.data(),
                                        0.0, st);
  MFEM_HIP_CHECK(hipGetLastError());
  MFEM_HIP_CHECK(hipMemcpyAsync(s_.data(), S_.data(), S_.bytes(), hipMemcpyDeviceToDevice, st));
  MFEM_HIP_CHECK(hipStreamSynchronize(st));
}

// Where the element Hessian phase writes. FP64 mode: Hpsi_ and Hinv_. FP32
// storage: Hpsi_ in FP32 for the stiffness blocks or the KKT product; with the
// Schur complement, FP64 Hloc_ for the local solve (sized here, on first
// use), which on tets it solves with by LU, and on triangles the FP64 inverse
// Hlinv_; with the KKT solver, the FP32 inverse for its preconditioner.
template <int DIM>
template <typename Real>
HessOut<Real> MixedStretch<DIM>::hess_out() {
  HessOut<Real> ho;
  ho.Hpsi = Hpsi_.template as<Real>();
  if (!Hpsi_.fp32()) {
    ho.Hinv = Hinv_.template as<Real>();
    return ho;
  }
  if (!schur_blocks_) {
    ho.Hinv = Hinv_.template as<Real>();
    return ho;
  }
  const size_t n = 21 * (size_t)this->mesh_->nT();
  const bool tet = this->mesh_->dev().dim == 3;
  if (Hloc_.size() != n) Hloc_.resize(n);
  if (!tet && Hlinv_.size() != n) Hlinv_.resize(n);
  ho.Hloc = Hloc_.data();
  if (!tet) ho.Hlinv = Hlinv_.data();
  return ho;
}

// update_rotations + update_derivatives + rhs (reference mixed_stretch.cpp:
// 126-226), fused into two element kernels.
template <int DIM>
void MixedStretch<DIM>::update(const double* x_full, double dt) {
  h2_ = dt * dt;
  auto* c = this->ctx_;
  auto m = this->mesh_->dev();
  if (m.dim == 2) {
    if (c->fp32()) {
      launch_mixed2d_polar_material<float>(m, x_full, R_.data(), S_.data(), s_.data(),
                                           dsdF_.f32(), hess_out<float>(), g_.data(), gl_.data(),
                                           h2_, c->stream);
      if (schur_blocks_)
        launch_mixed2d_blocks_rhs<float>(m, x_full, c->inertia(), dsdF_.f32(), Hpsi_.f32(),
                                       gl_.data(), c->blocks.f32(), c->fe.data(), c->stream);
    } else {
      launch_mixed2d_polar_material<double>(m, x_full, R_.data(), S_.data(), s_.data(),
                                            dsdF_.f64(), hess_out<double>(), g_.data(),
                                            gl_.data(), h2_, c->stream);
      if (schur_blocks_)
        launch_mixed2d_blocks_rhs<double>(m, x_full, c->inertia(), dsdF_.f64(), Hpsi_.f64(),
                                        gl_.data(), c->blocks.f64(), c->fe.data(), c->stream);
    }
    return;
  }
  if (c->fp32()) {
    launch_mixed_polar_material<float>(m, x_full, R_.data(), S_.data(), s_.data(),
                                       dsdF_.f32(), hess_out<float>(), g_.data(), gl_.data(), h2_,
                                       c->stream);
    if (schur_blocks_)
      launch_mixed_blocks_rhs<float>(m, x_full, c->inertia(), dsdF_.f32(), Hpsi_.f32(),
                                     gl_.data(), c->blocks.f32(), c->fe.data(), c->stream,
                                     mfma_blocks_);
  } else {
    launch_mixed_polar_material<double>(m, x_full, R_.data(), S_.data(), s_.data(),
                                        dsdF_.f64(), hess_out<double>(), g_.data(), gl_.data(),
                                        h2_, c->stream);
    if (schur_blocks_)
      launch_mixed_blocks_rhs<double>(m, x_full, c->inertia(), dsdF_.f64(),
                                      Hpsi_.f64(), gl_.data(), c->blocks.f64(),
                                      c->fe.data(), c->stream);
  }
}

template <int DIM>
void MixedStretch<DIM>::solve(const double* dx) {
  auto* c = this->ctx_;
  auto m = this->mesh_->dev();
  if (m.dim == 2) {
    if (c->fp32())
      launch_mixed2d_local_solve<float>(m, dx, dsdF_.f32(), local_hpsi(), local_hinv(),
                                        g_.data(), gl_.data(), la_.data(), ds_.data(),
                                        c->reducer.slot(RS_MAXDS), c->stream);
    else
      launch_mixed2d_local_solve<double>(m, dx, dsdF_.f64(), local_hpsi(), local_hinv(),
                                         g_.data(), gl_.data(), la_.data(), ds_.data(),
                                         c->reducer.slot(RS_MAXDS), c->stream);
    return;
  }
  if (c->fp32())
    launch_mixed_local_solve<float>(m, dx, dsdF_.f32(), local_hpsi(), local_hinv(),
                                    g_.data(), gl_.data(), la_.data(), ds_.data(),
                                    c->reducer.slot(RS_MAXDS), c->stream);
  else
    launch_mixed_local_solve<double>(m, dx, dsdF_.f64(), local_hpsi(), local_hinv(),
                                     g_.data(), gl_.data(), la_.data(), ds_.data(),
                                     c->reducer.slot(RS_MAXDS), c->stream);
}

template <int DIM>
int MixedStretch<DIM>::max_abs_delta_slot() const {
  return RS_MAXDS;
}

template <int DIM>
void MixedStretch<DIM>::post_solve() {
  la_.zero(this->ctx_->stream);
}

template <int DIM>
bool MixedStretch<DIM>::energy_launch(const double* x_full, const double* dx, double alpha,
                                      ReduceSlot out) {
  auto* c = this->ctx_;
  const auto m = this->mesh_->dev();
  if (m.dim == 2)
    launch_mixed2d_energy(m, x_full, dx, alpha, c->inertia(), s_.data(), ds_.data(), la_.data(),
                          h2_, out, c->stream);
  else
    launch_mixed_energy(m, x_full, dx, alpha, c->inertia(), R_.data(), s_.data(), ds_.data(),
                        la_.data(), h2_, out, c->stream);
  return true;
}

template <int DIM>
bool MixedStretch<DIM>::energy_launch2(const double* x_full, const double* dx, double alpha0,
                                       ReduceSlot out0, double alpha1, ReduceSlot out1) {
  auto* c = this->ctx_;
  const auto m = this->mesh_->dev();
  if (m.dim == 2) return Variable<DIM>::energy_launch2(x_full, dx, alpha0, out0, alpha1, out1);
  launch_mixed_energy2(m, x_full, dx, alpha0, alpha1, c->inertia(), R_.data(), s_.data(),
                       ds_.data(), la_.data(), h2_, out0, out1, c->stream);
  return true;
}

template <int DIM>
VariableFactory<DIM>::VariableFactory() {
  this->register_type(VAR_STRETCH, Stretch<DIM>::name(),
                      [](std::shared_ptr<Mesh> m, std::shared_ptr<SimConfig> c,
                         DeviceContext* ctx) -> std::unique_ptr<Variable<DIM>> {
                        return std::make_unique<Stretch<DIM>>(m, c, ctx);
                      });
}

template <int DIM>
MixedVariableFactory<DIM>::MixedVariableFactory() {
  this->register_type(VAR_MIXED_STRETCH, MixedStretch<DIM>::name(),
                      [](std::shared_ptr<Mesh> m, std::shared_ptr<SimConfig> c,
                         DeviceContext* ctx) -> std::unique_ptr<MixedVariable<DIM>> {
                        return std::make_unique<MixedStretch<DIM>>(m, c, ctx);
                      });
}

// ============================================================ linear system
// reference linear_system.h:18-40: lhs = PMP + sum lhs, rhs = sum rhs; the
// element kernels already folded the inertia residual into fe.
template <int DIM>
void SystemMatrixPD<DIM>::pre_solve(SimState<DIM>* state) {
  auto* c = state->ctx_.get();
  auto& m = *state->mesh_;
  const bool f32 = c->fp32();
  K.resize((size_t)m.nnzb() * 9, f32);
  invdiag.resize(3 * (size_t)m.nfree(), f32);
  rhs.resize(3 * (size_t)m.nfree());
  // K and rhs gathers in one launch, then the Jacobi diagonal of K
  if (f32) {
    launch_assemble_bsr_rhs<float>(m.nupper(), m.d_upper.data(), m.d_tpos.data(),
                                   m.d_slot_ptr.data(), m.d_slots.data(), m.d_mass.data(),
                                   c->blocks.f32(), K.f32(), m.nfree(), m.d_asm_rows.data(),
                                   m.d_inc_ptr.data(), m.d_inc.data(), c->fe.data(), rhs.data(),
                                   c->stream);
    launch_jacobi_diag<float, float>(m.nfree(), m.d_diag_block.data(), K.f32(),
                                     invdiag.f32(), c->stream);
  } else {
    launch_assemble_bsr_rhs<double>(m.nupper(), m.d_upper.data(), m.d_tpos.data(),
                                    m.d_slot_ptr.data(), m.d_slots.data(), m.d_mass.data(),
                                    c->blocks.f64(), K.f64(), m.nfree(), m.d_asm_rows.data(),
                                    m.d_inc_ptr.data(), m.d_inc.data(), c->fe.data(), rhs.data(),
                                    c->stream);
    launch_jacobi_diag<double, double>(m.nfree(), m.d_diag_block.data(), K.f64(),
                                       invdiag.f64(), c->stream);
  }
}

template <int DIM>
void SystemMatrixPD<DIM>::post_solve(SimState<DIM>* state, const double* dx) {
  auto& d = state->x_->delta();
  if (d.data() != dx)
    MFEM_HIP_CHECK(hipMemcpyAsync(d.data(), dx, d.bytes(), hipMemcpyDeviceToDevice,
                                  state->ctx_->stream));
  for (auto& var : state->mixed_vars_) var->solve(d.data());
}

// ============================================================ CG
template <int DIM>
PcgSolver<DIM>::PcgSolver(SimState<DIM>* state) : LinearSolver<DIM>(state) {
  pinned_ = std::make_unique<PinnedBuffer>(2 * sizeof(PcgCtrl) + 64,
                                           hipHostMallocMapped | hipHostMallocCoherent);
  ctrl_.resize(1);
  parts_.resize(2 * (size_t)kCgParts * kCgMaxBlocks);
  stamps_.resize(2 * 2 * (size_t)kCgMaxBlocks);
  int dev = 0, khz = 0;
  MFEM_HIP_CHECK(hipGetDevice(&dev));
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0)
    wall_khz_ = khz;
}

template <int DIM>
PcgSolver<DIM>::~PcgSolver() = default;

template <int DIM>
void PcgSolver<DIM>::force_cg_shape(int variant, int blocks) {
  if (variant < -1 || variant >= cg_variant_count())
    throw std::invalid_argument("CG kernel shape out of range");
  if (blocks != 0 && (blocks < 8 || blocks > kCgMaxBlocks || blocks % 8 != 0))
    throw std::invalid_argument("CG grid must be 0 or a multiple of 8 in [8, 4096]");
  cg_variant_ = variant;
  cg_blocks_ = variant < 0 ? 0 : blocks;
}

template <int DIM>
std::vector<unsigned long long> PcgSolver<DIM>::cg_stamps(hipStream_t st) const {
  const int nb = cg_blocks_ > 0 ? cg_blocks_ : pcg_default_blocks(3 * this->state_->mesh_->nfree());
  auto all = stamps_.to_host(st);
  std::vector<unsigned long long> out;
  for (int half = 0; half < 2; ++half)
    out.insert(out.end(), all.begin() + (size_t)half * 2 * kCgMaxBlocks,
               all.begin() + (size_t)half * 2 * kCgMaxBlocks + 2 * (size_t)nb);
  return out;
}

template <int DIM>
template <typename T>
PcgBuffers<T> PcgSolver<DIM>::buffers(const double* b, double* x) {
  auto& m = *this->state_->mesh_;
  PcgBuffers<T> B;
  B.n = 3 * m.nfree();
  B.rowptr = m.d_rowptr.data();
  B.col = m.d_col.data();
  B.K = system_.K.template as<T>();
  B.invdiag = system_.invdiag.template as<T>();
  B.b = b;
  B.x = x;
  B.state[0] = state0_.template as<T>();
  B.state[1] = state1_.template as<T>();
  B.ctrl = ctrl_.data();
  B.parts[0] = parts_.data();
  B.parts[1] = parts_.data() + (size_t)kCgParts * kCgMaxBlocks;
  if (this->profile_spmv) {
    B.stamps[0] = stamps_.data();
    B.stamps[1] = stamps_.data() + 2 * (size_t)kCgMaxBlocks;
  }
  B.variant = std::max(cg_variant_, 0);
  B.blocks = cg_blocks_ > 0 ? cg_blocks_ : pcg_default_blocks(B.n);
  return B;
}

template <int DIM>
template <typename T>
int PcgSolver<DIM>::run_cg(const double* b, double* x, double tol, int max_iters) {
  hipStream_t st = this->state_->ctx_->stream;
  PcgBuffers<T> B = buffers<T>(b, x);
  // Every CG kernel that does work writes its index into mapped pinned
  // memory, and the one that stops the solve writes the control block there.
  // The host keeps `ahead` kernels queued beyond the last one that began, so
  // the device never waits for it and at most that many kernels are queued
  // past the stop (they exit at once); no event or copy enters the stream.
  volatile PcgCtrl* mirror = pinned_->as<PcgCtrl>();
  mirror->done = 0;
  mirror->started = -1;
  B.mirror = pinned_->device<PcgCtrl>();
  launch_pcg_init(B, tol, max_iters, st);
  const long ahead = std::max(2, this->state_->config_->pcg_poll / 4);
  const long limit = (long)max_iters + 2;  // kernels 0 .. max_iters + 1
  long launched = 0;
  auto launch_to = [&](long n) {
    n = std::min(n, limit);
    if (launched >= n) return;
    for (; launched < n; ++launched) launch_pcg_iteration(B, (int)launched, st);
    MFEM_HIP_CHECK(hipGetLastError());
  };
  launch_to(ahead);
  while (!mirror->done) {
    if (launched < limit) {
      launch_to((long)mirror->started + 1 + ahead);
      _mm_pause();
      continue;
    }
    // all kernels queued: the last one stops the solve at the latest
    MFEM_HIP_CHECK(hipStreamSynchronize(st));
    if (!mirror->done) throw std::runtime_error("CG control block did not terminate");
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  last_.rr = mirror->rr;
  last_.b2 = mirror->b2;
  last_.threshold = mirror->threshold;
  last_.kern_ticks = mirror->kern_ticks;
  last_.iter = mirror->iter;
  last_.done = 1;
  last_.max_iters = mirror->max_iters;
  last_.stop_it = mirror->stop_it;
  last_.kern_launches = mirror->kern_launches;
  if (this->profile_spmv) {
    this->stats_.spmv_ms += last_.kern_ticks / wall_khz_;
    this->stats_.spmv_count += last_.kern_launches;
  }
  return last_.iter;
}

// Autotuner: time every plain SpMV shape on the assembled matrix (device
// events around `reps` back-to-back launches) and keep the fastest. Runs once
// per solver, on the first solve; a fixed SimConfig::spmv_variant skips it.
template <int DIM>
template <typename T>
void PcgSolver<DIM>::tune_spmv() {
  auto* c = this->state_->ctx_.get();
  auto& m = *this->state_->mesh_;
  hipStream_t st = c->stream;
  const int nv = spmv_variant_count();
  const int forced = this->state_->config_->spmv_variant;
  if (forced >= 0 && forced < nv) {
    variant_ = forced;
    return;
  }
  const int reps = 20;
  hipEvent_t a, b;
  MFEM_HIP_CHECK(hipEventCreate(&a));
  MFEM_HIP_CHECK(hipEventCreate(&b));
  tune_ms_.assign(nv, 0.0);
  const T* x = system_.invdiag.template as<T>();
  T* y = q_.template as<T>();
  for (int v = 0; v < nv; ++v) {
    launch_spmv<T, T>(m.nfree(), m.d_rowptr.data(), m.d_col.data(), system_.K.template as<T>(), x,
                      y, st, v);  // warm-up
    MFEM_HIP_CHECK(hipEventRecord(a, st));
    for (int r = 0; r < reps; ++r)
      launch_spmv<T, T>(m.nfree(), m.d_rowptr.data(), m.d_col.data(), system_.K.template as<T>(),
                        x, y, st, v);
    MFEM_HIP_CHECK(hipEventRecord(b, st));
    MFEM_HIP_CHECK(hipEventSynchronize(b));
    float ms = 0;
    MFEM_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    tune_ms_[v] = ms / reps;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  variant_ = (int)(std::min_element(tune_ms_.begin(), tune_ms_.end()) - tune_ms_.begin());
}

// CG kernel autotuner: runs a few iterations of a CG solve on the current
// system (tolerance 0, so it does not stop) with every shape on grids of 1,
// 2, 3, 4 and 6 workgroups per CU, timing back-to-back iterations with device
// events, and keeps the fastest pair. Systems below kTuneMinRows block rows
// keep shape 0 and the default grid.
template <int DIM>
template <typename T>
void PcgSolver<DIM>::tune_cg() {
  constexpr int kTuneMinRows = 20000;
  auto& m = *this->state_->mesh_;
  hipStream_t st = this->state_->ctx_->stream;
  cg_trials_.clear();
  cg_variant_ = 0;
  cg_blocks_ = 0;
  if (m.nfree() < kTuneMinRows) return;
  const int n = 3 * m.nfree();
  const int reps = 10;
  hipEvent_t a, b;
  MFEM_HIP_CHECK(hipEventCreate(&a));
  MFEM_HIP_CHECK(hipEventCreate(&b));
  PcgBuffers<T> B = buffers<T>(system_.rhs.data(), x_.data());
  B.stamps[0] = B.stamps[1] = nullptr;
  double best = 1e300;
  for (int v = 0; v < cg_variant_count(); ++v) {
    std::vector<int> grids;
    for (int per_cu : {1, 2, 3, 4, 6}) {
      const int g = cg_blocks_per_cu(v, n, per_cu);
      if (std::find(grids.begin(), grids.end(), g) == grids.end()) grids.push_back(g);
    }
    for (int g : grids) {
      B.variant = v;
      B.blocks = g;
      launch_pcg_init(B, 0.0, 1 << 30, st);
      launch_pcg_iteration(B, 0, st);
      launch_pcg_iteration(B, 1, st);
      MFEM_HIP_CHECK(hipEventRecord(a, st));
      for (int it = 2; it < 2 + reps; ++it) launch_pcg_iteration(B, it, st);
      MFEM_HIP_CHECK(hipEventRecord(b, st));
      MFEM_HIP_CHECK(hipEventSynchronize(b));
      float ms = 0;
      MFEM_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
      const double t = ms / reps;
      cg_trials_.push_back({v, g, t});
      if (t < best) {
        best = t;
        cg_variant_ = v;
        cg_blocks_ = g;
      }
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
}

// reference eigen_iterative_solver.h:20-29 (compute + solve, x0 = 0)
template <int DIM>
void PcgSolver<DIM>::solve() {
  auto* c = this->state_->ctx_.get();
  auto& m = *this->state_->mesh_;
  auto& cfg = *this->state_->config_;
  hipStream_t st = c->stream;
  system_.pre_solve(this->state_);
  const int n = 3 * m.nfree();
  const bool f32 = c->fp32();
  // the solution is written straight into the displacement step, so
  // post_solve copies nothing (x_ is the autotuner's scratch solution)
  auto& dx = this->state_->x_->delta();
  if ((long)dx.size() != (long)n) throw std::logic_error("displacement step has the wrong size");
  double* xo = dx.data();
  x_.resize(n);
  q_.resize(n, f32);
  state0_.resize(4 * (size_t)n, f32);
  state1_.resize(4 * (size_t)n, f32);
  this->stats_ = SolveStats();
  rf_ = Refinement();
  if (variant_ < 0) {
    if (f32) tune_spmv<float>();
    else tune_spmv<double>();
  }
  if (cg_variant_ < 0) {
    if (f32) tune_cg<float>();
    else tune_cg<double>();
  }
  if (!f32) {
    this->stats_.iterations = run_cg<double>(system_.rhs.data(), xo, cfg.itr_tol,
                                             cfg.max_iterative_solver_iters);
    this->stats_.refinements = 1;
    this->stats_.rel_residual = last_.b2 > 0 ? std::sqrt(last_.rr / last_.b2) : 0.0;
  } else {
    // FP32 CG sweeps inside FP64 defect correction on the true residual. The
    // first sweep solves K x = b from x = 0 (its CG setup yields |b|^2); each
    // further sweep solves K d = b - K x for the FP64 residual and adds d.
    // The loop stops on Eigen's rule for the FP64 residual, on the iteration
    // budget, or after max_refinements sweeps.
    r64_.resize(n);
    dx64_.resize(n);
    rf_.budget = cfg.max_iterative_solver_iters;
    if (cfg.max_refinements > 0) {
      // sqrt(threshold / |r|^2) = tol while r = b
      const double inner = std::min(std::max(cfg.inner_tol, 0.5 * cfg.itr_tol), 0.5);
      const int it = run_cg<float>(system_.rhs.data(), xo, inner, rf_.budget);
      rf_.budget -= std::max(it, 1);
      this->stats_.iterations += it;
      rf_.b2 = last_.b2;
      rf_.r2 = last_.rr;
      ++rf_.sweeps;
    } else {
      dx.zero(st);
    }
    rf_.thr = std::max(cfg.itr_tol * cfg.itr_tol * rf_.b2, 2.2250738585072014e-308);
    if (rf_.b2 > 0) {
      launch_residual<float>(m.nfree(), m.d_rowptr.data(), m.d_col.data(), system_.K.f32(), xo,
                             system_.rhs.data(), r64_.data(), c->reducer.slot(RS_RES), st);
      if (this->defer_) rf_.pending = true;
      else refine(xo);
    }
    this->stats_.refinements = rf_.sweeps;
    this->stats_.rel_residual = rf_.b2 > 0 ? std::sqrt(rf_.r2 / rf_.b2) : 0.0;
  }
  system_.post_solve(this->state_, xo);
}

// The defect-correction loop from the residual check on: reads |r|^2 of the
// last residual launch and runs further sweeps while Eigen's rule fails.
template <int DIM>
void PcgSolver<DIM>::refine(double* xo) {
  auto* c = this->state_->ctx_.get();
  auto& m = *this->state_->mesh_;
  auto& cfg = *this->state_->config_;
  hipStream_t st = c->stream;
  const int n = 3 * m.nfree();
  for (;;) {
    rf_.r2 = c->reducer.read(RS_RES, st);
    if (!(rf_.r2 >= rf_.thr) || rf_.budget <= 0 || rf_.sweeps >= cfg.max_refinements) break;
    const double inner = std::max(cfg.inner_tol, std::sqrt(rf_.thr / rf_.r2) * 0.5);
    const int it = run_cg<float>(r64_.data(), dx64_.data(), std::min(inner, 0.5), rf_.budget);
    rf_.budget -= std::max(it, 1);
    this->stats_.iterations += it;
    launch_axpy(n, 1.0, dx64_.data(), xo, st);
    ++rf_.sweeps;
    launch_residual<float>(m.nfree(), m.d_rowptr.data(), m.d_col.data(), system_.K.f32(), xo,
                           system_.rhs.data(), r64_.data(), c->reducer.slot(RS_RES), st);
  }
}

// A deferred solve ends here: the residual check that solve() queued is
// read now, behind whatever the caller queued meanwhile. More sweeps change
// the solution; post_solve then runs again and the caller's work has to be
// queued again.
template <int DIM>
bool PcgSolver<DIM>::settle() {
  if (!rf_.pending) return true;
  rf_.pending = false;
  const int sweeps = rf_.sweeps;
  double* xo = this->state_->x_->delta().data();
  refine(xo);
  this->stats_.refinements = rf_.sweeps;
  this->stats_.rel_residual = rf_.b2 > 0 ? std::sqrt(rf_.r2 / rf_.b2) : 0.0;
  if (rf_.sweeps == sweeps) return true;
  system_.post_solve(this->state_, xo);
  return false;
}

// ============================================================ direct
template <int DIM>
DirectSolver<DIM>::DirectSolver(SimState<DIM>* state, bool lu) : LinearSolver<DIM>(state), lu_(lu) {
  const long n = 3L * state->mesh_->nfree();
  if (n > kMaxRows)
    throw std::runtime_error("direct solvers factor the system densely on the device and take at "
                             "most " + std::to_string(kMaxRows) + " unknowns (this system has " +
                             std::to_string(n) + "); use \"eigen-pcg-diag\"");
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success)
    throw std::runtime_error("rocblas_create_handle failed");
  handle_ = h;
  if (rocblas_set_stream(h, state->ctx_->stream) != rocblas_status_success)
    throw std::runtime_error("rocblas_set_stream failed");
  info_.resize(1);
}

template <int DIM>
DirectSolver<DIM>::~DirectSolver() {
  if (handle_) (void)rocblas_destroy_handle((rocblas_handle)handle_);
}

template <int DIM>
void DirectSolver<DIM>::solve() {
  auto* c = this->state_->ctx_.get();
  auto& m = *this->state_->mesh_;
  hipStream_t st = c->stream;
  system_.pre_solve(this->state_);
  const int n = 3 * m.nfree();
  this->stats_ = SolveStats();
  dense_.resize((size_t)n * n);
  x_.resize(n);
  MFEM_HIP_CHECK(hipMemsetAsync(dense_.data(), 0, dense_.bytes(), st));
  if (c->fp32())
    launch_bsr_to_dense<float>(m.nfree(), m.d_rowptr.data(), m.d_col.data(), system_.K.f32(),
                               dense_.data(), n, st);
  else
    launch_bsr_to_dense<double>(m.nfree(), m.d_rowptr.data(), m.d_col.data(), system_.K.f64(),
                                dense_.data(), n, st);
  MFEM_HIP_CHECK(hipMemcpyAsync(x_.data(), system_.rhs.data(), (size_t)n * sizeof(double),
                                hipMemcpyDeviceToDevice, st));
  auto h = (rocblas_handle)handle_;
  rocblas_status rs;
  if (lu_) {
    ipiv_.resize(n);
    rs = rocsolver_dgetrf(h, n, n, dense_.data(), n, ipiv_.data(), info_.data());
    if (rs == rocblas_status_success)
      rs = rocsolver_dgetrs(h, rocblas_operation_none, n, 1, dense_.data(), n, ipiv_.data(),
                            x_.data(), n);
  } else {
    rs = rocsolver_dpotrf(h, rocblas_fill_lower, n, dense_.data(), n, info_.data());
    if (rs == rocblas_status_success)
      rs = rocsolver_dpotrs(h, rocblas_fill_lower, n, 1, dense_.data(), n, x_.data(), n);
  }
  if (rs != rocblas_status_success)
    throw std::runtime_error("rocSOLVER factorization call failed (status " +
                             std::to_string((int)rs) + ")");
  const int info = info_.to_host(st)[0];
  if (info != 0)
    throw std::runtime_error(std::string(lu_ ? "LU" : "Cholesky") +
                             " factorization of the system matrix failed at column " +
                             std::to_string(info));
  this->stats_.refinements = 1;
  system_.post_solve(this->state_, x_.data());
}

// ============================================================ IC-CG
template <int DIM>
IcPcgSolver<DIM>::IcPcgSolver(SimState<DIM>* state) : LinearSolver<DIM>(state) {
  pinned_ = std::make_unique<PinnedBuffer>(2 * sizeof(IcCtrl) + 64);
  ctrl_.resize(1);
  fail_.resize(1);
  for (auto& e : poll_ev_) MFEM_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

template <int DIM>
IcPcgSolver<DIM>::~IcPcgSolver() {
  for (auto e : poll_ev_)
    if (e) (void)hipEventDestroy(e);
  for (auto g : pair_graph_)
    if (g) (void)hipGraphExecDestroy(g);
}

// Pattern work, once per simulation (the BSR pattern is fixed): the
// colouring -- the smallest-last order (repeatedly remove a vertex of least
// remaining degree, lowest index on ties), then greedy first fit in its
// reverse (Matula and Beck 1983; 4 colours on the Kuhn beams, where first
// fit in vertex order needs 8) -- the colour order, and the lower and upper
// block lists of every row.
template <int DIM>
void IcPcgSolver<DIM>::setup() {
  auto& m = *this->state_->mesh_;
  hipStream_t st = this->state_->ctx_->stream;
  const int nb = m.nfree();
  const auto& rp = m.rowptr_;
  const auto& cl = m.col_;
  std::vector<int> deg(nb, 0), removal;
  std::vector<char> removed(nb, 0);
  for (int v = 0; v < nb; ++v)
    for (int k = rp[v]; k < rp[v + 1]; ++k) deg[v] += cl[k] != v;
  std::priority_queue<std::pair<int, int>, std::vector<std::pair<int, int>>,
                      std::greater<std::pair<int, int>>> heap;
  for (int v = 0; v < nb; ++v) heap.push({deg[v], v});
  while (!heap.empty()) {
    const auto [d, v] = heap.top();
    heap.pop();
    if (removed[v] || d != deg[v]) continue;  // stale entry
    removed[v] = 1;
    removal.push_back(v);
    for (int k = rp[v]; k < rp[v + 1]; ++k) {
      const int u = cl[k];
      if (u != v && !removed[u]) heap.push({--deg[u], u});
    }
  }
  color_.assign(nb, -1);
  std::vector<int> mark;
  int ncol = 0;
  for (auto it = removal.rbegin(); it != removal.rend(); ++it) {
    const int v = *it;
    for (int k = rp[v]; k < rp[v + 1]; ++k) {
      const int u = cl[k];
      if (u != v && color_[u] >= 0) {
        if ((int)mark.size() <= color_[u]) mark.resize(color_[u] + 1, -1);
        mark[color_[u]] = v;
      }
    }
    int c = 0;
    while (c < (int)mark.size() && mark[c] == v) ++c;
    color_[v] = c;
    ncol = std::max(ncol, c + 1);
  }
  color_ptr_.assign(ncol + 1, 0);
  for (int v = 0; v < nb; ++v) color_ptr_[color_[v] + 1]++;
  for (int c = 0; c < ncol; ++c) color_ptr_[c + 1] += color_ptr_[c];
  std::vector<int> rows(nb), pos(nb), fill(color_ptr_.begin(), color_ptr_.end() - 1);
  for (int v = 0; v < nb; ++v) rows[fill[color_[v]]++] = v;  // by (colour, index)
  for (int p = 0; p < nb; ++p) pos[rows[p]] = p;
  // lower lists by position, each sorted by the position of its columns
  std::vector<int> lo_ptr(nb + 1, 0), lo_col, lo_pos, lo_k;
  std::vector<std::array<int, 3>> ent;  // (position, column, K block)
  for (int p = 0; p < nb; ++p) {
    const int i = rows[p];
    ent.clear();
    for (int k = rp[i]; k < rp[i + 1]; ++k)
      if (pos[cl[k]] < p) ent.push_back({pos[cl[k]], cl[k], k});
    std::sort(ent.begin(), ent.end());
    for (const auto& e : ent) {
      lo_pos.push_back(e[0]);
      lo_col.push_back(e[1]);
      lo_k.push_back(e[2]);
    }
    lo_ptr[p + 1] = (int)lo_col.size();
  }
  // upper lists by position: position q gathers L_jq from the later rows j,
  // in increasing position of j
  std::vector<int> up_ptr(nb + 1, 0), up_col(lo_col.size()), up_l(lo_col.size());
  for (int q : lo_pos) up_ptr[q + 1]++;
  for (int p = 0; p < nb; ++p) up_ptr[p + 1] += up_ptr[p];
  std::vector<int> up_fill(up_ptr.begin(), up_ptr.end() - 1);
  for (int p = 0; p < nb; ++p)
    for (int a = lo_ptr[p]; a < lo_ptr[p + 1]; ++a) {
      const int q = lo_pos[a];
      up_col[up_fill[q]] = rows[p];
      up_l[up_fill[q]++] = a;
    }
  color_rows_.upload(rows, st);
  lo_ptr_.upload(lo_ptr, st);
  lo_col_.upload(lo_col, st);
  lo_pos_.upload(lo_pos, st);
  lo_k_.upload(lo_k, st);
  up_ptr_.upload(up_ptr, st);
  up_col_.upload(up_col, st);
  up_l_.upload(up_l, st);
  L_.resize(9 * std::max<size_t>(lo_col.size(), 1));
  Dinv_.resize(9 * (size_t)nb);
  scale_.resize(3 * (size_t)nb);
  vecs_.resize(8 * 3 * (size_t)nb);
  parts_.resize(ic_partials(nb, color_ptr_.data(), ncol));
  MFEM_HIP_CHECK(hipStreamSynchronize(st));
  ready_ = true;
}

template <int DIM>
IcFactorDev IcPcgSolver<DIM>::factor_dev() {
  auto& m = *this->state_->mesh_;
  IcFactorDev f;
  f.nfree = m.nfree();
  f.dim = m.dim();
  f.rowptr = m.d_rowptr.data();
  f.col = m.d_col.data();
  f.diag = m.d_diag_block.data();
  f.color_rows = color_rows_.data();
  f.lo_ptr = lo_ptr_.data();
  f.lo_col = lo_col_.data();
  f.lo_pos = lo_pos_.data();
  f.lo_k = lo_k_.data();
  f.up_ptr = up_ptr_.data();
  f.up_col = up_col_.data();
  f.up_l = up_l_.data();
  f.L = L_.data();
  f.Dinv = Dinv_.data();
  f.scale = scale_.data();
  f.fail = fail_.data();
  return f;
}

template <int DIM>
template <typename MT>
void IcPcgSolver<DIM>::run(const MT* K) {
  auto& cfg = *this->state_->config_;
  hipStream_t st = this->state_->ctx_->stream;
  const IcFactorDev f = factor_dev();
  const size_t n3 = 3 * (size_t)f.nfree;
  if (n3 == 0) {  // every vertex pinned: dx is empty
    last_ = IcCtrl{};
    attempts_ = 0;
    return;
  }
  auto* snap = pinned_->as<IcCtrl>();
  int* flag = reinterpret_cast<int*>(snap + 2);
  // scaling, then factorizations with the shift loop of Eigen's
  // IncompleteCholesky::factorize (initial shift 1e-3 - min diag when that
  // is not positive; on a failed pivot max(1e-3, 2 shift); 10 attempts)
  launch_ic_scale<MT>(f, K, parts_.data(), ctrl_.data(), st);
  MFEM_HIP_CHECK(hipMemcpyAsync(snap, ctrl_.data(), sizeof(IcCtrl), hipMemcpyDeviceToHost, st));
  MFEM_HIP_CHECK(hipStreamSynchronize(st));
  const double mindiag = snap[0].mindiag;
  shift_ = mindiag <= 0 ? 1e-3 - mindiag : 0.0;
  for (attempts_ = 1;; ++attempts_) {
    fail_.zero(st);
    launch_ic_factor<MT>(f, K, shift_, color_ptr_.data(), colors(), st);
    MFEM_HIP_CHECK(hipGetLastError());
    MFEM_HIP_CHECK(hipMemcpyAsync(flag, fail_.data(), sizeof(int), hipMemcpyDeviceToHost, st));
    MFEM_HIP_CHECK(hipStreamSynchronize(st));
    if (*flag == 0) break;
    if (attempts_ >= 10)
      throw std::runtime_error("eigen-pcg-IC: the incomplete Cholesky factorization broke down "
                               "with every shift up to " + std::to_string(shift_));
    shift_ = std::max(1e-3, 2 * shift_);
  }
  // CG: one chunk of iterations queued behind the one being polled
  IcVectors v;
  double* w = vecs_.data();
  v.b = system_.rhs.data();
  v.x = w;
  v.r = w + n3;
  v.p[0] = w + 2 * n3;
  v.p[1] = w + 3 * n3;
  v.q = w + 4 * n3;
  v.y = w + 5 * n3;
  v.zt = w + 6 * n3;
  v.z = w + 7 * n3;
  v.parts = parts_.data();
  v.ctrl = ctrl_.data();
  launch_ic_cg_init(f, v, cfg.itr_tol, cfg.max_iterative_solver_iters, color_ptr_.data(), colors(),
                    st);
  // the iteration pair as a graph (its kernels take only pointers that stay
  // fixed for the solver's lifetime, and the matrix, checked here)
  const int gi = std::is_same<MT, float>::value ? 1 : 0;
  if (!pair_graph_[gi] || pair_graph_K_[gi] != (const void*)K) {
    if (pair_graph_[gi]) MFEM_HIP_CHECK(hipGraphExecDestroy(pair_graph_[gi]));
    pair_graph_[gi] = nullptr;
    hipGraph_t g = nullptr;
    MFEM_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    launch_ic_cg_iteration<MT>(f, K, v, 0, color_ptr_.data(), colors(), st);
    launch_ic_cg_iteration<MT>(f, K, v, 1, color_ptr_.data(), colors(), st);
    MFEM_HIP_CHECK(hipStreamEndCapture(st, &g));
    MFEM_HIP_CHECK(hipGraphInstantiate(&pair_graph_[gi], g, nullptr, nullptr, 0));
    MFEM_HIP_CHECK(hipGraphDestroy(g));
    pair_graph_K_[gi] = K;
  }
  const int poll = std::max(1, (cfg.pcg_poll + 1) / 2);              // pairs per poll
  const long limit = (std::max(cfg.max_iterative_solver_iters, 0) + 1) / 2;  // pairs
  long launched = 0;
  int issued = 0, waited = 0;
  auto issue = [&] {
    const long chunk = std::min<long>(poll, limit - launched);
    for (long j = 0; j < chunk; ++j) MFEM_HIP_CHECK(hipGraphLaunch(pair_graph_[gi], st));
    launched += chunk;
    MFEM_HIP_CHECK(hipGetLastError());
    const int slot = issued & 1;
    MFEM_HIP_CHECK(hipMemcpyAsync(snap + slot, ctrl_.data(), sizeof(IcCtrl),
                                  hipMemcpyDeviceToHost, st));
    MFEM_HIP_CHECK(hipEventRecord(poll_ev_[slot], st));
    ++issued;
  };
  issue();
  while (true) {
    if (launched < limit && issued - waited < 2) issue();
    const int slot = waited & 1;
    MFEM_HIP_CHECK(hipEventSynchronize(poll_ev_[slot]));
    ++waited;
    if (snap[slot].done) {  // set at the latest by iteration max_iters
      last_ = snap[slot];
      break;
    }
    if (waited == issued) throw std::runtime_error("IC-CG control block did not terminate");
  }
}

// reference eigen_iterative_solver.h:20-29 (compute + solve, x0 = 0)
template <int DIM>
void IcPcgSolver<DIM>::solve() {
  system_.pre_solve(this->state_);
  if (!ready_) setup();
  this->stats_ = SolveStats();
  if (system_.K.fp32()) run<float>(system_.K.f32());
  else run<double>(system_.K.f64());
  this->stats_.iterations = last_.iter;
  this->stats_.refinements = 1;
  this->stats_.rel_residual = last_.b2 > 0 ? std::sqrt(last_.rr / last_.b2) : 0.0;
  system_.post_solve(this->state_, vecs_.data());
}

// ============================================================ MINRES
template <int DIM>
MinresSolver<DIM>::MinresSolver(SimState<DIM>* state) : LinearSolver<DIM>(state) {
  if (state->mixed_vars_.size() != 1 || !state->vars_.empty())
    throw std::runtime_error(
        "minres-indefinite-block needs exactly one mixed variable (\"mixed-stretch\") and no "
        "displacement-only elastic variable");
  mixed_ = dynamic_cast<MixedStretch<DIM>*>(state->mixed_vars_[0].get());
  if (!mixed_) throw std::runtime_error("minres-indefinite-block needs \"mixed-stretch\"");
  mixed_->schur_blocks_ = false;
  auto& m = *state->mesh_;
  hipStream_t st = state->ctx_->stream;
  n_ = 3L * m.nfree();
  N_ = n_ + 12L * m.nT();
  ns_ = m.dim() == 2 ? 3 : 6;
  Nref_ = (long)m.dim() * m.nfree() + 2L * ns_ * m.nT();
  // lumped mass (row sums of PMP) and the Jacobi scaling of the mass block
  std::vector<double> ml(m.nfree(), 0.0), md(m.nfree(), 0.0);
  const auto& mass = m.mass();
  for (int i = 0; i < m.nfree(); ++i) {
    for (int b = m.rowptr_[i]; b < m.rowptr_[i + 1]; ++b) ml[i] += mass[b];
    ml[i] = 1.0 / ml[i];
    md[i] = 1.0 / mass[m.diag_block()[i]];
  }
  mlump_inv_.upload(ml, st);
  mdinv_.upload(md, st);
  vecs_.resize(7 * (size_t)N_);
  sinv_.resize(21 * (size_t)m.nT());
  cheb_.resize(3 * (size_t)n_);
  parts_.resize(2 * (size_t)(kkt_tet_blocks(m.nT()) + kkt_vertex_blocks(m.nfree())) + 2);
  ctrl_.resize(1);
  pinned_ = std::make_unique<PinnedBuffer>(2 * sizeof(MinresCtrl));
  for (auto& e : poll_ev_) MFEM_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  MFEM_HIP_CHECK(hipStreamSynchronize(st));
}

template <int DIM>
MinresSolver<DIM>::~MinresSolver() {
  for (auto& e : poll_ev_)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : prof_ev_)
    if (e) (void)hipEventDestroy(e);
}

template <int DIM>
template <typename Real>
KktSystem<Real> MinresSolver<DIM>::system() {
  auto& m = *this->state_->mesh_;
  auto* c = this->state_->ctx_.get();
  KktSystem<Real> S;
  S.m = m.dev();
  S.rowptr = m.d_rowptr.data();
  S.col = m.d_col.data();
  S.mass = m.d_mass.data();
  S.inc_ptr = m.d_inc_ptr.data();
  S.inc = m.d_inc.data();
  S.dsdF = mixed_->dsdF_.template as<Real>();
  S.Hpsi = mixed_->Hpsi_.template as<Real>();
  S.Hinv = mixed_->Hinv_.template as<Real>();
  S.Sinv = sinv_.data();
  S.mlump_inv = mlump_inv_.data();
  S.mdinv = mdinv_.data();
  S.fe = c->fe.data();
  S.cheb_r = cheb_.data();
  S.cheb_d[0] = cheb_.data() + n_;
  S.cheb_d[1] = cheb_.data() + 2 * n_;
  const long np = kkt_tet_blocks(m.nT()) + kkt_vertex_blocks(m.nfree());
  S.parts[0] = parts_.data();
  S.parts[1] = parts_.data() + np;
  S.ctrl = ctrl_.data();
  return S;
}

// EigenIterativeSolver::solve (eigen_iterative_solver.h:20-30) with
// SystemMatrixIndefinite::pre_solve / post_solve (linear_system.h:67-99):
// MINRES from x = 0; the solution is written into dx, ds and lambda.
template <int DIM>
template <typename Real>
void MinresSolver<DIM>::run() {
  auto& m = *this->state_->mesh_;
  auto* c = this->state_->ctx_.get();
  auto& cfg = *this->state_->config_;
  hipStream_t st = c->stream;
  const KktSystem<Real> S = system<Real>();
  MinresBuffers B;
  auto vec = [&](int k) {
    double* base = vecs_.data() + (size_t)k * N_;
    return KktVec{base, base + n_, base + n_ + 6L * m.nT()};
  };
  B.v[0] = vec(0);
  B.v[1] = vec(1);
  B.w[0] = vec(2);
  B.w[1] = vec(3);
  for (int k = 0; k < 3; ++k) B.p[k] = vec(4 + k);
  B.x = KktVec{this->state_->x_->delta().data(), mixed_->ds_.data(), mixed_->la_.data()};
  // v_{-1}, p_{-1}, p_{-2} and x start at zero
  MFEM_HIP_CHECK(hipMemsetAsync(B.v[1].x, 0, N_ * sizeof(double), st));
  MFEM_HIP_CHECK(hipMemsetAsync(B.p[0].x, 0, 3 * N_ * sizeof(double), st));
  this->state_->x_->delta().zero(st);
  mixed_->ds_.zero(st);
  mixed_->la_.zero(st);
  launch_kkt_setup(S, st);
  launch_minres_init(S, this->state_->x_->value().data(), c->inertia(), mixed_->g_.data(),
                     mixed_->S_.data(), mixed_->s_.data(), B.v[0], B.w[0], cfg.itr_tol,
                     cfg.max_iterative_solver_iters, st);
  // iterations in chunks, two control-block snapshots in flight; kernels
  // queued after the stop exit at once
  const int poll = std::max(1, cfg.pcg_poll);
  const long limit = (long)cfg.max_iterative_solver_iters + 1;
  auto* snap = pinned_->as<MinresCtrl>();
  long launched = 0;
  int issued = 0, waited = 0;
  const bool prof = this->profile_spmv;
  if (prof && prof_ev_.empty()) {
    prof_ev_.resize(2 * kProfiledIters);
    for (auto& e : prof_ev_) MFEM_HIP_CHECK(hipEventCreate(&e));
  }
  auto issue = [&] {
    const long chunk = std::min<long>(poll, limit - launched);
    for (long j = 0; j < chunk; ++j) {
      const long k = launched + j;
      hipEvent_t* ev = prof && k < kProfiledIters ? &prof_ev_[2 * k] : nullptr;
      launch_minres_iteration(S, B, k, st, ev);
    }
    launched += chunk;
    MFEM_HIP_CHECK(hipGetLastError());
    const int slot = issued & 1;
    MFEM_HIP_CHECK(hipMemcpyAsync(snap + slot, ctrl_.data(), sizeof(MinresCtrl),
                                  hipMemcpyDeviceToHost, st));
    MFEM_HIP_CHECK(hipEventRecord(poll_ev_[slot], st));
    ++issued;
  };
  issue();
  while (true) {
    if (launched < limit && issued - waited < 2) issue();
    const int slot = waited & 1;
    MFEM_HIP_CHECK(hipEventSynchronize(poll_ev_[slot]));
    ++waited;
    if (snap[slot].done) {
      last_ = snap[slot];
      break;
    }
    if (waited == issued) throw std::runtime_error("MINRES control block did not terminate");
  }
  // one more launch applies the update of the last iteration (it is a no-op
  // when a queued launch already did)
  launch_minres_iteration(S, B, launched, st);
  if (last_.fail)
    throw std::runtime_error("minres-indefinite-block: preconditioner is not positive definite");
  if (prof) {
    // element-kernel time of the iterations that ran the product (the
    // launches queued after the stop exit at once and are not counted)
    const long n = std::min<long>({(long)kProfiledIters, launched, (long)last_.iter + 1});
    MFEM_HIP_CHECK(hipStreamSynchronize(st));
    for (long k = 0; k < n; ++k) {
      float ms = 0;
      MFEM_HIP_CHECK(hipEventElapsedTime(&ms, prof_ev_[2 * k], prof_ev_[2 * k + 1]));
      this->stats_.spmv_ms += ms;
      this->stats_.spmv_count += 1;
    }
  }
}

template <int DIM>
template <typename Real>
void MinresSolver<DIM>::apply_device(int which) {
  auto& m = *this->state_->mesh_;
  auto* c = this->state_->ctx_.get();
  hipStream_t st = c->stream;
  const KktSystem<Real> S = system<Real>();
  auto vec = [&](int k) {
    double* base = vecs_.data() + (size_t)k * N_;
    return KktVec{base, base + n_, base + n_ + 6L * m.nT()};
  };
  if (which == 2) {
    launch_minres_init(S, this->state_->x_->value().data(), c->inertia(), mixed_->g_.data(),
                       mixed_->S_.data(), mixed_->s_.data(), vec(1), vec(2), 1.0, 1, st);
  } else {
    launch_kkt_apply(S, which, vec(0), vec(1), vec(2), st);
  }
}

template <int DIM>
std::vector<double> MinresSolver<DIM>::apply(int which, const std::vector<double>& in) {
  auto& m = *this->state_->mesh_;
  auto* c = this->state_->ctx_.get();
  hipStream_t st = c->stream;
  if (which < 0 || which > 2) throw std::invalid_argument("kkt apply: which must be 0, 1 or 2");
  const int nT = m.nT(), D = m.dim(), nf = m.nfree();
  const long nr = (long)D * nf, ns = ns_;
  // reference layout [dx (D nfree); ds (ns nT); lambda (ns nT)], element
  // parts element-major -> device layout (3 per free vertex, element parts
  // structure of arrays)
  std::vector<double> dev(N_, 0.0), out(Nref_);
  if (which != 2) {
    if ((long)in.size() != Nref_) throw std::invalid_argument("kkt apply: wrong vector size");
    for (int f = 0; f < nf; ++f)
      for (int r = 0; r < D; ++r) dev[3L * f + r] = in[(long)D * f + r];
    for (int e = 0; e < nT; ++e)
      for (int i = 0; i < ns; ++i) {
        dev[n_ + (long)i * nT + e] = in[nr + ns * e + i];
        dev[n_ + 6L * nT + (long)i * nT + e] = in[nr + ns * nT + ns * e + i];
      }
    MFEM_HIP_CHECK(hipMemcpyAsync(vecs_.data(), dev.data(), N_ * sizeof(double),
                                  hipMemcpyHostToDevice, st));
  }
  if (c->fp32()) apply_device<float>(which);
  else apply_device<double>(which);
  MFEM_HIP_CHECK(hipMemcpyAsync(dev.data(), vecs_.data() + N_, N_ * sizeof(double),
                                hipMemcpyDeviceToHost, st));
  MFEM_HIP_CHECK(hipStreamSynchronize(st));
  for (int f = 0; f < nf; ++f)
    for (int r = 0; r < D; ++r) out[(long)D * f + r] = dev[3L * f + r];
  for (int e = 0; e < nT; ++e)
    for (int i = 0; i < ns; ++i) {
      out[nr + ns * e + i] = dev[n_ + (long)i * nT + e];
      out[nr + ns * nT + ns * e + i] = dev[n_ + 6L * nT + (long)i * nT + e];
    }
  return out;
}

template <int DIM>
void MinresSolver<DIM>::solve() {
  auto* c = this->state_->ctx_.get();
  this->stats_ = SolveStats();
  if (c->fp32()) run<float>();
  else run<double>();
  this->stats_.iterations = last_.iter;
  this->stats_.refinements = 1;
  this->stats_.rel_residual = estimated_error();
  // max |ds| into the slot the Newton decrement reads
  launch_reduce_maxabs(ns_ * this->state_->mesh_->nT(), mixed_->ds_.data(),
                       c->reducer.slot(RS_MAXDS), c->stream);
}

template <int DIM>
LinearSolverFactory<DIM>::LinearSolverFactory() {
  auto cholesky = [](SimState<DIM>* s) -> std::unique_ptr<LinearSolver<DIM>> {
    return std::make_unique<DirectSolver<DIM>>(s, false);
  };
  this->register_type(SOLVER_EIGEN_LLT, "eigen-llt", cholesky);
  this->register_type(SOLVER_EIGEN_LDLT, "eigen-ldlt", cholesky);
  this->register_type(SOLVER_CHOLMOD, "cholmod", cholesky);
  this->register_type(SOLVER_EIGEN_LU, "eigen-lu",
                      [](SimState<DIM>* s) -> std::unique_ptr<LinearSolver<DIM>> {
                        return std::make_unique<DirectSolver<DIM>>(s, true);
                      });
  this->register_type(SOLVER_EIGEN_CG_DIAG, "eigen-pcg-diag",
                      [](SimState<DIM>* s) -> std::unique_ptr<LinearSolver<DIM>> {
                        return std::make_unique<PcgSolver<DIM>>(s);
                      });
  this->register_type(SOLVER_EIGEN_CG_IC, "eigen-pcg-IC",
                      [](SimState<DIM>* s) -> std::unique_ptr<LinearSolver<DIM>> {
                        return std::make_unique<IcPcgSolver<DIM>>(s);
                      });
  this->register_type(SOLVER_MINRES_ID, "minres-indefinite-block",
                      [](SimState<DIM>* s) -> std::unique_ptr<LinearSolver<DIM>> {
                        return std::make_unique<MinresSolver<DIM>>(s);
                      });
}

// ============================================================ optimizer
template <int DIM>
void Optimizer<DIM>::reset() {}

template <int DIM>
NewtonOptimizer<DIM>::~NewtonOptimizer() {
  for (auto& e : ev_)
    if (e) (void)hipEventDestroy(e);
}

// reference newton_optimizer.cpp:127-140
template <int DIM>
void NewtonOptimizer<DIM>::reset() {
  Optimizer<DIM>::reset();
  auto& S = *this->state_;
  S.x_->reset();
  for (auto& v : S.vars_) v->reset();
  for (auto& v : S.mixed_vars_) v->reset();
  LinearSolverFactory<DIM> f;
  solver_ = f.create(S.config_->solver_type, &S);
  if (!solver_) throw std::runtime_error("unknown linear solver");
}

// reference newton_optimizer.cpp:92-110
template <int DIM>
void NewtonOptimizer<DIM>::update_system() {
  auto& S = *this->state_;
  const double dt = S.x_->integrator()->dt();
  const double* x = S.x_->value().data();
  for (auto& v : S.vars_) v->update(x, dt);
  for (auto& v : S.mixed_vars_) v->update(x, dt);
  if (S.vars_.empty() && S.mixed_vars_.empty()) {
    // no elastic variable: inertia-only residual and zero stiffness blocks
    auto* c = S.ctx_.get();
    const auto m = S.mesh_->dev();
    if (m.dim == 2) {
      if (c->fp32())
        launch_plain2d_update<float>(m, x, c->inertia(), 0.0, c->blocks.f32(), c->fe.data(),
                                     c->stream);
      else
        launch_plain2d_update<double>(m, x, c->inertia(), 0.0, c->blocks.f64(), c->fe.data(),
                                      c->stream);
    } else if (c->fp32()) {
      launch_plain_update<float>(m, x, c->inertia(), 0.0, c->blocks.f32(), c->fe.data(),
                                 c->stream);
    } else {
      launch_plain_update<double>(m, x, c->inertia(), 0.0, c->blocks.f64(), c->fe.data(),
                                  c->stream);
    }
  }
}

// reference newton_optimizer.cpp:113-124
template <int DIM>
void NewtonOptimizer<DIM>::substep(double& decrement) {
  auto& S = *this->state_;
  solver_->solve();
  double v[2] = {0, 0};
  S.ctx_->reducer.read_range(RS_MAXDX, launch_decrement(), v, S.ctx_->stream);
  decrement = std::max(v[0], v[1]);
}

template <int DIM>
int NewtonOptimizer<DIM>::launch_decrement() {
  auto& S = *this->state_;
  S.x_->launch_max_abs_delta();
  // the local solve already reduced max |ds| into its slot (RS_MAXDS)
  for (auto& v : S.mixed_vars_)
    if (v->max_abs_delta_slot() != RS_MAXDS) throw std::logic_error("unexpected max|ds| slot");
  return S.mixed_vars_.empty() ? 1 : 2;
}

// reference newton_optimizer.cpp:34-49: E(a) = inertia + sum of variable
// energies at (x + a dx, s + a ds); the element kernels include the inertia.
template <int DIM>
int NewtonOptimizer<DIM>::launch_energy(double alpha, int base) {
  auto& S = *this->state_;
  auto* c = S.ctx_.get();
  const double* x = S.x_->value().data();
  const double* dx = S.x_->delta().data();
  int k = 0;
  for (auto& v : S.mixed_vars_)
    if (v->energy_launch(x, dx, alpha, c->reducer.slot(base + k))) ++k;
  for (auto& v : S.vars_)
    if (k < 2 && v->energy_launch(x, dx, alpha, c->reducer.slot(base + k))) ++k;
  if (S.mixed_vars_.size() + S.vars_.size() > 2)
    throw std::runtime_error("at most two elastic variables are supported");
  if (k == 0) {  // no elastic variable: inertia only
    const auto m = S.mesh_->dev();
    if (m.dim == 2)
      launch_plain2d_energy(m, x, dx, alpha, c->inertia(), 0.0, c->reducer.slot(base), c->stream);
    else
      launch_plain_energy(m, x, dx, alpha, c->inertia(), 0.0, c->reducer.slot(base), c->stream);
    k = 1;
  }
  return k;
}

template <int DIM>
int NewtonOptimizer<DIM>::launch_energy2(double alpha0, int base0, double alpha1, int base1) {
  auto& S = *this->state_;
  auto* c = S.ctx_.get();
  const double* x = S.x_->value().data();
  const double* dx = S.x_->delta().data();
  auto& R = c->reducer;
  int k = 0;
  for (auto& v : S.mixed_vars_)
    if (v->energy_launch2(x, dx, alpha0, R.slot(base0 + k), alpha1, R.slot(base1 + k))) ++k;
  for (auto& v : S.vars_)
    if (k < 2 && v->energy_launch2(x, dx, alpha0, R.slot(base0 + k), alpha1, R.slot(base1 + k)))
      ++k;
  if (S.mixed_vars_.size() + S.vars_.size() > 2)
    throw std::runtime_error("at most two elastic variables are supported");
  if (k == 0) {
    launch_energy(alpha0, base0);
    return launch_energy(alpha1, base1);
  }
  return k;
}

template <int DIM>
double NewtonOptimizer<DIM>::energy(double alpha) {
  auto& S = *this->state_;
  double v[2] = {0, 0};
  const int k = launch_energy(alpha, RS_ENERGY);
  S.ctx_->reducer.read_range(RS_ENERGY, k, v, S.ctx_->stream);
  return k == 2 ? v[0] + v[1] : v[0];
}

// reference newton_optimizer.cpp:11-89 with linesearch_backtracking
// (linesearch.h:78-131, called with c = 0 and p = 0.5: accept on strict
// decrease; an exhausted search leaves the state unchanged). The decrement,
// E(0) and the first trial E(1) are launched together and read with one
// synchronisation.
template <int DIM>
void NewtonOptimizer<DIM>::step() {
  auto& S = *this->state_;
  auto* c = S.ctx_.get();
  hipStream_t st = c->stream;
  stats_ = StepStats();
  history_.clear();
  S.x_->pre_solve();
  for (auto& v : S.mixed_vars_) v->pre_solve();
  for (auto& v : S.vars_) v->pre_solve();
  int i = 0;
  double grad_norm = 0, E = 0, E_prev = 0, res = 0;
  const bool has_mixed = !S.mixed_vars_.empty();
  const int ls_iters = S.config_->ls_iters;
  // Phase timings come from events read once the step has ended: nothing in
  // the next iteration waits for the accepted step, so its first kernels are
  // queued behind it without a synchronisation in between.
  auto events = [&](int it) {
    while ((int)ev_.size() < 4 * (it + 1)) {
      hipEvent_t e = nullptr;
      MFEM_HIP_CHECK(hipEventCreate(&e));
      ev_.push_back(e);
    }
    return &ev_[4 * it];
  };
  auto phase_times = [&](int it) {
    const hipEvent_t* e = &ev_[4 * it];
    float t01 = 0, t12 = 0, t23 = 0;
    MFEM_HIP_CHECK(hipEventElapsedTime(&t01, e[0], e[1]));
    MFEM_HIP_CHECK(hipEventElapsedTime(&t12, e[1], e[2]));
    MFEM_HIP_CHECK(hipEventElapsedTime(&t23, e[2], e[3]));
    auto& h = history_[it];
    h.update_ms = t01;
    h.solve_ms = t12;
    h.ls_ms = t23;
    stats_.update_ms += t01;
    stats_.solve_ms += t12;
    stats_.ls_ms += t23;
  };
  const bool timing = this->phase_timing || this->sync_each_iteration || S.config_->show_timing;
  auto mark = [&](hipEvent_t* ev, int k) {
    if (ev) MFEM_HIP_CHECK(hipEventRecord(ev[k], st));
  };
  int timed = 0;  // iterations whose phase times are in stats_
  do {
    hipEvent_t* ev = timing ? events(i) : nullptr;
    mark(ev, 0);
    update_system();
    mark(ev, 1);
    // The decrement and the first line-search energies depend only on the
    // step: they are queued behind the solver's last residual check and
    // queued again in the rare case that the check asks for another sweep.
    solver_->solve_deferred();
    int nd = 0, ne = 0;
    do {
      nd = launch_decrement();
      mark(ev, 2);
      ne = ls_iters > 0 ? launch_energy2(0.0, RS_E0, 1.0, RS_E1) : launch_energy(0.0, RS_E0);
    } while (!solver_->settle());
    double v[6] = {0, 0, 0, 0, 0, 0};
    c->reducer.read_range(RS_MAXDX, ls_iters > 0 ? RS_E1 + 2 : RS_E0 + 2, v, st);
    grad_norm = nd == 2 ? std::max(v[0], v[1]) : v[0];
    const auto& ss = solver_->stats();
    stats_.pcg_iters += ss.iterations;
    stats_.refinements += ss.refinements;
    stats_.rel_residual = ss.rel_residual;
    stats_.spmv_ms += ss.spmv_ms;
    stats_.spmv_count += ss.spmv_count;
    E = ne == 2 ? v[RS_E0] + v[RS_E0 + 1] : v[RS_E0];
    res = std::fabs((E - E_prev) / (E + 1e-6));
    E_prev = E;
    // fx0 = energy(0) equals E (same state, deterministic kernels)
    const double fx0 = E;
    double alpha = 1.0;
    int iter = 0;
    if (ls_iters > 0) {
      double fx = ne == 2 ? v[RS_E1] + v[RS_E1 + 1] : v[RS_E1];
      while (!(fx < fx0)) {
        alpha *= 0.5;
        if (++iter >= ls_iters) break;
        fx = energy(alpha);
      }
    }
    if (iter < ls_iters) {
      double* s = has_mixed ? S.mixed_vars_[0]->value().data() : nullptr;
      const double* ds = has_mixed ? S.mixed_vars_[0]->delta().data() : nullptr;
      launch_accept_step(S.mesh_->dev(), S.x_->value().data(), S.x_->delta().data(), s, ds, alpha,
                         has_mixed ? 1 : 0, st);
    } else {
      stats_.ls_fail++;
    }
    stats_.ls_iters += iter;
    mark(ev, 3);
    stats_.alpha = alpha;
    history_.push_back({i + 1, E, res, grad_norm, alpha, 0.0, 0.0, 0.0, ss.iterations, iter});
    if (this->sync_each_iteration) {
      MFEM_HIP_CHECK(hipEventSynchronize(ev[3]));
      for (; timed <= i; ++timed) phase_times(timed);
    }
    ++i;
    this->callback(S);
  } while (i < S.config_->outer_steps && grad_norm > S.config_->newton_tol && res > 1e-12);
  if (timing) {
    MFEM_HIP_CHECK(hipEventSynchronize(ev_[4 * (i - 1) + 3]));
    for (; timed < i; ++timed) phase_times(timed);
  }
  stats_.newton_iters = i;
  stats_.decrement = grad_norm;
  stats_.energy = E;
  stats_.res = res;
  stats_.step_ms = stats_.update_ms + stats_.solve_ms + stats_.ls_ms;
  if (S.config_->show_data) print_history();
  // reference newton_optimizer.cpp:78-87: scripted boundary motion, then the
  // variables' post_solve
  S.mesh_->update_bcs(S.config_->h);
  S.x_->post_solve();
  for (auto& v : S.mixed_vars_) v->post_solve();
  for (auto& v : S.vars_) v->post_solve();
  MFEM_HIP_CHECK(hipGetLastError());
}

template <int DIM>
void NewtonOptimizer<DIM>::print_history() const {
  std::printf("%10s | %14s | %11s | %11s | %9s | %8s\n", "Iteration", "Energy", "Energy res",
              "Newton dec", "alpha", "solver");
  for (const auto& r : history_)
    std::printf("%10d | %14.8e | %11.4e | %11.4e | %9.3e | %8d\n", r.iteration, r.energy,
                r.energy_res, r.newton_dec, r.alpha, r.solver_iters);
  if (this->state_->config_->show_timing && !history_.empty()) {
    // the reference's Timer keys: "update", "global" (the solve), "LS"
    double tu = 0, tg = 0, tl = 0;
    for (const auto& r : history_) {
      tu += r.update_ms;
      tg += r.solve_ms;
      tl += r.ls_ms;
    }
    const double n = (double)history_.size();
    std::printf("Timing (in ms, device events):\n");
    std::printf("  [%10s]  Avg: %10.4f   Total: %10.4f\n", "LS", tl / n, tl);
    std::printf("  [%10s]  Avg: %10.4f   Total: %10.4f\n", "global", tg / n, tg);
    std::printf("  [%10s]  Avg: %10.4f   Total: %10.4f\n", "update", tu / n, tu);
  }
  std::fflush(stdout);
}

template <int DIM>
OptimizerFactory<DIM>::OptimizerFactory() {
  this->register_type(OPTIMIZER_NEWTON, NewtonOptimizer<DIM>::name(),
                      [](std::unique_ptr<SimState<DIM>> s) -> std::unique_ptr<Optimizer<DIM>> {
                        return std::make_unique<NewtonOptimizer<DIM>>(std::move(s));
                      });
}

// ============================================================ SimState
template <int DIM>
void SimState<DIM>::load_arrays(const SceneArrays& sc, std::shared_ptr<SimConfig> config,
                                const std::vector<std::string>& mixed_variables,
                                const std::vector<std::string>& variables) {
  config_ = std::move(config);
  if (sc.dim != 2 && sc.dim != 3) throw std::runtime_error("dim must be 2 or 3");
  const int nT = (int)(sc.T.size() / (sc.dim + 1));
  // per-element Lame parameters (one model per tet, as with per-tet
  // material ids); materials[0] supplies the density
  MaterialModelFactory mf;
  auto cfg0 = std::make_shared<MaterialConfig>();
  cfg0->density = sc.density;
  cfg0->material_model = nT > 0 ? (MaterialModelType)sc.mat[0] : MATERIAL_SNH;
  material_models_ = {std::shared_ptr<MaterialModel>(mf.create(cfg0->material_model, cfg0))};
  if (!material_models_[0]) throw std::runtime_error("unknown material type");
  if ((int)sc.mu.size() != nT || (int)sc.la.size() != nT || (int)sc.mat.size() != nT)
    throw std::runtime_error("per-element arrays must have nT entries");
  mesh_ = std::make_shared<Mesh>(sc.V, sc.T, material_models_[0], sc.mu, sc.la, sc.mat, sc.dim);
  mesh_->bc_config_ = sc.bc;
  if (!sc.v0.empty()) mesh_->initial_velocity_ = sc.v0;
  if (!sc.x0.empty()) {
    if (sc.x0.size() != mesh_->Vinit_.size()) throw std::runtime_error("x0 has the wrong size");
    mesh_->Vinit_ = mesh_->V_ = sc.x0;
  }
  fixed_override_ = sc.fixed;
  MixedVariableFactory<DIM> mvf;
  VariableFactory<DIM> vf;
  ctx_ = std::make_unique<DeviceContext>();
  ctx_->precision = config_->precision;
  mixed_vars_.clear();
  vars_.clear();
  for (const auto& n : mixed_variables) {
    if (!mvf.has_name(n)) throw std::runtime_error("unknown mixed variable '" + n + "'");
    config_->mixed_variables.insert(mvf.type_by_name(n));
  }
  for (const auto& n : variables) {
    if (!vf.has_name(n)) throw std::runtime_error("unknown variable '" + n + "'");
    config_->variables.insert(vf.type_by_name(n));
  }
  finish_setup();
}

template <int DIM>
void SimState<DIM>::finish_setup() {
  MixedVariableFactory<DIM> mvf;
  VariableFactory<DIM> vf;
  if (!ctx_) {
    ctx_ = std::make_unique<DeviceContext>();
    ctx_->precision = config_->precision;
  }
  // FP32 storage pays off only in the iterative solvers' sweeps over K; the
  // direct solvers factor K densely in FP64, and an FP32-rounded K can stop
  // being positive definite when element stiffnesses span many decades
  // (E = 6e4 and 1e14 in the reference's gummy-bear scene). With a direct
  // solver the mixed mode therefore stores everything in FP64.
  const auto solver = config_->solver_type;
  if (solver == SOLVER_EIGEN_LLT || solver == SOLVER_EIGEN_LDLT || solver == SOLVER_EIGEN_LU ||
      solver == SOLVER_CHOLMOD)
    ctx_->precision = Precision::FP64;
  if (!ctx_->stream) {
    MFEM_HIP_CHECK(hipStreamCreateWithFlags(&ctx_->stream, hipStreamNonBlocking));
    ctx_->own_stream = true;
  }
  mesh_->init(ctx_->stream, fixed_override_);
  const int nT = mesh_->nT();
  ctx_->reducer.init(std::max<long>(std::max(nT, 3 * mesh_->nV()), 1));
  ctx_->blocks.resize(90 * (size_t)nT, ctx_->fp32());
  ctx_->fe.resize(12 * (size_t)nT);
  x_ = std::make_unique<Displacement<DIM>>(mesh_, config_, ctx_.get());
  mixed_vars_.clear();
  vars_.clear();
  for (VariableType t : config_->mixed_variables) {
    auto v = mvf.create(t, mesh_, config_, ctx_.get());
    if (v) mixed_vars_.push_back(std::move(v));
  }
  for (VariableType t : config_->variables) {
    auto v = vf.create(t, mesh_, config_, ctx_.get());
    if (v) vars_.push_back(std::move(v));
  }
  if (mixed_vars_.size() + vars_.size() > 1)
    throw std::runtime_error("the device backend supports one elastic variable per simulation");
}

template <int DIM>
bool SimState<DIM>::load(const std::string& json_file, int dim) {
  if (std::filesystem::path(json_file).extension() != ".json") {
    std::cerr << "File: " << json_file << " needs to be json" << std::endl;
    return false;
  }
  Json args;
  try {
    args = Json::parse_file(json_file);
  } catch (const std::exception& e) {
    std::cerr << e.what() << std::endl;
    return false;
  }
  return load(args, std::filesystem::path(json_file).parent_path().string(), dim);
}

// reference simulation_state.cpp:344-366
template <int DIM>
void SimState<DIM>::load_params(const Json& a) {
  if (a.has("body_force")) {
    auto v = a["body_force"].as_numbers();
    for (int i = 0; i < 3 && i < (int)v.size(); ++i) config_->ext[i] = (float)v[i];
  }
  if (a.has("dt")) config_->h = a["dt"].as_number();
  if (a.has("print_timing")) config_->show_timing = a["print_timing"].as_bool();
  if (a.has("print_stats")) config_->show_data = a["print_stats"].as_bool();
  if (a.has("enable_ccd")) config_->enable_ccd = a["enable_ccd"].as_bool();
  if (a.has("dhat")) config_->dhat = a["dhat"].as_number();
  if (a.has("kappa")) config_->kappa = a["kappa"].as_number();
  if (a.has("max_newton_iterations")) config_->outer_steps = a["max_newton_iterations"].as_int();
  if (a.has("max_linesearch_iterations")) config_->ls_iters = a["max_linesearch_iterations"].as_int();
  if (a.has("max_iterative_solver_iters"))
    config_->max_iterative_solver_iters = a["max_iterative_solver_iters"].as_int();
  if (a.has("iterative_solver_tolerance")) config_->itr_tol = a["iterative_solver_tolerance"].as_number();
  // device options (no counterpart in the reference)
  if (a.has("newton_tol")) config_->newton_tol = a["newton_tol"].as_number();
  if (a.has("precision"))
    config_->precision = a["precision"].as_string() == "mixed" ? Precision::MIXED : Precision::FP64;
  if (a.has("pcg_poll")) config_->pcg_poll = a["pcg_poll"].as_int();
}

// reference simulation_state.cpp:92-341
template <int DIM>
bool SimState<DIM>::load(const Json& args, const std::string& base_dir, int dim) {
  if (dim != 2 && dim != 3) throw std::runtime_error("scene dimension must be 2 or 3");
  config_ = std::make_shared<SimConfig>();
  load_params(args);
  if (args.has("optimizer")) {
    OptimizerFactory<DIM> of;
    config_->optimizer = of.type_by_name(args["optimizer"].as_string());
  }
  if (args.has("time_integrator")) {
    IntegratorFactory itf;
    config_->ti_type = itf.type_by_name(args["time_integrator"].as_string());
  }
  MaterialModelFactory mf;
  std::vector<std::shared_ptr<MaterialModel>> materials;
  if (args.has("material_models")) {
    for (const auto& obj : args["material_models"].arr) {
      auto cfg = std::make_shared<MaterialConfig>();
      if (obj.has("youngs_modulus")) cfg->ym = obj["youngs_modulus"].as_number();
      if (obj.has("poissons_ratio")) cfg->pr = obj["poissons_ratio"].as_number();
      if (obj.has("density")) cfg->density = obj["density"].as_number();
      if (obj.has("energy")) cfg->material_model = mf.type_by_name(obj["energy"].as_string());
      Enu_to_lame(cfg->ym, cfg->pr, cfg->la, cfg->mu);
      materials.push_back(std::shared_ptr<MaterialModel>(mf.create(cfg->material_model, cfg)));
    }
  } else {
    auto cfg = std::make_shared<MaterialConfig>();
    materials.push_back(std::shared_ptr<MaterialModel>(mf.create(cfg->material_model, cfg)));
  }
  material_models_ = materials;
  // relative paths: against the scene file's folder, else as given (the
  // reference opens them relative to the working directory,
  // simulation_state.cpp:50-67)
  auto resolve = [&](const std::string& p) {
    std::filesystem::path q(p);
    if (q.is_absolute()) return p;
    const std::string here = (std::filesystem::path(base_dir) / q).string();
    return std::filesystem::exists(here) || !std::filesystem::exists(q) ? here : p;
  };
  // objects: the device backend takes one object per simulation
  std::vector<double> V;
  std::vector<int> T;
  std::vector<int> mat_ids;
  BoundaryConditionConfig bc;
  std::vector<double> v0;
  if (!args.has("objects") || args["objects"].size() != 1) {
    std::cerr << "the device backend loads exactly one object per scene" << std::endl;
    return false;
  }
  const Json& obj = args["objects"][0];
  if (!obj.has("path")) {
    std::cerr << "Object missing path!" << std::endl;
    return false;
  }
  if (dim == 3) {
    read_mesh_file(resolve(obj["path"].as_string()), V, T);
    // 3D meshes are scaled by 1/max coefficient (reference :62-68)
    double fac = *std::max_element(V.begin(), V.end());
    for (auto& v : V) v /= fac;
  } else {
    // igl::read_triangle_mesh + remove_unreferenced, z truncated (:46-60)
    read_obj_file(resolve(obj["path"].as_string()), V, T);
    remove_unreferenced(V, T, 3);
    for (size_t i = 2; i < V.size(); i += 3) V[i] = 0.0;
  }
  const int nV = (int)(V.size() / 3);
  // offset, then the DIM x DIM transformation (column-major) about the
  // centroid (reference :187-198)
  std::vector<double> offset = {0, 0, 0};
  if (obj.has("offset")) offset = obj["offset"].as_numbers();
  for (int i = 0; i < nV; ++i)
    for (int d = 0; d < dim && d < (int)offset.size(); ++d) V[3 * i + d] += offset[d];
  if (obj.has("transformation")) {
    auto tr = obj["transformation"].as_numbers();
    if ((int)tr.size() != dim * dim) throw std::runtime_error("transformation must have DIM^2 entries");
    double cen[3] = {0, 0, 0};
    for (int i = 0; i < nV; ++i)
      for (int d = 0; d < dim; ++d) cen[d] += V[3 * i + d] / nV;
    for (int i = 0; i < nV; ++i) {
      double p[3];
      for (int d = 0; d < dim; ++d) p[d] = V[3 * i + d] - cen[d];
      for (int r = 0; r < dim; ++r) {
        double a = 0;
        for (int c = 0; c < dim; ++c) a += tr[r + dim * c] * p[c];
        V[3 * i + r] = a + cen[r];
      }
    }
  }
  const int nT = (int)(T.size() / (dim + 1));
  int idx = 0;
  if (obj.has("material_ids")) {
    if (dim == 3) {
      int rows, cols;
      auto ids = read_dmat(resolve(obj["material_ids"].as_string()), rows, cols);
      if ((long)rows * cols < nT) throw std::runtime_error("material_ids has too few entries");
      mat_ids.resize(nT);
      for (int e = 0; e < nT; ++e) {
        mat_ids[e] = (int)ids[e];
        if (mat_ids[e] < 0 || mat_ids[e] >= (int)materials.size())
          throw std::runtime_error("material id out of range");
      }
    } else {
      // the reference's 2D branch keeps materials[0] (:200-205)
      std::cerr << "Missing heterogeneous 2Dtrimesh support!" << std::endl;
    }
  } else if (obj.has("material_index")) {
    idx = obj["material_index"].as_int();
    if (idx < 0 || idx >= (int)materials.size()) throw std::runtime_error("material_index out of range");
  }
  if (mat_ids.empty()) mat_ids.assign(nT, idx);
  if (obj.has("initial_velocity")) {
    auto v = obj["initial_velocity"].as_numbers();
    v0.assign(V.size(), 0.0);
    for (int i = 0; i < nV; ++i)
      for (int d = 0; d < dim && d < (int)v.size(); ++d) v0[3 * i + d] = v[d];
  }
  if (obj.has("boundary_condition")) {
    const Json& b = obj["boundary_condition"];
    static const std::pair<const char*, BCScriptType> names[] = {
        {"null", BC_NULL}, {"scale", BC_SCALEF}, {"randomize", BC_RANDOM},
        {"onepoint", BC_ONEPOINT}, {"hang", BC_HANG}, {"hangends", BC_HANGENDS},
        {"bend", BC_BEND}, {"stretch", BC_STRETCH}, {"twist", BC_TWIST}};
    if (b.has("type")) {
      std::string t = b["type"].as_string();
      bc.type = BC_NULL;
      bool found = false;
      for (auto& nm : names)
        if (t == nm.first) { bc.type = nm.second; found = true; }
      if (!found) std::cerr << "Factory type_by_name: type not found" << std::endl;
    }
    if (b.has("axis")) bc.axis = b["axis"].as_int();
    if (b.has("ratio")) bc.ratio = b["ratio"].as_number();
    if (b.has("velocity")) bc.velocity = b["velocity"].as_number();
  }
  if (dim == 3) {
    mesh_ = std::make_shared<Mesh>(V, T, mat_ids, materials);
  } else {
    // per-element parameters of the object's material; materials[0] gives
    // the density (tri2d_mesh.cpp:36-41)
    const auto& mc = *materials[mat_ids.empty() ? 0 : mat_ids[0]]->config();
    std::vector<double> mu(nT, mc.mu), la(nT, mc.la);
    std::vector<int> mt(nT, (int)mc.material_model);
    mesh_ = std::make_shared<Mesh>(V, T, materials[0], mu, la, mt, 2);
  }
  mesh_->bc_config_ = bc;
  if (!v0.empty()) mesh_->initial_velocity_ = v0;
  if (args.has("initial_state")) {
    // DMAT start positions and velocities, nV x DIM (:263-308)
    const Json& is = args["initial_state"];
    int rows, cols;
    if (is.has("x_path")) {
      auto x = read_dmat(resolve(is["x_path"].as_string()), rows, cols);
      if (rows != nV || cols != dim) throw std::runtime_error("initial_state x_path has wrong shape");
      for (int i = 0; i < nV; ++i)
        for (int d = 0; d < dim; ++d) mesh_->Vinit_[3 * i + d] = mesh_->V_[3 * i + d] = x[i + nV * d];
    }
    if (is.has("v_path")) {
      auto v = read_dmat(resolve(is["v_path"].as_string()), rows, cols);
      if (rows != nV || cols != dim) throw std::runtime_error("initial_state v_path has wrong shape");
      for (int i = 0; i < nV; ++i)
        for (int d = 0; d < dim; ++d) mesh_->initial_velocity_[3 * i + d] = v[i + nV * d];
    }
  }
  MixedVariableFactory<DIM> mvf;
  VariableFactory<DIM> vf;
  config_->mixed_variables.clear();
  config_->variables.clear();
  if (args.has("mixed_variables"))
    for (const auto& n : args["mixed_variables"].as_strings())
      config_->mixed_variables.insert(mvf.type_by_name(n));
  if (args.has("variables"))
    for (const auto& n : args["variables"].as_strings()) config_->variables.insert(vf.type_by_name(n));
  if (args.has("linear_solver")) {
    LinearSolverFactory<DIM> lf;
    config_->solver_type = lf.type_by_name(args["linear_solver"].as_string());
  }
  fixed_override_.clear();
  ctx_.reset();
  finish_setup();
  return true;
}

// ============================================================ file formats
void read_mesh_file(const std::string& path, std::vector<double>& V, std::vector<int>& T) {
  std::ifstream f(path);
  if (!f.good()) throw std::runtime_error("Unable to open mesh: " + path);
  std::vector<std::string> tok;
  std::string line;
  while (std::getline(f, line)) {
    auto h = line.find('#');
    if (h != std::string::npos) line = line.substr(0, h);
    std::istringstream ss(line);
    std::string t;
    while (ss >> t) tok.push_back(t);
  }
  V.clear();
  T.clear();
  for (size_t i = 0; i < tok.size();) {
    if (tok[i] == "Vertices") {
      long n = std::stol(tok[i + 1]);
      for (long k = 0; k < n; ++k)
        for (int d = 0; d < 3; ++d) V.push_back(std::stod(tok[i + 2 + 4 * k + d]));
      i += 2 + 4 * n;
    } else if (tok[i] == "Tetrahedra") {
      long n = std::stol(tok[i + 1]);
      for (long k = 0; k < n; ++k)
        for (int d = 0; d < 4; ++d) T.push_back(std::stoi(tok[i + 2 + 5 * k + d]) - 1);
      i += 2 + 5 * n;
    } else if (tok[i] == "Triangles") {
      i += 2 + 4 * std::stol(tok[i + 1]);
    } else if (tok[i] == "Edges") {
      i += 2 + 3 * std::stol(tok[i + 1]);
    } else {
      ++i;
    }
  }
  if (V.empty() || T.empty()) throw std::runtime_error(path + ": missing Vertices or Tetrahedra");
}

std::vector<double> read_dmat(const std::string& path, int& rows, int& cols) {
  std::ifstream f(path);
  if (!f.good()) throw std::runtime_error("Unable to open dmat: " + path);
  f >> cols >> rows;
  std::vector<double> v((size_t)rows * cols);
  for (auto& x : v) f >> x;
  if (!f) throw std::runtime_error(path + ": short dmat");
  return v;  // column-major rows x cols
}

void read_obj_file(const std::string& path, std::vector<double>& V, std::vector<int>& T) {
  std::ifstream f(path);
  if (!f.good()) throw std::runtime_error("Unable to open mesh: " + path);
  V.clear();
  T.clear();
  std::string line;
  std::vector<int> poly;
  while (std::getline(f, line)) {
    std::istringstream ss(line);
    std::string key;
    if (!(ss >> key)) continue;
    if (key == "v") {
      double x = 0, y = 0, z = 0;
      ss >> x >> y;
      if (!(ss >> z)) z = 0;
      V.push_back(x);
      V.push_back(y);
      V.push_back(z);
    } else if (key == "f") {
      poly.clear();
      std::string tok;
      while (ss >> tok) {
        const int k = std::stoi(tok.substr(0, tok.find('/')));  // v, v/vt, v/vt/vn, v//vn
        const int nv = (int)(V.size() / 3);
        const int id = k > 0 ? k - 1 : nv + k;  // negative: relative to the last vertex
        if (k == 0 || id < 0) throw std::runtime_error(path + ": bad face index");
        poly.push_back(id);
      }
      for (size_t c = 1; c + 1 < poly.size(); ++c) {  // fan from the first corner
        T.push_back(poly[0]);
        T.push_back(poly[c]);
        T.push_back(poly[c + 1]);
      }
    }
  }
  const int nv = (int)(V.size() / 3);
  for (int v : T)
    if (v >= nv) throw std::runtime_error(path + ": face index past the vertex list");
  if (V.empty() || T.empty()) throw std::runtime_error(path + ": no vertices or faces");
}

void remove_unreferenced(std::vector<double>& V, std::vector<int>& T, int nodes) {
  (void)nodes;
  const int nv = (int)(V.size() / 3);
  std::vector<int> used(nv, 0), map(nv, -1);
  for (int v : T) used[v] = 1;
  std::vector<double> NV;
  for (int i = 0, k = 0; i < nv; ++i)
    if (used[i]) {
      map[i] = k++;
      NV.insert(NV.end(), V.begin() + 3 * i, V.begin() + 3 * i + 3);
    }
  for (int& v : T) v = map[v];
  V.swap(NV);
}

template class BDF<1>;
template class BDF<2>;
template class BDF<3>;
template class BDF<4>;
template class BDF<5>;
template class BDF<6>;
template class Displacement<3>;
template class Stretch<3>;
template class MixedStretch<3>;
template class VariableFactory<3>;
template class MixedVariableFactory<3>;
template class SystemMatrixPD<3>;
template class PcgSolver<3>;
template class DirectSolver<3>;
template class IcPcgSolver<3>;
template class MinresSolver<3>;
template class LinearSolverFactory<3>;
template class Optimizer<3>;
template class NewtonOptimizer<3>;
template class OptimizerFactory<3>;
template struct SimState<3>;

}  // namespace mfem
