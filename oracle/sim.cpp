// ORACLE — test infrastructure only. Not part of the product library.
//
// FP64 CPU restatement of one timestep of the reference (tytrusty/mfem):
//   mesh setup        reference src/mesh/mesh.cpp:44-95, tet_mesh.cpp:16-80,
//                     tri2d_mesh.cpp:15-77, utils/pinning_matrix.cpp:5-34
//   boundary groups   boundary_conditions/boundary_condition.cpp:5-21,
//                     fixed_boundary_conditions.h:75-166
//   Assembler         utils/sparse_utils.cpp:19-203 (free-free pairs sorted by
//                     global (row, col); duplicates summed in sorted order)
//   Displacement      variables/displacement.cpp:21-98
//   Stretch (plain)   variables/stretch.cpp:15-89
//   MixedStretch      variables/mixed_stretch.cpp:83-332
//   SystemMatrixPD    linear_solvers/linear_system.h:18-40
//   Eigen CG(Jacobi)  linear_solver_factory.cpp:144-154 (x0 = 0, stop at
//                     |r| < tol |b|, Jacobi preconditioner)
//   Newton + LS       optimizers/newton_optimizer.cpp:11-140, linesearch.h:78-131
//   BDF1..6           time_integrators/BDF.cpp:7-93, BDF.h:19-26
// Element-local loops run under OpenMP (as the reference's #pragma omp loops);
// every reduction here is serial and therefore deterministic.
#include <algorithm>
#include <array>
#include <cstdio>
#include <cstring>
#include <deque>
#include <numeric>
#include <queue>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"

using namespace orc;

namespace {

// Free-free vertex-pair scatter map (reference sparse_utils.cpp:19-138).
// Duplicate (row, col) entries are kept in (element, local j, local k) order,
// a total order the reference leaves to std::sort.
struct Assembler {
  int nodes = 0;  // vertices per element
  int dim = 0;
  std::vector<int> element_ids;
  std::vector<std::pair<int, int>> global_pairs, local_pairs;
  std::vector<int> multiplicity, row_offsets, offsets;
  // CSR of the scalar matrix (full DIMxDIM blocks)
  int n = 0;
  std::vector<int> rowptr, colind;
  void build(const std::vector<int>& T, int nT, int nodes_, int dim_,
             const std::vector<int>& free_map, int nfree) {
    nodes = nodes_; dim = dim_;
    std::vector<int> eids; std::vector<std::pair<int, int>> gp, lp;
    for (int i = 0; i < nT; ++i)
      for (int j = 0; j < nodes; ++j) {
        int vj = T[i * nodes + j];
        if (vj < 0 || free_map[vj] < 0) continue;
        for (int k = 0; k < nodes; ++k) {
          int vk = T[i * nodes + k];
          if (vk < 0 || free_map[vk] < 0) continue;
          eids.push_back(i);
          gp.push_back({free_map[vj], free_map[vk]});
          lp.push_back({j, k});
        }
      }
    std::vector<int> ids(eids.size());
    std::iota(ids.begin(), ids.end(), 0);
    std::stable_sort(ids.begin(), ids.end(), [&](int a, int b) {
      return gp[a] < gp[b];
    });
    element_ids.resize(ids.size()); global_pairs.resize(ids.size());
    local_pairs.resize(ids.size());
    for (size_t i = 0; i < ids.size(); ++i) {
      element_ids[i] = eids[ids[i]];
      global_pairs[i] = gp[ids[i]];
      local_pairs[i] = lp[ids[i]];
    }
    multiplicity.assign(ids.size(), 1);
    row_offsets.clear(); offsets.clear();
    size_t i = 0; int curr_row = -1;
    while (i < multiplicity.size()) {
      auto cp = global_pairs[i];
      if (cp.first != curr_row) { row_offsets.push_back((int)i); curr_row = cp.first; }
      while (i + multiplicity[i] < multiplicity.size() &&
             global_pairs[i + multiplicity[i]] == cp) ++multiplicity[i];
      offsets.push_back((int)i);
      i += multiplicity[i];
    }
    row_offsets.push_back((int)multiplicity.size());
    // scalar CSR pattern: row (dim*br + r) lists blocks in column order
    n = dim * nfree;
    rowptr.assign(n + 1, 0);
    int nb_rows = (int)row_offsets.size() - 1;
    std::vector<std::vector<int>> bcols(nfree);
    for (int ii = 0; ii < nb_rows; ++ii) {
      int beg = row_offsets[ii], end = row_offsets[ii + 1];
      int br = global_pairs[beg].first;
      while (beg < end) {
        bcols[br].push_back(global_pairs[beg].second);
        beg += multiplicity[beg];
      }
    }
    for (int br = 0; br < nfree; ++br)
      for (int r = 0; r < dim; ++r)
        rowptr[dim * br + r + 1] = (int)bcols[br].size() * dim;
    for (int k = 0; k < n; ++k) rowptr[k + 1] += rowptr[k];
    colind.assign(rowptr[n], 0);
    for (int br = 0; br < nfree; ++br)
      for (int r = 0; r < dim; ++r) {
        int p = rowptr[dim * br + r];
        for (int bc : bcols[br])
          for (int c = 0; c < dim; ++c) colind[p++] = dim * bc + c;
      }
  }
  // blocks: per element, a (nodes*dim)^2 column-major matrix
  // (reference sparse_utils.cpp:141-203)
  void update_matrix(const std::vector<double>& blocks, std::vector<double>& vals) const {
    int M = nodes * dim;
    vals.assign(rowptr[n], 0.0);
    int nb_rows = (int)row_offsets.size() - 1;
#pragma omp parallel for schedule(static)
    for (int ii = 0; ii < nb_rows; ++ii) {
      int beg = row_offsets[ii], end = row_offsets[ii + 1];
      int block_col = 0;
      while (beg < end) {
        int nm = multiplicity[beg];
        auto g = global_pairs[beg];
        double lb[9] = {0};
        for (int q = 0; q < nm; ++q) {
          int e = element_ids[beg + q];
          auto l = local_pairs[beg + q];
          const double* B = &blocks[(size_t)e * M * M];
          for (int r = 0; r < dim; ++r)
            for (int c = 0; c < dim; ++c)
              lb[r * 3 + c] += B[(dim * l.first + r) + M * (dim * l.second + c)];
        }
        for (int r = 0; r < dim; ++r) {
          int start = rowptr[dim * g.first + r] + dim * block_col;
          for (int c = 0; c < dim; ++c) vals[start + c] = lb[r * 3 + c];
        }
        beg += nm;
        ++block_col;
      }
    }
  }
};

double dotv(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

}  // namespace

struct OracleSim {
  oracle_params p;
  int dim, nodes, nV, nT;
  std::vector<double> Vref, Vcur;
  std::vector<int> T, mat, fixed, free_map;
  std::vector<double> mu, la;
  int nfree = 0, n = 0;
  double h = 0;  // integrator dt() = beta * h
  // mesh quantities
  std::vector<double> vols, dX;  // dX[e][k][c], nodes x dim per element
  Assembler asmb;
  std::vector<double> PMP;  // values on asmb pattern
  // displacement
  std::vector<double> x, b, dx, f_ext, grad_x_disp;
  std::deque<std::vector<double>> x_prevs, v_prevs;
  std::vector<double> alphas;
  double beta = 1;
  // mixed stretch (3D)
  std::vector<M3> R;
  std::vector<Vec<6>> S, g, gl;
  std::vector<double> s, lam, ds;
  std::vector<Mat<9, 6>> dSdF;
  std::vector<Mat<6, 6>> Hm, Hinv, Hloc;
  std::vector<double> Aloc;  // per element (nodes*dim)^2
  std::vector<double> grad_x_mixed, grad_s;
  // mixed stretch (2D): 3-vector s, 2x2 R, 4x3 dSdF, 3x3 Hessians
  std::vector<Mat<2, 2>> R2;
  std::vector<Vec<3>> S3, g3, gl3;
  std::vector<Mat<4, 3>> dSdF2;
  std::vector<Mat<3, 3>> Hm3, Hinv3, Hloc3;
  // plain stretch
  std::vector<double> Hp, gp, grad_plain;
  // system
  std::vector<double> lhs, rhs;
  oracle_stats st{};
  int pcg_iters_last = 0;
  double pcg_err_last = 0;
  // per Newton iteration of the last step: iteration, energy, energy res,
  // Newton dec, alpha (the reference's OptimizerData rows)
  std::vector<std::array<double, 5>> hist;
  double ic_shift_last = 0;  // "eigen-pcg-IC": diagonal shift and colours of the last factor
  int ic_colors_last = 0;
  // indefinite (KKT) solve: preconditioner state and the reference's
  // refresh counter (preconditioners.h:150-158)
  // boundary-condition script: end groups, twist center, bend centers
  std::vector<int> bc_g[2];
  double bc_center[3] = {0, 0, 0};
  double bc_gc[2][3] = {{0, 0, 0}, {0, 0, 0}};
  int kkt_solves = 0;
  bool kkt_mass_ready = false;
  std::vector<double> mlump, mdinv;   // per free dof: row sum and 1/diag of PMP
  std::vector<double> mchol;          // dense Cholesky of PMP (exact preconditioner)
  std::vector<double> lchol;          // dense Cholesky of L (exact preconditioner)
  std::vector<double> linv_blk;       // element blocks of L, inverted (ns x ns row-major each)
};

namespace {

int mat_dim(const OracleSim* S_) { return S_->dim * S_->dim; }

// x_full = P^T x + b  (reference displacement.h:49-53)
void unproject(const OracleSim* S_, const std::vector<double>& xf, std::vector<double>& full) {
  full = S_->b;
  for (int v = 0; v < S_->nV; ++v) {
    int f = S_->free_map[v];
    if (f < 0) continue;
    for (int d = 0; d < S_->dim; ++d) full[S_->dim * v + d] += xf[S_->dim * f + d];
  }
}

// F_e = Jloc_e q_e  (column-major vec; reference tet_mesh.cpp:16-27)
void def_grad(const OracleSim* S_, const std::vector<double>& full, int e, double* F) {
  int D = S_->dim;
  for (int i = 0; i < D * D; ++i) F[i] = 0;
  for (int k = 0; k < S_->nodes; ++k) {
    int v = S_->T[e * S_->nodes + k];
    const double* dx = &S_->dX[((size_t)e * S_->nodes + k) * D];
    for (int c = 0; c < D; ++c)
      for (int r = 0; r < D; ++r) F[r + D * c] += full[D * v + r] * dx[c];
  }
}

// out += sign * PJW * tmp, tmp holding one DxD column-major vector per
// element (reference mesh.cpp:63: PJW = P J^T W).
void apply_PJW(const OracleSim* S_, const std::vector<double>& tmp, std::vector<double>& out, double sign) {
  int D = S_->dim, M = D * D;
  std::vector<double> full(S_->nV * D, 0.0);
  for (int e = 0; e < S_->nT; ++e) {
    double vol = S_->vols[e];
    for (int k = 0; k < S_->nodes; ++k) {
      int v = S_->T[e * S_->nodes + k];
      const double* dx = &S_->dX[((size_t)e * S_->nodes + k) * D];
      for (int r = 0; r < D; ++r) {
        double a = 0;
        for (int c = 0; c < D; ++c) a += dx[c] * tmp[(size_t)M * e + r + D * c];
        full[D * v + r] += vol * a;
      }
    }
  }
  for (int v = 0; v < S_->nV; ++v) {
    int f = S_->free_map[v];
    if (f < 0) continue;
    for (int r = 0; r < D; ++r) out[D * f + r] += sign * full[D * v + r];
  }
}

// full mass product  y = M d  (consistent linear-element mass)
void mass_mult_full(const OracleSim* S_, const std::vector<double>& d, std::vector<double>& y) {
  int D = S_->dim, N = S_->nodes;
  y.assign(S_->nV * D, 0.0);
  double denom = (D == 3) ? 20.0 : 12.0;
  for (int e = 0; e < S_->nT; ++e) {
    double m = S_->p.density * S_->vols[e] / denom;
    for (int a = 0; a < N; ++a) {
      int va = S_->T[e * N + a];
      for (int bb = 0; bb < N; ++bb) {
        int vb = S_->T[e * N + bb];
        double c = m * (a == bb ? 2.0 : 1.0);
        for (int r = 0; r < D; ++r) y[D * va + r] += c * d[D * vb + r];
      }
    }
  }
}

std::vector<double> x_tilde(const OracleSim* S_) {
  size_t N = S_->x_prevs.front().size();
  std::vector<double> wx(N, 0.0), wv(N, 0.0);
  for (size_t i = 0; i < S_->x_prevs.size(); ++i)
    for (size_t k = 0; k < N; ++k) {
      wx[k] += S_->alphas[i] * S_->x_prevs[i][k];
      wv[k] += S_->alphas[i] * S_->v_prevs[i][k];
    }
  for (size_t k = 0; k < N; ++k) wx[k] += S_->h * wv[k];
  return wx;
}

// d = P^T x + b - x_tilde - h^2 f_ext  (reference displacement.cpp:21-31,64-73)
std::vector<double> inertia_diff(const OracleSim* S_, const std::vector<double>& xf) {
  std::vector<double> full;
  unproject(S_, xf, full);
  std::vector<double> xt = x_tilde(S_);
  double h2 = S_->h * S_->h;
  for (size_t k = 0; k < full.size(); ++k) full[k] = full[k] - xt[k] - h2 * S_->f_ext[k];
  return full;
}

double disp_energy(const OracleSim* S_, const std::vector<double>& xf) {
  std::vector<double> d = inertia_diff(S_, xf), Md;
  mass_mult_full(S_, d, Md);
  return 0.5 * dotv(d, Md);
}

void disp_gradient(OracleSim* S_) {
  std::vector<double> d = inertia_diff(S_, S_->x), Md;
  mass_mult_full(S_, d, Md);
  S_->grad_x_disp.assign(S_->n, 0.0);
  for (int v = 0; v < S_->nV; ++v) {
    int f = S_->free_map[v];
    if (f < 0) continue;
    for (int r = 0; r < S_->dim; ++r) S_->grad_x_disp[S_->dim * f + r] = Md[S_->dim * v + r];
  }
}

const double kSym[6] = {1, 1, 1, 2, 2, 2};
const double kSyminv[6] = {1, 1, 1, 0.5, 0.5, 0.5};

// ---------------- MixedStretch (3D) ----------------
void mixed_update_rotations(OracleSim* S_, const std::vector<double>& full) {
#pragma omp parallel for schedule(static)
  for (int e = 0; e < S_->nT; ++e) {
    double f[9]; def_grad(S_, full, e, f);
    M3 F; for (int i = 0; i < 9; ++i) F[i] = f[i];
    Mat<6, 9> dsdF;
    polar_svd3(S_->R[e], S_->S[e], F, true, &dsdF, nullptr);
    S_->dSdF[e] = dsdF.T() * diag<6>(kSym);
  }
}

void mixed_update_derivatives(OracleSim* S_) {
  double h2 = S_->h * S_->h;
  int M = 12;
#pragma omp parallel for schedule(static)
  for (int e = 0; e < S_->nT; ++e) {
    double vol = S_->vols[e];
    Vec<6> si; for (int i = 0; i < 6; ++i) si[i] = S_->s[6 * e + i];
    Vec<6> gg; Mat<6, 6> HH;
    material_s6(S_->mat[e], S_->mu[e], S_->la[e], si, true, nullptr, &gg, &HH);
    Mat<6, 6> H = h2 * HH;
    S_->Hinv[e] = lu_inverse<6>(H);
    S_->g[e] = h2 * gg;
    S_->Hm[e] = (1.0 / vol) * (diag<6>(kSyminv) * H * diag<6>(kSyminv));
    S_->Hloc[e] = vol * H;
  }
#pragma omp parallel for schedule(static)
  for (int e = 0; e < S_->nT; ++e) {
    double vol = S_->vols[e];
    Mat<9, 9> inner = S_->dSdF[e] * S_->Hm[e] * S_->dSdF[e].T();
    // Jloc^T inner Jloc, Jloc(r + 3c, 3k + r) = dX(k, c)
    double* A = &S_->Aloc[(size_t)e * M * M];
    const double* dx = &S_->dX[(size_t)e * 12];
    for (int k = 0; k < 4; ++k)
      for (int r = 0; r < 3; ++r)
        for (int l = 0; l < 4; ++l)
          for (int q = 0; q < 3; ++q) {
            double a = 0;
            for (int c = 0; c < 3; ++c)
              for (int c2 = 0; c2 < 3; ++c2)
                a += dx[3 * k + c] * inner(r + 3 * c, q + 3 * c2) * dx[3 * l + c2];
            A[(3 * k + r) + M * (3 * l + q)] = a * vol * vol;
          }
  }
  S_->asmb.update_matrix(S_->Aloc, S_->lhs);  // temporarily holds A
  // grad_x = -PJW (dSdF la)
  std::vector<double> tmp(9 * S_->nT);
  for (int e = 0; e < S_->nT; ++e) {
    Vec<6> l; for (int i = 0; i < 6; ++i) l[i] = S_->lam[6 * e + i];
    Vec<9> t = S_->dSdF[e] * l;
    for (int i = 0; i < 9; ++i) tmp[9 * e + i] = t[i];
  }
  S_->grad_x_mixed.assign(S_->n, 0.0);
  apply_PJW(S_, tmp, S_->grad_x_mixed, -1.0);
  S_->grad_s.resize(6 * S_->nT);
  for (int e = 0; e < S_->nT; ++e)
    for (int i = 0; i < 6; ++i) S_->grad_s[6 * e + i] = S_->vols[e] * S_->g[e][i];
}

// rhs of the mixed variable (reference mixed_stretch.cpp:207-226)
std::vector<double> mixed_rhs(OracleSim* S_) {
  std::vector<double> out(S_->n, 0.0), tmp(9 * S_->nT);
  for (int e = 0; e < S_->nT; ++e) {
    double vol = S_->vols[e];
    Vec<6> d; for (int i = 0; i < 6; ++i) d[i] = S_->S[e][i] - S_->s[6 * e + i];
    S_->gl[e] = vol * (S_->Hm[e] * (diag<6>(kSym) * d)) + diag<6>(kSyminv) * S_->g[e];
    Vec<9> t = S_->dSdF[e] * S_->gl[e];
    for (int i = 0; i < 9; ++i) tmp[9 * e + i] = t[i];
  }
  apply_PJW(S_, tmp, out, -1.0);
  return out;
}

// local solve (reference mixed_stretch.cpp:244-259)
void mixed_solve(OracleSim* S_, const std::vector<double>& dxv) {
  // Jdx = -PJW^T dx
  std::vector<double> full(S_->nV * 3, 0.0);
  for (int v = 0; v < S_->nV; ++v) {
    int f = S_->free_map[v];
    if (f < 0) continue;
    for (int r = 0; r < 3; ++r) full[3 * v + r] = dxv[3 * f + r];
  }
#pragma omp parallel for schedule(static)
  for (int e = 0; e < S_->nT; ++e) {
    double Fd[9]; def_grad(S_, full, e, Fd);
    Vec<9> jdx; for (int i = 0; i < 9; ++i) jdx[i] = -S_->vols[e] * Fd[i];
    Vec<6> l = -S_->gl[e] + S_->Hm[e] * (S_->dSdF[e].T() * jdx);
    Vec<6> ls = diag<6>(kSym) * l + S_->g[e];
    Vec<6> d = -(S_->Hinv[e] * ls);
    for (int i = 0; i < 6; ++i) { S_->lam[6 * e + i] = l[i]; S_->ds[6 * e + i] = d[i]; }
  }
}

double mixed_energy(const OracleSim* S_, const std::vector<double>& sv) {
  double e = 0;
  for (int i = 0; i < S_->nT; ++i) {
    Vec<6> si; for (int k = 0; k < 6; ++k) si[k] = sv[6 * i + k];
    double E;
    material_s6(S_->mat[i], S_->mu[i], S_->la[i], si, false, &E, nullptr, nullptr);
    e += E * S_->vols[i];
  }
  return e;
}

double mixed_constraint_value(const OracleSim* S_, const std::vector<double>& full,
                              const std::vector<double>& sv) {
  std::vector<double> per(S_->nT);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < S_->nT; ++i) {
    double f[9]; def_grad(S_, full, i, f);
    M3 F; for (int k = 0; k < 9; ++k) F[k] = f[k];
    M3 Rc = S_->R[i];
    Vec<6> stmp;
    polar_svd3(Rc, stmp, F, false, nullptr, nullptr);
    double el = 0;
    for (int k = 0; k < 6; ++k)
      el += S_->lam[6 * i + k] * kSym[k] * (stmp[k] - sv[6 * i + k]);
    per[i] = el * S_->vols[i];
  }
  double e = 0;
  for (int i = 0; i < S_->nT; ++i) e += per[i];
  return e;
}

// ---------------- MixedStretch (2D) ----------------
// The DIM = 2 instantiation of mixed_stretch.cpp:134-332: s = (S00, S11, S01),
// Sym = diag(1, 1, 2), dSdF = dsdF^T Sym (4 x 3), the same element formulas as
// the 3D path with 3x3 material Hessians and 6x6 triangle blocks.
const double kSym3[3] = {1, 1, 2};
const double kSyminv3[3] = {1, 1, 0.5};

Mat<2, 2> def_grad2(const OracleSim* S_, const std::vector<double>& full, int e) {
  double f[4]; def_grad(S_, full, e, f);
  Mat<2, 2> F; for (int i = 0; i < 4; ++i) F[i] = f[i];
  return F;
}

void mixed2_update_rotations(OracleSim* S_, const std::vector<double>& full) {
#pragma omp parallel for schedule(static)
  for (int e = 0; e < S_->nT; ++e) {
    Mat<3, 4> dsdF;
    polar2d(S_->R2[e], S_->S3[e], def_grad2(S_, full, e), true, &dsdF);
    S_->dSdF2[e] = dsdF.T() * diag<3>(kSym3);
  }
}

void mixed2_update_derivatives(OracleSim* S_) {
  const double h2 = S_->h * S_->h;
  const int M = 6;
#pragma omp parallel for schedule(static)
  for (int e = 0; e < S_->nT; ++e) {
    const double vol = S_->vols[e];
    Vec<3> si; for (int i = 0; i < 3; ++i) si[i] = S_->s[3 * e + i];
    Vec<3> gg; Mat<3, 3> HH;
    material_s3(S_->mat[e], S_->mu[e], S_->la[e], si, true, nullptr, &gg, &HH);
    Mat<3, 3> H = h2 * HH;
    S_->Hinv3[e] = lu_inverse<3>(H);
    S_->g3[e] = h2 * gg;
    S_->Hm3[e] = (1.0 / vol) * (diag<3>(kSyminv3) * H * diag<3>(kSyminv3));
    S_->Hloc3[e] = vol * H;
    Mat<4, 4> inner = S_->dSdF2[e] * S_->Hm3[e] * S_->dSdF2[e].T();
    double* A = &S_->Aloc[(size_t)e * M * M];
    const double* dx = &S_->dX[(size_t)e * 6];
    for (int k = 0; k < 3; ++k)
      for (int r = 0; r < 2; ++r)
        for (int l = 0; l < 3; ++l)
          for (int q = 0; q < 2; ++q) {
            double a = 0;
            for (int c = 0; c < 2; ++c)
              for (int c2 = 0; c2 < 2; ++c2)
                a += dx[2 * k + c] * inner(r + 2 * c, q + 2 * c2) * dx[2 * l + c2];
            A[(2 * k + r) + M * (2 * l + q)] = a * vol * vol;
          }
  }
  S_->asmb.update_matrix(S_->Aloc, S_->lhs);
}

std::vector<double> mixed2_rhs(OracleSim* S_) {
  std::vector<double> out(S_->n, 0.0), tmp(4 * (size_t)S_->nT);
  for (int e = 0; e < S_->nT; ++e) {
    const double vol = S_->vols[e];
    Vec<3> d; for (int i = 0; i < 3; ++i) d[i] = S_->S3[e][i] - S_->s[3 * e + i];
    S_->gl3[e] = vol * (S_->Hm3[e] * (diag<3>(kSym3) * d)) + diag<3>(kSyminv3) * S_->g3[e];
    Vec<4> t = S_->dSdF2[e] * S_->gl3[e];
    for (int i = 0; i < 4; ++i) tmp[4 * e + i] = t[i];
  }
  apply_PJW(S_, tmp, out, -1.0);
  return out;
}

void mixed2_solve(OracleSim* S_, const std::vector<double>& dxv) {
  std::vector<double> full(S_->nV * 2, 0.0);
  for (int v = 0; v < S_->nV; ++v) {
    const int f = S_->free_map[v];
    if (f < 0) continue;
    for (int r = 0; r < 2; ++r) full[2 * v + r] = dxv[2 * f + r];
  }
#pragma omp parallel for schedule(static)
  for (int e = 0; e < S_->nT; ++e) {
    Mat<2, 2> Fd = def_grad2(S_, full, e);
    Vec<4> jdx; for (int i = 0; i < 4; ++i) jdx[i] = -S_->vols[e] * Fd[i];
    Vec<3> l = -S_->gl3[e] + S_->Hm3[e] * (S_->dSdF2[e].T() * jdx);
    Vec<3> ls = diag<3>(kSym3) * l + S_->g3[e];
    Vec<3> d = -(S_->Hinv3[e] * ls);
    for (int i = 0; i < 3; ++i) { S_->lam[3 * e + i] = l[i]; S_->ds[3 * e + i] = d[i]; }
  }
}

double mixed2_energy(const OracleSim* S_, const std::vector<double>& sv) {
  double e = 0;
  for (int i = 0; i < S_->nT; ++i) {
    Vec<3> si; for (int k = 0; k < 3; ++k) si[k] = sv[3 * i + k];
    double E;
    material_s3(S_->mat[i], S_->mu[i], S_->la[i], si, false, &E, nullptr, nullptr);
    e += E * S_->vols[i];
  }
  return e;
}

double mixed2_constraint_value(const OracleSim* S_, const std::vector<double>& full,
                               const std::vector<double>& sv) {
  std::vector<double> per(S_->nT);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < S_->nT; ++i) {
    Mat<2, 2> R; Vec<3> st;
    polar2d(R, st, def_grad2(S_, full, i), false, nullptr);
    double el = 0;
    for (int k = 0; k < 3; ++k) el += S_->lam[3 * i + k] * kSym3[k] * (st[k] - sv[3 * i + k]);
    per[i] = el * S_->vols[i];
  }
  double e = 0;
  for (int i = 0; i < S_->nT; ++i) e += per[i];
  return e;
}

// ---------------- Stretch (plain) ----------------
void plain_update(OracleSim* S_, const std::vector<double>& full) {
  int D = S_->dim, N = S_->nodes, M = D * N;
  double h2 = S_->h * S_->h;
#pragma omp parallel for schedule(static)
  for (int e = 0; e < S_->nT; ++e) {
    double f[9]; def_grad(S_, full, e, f);
    double vol = S_->vols[e];
    double Hf[81], gf[9];
    if (D == 3) {
      Vec<9> fv; for (int i = 0; i < 9; ++i) fv[i] = f[i];
      Vec<9> gg; Mat<9, 9> HH;
      material_f9(S_->mat[e], S_->mu[e], S_->la[e], fv, nullptr, &gg, &HH);
      for (int i = 0; i < 81; ++i) Hf[i] = HH[i];
      for (int i = 0; i < 9; ++i) gf[i] = gg[i];
    } else {
      Vec<4> fv; for (int i = 0; i < 4; ++i) fv[i] = f[i];
      Vec<4> gg; Mat<4, 4> HH;
      snh_f4(S_->mu[e], S_->la[e], fv, nullptr, &gg, &HH);
      for (int i = 0; i < 16; ++i) Hf[i] = HH[i];
      for (int i = 0; i < 4; ++i) gf[i] = gg[i];
    }
    const double* dx = &S_->dX[(size_t)e * N * D];
    double* A = &S_->Hp[(size_t)e * M * M];
    int MM = D * D;
    for (int k = 0; k < N; ++k)
      for (int r = 0; r < D; ++r) {
        for (int l = 0; l < N; ++l)
          for (int q = 0; q < D; ++q) {
            double a = 0;
            for (int c = 0; c < D; ++c)
              for (int c2 = 0; c2 < D; ++c2)
                a += dx[D * k + c] * Hf[(r + D * c) + MM * (q + D * c2)] * dx[D * l + c2];
            A[(D * k + r) + M * (D * l + q)] = a * vol * h2;
          }
        double a = 0;
        for (int c = 0; c < D; ++c) a += dx[D * k + c] * gf[r + D * c];
        S_->gp[(size_t)e * M + D * k + r] = a * vol * h2;
      }
  }
  // VecAssembler::assemble (reference sparse_utils.cpp:285-323)
  S_->grad_plain.assign(S_->n, 0.0);
  std::vector<double> full_g(S_->nV * D, 0.0);
  for (int e = 0; e < S_->nT; ++e)
    for (int k = 0; k < N; ++k) {
      int v = S_->T[e * N + k];
      for (int r = 0; r < D; ++r) full_g[D * v + r] += S_->gp[(size_t)e * M + D * k + r];
    }
  for (int v = 0; v < S_->nV; ++v) {
    int f = S_->free_map[v];
    if (f < 0) continue;
    for (int r = 0; r < D; ++r) S_->grad_plain[D * f + r] = full_g[D * v + r];
  }
}

double plain_energy(const OracleSim* S_, const std::vector<double>& full) {
  std::vector<double> per(S_->nT);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < S_->nT; ++e) {
    double f[9]; def_grad(S_, full, e, f);
    double E;
    if (S_->dim == 3) {
      Vec<9> fv; for (int i = 0; i < 9; ++i) fv[i] = f[i];
      material_f9(S_->mat[e], S_->mu[e], S_->la[e], fv, &E, nullptr, nullptr);
    } else {
      Vec<4> fv; for (int i = 0; i < 4; ++i) fv[i] = f[i];
      snh_f4(S_->mu[e], S_->la[e], fv, &E, nullptr, nullptr);
    }
    per[e] = E * S_->vols[e];
  }
  double s = 0;
  for (int e = 0; e < S_->nT; ++e) s += per[e];
  return s;
}

// ---------------- linear solvers ----------------
void csr_mult(const OracleSim* S_, const std::vector<double>& vals, const std::vector<double>& xv,
              std::vector<double>& y) {
  const auto& rp = S_->asmb.rowptr;
  const auto& ci = S_->asmb.colind;
  y.assign(S_->n, 0.0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < S_->n; ++i) {
    double a = 0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) a += vals[k] * xv[ci[k]];
    y[i] = a;
  }
}

// Conjugate gradient with Jacobi preconditioner and Eigen's stopping rule.
int pcg(OracleSim* S_, std::vector<double>& xo) {
  int n = S_->n;
  const auto& A = S_->lhs;
  const auto& bb = S_->rhs;
  xo.assign(n, 0.0);
  std::vector<double> invd(n, 1.0);
  for (int i = 0; i < n; ++i)
    for (int k = S_->asmb.rowptr[i]; k < S_->asmb.rowptr[i + 1]; ++k)
      if (S_->asmb.colind[k] == i && A[k] != 0.0) invd[i] = 1.0 / A[k];
  double rhsNorm2 = dotv(bb, bb);
  if (rhsNorm2 == 0) { S_->pcg_iters_last = 0; S_->pcg_err_last = 0; return 0; }
  double tol = S_->p.itr_tol;
  double threshold = std::max(tol * tol * rhsNorm2, 2.2250738585072014e-308);
  std::vector<double> r = bb, z(n), p(n), tmp(n);
  double rn2 = dotv(r, r);
  if (rn2 < threshold) {
    S_->pcg_iters_last = 0; S_->pcg_err_last = std::sqrt(rn2 / rhsNorm2); return 0;
  }
  for (int i = 0; i < n; ++i) p[i] = invd[i] * r[i];
  double absNew = dotv(r, p);
  int it = 0;
  while (it < S_->p.max_iters) {
    csr_mult(S_, A, p, tmp);
    double alpha = absNew / dotv(p, tmp);
    for (int i = 0; i < n; ++i) { xo[i] += alpha * p[i]; r[i] -= alpha * tmp[i]; }
    rn2 = dotv(r, r);
    if (rn2 < threshold) break;
    for (int i = 0; i < n; ++i) z[i] = invd[i] * r[i];
    double absOld = absNew;
    absNew = dotv(r, z);
    double beta = absNew / absOld;
    for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    ++it;
  }
  S_->pcg_iters_last = it;
  S_->pcg_err_last = std::sqrt(rn2 / rhsNorm2);
  return 0;
}

// ---------------- "eigen-pcg-IC" ----------------
// The reference runs Eigen 3.4's ConjugateGradient with IncompleteCholesky
// (linear_solver_factory.cpp:87-97; Eigen is not vendored in the reference).
// That preconditioner (Eigen/src/IterativeLinearSolvers/IncompleteCholesky.h,
// after Lin & More, SIAM J. Sci. Comput. 21(1), 1999) orders the matrix by
// AMD, scales it symmetrically by 1/sqrt(|column|_2), shifts the diagonal by
// 1e-3 - min diag when that is not positive, factors with as many kept
// entries per column as the column had, and on a non-positive pivot restarts
// with the shift max(1e-3, 2 shift), at most 10 attempts. The device keeps the
// scaling and the shift loop and replaces the ordering and fill rule with
// IC(0) on the 3x3 vertex blocks in smallest-last multicolour order (vertices of one
// colour share no matrix entry, so a colour is one parallel step of the
// factorization and of both triangular solves). This is that algorithm, in
// scalar form, as the parity reference for the device.
struct IcFactor {
  int n = 0, bs = 1, nb = 0, ncolors = 0, attempts = 0;
  std::vector<int> color, order, pos;  // vertex colour; vertex order; vertex -> position
  std::vector<double> scale;           // n
  double shift = 0;
  // L by rows in the permuted scalar numbering p = pos[v] bs + d: strictly
  // lower entries (sorted) in lptr/lcol/lval, the diagonal in ldiag
  std::vector<int> lptr, lcol;
  std::vector<double> lval, ldiag;
};

// Colouring of the vertex graph of a block pattern: the smallest-last
// (degeneracy) order -- repeatedly remove a vertex of least remaining degree,
// the lowest index among ties -- then greedy first fit in the reverse of the
// removal order (Matula and Beck, J. ACM 30(3), 1983). On the Kuhn beams it
// gives 4 colours (first fit in vertex order gives 8). Returns the number of
// colours.
int ic_color(int nb, const std::vector<int>& browptr, const std::vector<int>& bcol,
             std::vector<int>& color) {
  std::vector<int> deg(nb, 0), order;
  std::vector<char> removed(nb, 0);
  for (int v = 0; v < nb; ++v)
    for (int k = browptr[v]; k < browptr[v + 1]; ++k) deg[v] += bcol[k] != v;
  std::priority_queue<std::pair<int, int>, std::vector<std::pair<int, int>>,
                      std::greater<std::pair<int, int>>> heap;
  for (int v = 0; v < nb; ++v) heap.push({deg[v], v});
  while (!heap.empty()) {
    const auto [d, v] = heap.top();
    heap.pop();
    if (removed[v] || d != deg[v]) continue;  // stale entry
    removed[v] = 1;
    order.push_back(v);
    for (int k = browptr[v]; k < browptr[v + 1]; ++k) {
      const int u = bcol[k];
      if (u != v && !removed[u]) heap.push({--deg[u], u});
    }
  }
  color.assign(nb, -1);
  std::vector<int> mark;  // mark[c] == v: colour c taken by a neighbour of v
  int ncol = 0;
  for (auto it = order.rbegin(); it != order.rend(); ++it) {
    const int v = *it;
    for (int k = browptr[v]; k < browptr[v + 1]; ++k) {
      const int u = bcol[k];
      if (u != v && color[u] >= 0) {
        if ((int)mark.size() <= color[u]) mark.resize(color[u] + 1, -1);
        mark[color[u]] = v;
      }
    }
    int c = 0;
    while (c < (int)mark.size() && mark[c] == v) ++c;
    color[v] = c;
    ncol = std::max(ncol, c + 1);
  }
  return ncol;
}

// Block pattern of a scalar CSR whose rows come in groups of bs (every
// block dense in the pattern, as the assembled K is).
void ic_block_pattern(int n, int bs, const int* rowptr, const int* colind, std::vector<int>& bp,
                      std::vector<int>& bc) {
  const int nb = n / bs;
  bp.assign(nb + 1, 0);
  bc.clear();
  for (int v = 0; v < nb; ++v) {
    const int r = v * bs;
    for (int k = rowptr[r]; k < rowptr[r + 1]; k += bs) bc.push_back(colind[k] / bs);
    bp[v + 1] = (int)bc.size();
  }
}

// One IC(0) attempt of the scaled, shifted matrix; false on a non-positive
// pivot. Up-looking: row p of L from the rows before it, within the pattern.
bool ic_attempt(IcFactor& f, const int* rowptr, const int* colind, const double* vals) {
  const int n = f.n, bs = f.bs;
  std::vector<double> w(n, 0.0);
  std::vector<char> in(n, 0);
  std::vector<int> cols;
  f.lptr.assign(n + 1, 0);
  f.lcol.clear();
  f.lval.clear();
  f.ldiag.assign(n, 0.0);
  for (int p = 0; p < n; ++p) {
    const int i = f.order[p / bs] * bs + p % bs;  // original scalar row
    double diag = 0;
    cols.clear();
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) {
      const int j = colind[k];
      const int q = f.pos[j / bs] * bs + j % bs;
      const double a = vals[k] * f.scale[i] * f.scale[j];
      if (q == p) diag = a + f.shift;
      else if (q < p) { w[q] = a; in[q] = 1; cols.push_back(q); }
    }
    std::sort(cols.begin(), cols.end());
    for (int q : cols) {
      double s = w[q];
      for (int k = f.lptr[q]; k < f.lptr[q + 1]; ++k)
        if (in[f.lcol[k]]) s -= f.lval[k] * w[f.lcol[k]];
      w[q] = s / f.ldiag[q];
    }
    for (int q : cols) diag -= w[q] * w[q];
    if (!(diag > 0)) {
      for (int q : cols) { w[q] = 0; in[q] = 0; }
      return false;
    }
    f.ldiag[p] = std::sqrt(diag);
    for (int q : cols) {
      f.lcol.push_back(q);
      f.lval.push_back(w[q]);
      w[q] = 0;
      in[q] = 0;
    }
    f.lptr[p + 1] = (int)f.lcol.size();
  }
  return true;
}

// Colouring, ordering, scaling and the shift loop. Returns 0, or -5 when ten
// attempts break down (Eigen then reports NumericalIssue).
int ic_factor(IcFactor& f, int n, int bs, const int* rowptr, const int* colind,
              const double* vals) {
  f.n = n;
  f.bs = bs;
  f.nb = n / bs;
  std::vector<int> bp, bc;
  ic_block_pattern(n, bs, rowptr, colind, bp, bc);
  f.ncolors = ic_color(f.nb, bp, bc, f.color);
  f.order.clear();
  for (int c = 0; c < f.ncolors; ++c)
    for (int v = 0; v < f.nb; ++v)
      if (f.color[v] == c) f.order.push_back(v);
  f.pos.assign(f.nb, 0);
  for (int p = 0; p < f.nb; ++p) f.pos[f.order[p]] = p;
  // scale_j = |column j|_2^(-1/2) (1 when the column is zero)
  f.scale.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double s = 0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) s += vals[k] * vals[k];
    s = std::sqrt(std::sqrt(s));
    f.scale[i] = s > 2.2250738585072014e-308 ? 1.0 / s : 1.0;
  }
  double mindiag = 1.79769313486231570e308;
  for (int i = 0; i < n; ++i)
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k)
      if (colind[k] == i) mindiag = std::min(mindiag, vals[k] * f.scale[i] * f.scale[i]);
  f.shift = mindiag <= 0 ? 1e-3 - mindiag : 0.0;
  for (f.attempts = 1;; ++f.attempts) {
    if (ic_attempt(f, rowptr, colind, vals)) return 0;
    if (f.attempts >= 10) return -5;
    f.shift = std::max(1e-3, 2 * f.shift);
  }
}

// z = S P^T L^-T L^-1 P S r
void ic_apply(const IcFactor& f, const double* r, double* z) {
  const int n = f.n, bs = f.bs;
  std::vector<double> y(n);
  for (int p = 0; p < n; ++p) {
    const int i = f.order[p / bs] * bs + p % bs;
    double s = f.scale[i] * r[i];
    for (int k = f.lptr[p]; k < f.lptr[p + 1]; ++k) s -= f.lval[k] * y[f.lcol[k]];
    y[p] = s / f.ldiag[p];
  }
  for (int p = n - 1; p >= 0; --p) {
    y[p] /= f.ldiag[p];
    for (int k = f.lptr[p]; k < f.lptr[p + 1]; ++k) y[f.lcol[k]] -= f.lval[k] * y[p];
  }
  for (int p = 0; p < n; ++p) {
    const int i = f.order[p / bs] * bs + p % bs;
    z[i] = f.scale[i] * y[p];
  }
}

// Eigen's ConjugateGradient loop (x0 = 0) with a preconditioner callback.
template <class Mult, class Pre>
int cg_loop(int n, Mult mult, Pre pre, const double* bb, double* xo, double tol, int max_iters,
            double* err) {
  std::vector<double> b(bb, bb + n);
  std::fill(xo, xo + n, 0.0);
  const double rhsNorm2 = dotv(b, b);
  if (rhsNorm2 == 0) { *err = 0; return 0; }
  const double threshold = std::max(tol * tol * rhsNorm2, 2.2250738585072014e-308);
  std::vector<double> r = b, z(n), p(n), tmp(n);
  double rn2 = dotv(r, r);
  if (rn2 < threshold) { *err = std::sqrt(rn2 / rhsNorm2); return 0; }
  pre(r.data(), p.data());
  double absNew = dotv(r, p);
  int it = 0;
  while (it < max_iters) {
    mult(p, tmp);
    const double alpha = absNew / dotv(p, tmp);
    for (int i = 0; i < n; ++i) { xo[i] += alpha * p[i]; r[i] -= alpha * tmp[i]; }
    rn2 = dotv(r, r);
    if (rn2 < threshold) break;
    pre(r.data(), z.data());
    const double absOld = absNew;
    absNew = dotv(r, z);
    const double beta = absNew / absOld;
    for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    ++it;
  }
  *err = std::sqrt(rn2 / rhsNorm2);
  return it;
}

int ic_pcg(OracleSim* S_, std::vector<double>& xo) {
  const int n = S_->n;
  IcFactor f;
  int rc = ic_factor(f, n, S_->dim, S_->asmb.rowptr.data(), S_->asmb.colind.data(),
                     S_->lhs.data());
  if (rc != 0) return rc;
  S_->ic_shift_last = f.shift;
  S_->ic_colors_last = f.ncolors;
  xo.assign(n, 0.0);
  auto mult = [&](const std::vector<double>& v, std::vector<double>& y) {
    csr_mult(S_, S_->lhs, v, y);
  };
  auto pre = [&](const double* r, double* z) { ic_apply(f, r, z); };
  S_->pcg_iters_last = cg_loop(n, mult, pre, S_->rhs.data(), xo.data(), S_->p.itr_tol,
                               S_->p.max_iters, &S_->pcg_err_last);
  return 0;
}

// dense Cholesky (stands in for the reference's SimplicialLLT / CHOLMOD;
// exact-arithmetic solution is identical). Returns -1 if not PD.
int dense_llt(OracleSim* S_, std::vector<double>& xo) {
  int n = S_->n;
  std::vector<double> L((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int k = S_->asmb.rowptr[i]; k < S_->asmb.rowptr[i + 1]; ++k)
      L[(size_t)i * n + S_->asmb.colind[k]] = S_->lhs[k];
  for (int j = 0; j < n; ++j) {
    double d = L[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) d -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
    if (!(d > 0)) return -1;
    d = std::sqrt(d);
    L[(size_t)j * n + j] = d;
#pragma omp parallel for schedule(static)
    for (int i = j + 1; i < n; ++i) {
      double s = L[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) s -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
      L[(size_t)i * n + j] = s / d;
    }
  }
  std::vector<double> y(n);
  for (int i = 0; i < n; ++i) {
    double s = S_->rhs[i];
    for (int k = 0; k < i; ++k) s -= L[(size_t)i * n + k] * y[k];
    y[i] = s / L[(size_t)i * n + i];
  }
  xo.assign(n, 0.0);
  for (int i = n - 1; i >= 0; --i) {
    double s = y[i];
    for (int k = i + 1; k < n; ++k) s -= L[(size_t)k * n + i] * xo[k];
    xo[i] = s / L[(size_t)i * n + i];
  }
  S_->pcg_iters_last = 0;
  return 0;
}

// Sparse direct solve for larger systems: the same Cholesky factorization of
// K (the reference's direct solvers, eigen_solver.h:17-31, are Cholesky
// factorizations; CHOLMOD and SimplicialLLT order the unknowns to limit fill),
// here with a reverse Cuthill-McKee vertex ordering (Cuthill and McKee 1969,
// George and Liu 1981) and a row-oriented envelope (profile) factor. The
// ordering changes only the rounding, not the solution.
//
// RCM on the vertex graph of a bs x bs-blocked pattern: breadth-first search
// from a minimum-degree vertex of each component, neighbours in increasing
// degree, then the order reversed. Returns perm[new] = old block row.
std::vector<int> rcm_order(int nb, int bs, const int* rowptr, const int* colind) {
  std::vector<std::vector<int>> adj(nb);
  for (int i = 0; i < nb; ++i) {
    for (int k = rowptr[bs * i]; k < rowptr[bs * i + 1]; ++k) {
      const int j = colind[k] / bs;
      if (j != i && (adj[i].empty() || adj[i].back() != j)) adj[i].push_back(j);
    }
    std::sort(adj[i].begin(), adj[i].end());
    adj[i].erase(std::unique(adj[i].begin(), adj[i].end()), adj[i].end());
  }
  std::vector<int> order;
  order.reserve(nb);
  std::vector<char> seen(nb, 0);
  std::vector<int> by_degree(nb);
  for (int i = 0; i < nb; ++i) by_degree[i] = i;
  std::stable_sort(by_degree.begin(), by_degree.end(),
                   [&](int a, int b) { return adj[a].size() < adj[b].size(); });
  for (int start : by_degree) {
    if (seen[start]) continue;
    size_t head = order.size();
    order.push_back(start);
    seen[start] = 1;
    while (head < order.size()) {
      const int v = order[head++];
      std::vector<int> nb_v;
      for (int w : adj[v])
        if (!seen[w]) {
          seen[w] = 1;
          nb_v.push_back(w);
        }
      std::stable_sort(nb_v.begin(), nb_v.end(),
                       [&](int a, int b) { return adj[a].size() < adj[b].size(); });
      order.insert(order.end(), nb_v.begin(), nb_v.end());
    }
  }
  std::reverse(order.begin(), order.end());
  return order;
}

// Solve A x = b (A symmetric positive definite, scalar CSR with bs x bs
// blocks) by envelope Cholesky in RCM order. Returns -1 if A is not PD.
int envelope_llt_solve(int n, int bs, const int* rowptr, const int* colind, const double* vals,
                       const double* b, double* x) {
  const int nb = n / bs;
  const std::vector<int> perm = rcm_order(nb, bs, rowptr, colind);
  std::vector<int> inv(nb);
  for (int i = 0; i < nb; ++i) inv[perm[i]] = i;
  auto newidx = [&](int old) { return bs * inv[old / bs] + old % bs; };
  // first column of every permuted row, then the row offsets of L
  std::vector<int> first(n);
  for (int i = 0; i < n; ++i) {
    const int o = bs * perm[i / bs] + i % bs;
    int f = i;
    for (int k = rowptr[o]; k < rowptr[o + 1]; ++k) f = std::min(f, newidx(colind[k]));
    first[i] = f;
  }
  std::vector<long> off(n + 1, 0);
  for (int i = 0; i < n; ++i) off[i + 1] = off[i] + (i - first[i] + 1);
  std::vector<double> L(off[n], 0.0);
  auto at = [&](int i, int j) -> double& { return L[off[i] + (j - first[i])]; };
  for (int i = 0; i < n; ++i) {
    const int o = bs * perm[i / bs] + i % bs;
    for (int k = rowptr[o]; k < rowptr[o + 1]; ++k) {
      const int j = newidx(colind[k]);
      if (j <= i) at(i, j) = vals[k];
    }
  }
  for (int i = 0; i < n; ++i) {
    const int fi = first[i];
    double* Li = &L[off[i]];
    for (int j = fi; j < i; ++j) {
      const int k0 = std::max(fi, first[j]);
      const double* Lj = &L[off[j]];
      double s = Li[j - fi];
      for (int k = k0; k < j; ++k) s -= Li[k - fi] * Lj[k - first[j]];
      Li[j - fi] = s / Lj[j - first[j]];
    }
    double d = Li[i - fi];
    for (int k = fi; k < i; ++k) d -= Li[k - fi] * Li[k - fi];
    if (!(d > 0)) return -1;
    Li[i - fi] = std::sqrt(d);
  }
  std::vector<double> y(n);
  for (int i = 0; i < n; ++i) {
    double s = b[bs * perm[i / bs] + i % bs];
    for (int k = first[i]; k < i; ++k) s -= at(i, k) * y[k];
    y[i] = s / at(i, i);
  }
  for (int i = n - 1; i >= 0; --i) {
    const double xi = y[i] / at(i, i);
    y[i] = xi;
    for (int k = first[i]; k < i; ++k) y[k] -= at(i, k) * xi;
  }
  for (int i = 0; i < n; ++i) x[bs * perm[i / bs] + i % bs] = y[i];
  return 0;
}

// Direct solve of the Newton system: dense LLT for small systems (the
// golden trajectories were recorded with it), the RCM envelope factor above
// beyond 3000 unknowns.
int direct_solve(OracleSim* S_, std::vector<double>& xo) {
  if (S_->n <= 3000) return dense_llt(S_, xo);
  xo.assign(S_->n, 0.0);
  S_->pcg_iters_last = 0;
  return envelope_llt_solve(S_->n, S_->dim, S_->asmb.rowptr.data(), S_->asmb.colind.data(),
                            S_->lhs.data(), S_->rhs.data(), xo.data());
}

// ---------------- indefinite (KKT) solve: "minres-indefinite-block" -------
// Unknowns [dx (n); ds (ns nT); lambda (ns nT)], ns = 6 in 3D and 3 in 2D,
// element-major, as SystemMatrixIndefinite lays them out (reference
// linear_system.h:60-99; the factory registers it for DIM = 2 and 3,
// linear_solver_factory.cpp:117-133, 150-151).

// In-place dense Cholesky (lower, row-major n x n); false if not PD.
bool dense_chol(std::vector<double>& A, int n) {
  for (int j = 0; j < n; ++j) {
    double d = A[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) d -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
    if (!(d > 0)) return false;
    d = std::sqrt(d);
    A[(size_t)j * n + j] = d;
#pragma omp parallel for schedule(static)
    for (int i = j + 1; i < n; ++i) {
      double s = A[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) s -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
      A[(size_t)i * n + j] = s / d;
    }
  }
  return true;
}

void dense_chol_solve(const std::vector<double>& L, int n, const double* b, double* x) {
  std::vector<double> y(n);
  for (int i = 0; i < n; ++i) {
    double s = b[i];
    for (int k = 0; k < i; ++k) s -= L[(size_t)i * n + k] * y[k];
    y[i] = s / L[(size_t)i * n + i];
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = y[i];
    for (int k = i + 1; k < n; ++k) s -= L[(size_t)k * n + i] * x[k];
    x[i] = s / L[(size_t)i * n + i];
  }
}

// The mixed variable's element data per dimension: 3D 6-vectors with a 9 x 6
// dSdF, 2D 3-vectors with a 4 x 3 dSdF (the DIM = 3 and DIM = 2 instances of
// MixedStretch).
template <int D> struct MixedView;
template <> struct MixedView<3> {
  static constexpr int NS = 6, NF = 9, NC = 4;
  static double sym(int i) { return kSym[i]; }
  static const Mat<9, 6>& dSdF(const OracleSim* S_, int e) { return S_->dSdF[e]; }
  static const Mat<6, 6>& Hloc(const OracleSim* S_, int e) { return S_->Hloc[e]; }
  static const Mat<6, 6>& Hinv(const OracleSim* S_, int e) { return S_->Hinv[e]; }
  static double g(const OracleSim* S_, int e, int i) { return S_->g[e][i]; }
  static double S(const OracleSim* S_, int e, int i) { return S_->S[e][i]; }
};
template <> struct MixedView<2> {
  static constexpr int NS = 3, NF = 4, NC = 3;
  static double sym(int i) { return kSym3[i]; }
  static const Mat<4, 3>& dSdF(const OracleSim* S_, int e) { return S_->dSdF2[e]; }
  static const Mat<3, 3>& Hloc(const OracleSim* S_, int e) { return S_->Hloc3[e]; }
  static const Mat<3, 3>& Hinv(const OracleSim* S_, int e) { return S_->Hinv3[e]; }
  static double g(const OracleSim* S_, int e, int i) { return S_->g3[e][i]; }
  static double S(const OracleSim* S_, int e, int i) { return S_->S3[e][i]; }
};

int kkt_ns(const OracleSim* S_) { return S_->dim == 2 ? 3 : 6; }

// rhs = [-grad_x; -vol g; vol Sym (S - s)]: linear_system.h:67-86 with
// Displacement::gradient (displacement.cpp:64-73), grad_ = vol g
// (mixed_stretch.cpp:195-203) and grad_la_ = -vol Sym (S - s)
// (mixed_stretch.cpp:127-131, 313-332).
template <int D>
void kkt_rhs_t(OracleSim* S_, std::vector<double>& b) {
  using V = MixedView<D>;
  const int n = S_->n, nT = S_->nT, ns = V::NS;
  b.assign((size_t)n + 2 * ns * (size_t)nT, 0.0);
  disp_gradient(S_);
  for (int i = 0; i < n; ++i) b[i] = -S_->grad_x_disp[i];
  for (int e = 0; e < nT; ++e) {
    const double vol = S_->vols[e];
    for (int i = 0; i < ns; ++i) {
      b[n + (size_t)ns * e + i] = -vol * V::g(S_, e, i);
      b[n + (size_t)ns * nT + (size_t)ns * e + i] =
          vol * V::sym(i) * (V::S(S_, e, i) - S_->s[(size_t)ns * e + i]);
    }
  }
}

void kkt_rhs(OracleSim* S_, std::vector<double>& b) {
  if (S_->dim == 2) kkt_rhs_t<2>(S_, b);
  else kkt_rhs_t<3>(S_, b);
}

// y = K w with K = [[PMP, 0, Gx], [0, Hloc, C], [Gx^T, C, 0]],
// Gx = -PJW diag(dSdF) (jacobian_x), C = vol Sym (jacobian_mixed), Hloc =
// vol h^2 Hpsi: BlockMatrix::scaleAndAddTo (block_matrix.h:74-125) with
// Displacement::product_hessian (displacement.h:59-63) and
// MixedStretch::product_hessian / product_jacobian_x / product_jacobian_mixed
// (mixed_stretch.cpp:373-434), accumulated in the same order.
template <int D>
void kkt_product_t(const OracleSim* S_, const double* w, double* y) {
  using V = MixedView<D>;
  constexpr int ns = V::NS, nf = V::NF;
  const int n = S_->n, nT = S_->nT;
  const double* wx = w;
  const double* ws = w + n;
  const double* wl = w + n + (size_t)ns * nT;
  double* yx = y;
  double* ys = y + n;
  double* yl = y + n + (size_t)ns * nT;
  std::fill(y, y + (size_t)n + 2 * ns * (size_t)nT, 0.0);
  // x rows: PMP dx, then -= PJW (dSdF la)
  {
    std::vector<double> xv(wx, wx + n), t;
    csr_mult(S_, S_->PMP, xv, t);
    for (int i = 0; i < n; ++i) yx[i] += t[i];
    std::vector<double> tmp((size_t)nf * nT);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < nT; ++e) {
      Vec<ns> l; for (int i = 0; i < ns; ++i) l[i] = wl[(size_t)ns * e + i];
      Vec<nf> v = V::dSdF(S_, e) * l;
      for (int i = 0; i < nf; ++i) tmp[(size_t)nf * e + i] = v[i];
    }
    std::vector<double> acc(n, 0.0);
    apply_PJW(S_, tmp, acc, -1.0);
    for (int i = 0; i < n; ++i) yx[i] += acc[i];
  }
  // s rows: Hloc ds + vol Sym la
#pragma omp parallel for schedule(static)
  for (int e = 0; e < nT; ++e) {
    const double vol = S_->vols[e];
    for (int i = 0; i < ns; ++i) {
      double a = 0;
      for (int j = 0; j < ns; ++j) a += V::Hloc(S_, e)(i, j) * ws[(size_t)ns * e + j];
      ys[(size_t)ns * e + i] += a;
    }
    for (int i = 0; i < ns; ++i) ys[(size_t)ns * e + i] += vol * V::sym(i) * wl[(size_t)ns * e + i];
  }
  // la rows: -= dSdF^T (PJW^T dx), then += vol Sym ds
  {
    std::vector<double> full((size_t)S_->nV * D, 0.0);
    for (int v = 0; v < S_->nV; ++v) {
      int f = S_->free_map[v];
      if (f < 0) continue;
      for (int r = 0; r < D; ++r) full[(size_t)D * v + r] = wx[(size_t)D * f + r];
    }
#pragma omp parallel for schedule(static)
    for (int e = 0; e < nT; ++e) {
      double Fd[9]; def_grad(S_, full, e, Fd);
      Vec<nf> t; for (int i = 0; i < nf; ++i) t[i] = S_->vols[e] * Fd[i];
      Vec<ns> u = V::dSdF(S_, e).T() * t;
      for (int i = 0; i < ns; ++i) yl[(size_t)ns * e + i] -= u[i];
    }
    for (int e = 0; e < nT; ++e)
      for (int i = 0; i < ns; ++i)
        yl[(size_t)ns * e + i] += S_->vols[e] * V::sym(i) * ws[(size_t)ns * e + i];
  }
}

void kkt_product(const OracleSim* S_, const double* w, double* y) {
  if (S_->dim == 2) kkt_product_t<2>(S_, w, y);
  else kkt_product_t<3>(S_, w, y);
}

// Chebyshev steps of the mass block: kMassCheb steps of Chebyshev
// acceleration of Jacobi on PMP (Saad, Iterative Methods for Sparse Linear
// Systems, 2nd ed., Alg. 12.1) over the interval [1/2, 5/2], which holds the
// spectrum of diag(M)^-1 M for linear tets and triangles (Wathen 1987: the
// scaled element matrix (1 + delta_ab)/2 has eigenvalues 1/2 and 5/2 on a
// tet, 1/2 and 2 on a triangle, and a principal submatrix stays inside). The result is p(D^-1 M) D^-1 b for a fixed polynomial p,
// positive on the interval, so the operator is symmetric positive definite.
constexpr int kMassCheb = 4;

void mass_cheb(const OracleSim* S_, const double* b, double* z) {
  const int n = S_->n;
  const double theta = 1.5, delta = 1.0, sigma = theta / delta;
  std::vector<double> r(b, b + n), d(n), Ad;
  double rho = 1.0 / sigma;
  for (int i = 0; i < n; ++i) { z[i] = 0.0; d[i] = S_->mdinv[i] * r[i] / theta; }
  for (int k = 1; k <= kMassCheb; ++k) {
    for (int i = 0; i < n; ++i) z[i] += d[i];
    if (k == kMassCheb) break;
    csr_mult(S_, S_->PMP, d, Ad);
    const double rho1 = 1.0 / (2.0 * sigma - rho);
    for (int i = 0; i < n; ++i) {
      r[i] -= Ad[i];
      d[i] = rho1 * rho * d[i] + (2.0 * rho1 / delta) * S_->mdinv[i] * r[i];
    }
    rho = rho1;
  }
}

// Mass-only preconditioner data (the reference's init(), preconditioners.h:81-100):
// lumped mass Mlump = PMP 1, and for the GPU form the Jacobi scaling.
void kkt_mass_setup(OracleSim* S_, bool exact) {
  const int n = S_->n;
  if (!S_->kkt_mass_ready) {
    S_->mlump.assign(n, 0.0);
    S_->mdinv.assign(n, 0.0);
    for (int i = 0; i < n; ++i)
      for (int k = S_->asmb.rowptr[i]; k < S_->asmb.rowptr[i + 1]; ++k) {
        S_->mlump[i] += S_->PMP[k];
        if (S_->asmb.colind[k] == i) S_->mdinv[i] = 1.0 / S_->PMP[k];
      }
    S_->kkt_mass_ready = true;
  }
  if (exact && S_->mchol.empty()) {
    S_->mchol.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
      for (int k = S_->asmb.rowptr[i]; k < S_->asmb.rowptr[i + 1]; ++k)
        S_->mchol[(size_t)i * n + S_->asmb.colind[k]] = S_->PMP[k];
    dense_chol(S_->mchol, n);
  }
}

// Gx column block of element e restricted to its corners: G_e(D k + r, j) =
// -vol dX(k, c) dSdF(r + D c, j) for free corner k (Gx = -PJW diag(dSdF),
// mixed_stretch.cpp:342-348).
template <int D>
void gx_block(const OracleSim* S_, int e, double G[12][6]) {
  using V = MixedView<D>;
  const double vol = S_->vols[e];
  const double* dx = &S_->dX[(size_t)e * V::NC * D];
  for (int k = 0; k < V::NC; ++k)
    for (int r = 0; r < D; ++r)
      for (int j = 0; j < V::NS; ++j) {
        double a = 0;
        for (int c = 0; c < D; ++c) a += dx[D * k + c] * V::dSdF(S_, e)(r + D * c, j);
        G[D * k + r][j] = -vol * a;
      }
}

// The lambda block. The reference factors L = Gx^T Mlump^-1 Gx + 1e-8 I
// whole (preconditioners.h:102-114, 124-146; exact = true here: dense
// Cholesky of the SPD L, the same solution as its SimplicialLDLT). The GPU
// form is the block-diagonal preconditioner of Murphy, Golub and Wathen
// (SIAM J. Sci. Comput. 21, 2000) with the Schur complement
// Gx^T M^-1 Gx + C Hloc^-1 C replaced by its element blocks, lumped mass in
// place of M: S_e = Gx_e^T Mlump^-1 Gx_e + vol Sym Hpsi^-1 Sym, inverted.
template <int D>
void kkt_refresh_L_t(OracleSim* S_, bool exact) {
  using V = MixedView<D>;
  constexpr int ns = V::NS, nc = V::NC;
  const int nT = S_->nT;
  S_->linv_blk.assign((size_t)nT * ns * ns, 0.0);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < nT; ++e) {
    double G[12][6];
    gx_block<D>(S_, e, G);
    Mat<ns, ns> Le;
    for (int i = 0; i < ns; ++i)
      for (int j = 0; j < ns; ++j) {
        double a = 0;
        for (int k = 0; k < nc; ++k) {
          int f = S_->free_map[S_->T[(size_t)nc * e + k]];
          if (f < 0) continue;
          for (int r = 0; r < D; ++r)
            a += G[D * k + r][i] * G[D * k + r][j] / S_->mlump[(size_t)D * f + r];
        }
        Le(i, j) = a + (exact ? (i == j ? 1e-8 : 0.0)
                              : S_->vols[e] * V::sym(i) * V::Hinv(S_, e)(i, j) * V::sym(j));
      }
    const Mat<ns, ns> Li = lu_inverse<ns>(Le);
    for (int i = 0; i < ns; ++i)
      for (int j = 0; j < ns; ++j) S_->linv_blk[((size_t)e * ns + i) * ns + j] = Li(i, j);
  }
  if (!exact) return;
  const int m = ns * nT;
  // dense Gx (n x ns nT), then L
  const int n = S_->n;
  std::vector<double> Gx((size_t)n * m, 0.0);
  for (int e = 0; e < nT; ++e) {
    double G[12][6];
    gx_block<D>(S_, e, G);
    for (int k = 0; k < nc; ++k) {
      int f = S_->free_map[S_->T[(size_t)nc * e + k]];
      if (f < 0) continue;
      for (int r = 0; r < D; ++r)
        for (int j = 0; j < ns; ++j) Gx[(size_t)(D * f + r) * m + ns * e + j] += G[D * k + r][j];
    }
  }
  S_->lchol.assign((size_t)m * m, 0.0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < m; ++i)
    for (int row = 0; row < n; ++row) {
      double gi = Gx[(size_t)row * m + i];
      if (gi == 0.0) continue;
      double w = gi / S_->mlump[row];
      for (int j = 0; j < m; ++j) S_->lchol[(size_t)i * m + j] += w * Gx[(size_t)row * m + j];
    }
  for (int i = 0; i < m; ++i) S_->lchol[(size_t)i * m + i] += 1e-8;
  dense_chol(S_->lchol, m);
}

void kkt_refresh_L(OracleSim* S_, bool exact) {
  if (S_->dim == 2) kkt_refresh_L_t<2>(S_, exact);
  else kkt_refresh_L_t<3>(S_, exact);
}

// z = P b. exact: the reference's P = diag(PMP^-1, vol Hpsi^-1, L^-1)
// (BlockDiagonalPreconditioner::_solve_impl, preconditioners.h:160-192, with
// product_hessian_inv, mixed_stretch.cpp:385-396; Hpsi here is h^2 times the
// material Hessian). GPU form: diag(Chebyshev(PMP), Hloc^-1 = Hpsi^-1 / vol,
// element blocks S_e^-1).
template <int D>
void kkt_precond_t(const OracleSim* S_, bool exact, const double* b, double* z) {
  using V = MixedView<D>;
  constexpr int ns = V::NS;
  const int n = S_->n, nT = S_->nT;
  if (exact) dense_chol_solve(S_->mchol, n, b, z);
  else mass_cheb(S_, b, z);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < nT; ++e) {
    const double vol = S_->vols[e];
    for (int i = 0; i < ns; ++i) {
      double a = 0;
      for (int j = 0; j < ns; ++j) a += V::Hinv(S_, e)(i, j) * b[n + (size_t)ns * e + j];
      z[n + (size_t)ns * e + i] = exact ? vol * a : a / vol;
    }
  }
  const size_t lo = n + (size_t)ns * nT;
  if (exact) {
    dense_chol_solve(S_->lchol, ns * nT, b + lo, z + lo);
  } else {
#pragma omp parallel for schedule(static)
    for (int e = 0; e < nT; ++e)
      for (int i = 0; i < ns; ++i) {
        double a = 0;
        for (int j = 0; j < ns; ++j)
          a += S_->linv_blk[((size_t)e * ns + i) * ns + j] * b[lo + (size_t)ns * e + j];
        z[lo + (size_t)ns * e + i] = a;
      }
  }
}

void kkt_precond(const OracleSim* S_, bool exact, const double* b, double* z) {
  if (S_->dim == 2) kkt_precond_t<2>(S_, exact, b, z);
  else kkt_precond_t<3>(S_, exact, b, z);
}

// Preconditioned MINRES of Eigen 3.4 (unsupported/Eigen/src/IterativeSolvers/
// MINRES.h, the Lanczos variant A(2,7) of Paige 1972 with Givens QR), with
// x0 = 0 (IterativeSolverBase::_solve_impl), the stop test on the estimated
// residual |b|^2 prod s_i^2 < tol^2 |b|^2 and iters counting the iterations
// completed without stopping. Returns the estimated relative residual.
template <class Op, class Pre>
double minres(int N, Op op, Pre pre, const double* b, double* x, int max_iters, double tol,
              int* iters) {
  std::fill(x, x + N, 0.0);
  double rhs2 = 0;
  for (int i = 0; i < N; ++i) rhs2 += b[i] * b[i];
  *iters = 0;
  if (rhs2 == 0) return 0.0;
  const double threshold2 = tol * tol * rhs2;
  std::vector<double> v_old(N), v(N, 0.0), v_new(b, b + N), w(N), w_new(N), Aw(N);
  std::vector<double> p_oold(N), p_old(N, 0.0), p(N, 0.0);
  // v_new = b - K x0 with x0 = 0
  double res2 = rhs2;
  pre(v_new.data(), w_new.data());
  double beta_new2 = 0;
  for (int i = 0; i < N; ++i) beta_new2 += v_new[i] * w_new[i];
  if (!(beta_new2 >= 0)) return -1.0;  // preconditioner not positive definite
  double beta_new = std::sqrt(beta_new2);
  const double beta_one = beta_new;
  double c = 1.0, c_old = 1.0, s = 0.0, s_old = 0.0, eta = 1.0;
  int it = 0;
  while (it < max_iters) {
    const double beta = beta_new;
    v_old = v;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) { v_new[i] /= beta_new; w_new[i] /= beta_new; }
    v = v_new;
    w = w_new;
    op(w.data(), Aw.data());
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) v_new[i] = Aw[i] - beta * v_old[i];
    double alpha = 0;
    for (int i = 0; i < N; ++i) alpha += v_new[i] * w[i];
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) v_new[i] -= alpha * v[i];
    pre(v_new.data(), w_new.data());
    beta_new2 = 0;
    for (int i = 0; i < N; ++i) beta_new2 += v_new[i] * w_new[i];
    if (!(beta_new2 >= 0)) return -1.0;
    beta_new = std::sqrt(beta_new2);
    // Givens rotation (s, c, s_old, c_old still from the previous iteration)
    const double r2 = s * alpha + c * c_old * beta;
    const double r3 = s_old * beta;
    const double r1_hat = c * alpha - c_old * s * beta;
    const double r1 = std::sqrt(r1_hat * r1_hat + beta_new * beta_new);
    c_old = c;
    s_old = s;
    c = r1_hat / r1;
    s = beta_new / r1;
    p_oold = p_old;
    p_old = p;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
      p[i] = (w[i] - r2 * p_old[i] - r3 * p_oold[i]) / r1;
      x[i] += beta_one * c * eta * p[i];
    }
    res2 *= s * s;
    if (res2 < threshold2) break;
    eta = -s * eta;
    ++it;
  }
  *iters = it;
  return std::sqrt(res2 / rhs2);
}

// EigenIterativeSolver<MINRES<BlockMatrix, ..., BlockDiagonalPreconditioner>>::
// solve (eigen_iterative_solver.h:20-30) and SystemMatrixIndefinite::post_solve
// (linear_system.h:88-99): dx, ds and lambda straight from the KKT solution.
// The reference refactors L on every max_newton-th solve (compute(),
// preconditioners.h:150-158; the counter here is per simulation); the GPU
// form's element blocks are rebuilt on every solve.
int kkt_solve(OracleSim* S_) {
  const bool exact = S_->p.solver == 3;
  kkt_mass_setup(S_, exact);
  if (S_->kkt_solves == 0 || !exact) kkt_refresh_L(S_, exact);
  S_->kkt_solves = (S_->kkt_solves + 1) % std::max(1, S_->p.max_newton);
  std::vector<double> b;
  kkt_rhs(S_, b);
  const int n = S_->n, nT = S_->nT, ns = kkt_ns(S_), N = n + 2 * ns * nT;
  std::vector<double> sol(N);
  int iters = 0;
  double err = minres(
      N, [&](const double* w, double* y) { kkt_product(S_, w, y); },
      [&](const double* r, double* z) { kkt_precond(S_, exact, r, z); }, b.data(), sol.data(),
      S_->p.max_iters, S_->p.itr_tol, &iters);
  if (err < 0) return -2;
  S_->pcg_iters_last = iters;
  S_->pcg_err_last = err;
  S_->dx.assign(sol.begin(), sol.begin() + n);
  S_->ds.assign(sol.begin() + n, sol.begin() + n + (size_t)ns * nT);
  S_->lam.assign(sol.begin() + n + (size_t)ns * nT, sol.end());
  return 0;
}

// ---------------- Newton optimizer ----------------
void update_system(OracleSim* S_) {
  std::vector<double> full;
  unproject(S_, S_->x, full);
  if (S_->p.mixed && S_->dim == 2) {
    mixed2_update_rotations(S_, full);
    mixed2_update_derivatives(S_);
  } else if (S_->p.mixed) {
    mixed_update_rotations(S_, full);
    mixed_update_derivatives(S_);
  } else {
    plain_update(S_, full);
  }
}

int substep(OracleSim* S_, double& decrement) {
  if (S_->p.solver == 2 || S_->p.solver == 3) {
    // indefinite system (mixed scenes only)
    if (!S_->p.mixed) return -3;
    int rc = kkt_solve(S_);
    if (rc != 0) return rc;
    S_->st.pcg_iters_total += S_->pcg_iters_last;
    decrement = 0;
    for (double v : S_->dx) decrement = std::max(decrement, std::fabs(v));
    for (double v : S_->ds) decrement = std::max(decrement, std::fabs(v));
    return 0;
  }
  // SystemMatrixPD::pre_solve: lhs = PMP + var lhs; rhs = x rhs + var rhs
  std::vector<double> Avals;
  if (S_->p.mixed) {
    Avals = S_->lhs;  // assembled in mixed_update_derivatives
  } else {
    S_->asmb.update_matrix(S_->Hp, Avals);
  }
  S_->lhs = S_->PMP;
  for (size_t k = 0; k < S_->lhs.size(); ++k) S_->lhs[k] += Avals[k];
  disp_gradient(S_);
  S_->rhs.assign(S_->n, 0.0);
  for (int i = 0; i < S_->n; ++i) S_->rhs[i] = -S_->grad_x_disp[i];
  if (S_->p.mixed) {
    std::vector<double> mr = S_->dim == 2 ? mixed2_rhs(S_) : mixed_rhs(S_);
    for (int i = 0; i < S_->n; ++i) S_->rhs[i] += mr[i];
  } else {
    for (int i = 0; i < S_->n; ++i) S_->rhs[i] += -S_->grad_plain[i];
  }
  std::vector<double> sol;
  int rc = S_->p.solver == 0 ? direct_solve(S_, sol) : S_->p.solver == 4 ? ic_pcg(S_, sol) : pcg(S_, sol);
  if (rc != 0) return rc;
  S_->st.pcg_iters_total += S_->pcg_iters_last;
  // post_solve
  S_->dx = sol;
  if (S_->p.mixed && S_->dim == 2) mixed2_solve(S_, sol);
  else if (S_->p.mixed) mixed_solve(S_, sol);
  decrement = 0;
  for (double v : S_->dx) decrement = std::max(decrement, std::fabs(v));
  if (S_->p.mixed)
    for (double v : S_->ds) decrement = std::max(decrement, std::fabs(v));
  return 0;
}

double energy_func(const OracleSim* S_, double a) {
  double h2 = S_->h * S_->h;
  std::vector<double> x0(S_->n);
  for (int i = 0; i < S_->n; ++i) x0[i] = S_->x[i] + a * S_->dx[i];
  double val = disp_energy(S_, x0);
  std::vector<double> full;
  unproject(S_, x0, full);
  if (S_->p.mixed) {
    std::vector<double> si(S_->s.size());
    for (size_t i = 0; i < si.size(); ++i) si[i] = S_->s[i] + a * S_->ds[i];
    if (S_->dim == 2)
      val += h2 * mixed2_energy(S_, si) - mixed2_constraint_value(S_, full, si);
    else
      val += h2 * mixed_energy(S_, si) - mixed_constraint_value(S_, full, si);
  } else {
    val += h2 * plain_energy(S_, full);
  }
  return val;
}

// Moves the pinned values b of the end groups: stretch translates group i by
// (-1)^i v dt along the axis; twist rotates it by (-1)^i v pi dt about the x
// axis through the start bounding-box center (about z in 2D); bend about the
// z axis through the group's last vertex. Rotations act on the current b.
void bc_step(OracleSim* S_, double dt) {
  const int t = S_->p.bc_type, D = S_->dim;
  if (t != 4 && t != 7 && t != 8) return;
  for (int gi = 0; gi < 2; ++gi) {
    const double sgn = gi == 0 ? 1.0 : -1.0;
    if (t == 4) {
      for (int j : S_->bc_g[gi])
        if (S_->fixed[j]) S_->b[D * j + S_->p.bc_axis] += sgn * S_->p.bc_velocity * dt;
      continue;
    }
    const double a = sgn * S_->p.bc_velocity * M_PI * dt;
    const double c = std::cos(a), s = std::sin(a);
    const double* o = t == 8 ? S_->bc_center : S_->bc_gc[gi];
    const int p0 = (t == 8 && D == 3) ? 1 : 0, p1 = p0 + 1;
    for (int j : S_->bc_g[gi]) {
      if (S_->fixed[j] == 0) continue;
      const double u = S_->b[D * j + p0] - o[p0], w = S_->b[D * j + p1] - o[p1];
      S_->b[D * j + p0] = c * u - s * w + o[p0];
      S_->b[D * j + p1] = s * u + c * w + o[p1];
    }
  }
}

int step(OracleSim* S_) {
  int i = 0;
  double grad_norm = 0, E = 0, E_prev = 0, res = 0;
  S_->st.newton_iters = 0;
  S_->st.pcg_iters_total = 0;
  S_->st.ls_fail = 0;
  S_->hist.clear();
  do {
    update_system(S_);
    int rc = substep(S_, grad_norm);
    if (rc != 0) return rc;
    double alpha = 1.0;
    E = energy_func(S_, 0.0);
    res = std::fabs((E - E_prev) / (E + 1e-6));
    E_prev = E;
    // linesearch_backtracking with c = 0, p = 0.5
    double fx0 = energy_func(S_, 0.0);
    int iter = 0;
    while (iter < S_->p.ls_iters) {
      double fx = energy_func(S_, alpha);
      if (fx < fx0 + (alpha * 0.0) * 0.0) break;
      alpha *= 0.5;
      ++iter;
    }
    if (iter < S_->p.ls_iters) {
      for (int k = 0; k < S_->n; ++k) S_->x[k] += alpha * S_->dx[k];
      if (S_->p.mixed)
        for (size_t k = 0; k < S_->s.size(); ++k) S_->s[k] += alpha * S_->ds[k];
    } else {
      S_->st.ls_fail++;
    }
    S_->st.last_alpha = alpha;
    // OptimizerData rows (newton_optimizer.cpp:60-66)
    S_->hist.push_back({(double)(i + 1), E, res, grad_norm, alpha});
    ++i;
  } while (i < S_->p.max_newton && grad_norm > S_->p.newton_tol && res > 1e-12);
  S_->st.newton_iters = i;
  S_->st.last_decrement = grad_norm;
  S_->st.last_energy = E;
  S_->st.last_res = res;
  // update_bcs (newton_optimizer.cpp:78-79): the scripted end groups move
  // (stretch_boundary_condition.h, twist_..., bend_...; group i with
  // (-1)^i velocity); then Displacement::post_solve takes the pinned values
  bc_step(S_, S_->p.dt);
  std::vector<double> full;
  unproject(S_, S_->x, full);
  {
    size_t N = full.size();
    std::vector<double> wx(N, 0.0);
    for (size_t k = 0; k < S_->x_prevs.size(); ++k)
      for (size_t q = 0; q < N; ++q) wx[q] += S_->alphas[k] * S_->x_prevs[k][q];
    std::vector<double> v(N);
    for (size_t q = 0; q < N; ++q) v[q] = (full[q] - wx[q]) / S_->h;
    S_->x_prevs.push_front(full);
    S_->v_prevs.push_front(v);
    if ((int)S_->x_prevs.size() > S_->p.bdf_order) {
      S_->x_prevs.pop_back();
      S_->v_prevs.pop_back();
    }
  }
  S_->Vcur = full;
  if (S_->p.mixed) std::fill(S_->lam.begin(), S_->lam.end(), 0.0);
  return 0;
}

void bdf_tables(int I, std::vector<double>& a, double& beta) {
  switch (I) {
    case 1: a = {1.0}; beta = 1.0; break;
    case 2: a = {4.0 / 3.0, -1.0 / 3.0}; beta = 2.0 / 3.0; break;
    case 3: a = {18.0 / 11.0, -9.0 / 11.0, 2.0 / 11.0}; beta = 6.0 / 11.0; break;
    case 4: a = {48.0 / 25.0, -36.0 / 25.0, 16.0 / 25.0, -3.0 / 25.0}; beta = 12.0 / 25.0; break;
    case 5: a = {300.0 / 137.0, -300.0 / 137.0, 200.0 / 137.0, -75.0 / 137.0, 12.0 / 137.0};
            beta = 60.0 / 137.0; break;
    default: a = {360.0 / 147.0, -450.0 / 147.0, 400.0 / 147.0, -225.0 / 147.0, 72.0 / 147.0,
                  -10.0 / 147.0}; beta = 60.0 / 147.0; break;
  }
}

}  // namespace

extern "C" {

// the BDF table of `order` (1-6): a[0 .. order-1] and beta; returns order
int oracle_bdf_tables(int order, double* a, double* beta) {
  std::vector<double> av;
  bdf_tables(order, av, *beta);
  std::copy(av.begin(), av.end(), a);
  return (int)av.size();
}

int oracle_newton_procrustes(double* R, const double* A, const double* B, int cg,
                             double* dRdF, double tol, int max_iter, int* iters) {
  M3 Rm, Am, Bm;
  for (int i = 0; i < 9; ++i) { Rm[i] = R[i]; Am[i] = A[i]; Bm[i] = B[i]; }
  Mat<9, 9> d;
  int rc = newton_procrustes(Rm, Am, Bm, cg != 0, &d, tol, max_iter, iters);
  for (int i = 0; i < 9; ++i) R[i] = Rm[i];
  if (cg && dRdF) for (int i = 0; i < 81; ++i) dRdF[i] = d[i];
  return rc;
}

int oracle_polar_svd3(double* R, double* s, const double* F, int grad, double* dsdF,
                      int* iters) {
  M3 Rm, Fm;
  for (int i = 0; i < 9; ++i) { Rm[i] = R[i]; Fm[i] = F[i]; }
  Vec<6> sv; Mat<6, 9> d;
  int rc = polar_svd3(Rm, sv, Fm, grad != 0, &d, iters);
  for (int i = 0; i < 9; ++i) R[i] = Rm[i];
  for (int i = 0; i < 6; ++i) s[i] = sv[i];
  if (grad && dsdF) for (int i = 0; i < 54; ++i) dsdF[i] = d[i];
  return rc;
}

void oracle_material_s6(int type, double mu, double la, const double* s, int psd,
                        double* E, double* g, double* H) {
  Vec<6> sv; for (int i = 0; i < 6; ++i) sv[i] = s[i];
  Vec<6> gg; Mat<6, 6> HH;
  material_s6(type, mu, la, sv, psd != 0, E, &gg, &HH);
  if (g) for (int i = 0; i < 6; ++i) g[i] = gg[i];
  if (H) for (int i = 0; i < 36; ++i) H[i] = HH[i];
}

void oracle_snh_f9(double mu, double la, const double* F, double* E, double* g, double* H) {
  Vec<9> f; for (int i = 0; i < 9; ++i) f[i] = F[i];
  Vec<9> gg; Mat<9, 9> HH;
  snh_f9(mu, la, f, E, &gg, H ? &HH : nullptr);
  if (g) for (int i = 0; i < 9; ++i) g[i] = gg[i];
  if (H) for (int i = 0; i < 81; ++i) H[i] = HH[i];
}

int oracle_polar2d(double* R, double* s, const double* F, int grad, double* dsdF) {
  Mat<2, 2> Rm, Fm; for (int i = 0; i < 4; ++i) Fm[i] = F[i];
  Vec<3> sv; Mat<3, 4> d;
  int rc = polar2d(Rm, sv, Fm, grad != 0, &d);
  for (int i = 0; i < 4; ++i) R[i] = Rm[i];
  for (int i = 0; i < 3; ++i) s[i] = sv[i];
  if (grad && dsdF) for (int i = 0; i < 12; ++i) dsdF[i] = d[i];
  return rc;
}

int oracle_material_s3(int type, double mu, double la, const double* s, int psd_fix, double* E,
                       double* g, double* H) {
  Vec<3> sv; for (int i = 0; i < 3; ++i) sv[i] = s[i];
  Vec<3> gg; Mat<3, 3> HH;
  if (!material_s3(type, mu, la, sv, psd_fix != 0, E, g ? &gg : nullptr, H ? &HH : nullptr))
    return 0;
  if (g) for (int i = 0; i < 3; ++i) g[i] = gg[i];
  if (H) for (int i = 0; i < 9; ++i) H[i] = HH[i];
  return 1;
}

int oracle_material_f9(int type, double mu, double la, const double* F, double* E, double* g,
                       double* H) {
  Vec<9> f; for (int i = 0; i < 9; ++i) f[i] = F[i];
  Vec<9> gg; Mat<9, 9> HH;
  if (!material_f9(type, mu, la, f, E, g ? &gg : nullptr, H ? &HH : nullptr)) return 0;
  if (g) for (int i = 0; i < 9; ++i) g[i] = gg[i];
  if (H) for (int i = 0; i < 81; ++i) H[i] = HH[i];
  return 1;
}

void oracle_snh_f4(double mu, double la, const double* F, double* E, double* g, double* H) {
  Vec<4> f; for (int i = 0; i < 4; ++i) f[i] = F[i];
  Vec<4> gg; Mat<4, 4> HH;
  snh_f4(mu, la, f, E, &gg, &HH);
  if (g) for (int i = 0; i < 4; ++i) g[i] = gg[i];
  if (H) for (int i = 0; i < 16; ++i) H[i] = HH[i];
}

void oracle_psd_fix(int n, double* A, double tol) {
  if (n == 6) { Mat<6, 6> M; for (int i = 0; i < 36; ++i) M[i] = A[i]; simple_psd_fix<6>(M, tol); for (int i = 0; i < 36; ++i) A[i] = M[i]; }
  else if (n == 9) { Mat<9, 9> M; for (int i = 0; i < 81; ++i) M[i] = A[i]; simple_psd_fix<9>(M, tol); for (int i = 0; i < 81; ++i) A[i] = M[i]; }
  else if (n == 3) { Mat<3, 3> M; for (int i = 0; i < 9; ++i) M[i] = A[i]; simple_psd_fix<3>(M, tol); for (int i = 0; i < 9; ++i) A[i] = M[i]; }
}

// Boundary groups and pinned sets (reference boundary_condition.cpp:5-21,
// fixed_boundary_conditions.h:75-166, bend/stretch/twist headers).
// type: 0 null, 1 scale, 2 hang, 3 hangends, 4 stretch, 7 bend, 8 twist,
//       13 onepoint, 14 randomize (reference config.h:70-87 numbering).
void oracle_bc_fixed(int type, const double* V, int nV, int dim, int axis, double ratio,
                     int* fixed) {
  std::vector<double> mn(dim, 1e300), mx(dim, -1e300);
  for (int i = 0; i < nV; ++i)
    for (int d = 0; d < dim; ++d) {
      mn[d] = std::min(mn[d], V[i * dim + d]);
      mx[d] = std::max(mx[d], V[i * dim + d]);
    }
  double range = mx[axis] - mn[axis];
  std::vector<int> g0, g1;
  for (int i = 0; i < nV; ++i) {
    double v = V[i * dim + axis];
    if (v < mn[axis] + range * ratio) g0.push_back(i);
    else if (v > mx[axis] - range * ratio) g1.push_back(i);
  }
  for (int i = 0; i < nV; ++i) fixed[i] = 0;
  switch (type) {
    case 2: if (!g0.empty()) fixed[g0.back()] = 1; if (!g1.empty()) fixed[g1.back()] = 1; break;
    case 3: for (int i : g1) fixed[i] = 1; break;
    case 4: case 7: case 8: for (int i : g0) fixed[i] = 1; for (int i : g1) fixed[i] = 1; break;
    case 13: if (nV > 0) fixed[0] = 1; break;
    default: break;
  }
}

// Per-tet |det Dm|/6 and shape-function gradients dX[e][k][c] = dphi_k/dX_c
// (phi_1..3 rows of Dm^{-1}; phi_0 = -sum).
void oracle_tet_geometry(const double* V, const int* T, int nT, double* vols, double* dX) {
  for (int e = 0; e < nT; ++e) {
    const double* X0 = &V[3 * T[4 * e + 0]];
    M3 Dm;
    for (int k = 1; k < 4; ++k) {
      const double* Xk = &V[3 * T[4 * e + k]];
      for (int r = 0; r < 3; ++r) Dm(r, k - 1) = Xk[r] - X0[r];
    }
    vols[e] = std::fabs(det3(Dm)) / 6.0;
    M3 Di = inv3(Dm);
    for (int c = 0; c < 3; ++c) {
      double s = 0;
      for (int k = 1; k < 4; ++k) { dX[e * 12 + 3 * k + c] = Di(k - 1, c); s += Di(k - 1, c); }
      dX[e * 12 + c] = -s;
    }
  }
}

void* oracle_sim_create(const oracle_params* p) {
  OracleSim* S_ = new OracleSim();
  S_->p = *p;
#ifdef _OPENMP
  if (p->threads > 0) omp_set_num_threads(p->threads);
#endif
  int D = p->dim;
  S_->dim = D; S_->nodes = D + 1; S_->nV = p->nV; S_->nT = p->nT;
  S_->Vref.assign(p->V, p->V + (size_t)p->nV * D);
  S_->Vcur = S_->Vref;
  if (p->x0) S_->Vcur.assign(p->x0, p->x0 + (size_t)p->nV * D);
  S_->T.assign(p->T, p->T + (size_t)p->nT * S_->nodes);
  S_->mu.assign(p->mu, p->mu + p->nT);
  S_->la.assign(p->la, p->la + p->nT);
  S_->mat.assign(p->mat, p->mat + p->nT);
  S_->fixed.assign(p->fixed, p->fixed + p->nV);
  S_->free_map.assign(p->nV, -1);
  int curr = 0;
  for (int i = 0; i < p->nV; ++i) if (!S_->fixed[i]) S_->free_map[i] = curr++;
  S_->nfree = curr; S_->n = D * curr;
  // geometry
  S_->vols.resize(p->nT); S_->dX.resize((size_t)p->nT * S_->nodes * D);
  if (D == 3) {
    oracle_tet_geometry(p->V, p->T, p->nT, S_->vols.data(), S_->dX.data());
  } else {
    // doublearea (signed, reference tri2d_mesh.cpp:32-34) and barycentric gradients
    for (int e = 0; e < p->nT; ++e) {
      const double* X0 = &p->V[2 * p->T[3 * e]];
      const double* X1 = &p->V[2 * p->T[3 * e + 1]];
      const double* X2 = &p->V[2 * p->T[3 * e + 2]];
      double a = X1[0] - X0[0], bq = X2[0] - X0[0], c = X1[1] - X0[1], d = X2[1] - X0[1];
      double det = a * d - bq * c;
      S_->vols[e] = det;
      // Dm^{-1} = [d -b; -c a]/det, rows are grad phi_1, grad phi_2
      double g1x = d / det, g1y = -bq / det, g2x = -c / det, g2y = a / det;
      double* o = &S_->dX[(size_t)e * 6];
      o[2] = g1x; o[3] = g1y; o[4] = g2x; o[5] = g2y;
      o[0] = -(g1x + g2x); o[1] = -(g1y + g2y);
    }
  }
  S_->asmb.build(S_->T, p->nT, S_->nodes, D, S_->free_map, S_->nfree);
  // PMP on the assembler pattern: element mass blocks m_ab * I
  {
    int M = S_->nodes * D;
    std::vector<double> mb((size_t)p->nT * M * M, 0.0);
    double denom = (D == 3) ? 20.0 : 12.0;
    for (int e = 0; e < p->nT; ++e) {
      double m = p->density * S_->vols[e] / denom;
      for (int a = 0; a < S_->nodes; ++a)
        for (int bb = 0; bb < S_->nodes; ++bb)
          for (int r = 0; r < D; ++r)
            mb[(size_t)e * M * M + (D * a + r) + M * (D * bb + r)] = m * (a == bb ? 2.0 : 1.0);
    }
    S_->asmb.update_matrix(mb, S_->PMP);
  }
  // boundary-condition script: groups from the rest positions (the
  // BoundaryCondition constructor, boundary_condition.cpp:5-21), centers from
  // the start positions (init(V) of twist_/bend_/fixed_boundary_conditions.h)
  if (p->nV > 0) {
    const int ax = p->bc_axis;
    std::vector<double> mn(D, 1e300), mx(D, -1e300);
    for (int i = 0; i < p->nV; ++i)
      for (int d = 0; d < D; ++d) {
        mn[d] = std::min(mn[d], p->V[i * D + d]);
        mx[d] = std::max(mx[d], p->V[i * D + d]);
      }
    const double range = mx[ax] - mn[ax];
    for (int i = 0; i < p->nV; ++i) {
      const double v = p->V[i * D + ax];
      if (v < mn[ax] + range * p->bc_ratio) S_->bc_g[0].push_back(i);
      else if (v > mx[ax] - range * p->bc_ratio) S_->bc_g[1].push_back(i);
    }
    std::fill(mn.begin(), mn.end(), 1e300);
    std::fill(mx.begin(), mx.end(), -1e300);
    for (int i = 0; i < p->nV; ++i)
      for (int d = 0; d < D; ++d) {
        mn[d] = std::min(mn[d], S_->Vcur[i * D + d]);
        mx[d] = std::max(mx[d], S_->Vcur[i * D + d]);
      }
    for (int d = 0; d < D; ++d) S_->bc_center[d] = 0.5 * (mn[d] + mx[d]);
    for (int gi = 0; gi < 2; ++gi)
      if (!S_->bc_g[gi].empty())
        for (int d = 0; d < D; ++d) S_->bc_gc[gi][d] = S_->Vcur[S_->bc_g[gi].back() * D + d];
    if (p->bc_type == 1)  // ScaleBC::init
      for (int i = 0; i < p->nV; ++i)
        for (int d = 0; d < D; ++d)
          S_->Vcur[i * D + d] = (S_->Vcur[i * D + d] - S_->bc_center[d]) * 1.5 + S_->bc_center[d];
  }
  // Displacement::reset
  bdf_tables(p->bdf_order, S_->alphas, S_->beta);
  S_->h = S_->beta * p->dt;
  std::vector<double> x0 = S_->Vcur;
  std::vector<double> v0(x0.size(), 0.0);
  if (p->v0) v0.assign(p->v0, p->v0 + x0.size());
  for (int k = 0; k < p->bdf_order; ++k) { S_->x_prevs.push_front(x0); S_->v_prevs.push_front(v0); }
  S_->b.assign(x0.size(), 0.0);
  S_->x.assign(S_->n, 0.0);
  for (int v = 0; v < p->nV; ++v) {
    int f = S_->free_map[v];
    for (int d = 0; d < D; ++d) {
      if (f < 0) S_->b[D * v + d] = x0[D * v + d];
      else S_->x[D * f + d] = x0[D * v + d];
    }
  }
  S_->dx.assign(S_->n, 0.0);
  S_->f_ext.resize(x0.size());
  for (int v = 0; v < p->nV; ++v)
    for (int d = 0; d < D; ++d) S_->f_ext[D * v + d] = (double)(float)p->gravity[d];
  // variables
  int M = S_->nodes * D;
  if (p->mixed && D == 2) {
    const int nT = p->nT;
    for (int e = 0; e < nT; ++e)
      if (!material_has_s3(S_->mat[e])) {
        delete S_;
        return nullptr;  // no 3-vector (2D mixed) API
      }
    S_->R2.assign(nT, Mat<2, 2>::Identity());
    S_->S3.assign(nT, Vec<3>()); S_->g3.assign(nT, Vec<3>()); S_->gl3.assign(nT, Vec<3>());
    S_->dSdF2.assign(nT, Mat<4, 3>());
    S_->Hm3.assign(nT, Mat<3, 3>::Identity()); S_->Hinv3.assign(nT, Mat<3, 3>::Identity());
    S_->Hloc3.assign(nT, Mat<3, 3>());
    S_->s.assign(3 * nT, 0.0); S_->lam.assign(3 * nT, 0.0); S_->ds.assign(3 * nT, 0.0);
    S_->Aloc.assign((size_t)nT * M * M, 0.0);
    for (int e = 0; e < nT; ++e) {
      Mat<3, 4> dsdF;
      polar2d(S_->R2[e], S_->S3[e], def_grad2(S_, S_->Vcur, e), true, &dsdF);
      S_->dSdF2[e] = dsdF.T() * diag<3>(kSym3);
      for (int i = 0; i < 3; ++i) S_->s[3 * e + i] = S_->S3[e][i];
    }
  } else if (p->mixed) {
    int nT = p->nT;
    S_->R.assign(nT, M3::Identity());
    S_->S.assign(nT, Vec<6>()); S_->g.assign(nT, Vec<6>()); S_->gl.assign(nT, Vec<6>());
    S_->s.assign(6 * nT, 0.0); S_->lam.assign(6 * nT, 0.0); S_->ds.assign(6 * nT, 0.0);
    S_->dSdF.assign(nT, Mat<9, 6>());
    S_->Hm.assign(nT, Mat<6, 6>::Identity()); S_->Hinv.assign(nT, Mat<6, 6>::Identity());
    S_->Hloc.assign(nT, Mat<6, 6>());
    S_->Aloc.assign((size_t)nT * M * M, 0.0);
    std::vector<double> full = S_->Vcur;
    for (int e = 0; e < nT; ++e) {
      double f[9]; def_grad(S_, full, e, f);
      M3 F; for (int i = 0; i < 9; ++i) F[i] = f[i];
      Mat<6, 9> dsdF;
      polar_svd3(S_->R[e], S_->S[e], F, true, &dsdF, nullptr);
      S_->dSdF[e] = dsdF.T() * diag<6>(kSym);
      for (int i = 0; i < 6; ++i) S_->s[6 * e + i] = S_->S[e][i];
    }
  } else {
    for (int e = 0; e < p->nT; ++e)
      if (D == 3 ? !material_has_f9(S_->mat[e]) : S_->mat[e] != MAT_SNH) {
        delete S_;
        return nullptr;  // no 9-vector (3D) or 4-vector (2D, SNH only) API
      }
    S_->Hp.assign((size_t)p->nT * M * M, 0.0);
    S_->gp.assign((size_t)p->nT * M, 0.0);
  }
  return S_;
}

void oracle_sim_destroy(void* sim) { delete (OracleSim*)sim; }

int oracle_sim_step(void* sim) { return step((OracleSim*)sim); }

int oracle_sim_update_system(void* sim) { update_system((OracleSim*)sim); return 0; }

int oracle_sim_solve(void* sim) {
  double dec; return substep((OracleSim*)sim, dec);
}

double oracle_sim_energy(void* sim, double alpha) { return energy_func((OracleSim*)sim, alpha); }

void oracle_sim_stats(void* sim, oracle_stats* st) { *st = ((OracleSim*)sim)->st; }

void oracle_sim_get_positions(void* sim, double* out) {
  OracleSim* S_ = (OracleSim*)sim;
  std::copy(S_->Vcur.begin(), S_->Vcur.end(), out);
}

long oracle_sim_size(void* sim, const char* name) {
  OracleSim* S_ = (OracleSim*)sim;
  std::string k(name);
  long nT = S_->nT;
  const bool two = S_->dim == 2;
  if (k == "R") return (two ? 4 : 9) * nT;
  if (k == "S" || k == "s" || k == "lambda" || k == "ds" || k == "g" || k == "gl")
    return (two ? 3 : 6) * nT;
  if (k == "dSdF") return (two ? 12 : 54) * nT;
  if (k == "H" || k == "Hinv") return (two ? 9 : 36) * nT;
  if (k == "Aloc") return (long)S_->Aloc.size();
  if (k == "Hplain") return (long)S_->Hp.size();
  if (k == "gplain") return (long)S_->gp.size();
  if (k == "lhs") return (long)S_->lhs.size();
  if (k == "PMP") return (long)S_->PMP.size();
  if (k == "rhs" || k == "dx" || k == "x") return S_->n;
  if (k == "vols") return nT;
  if (k == "dX") return (long)S_->dX.size();
  if (k == "positions") return (long)S_->Vcur.size();
  if (k == "x_tilde") return (long)S_->Vcur.size();
  if (k == "ic_shift" || k == "pcg_error") return 1;
  return -1;
}

int oracle_sim_get(void* sim, const char* name, double* out) {
  OracleSim* S_ = (OracleSim*)sim;
  std::string k(name);
  int nT = S_->nT;
  auto put6 = [&](const std::vector<Vec<6>>& v) {
    for (int e = 0; e < nT; ++e) for (int i = 0; i < 6; ++i) out[6 * e + i] = v[e][i];
  };
  if (S_->dim == 2 && S_->p.mixed) {
    auto put = [&](const auto& v, int w) {
      for (int e = 0; e < nT; ++e) for (int i = 0; i < w; ++i) out[(size_t)w * e + i] = v[e][i];
    };
    if (k == "R") { put(S_->R2, 4); return 0; }
    if (k == "S") { put(S_->S3, 3); return 0; }
    if (k == "g") { put(S_->g3, 3); return 0; }
    if (k == "gl") { put(S_->gl3, 3); return 0; }
    if (k == "dSdF") { put(S_->dSdF2, 12); return 0; }
    if (k == "H") { put(S_->Hm3, 9); return 0; }
    if (k == "Hinv") { put(S_->Hinv3, 9); return 0; }
  }
  if (k == "R") { for (int e = 0; e < nT; ++e) for (int i = 0; i < 9; ++i) out[9 * e + i] = S_->R[e][i]; return 0; }
  if (k == "S") { put6(S_->S); return 0; }
  if (k == "g") { put6(S_->g); return 0; }
  if (k == "gl") { put6(S_->gl); return 0; }
  if (k == "s") { std::copy(S_->s.begin(), S_->s.end(), out); return 0; }
  if (k == "lambda") { std::copy(S_->lam.begin(), S_->lam.end(), out); return 0; }
  if (k == "ds") { std::copy(S_->ds.begin(), S_->ds.end(), out); return 0; }
  if (k == "dSdF") { for (int e = 0; e < nT; ++e) for (int i = 0; i < 54; ++i) out[54 * e + i] = S_->dSdF[e][i]; return 0; }
  if (k == "H") { for (int e = 0; e < nT; ++e) for (int i = 0; i < 36; ++i) out[36 * e + i] = S_->Hm[e][i]; return 0; }
  if (k == "Hinv") { for (int e = 0; e < nT; ++e) for (int i = 0; i < 36; ++i) out[36 * e + i] = S_->Hinv[e][i]; return 0; }
  if (k == "Aloc") { std::copy(S_->Aloc.begin(), S_->Aloc.end(), out); return 0; }
  if (k == "Hplain") { std::copy(S_->Hp.begin(), S_->Hp.end(), out); return 0; }
  if (k == "gplain") { std::copy(S_->gp.begin(), S_->gp.end(), out); return 0; }
  if (k == "lhs") { std::copy(S_->lhs.begin(), S_->lhs.end(), out); return 0; }
  if (k == "PMP") { std::copy(S_->PMP.begin(), S_->PMP.end(), out); return 0; }
  if (k == "rhs") { std::copy(S_->rhs.begin(), S_->rhs.end(), out); return 0; }
  if (k == "dx") { std::copy(S_->dx.begin(), S_->dx.end(), out); return 0; }
  if (k == "x") { std::copy(S_->x.begin(), S_->x.end(), out); return 0; }
  if (k == "vols") { std::copy(S_->vols.begin(), S_->vols.end(), out); return 0; }
  if (k == "dX") { std::copy(S_->dX.begin(), S_->dX.end(), out); return 0; }
  if (k == "positions") { std::copy(S_->Vcur.begin(), S_->Vcur.end(), out); return 0; }
  if (k == "x_tilde") { auto v = x_tilde(S_); std::copy(v.begin(), v.end(), out); return 0; }
  if (k == "ic_shift") { *out = S_->ic_shift_last; return 0; }
  if (k == "pcg_error") { *out = S_->pcg_err_last; return 0; }
  return -1;
}

long oracle_kkt_size(void* sim) {
  OracleSim* S_ = (OracleSim*)sim;
  return (long)S_->n + 2L * kkt_ns(S_) * S_->nT;
}

int oracle_kkt_rhs(void* sim, double* out) {
  OracleSim* S_ = (OracleSim*)sim;
  if (!S_->p.mixed) return -3;
  std::vector<double> b;
  kkt_rhs(S_, b);
  std::copy(b.begin(), b.end(), out);
  return 0;
}

int oracle_kkt_product(void* sim, const double* w, double* y) {
  OracleSim* S_ = (OracleSim*)sim;
  if (!S_->p.mixed) return -3;
  kkt_product(S_, w, y);
  return 0;
}

int oracle_kkt_precond(void* sim, int exact, const double* b, double* z) {
  OracleSim* S_ = (OracleSim*)sim;
  if (!S_->p.mixed) return -3;
  kkt_mass_setup(S_, exact != 0);
  kkt_refresh_L(S_, exact != 0);
  kkt_precond(S_, exact != 0, b, z);
  return 0;
}

double oracle_minres_dense(int N, const double* A, const double* P, const double* b, double* x,
                           int max_iters, double tol, int* iters) {
  auto mv = [N](const double* M, const double* v, double* y) {
    for (int i = 0; i < N; ++i) {
      double a = 0;
      for (int j = 0; j < N; ++j) a += M[(size_t)i * N + j] * v[j];
      y[i] = a;
    }
  };
  return minres(
      N, [&](const double* w, double* y) { mv(A, w, y); },
      [&](const double* r, double* z) { mv(P, r, z); }, b, x, max_iters, tol, iters);
}

int oracle_sim_get_int(void* sim, const char* name, int* out) {
  OracleSim* S_ = (OracleSim*)sim;
  std::string k(name);
  if (k == "rowptr") { std::copy(S_->asmb.rowptr.begin(), S_->asmb.rowptr.end(), out); return 0; }
  if (k == "colind") { std::copy(S_->asmb.colind.begin(), S_->asmb.colind.end(), out); return 0; }
  if (k == "free_map") { std::copy(S_->free_map.begin(), S_->free_map.end(), out); return 0; }
  if (k == "pcg_iters") { *out = S_->pcg_iters_last; return 0; }
  if (k == "ic_colors") { *out = S_->ic_colors_last; return 0; }
  return -1;
}

// The last step's Newton iterations, 5 values each (iteration, energy,
// energy res, Newton dec, alpha); returns the row count (out may be NULL).
int oracle_sim_history(void* sim, double* out) {
  const auto& h = ((OracleSim*)sim)->hist;
  if (out)
    for (size_t k = 0; k < h.size(); ++k) std::copy(h[k].begin(), h[k].end(), out + 5 * k);
  return (int)h.size();
}

// "eigen-pcg-IC" preconditioner on a scalar CSR matrix with bs x bs vertex
// blocks: the factor as a dense n x n matrix in the original numbering
// (L_orig = P^T L P, lower triangular in the colour order), the scaling,
// the final shift, the vertex colours and the number of attempts.
int oracle_ic_factor_dense(int n, int bs, const int* rowptr, const int* colind, const double* vals,
                           double* L, double* scale, double* shift, int* color, int* attempts) {
  IcFactor f;
  int rc = ic_factor(f, n, bs, rowptr, colind, vals);
  *attempts = f.attempts;
  *shift = f.shift;
  for (int v = 0; v < f.nb; ++v) color[v] = f.color[v];
  if (rc != 0) return rc;
  std::copy(f.scale.begin(), f.scale.end(), scale);
  std::fill(L, L + (size_t)n * n, 0.0);
  auto orig = [&](int p) { return f.order[p / bs] * bs + p % bs; };
  for (int p = 0; p < n; ++p) {
    L[(size_t)orig(p) * n + orig(p)] = f.ldiag[p];
    for (int k = f.lptr[p]; k < f.lptr[p + 1]; ++k)
      L[(size_t)orig(p) * n + orig(f.lcol[k])] = f.lval[k];
  }
  return 0;
}

// Eigen's CG with its default DiagonalPreconditioner (reference
// eigen_iterative_solver.h:14-29, "eigen-pcg-diag") on a scalar CSR matrix:
// x0 = 0, z = D^-1 r with D^-1_jj = 1 / a_jj, or 1 where the diagonal entry is
// zero or absent. Returns the iteration count; err = |r| / |b|.
int oracle_jacobi_pcg(int n, const int* rowptr, const int* colind, const double* vals,
                      const double* b, double* x, double tol, int max_iters, double* err) {
  std::vector<double> invd(n, 1.0);
  for (int i = 0; i < n; ++i)
    for (int k = rowptr[i]; k < rowptr[i + 1]; ++k)
      if (colind[k] == i) {
        if (vals[k] != 0.0) invd[i] = 1.0 / vals[k];
        break;
      }
  auto mult = [&](const std::vector<double>& v, std::vector<double>& y) {
    y.assign(n, 0.0);
    for (int i = 0; i < n; ++i) {
      double a = 0;
      for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) a += vals[k] * v[colind[k]];
      y[i] = a;
    }
  };
  auto pre = [&](const double* r, double* z) {
    for (int i = 0; i < n; ++i) z[i] = invd[i] * r[i];
  };
  return cg_loop(n, mult, pre, b, x, tol, max_iters, err);
}

// The envelope Cholesky solve of a scalar CSR matrix with bs x bs blocks
// (0, or -1 when the matrix is not positive definite); perm receives the RCM
// vertex order (n / bs entries) when not NULL.
int oracle_envelope_solve(int n, int bs, const int* rowptr, const int* colind, const double* vals,
                          const double* b, double* x, int* perm) {
  if (perm) {
    const auto p = rcm_order(n / bs, bs, rowptr, colind);
    std::copy(p.begin(), p.end(), perm);
  }
  return envelope_llt_solve(n, bs, rowptr, colind, vals, b, x);
}

// Eigen's CG with that preconditioner on a scalar CSR matrix (x0 = 0).
// Returns the iteration count, or a negative code when the factor fails.
int oracle_ic_pcg(int n, int bs, const int* rowptr, const int* colind, const double* vals,
                  const double* b, double* x, double tol, int max_iters, double* err) {
  IcFactor f;
  int rc = ic_factor(f, n, bs, rowptr, colind, vals);
  if (rc != 0) return rc;
  auto mult = [&](const std::vector<double>& v, std::vector<double>& y) {
    y.assign(n, 0.0);
    for (int i = 0; i < n; ++i) {
      double a = 0;
      for (int k = rowptr[i]; k < rowptr[i + 1]; ++k) a += vals[k] * v[colind[k]];
      y[i] = a;
    }
  };
  auto pre = [&](const double* r, double* z) { ic_apply(f, r, z); };
  return cg_loop(n, mult, pre, b, x, tol, max_iters, err);
}

}  // extern "C"
