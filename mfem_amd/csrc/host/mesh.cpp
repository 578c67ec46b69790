#include "mesh.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>

namespace mfem {

Mesh::Mesh(const std::vector<double>& V, const std::vector<int>& T,
           const std::vector<int>& material_ids,
           const std::vector<std::shared_ptr<MaterialModel>>& materials)
    : V_(V), Vref_(V), Vinit_(V), T_(T) {
  if (materials.empty()) throw std::runtime_error("Mesh: no materials");
  material_ = materials[0];
  nV_ = (int)(V.size() / 3);
  nT_ = (int)(T.size() / 4);
  if ((int)material_ids.size() != nT_) throw std::runtime_error("Mesh: material_ids size");
  for (int i = 0; i < nT_; ++i) {
    int id = material_ids[i];
    if (id < 0 || id >= (int)materials.size()) throw std::runtime_error("Mesh: bad material id");
    elements_.emplace_back(materials[id]);
  }
  initial_velocity_.assign(V.size(), 0.0);
}

Mesh::Mesh(const std::vector<double>& V, const std::vector<int>& T,
           std::shared_ptr<MaterialModel> material)
    : V_(V), Vref_(V), Vinit_(V), T_(T), material_(material) {
  nV_ = (int)(V.size() / 3);
  nT_ = (int)(T.size() / 4);
  for (int i = 0; i < nT_; ++i) elements_.emplace_back(material);
  initial_velocity_.assign(V.size(), 0.0);
}

Mesh::Mesh(const std::vector<double>& V, const std::vector<int>& T,
           std::shared_ptr<MaterialModel> material, const std::vector<double>& mu,
           const std::vector<double>& la, const std::vector<int>& mat, int dim)
    : V_(V), Vref_(V), Vinit_(V), T_(T), material_(material), elem_mu_(mu), elem_la_(la),
      elem_mat_(mat) {
  if (dim != 2 && dim != 3) throw std::runtime_error("Mesh: dim must be 2 or 3");
  dim_ = dim;
  nodes_ = dim + 1;
  nV_ = (int)(V.size() / 3);
  nT_ = (int)(T.size() / nodes_);
  if ((int)mu.size() != nT_ || (int)la.size() != nT_ || (int)mat.size() != nT_)
    throw std::runtime_error("Mesh: per-element parameter arrays must have nT entries");
  initial_velocity_.assign(V.size(), 0.0);
}

std::vector<int> boundary_fixed(const BoundaryConditionConfig& bc, const std::vector<double>& V,
                                int dim) {
  const int nV = (int)(V.size() / dim);
  std::vector<int> fixed(nV, 0);
  if (nV == 0) return fixed;
  const int ax = bc.axis;
  double mn = V[ax], mx = V[ax];
  for (int i = 0; i < nV; ++i) {
    mn = std::min(mn, V[i * dim + ax]);
    mx = std::max(mx, V[i * dim + ax]);
  }
  const double range = mx - mn;
  std::vector<int> g0, g1;
  for (int i = 0; i < nV; ++i) {
    double v = V[i * dim + ax];
    if (v < mn + range * bc.ratio) g0.push_back(i);
    else if (v > mx - range * bc.ratio) g1.push_back(i);
  }
  switch (bc.type) {
    case BC_HANG:
      if (!g0.empty()) fixed[g0.back()] = 1;
      if (!g1.empty()) fixed[g1.back()] = 1;
      break;
    case BC_HANGENDS:
      for (int i : g1) fixed[i] = 1;
      break;
    case BC_STRETCH:
    case BC_BEND:
    case BC_TWIST:
      for (int i : g0) fixed[i] = 1;
      for (int i : g1) fixed[i] = 1;
      break;
    case BC_ONEPOINT:
      fixed[0] = 1;
      break;
    default:
      break;
  }
  return fixed;
}

void Mesh::init(hipStream_t st, const std::vector<int>& fixed_override) {
  V_ = Vinit_;
  for (int v : T_)
    if (v < 0 || v >= nV_) throw std::runtime_error("Mesh: tet vertex index out of range");
  // pins from the rest positions: the reference builds its boundary
  // condition on Vref_ (mesh.cpp:84) and only init() sees the start positions
  is_fixed_ = fixed_override.empty() ? boundary_fixed(bc_config_, Vref_, 3) : fixed_override;
  if ((int)is_fixed_.size() != nV_) throw std::runtime_error("Mesh: pin mask size");
  // boundary-condition scripts (BoundaryCondition constructor + init,
  // boundary_condition.cpp:5-21 and the *_boundary_condition.h headers)
  bc_groups_.assign(2, {});
  bc_group_center_.clear();
  if (nV_ > 0) {
    // groups from the rest positions (the constructor's V), centers from the
    // start positions (init's V)
    const int ax = bc_config_.axis;
    auto bounds = [&](const std::vector<double>& X, double mn[3], double mx[3]) {
      for (int d = 0; d < 3; ++d) mn[d] = mx[d] = X[d];
      for (int i = 0; i < nV_; ++i)
        for (int d = 0; d < 3; ++d) {
          mn[d] = std::min(mn[d], X[3 * i + d]);
          mx[d] = std::max(mx[d], X[3 * i + d]);
        }
    };
    double mn[3], mx[3];
    bounds(Vref_, mn, mx);
    const double range = mx[ax] - mn[ax];
    for (int i = 0; i < nV_; ++i) {
      const double v = Vref_[3 * i + ax];
      if (v < mn[ax] + range * bc_config_.ratio) bc_groups_[0].push_back(i);
      else if (v > mx[ax] - range * bc_config_.ratio) bc_groups_[1].push_back(i);
    }
    bounds(V_, mn, mx);
    for (int d = 0; d < 3; ++d) bc_center_[d] = 0.5 * (mn[d] + mx[d]);
    for (const auto& g : bc_groups_) {
      std::array<double, 3> c{0, 0, 0};
      if (!g.empty())
        for (int d = 0; d < 3; ++d) c[d] = V_[3 * g.back() + d];
      bc_group_center_.push_back(c);
    }
    if (bc_config_.type == BC_SCALEF) {
      // ScaleBC::init: the start configuration is the mesh scaled by 1.5
      // about its bounding-box center (no pins)
      for (int i = 0; i < nV_; ++i)
        for (int d = 0; d < dim_; ++d)
          V_[3 * i + d] = (V_[3 * i + d] - bc_center_[d]) * 1.5 + bc_center_[d];
    } else if (bc_config_.type == BC_RANDOM) {
      throw std::runtime_error("the \"randomize\" boundary condition (C rand() positions) has no "
                               "reproducible device counterpart");
    }
  }
  pinned_.clear();
  for (int i = 0; i < nV_; ++i)
    if (is_fixed_[i]) pinned_.push_back(i);
  free_map_.assign(nV_, -1);
  nfree_ = 0;
  for (int i = 0; i < nV_; ++i)
    if (!is_fixed_[i]) free_map_[i] = nfree_++;
  vols_.assign(nT_, 0.0);
  dX_.assign((size_t)nT_ * 12, 0.0);
  if (dim_ == 2) {
    // triangles (reference tri2d_mesh.cpp:15-41): vols = igl::doublearea,
    // signed (twice the area), dphi/dX = rows of Dm^{-1} (phi_0 = -sum), in
    // rows 3k + c of the 12-value element layout
    for (int e = 0; e < nT_; ++e) {
      const double* X0 = &Vref_[3 * T_[3 * e]];
      const double* X1 = &Vref_[3 * T_[3 * e + 1]];
      const double* X2 = &Vref_[3 * T_[3 * e + 2]];
      const double a = X1[0] - X0[0], b = X2[0] - X0[0], c = X1[1] - X0[1], d = X2[1] - X0[1];
      const double det = a * d - b * c;
      if (det == 0.0) throw std::runtime_error("Mesh: degenerate triangle " + std::to_string(e));
      vols_[e] = det;
      const double g1x = d / det, g1y = -b / det, g2x = -c / det, g2y = a / det;
      double* o = &dX_[(size_t)e * 12];
      o[3] = g1x; o[4] = g1y; o[6] = g2x; o[7] = g2y;
      o[0] = -(g1x + g2x); o[1] = -(g1y + g2y);
    }
  }
  // tets: volumes |det Dm| / 6 and dphi/dX = rows of Dm^{-1} (phi_0 = -sum)
  for (int e = 0; e < (dim_ == 3 ? nT_ : 0); ++e) {
    const double* X0 = &Vref_[3 * T_[4 * e]];
    double D[9];
    for (int k = 1; k < 4; ++k) {
      const double* Xk = &Vref_[3 * T_[4 * e + k]];
      for (int r = 0; r < 3; ++r) D[r + 3 * (k - 1)] = Xk[r] - X0[r];
    }
    double c00 = D[4] * D[8] - D[7] * D[5], c01 = D[7] * D[2] - D[1] * D[8],
           c02 = D[1] * D[5] - D[4] * D[2];
    double det = D[0] * c00 + D[3] * c01 + D[6] * c02;
    if (det == 0.0) throw std::runtime_error("Mesh: degenerate tetrahedron " + std::to_string(e));
    vols_[e] = std::fabs(det) / 6.0;
    // inverse (row i of Dm^{-1} is grad phi_{i+1})
    double inv[9];
    inv[0] = c00 / det;
    inv[3] = (D[6] * D[5] - D[3] * D[8]) / det;
    inv[6] = (D[3] * D[7] - D[6] * D[4]) / det;
    inv[1] = c01 / det;
    inv[4] = (D[0] * D[8] - D[6] * D[2]) / det;
    inv[7] = (D[6] * D[1] - D[0] * D[7]) / det;
    inv[2] = c02 / det;
    inv[5] = (D[3] * D[2] - D[0] * D[5]) / det;
    inv[8] = (D[0] * D[4] - D[3] * D[1]) / det;
    for (int c = 0; c < 3; ++c) {
      double s = 0;
      for (int k = 1; k < 4; ++k) {
        double v = inv[(k - 1) + 3 * c];
        dX_[(size_t)e * 12 + 3 * k + c] = v;
        s += v;
      }
      dX_[(size_t)e * 12 + c] = -s;
    }
  }
  build_maps();
  // upload
  std::vector<int4> T4(nT_);
  const int N = nodes_;
  for (int e = 0; e < nT_; ++e)
    T4[e] = make_int4(T_[N * e], T_[N * e + 1], T_[N * e + 2], N == 4 ? T_[N * e + 3] : T_[N * e]);
  d_T.upload(T4, st);
  std::vector<double> dXs((size_t)nT_ * 12);
  for (int e = 0; e < nT_; ++e)
    for (int i = 0; i < 12; ++i) dXs[(size_t)i * nT_ + e] = dX_[(size_t)e * 12 + i];
  d_dX.upload(dXs, st);
  d_vol.upload(vols_, st);
  std::vector<double> mu(nT_), la(nT_);
  std::vector<int> mat(nT_);
  for (int e = 0; e < nT_; ++e) {
    MaterialModelType t;
    if (!elem_mat_.empty()) {
      t = (MaterialModelType)elem_mat_[e];
      mu[e] = elem_mu_[e];
      la[e] = elem_la_[e];
    } else {
      const auto& c = elements_[e].material_->config();
      t = c->material_model;
      mu[e] = c->mu;
      la[e] = c->la;
    }
    if (!material_supported_on_device(t))
      throw std::runtime_error("material '" + MaterialModelFactory().name_by_type(t) +
                               "' has no device kernel");
    mat[e] = (int)t;
  }
  elem_type_ = mat;
  d_mu.upload(mu, st);
  d_la.upload(la, st);
  d_mat.upload(mat, st);
  d_free_map.upload(free_map_, st);
  d_rowptr.upload(rowptr_, st);
  d_col.upload(col_, st);
  d_slot_ptr.upload(slot_ptr_, st);
  d_slots.upload(slots_, st);
  d_upper.upload(upper_, st);
  d_tpos.upload(tpos_, st);
  d_diag_block.upload(diag_block_, st);
  d_inc_ptr.upload(inc_ptr_, st);
  d_inc.upload(inc_, st);
  d_asm_rows.upload(asm_rows_, st);
  d_mass.upload(mass_, st);
  MFEM_HIP_CHECK(hipStreamSynchronize(st));
}

// Scatter maps (reference sparse_utils.cpp:19-138): free-free vertex pairs of
// every tet, sorted by (row, col) with ties in (tet, local pair) order.
void Mesh::build_maps() {
  // vertex -> (tet, corner) incidences of free vertices, in tet order
  inc_ptr_.assign(nfree_ + 1, 0);
  const int N = nodes_;
  for (int e = 0; e < nT_; ++e)
    for (int a = 0; a < N; ++a) {
      int f = free_map_[T_[N * e + a]];
      if (f >= 0) inc_ptr_[f + 1]++;
    }
  for (int i = 0; i < nfree_; ++i) inc_ptr_[i + 1] += inc_ptr_[i];
  inc_.assign(inc_ptr_[nfree_], 0);
  {
    std::vector<int> fill(inc_ptr_.begin(), inc_ptr_.end() - 1);
    for (int e = 0; e < nT_; ++e)
      for (int a = 0; a < N; ++a) {
        int f = free_map_[T_[N * e + a]];
        if (f >= 0) inc_[fill[f]++] = (e << 2) | a;
      }
  }
  if (nT_ >= (1 << 27)) throw std::runtime_error("Mesh: too many tets for 32-bit slot ids");
  // consistent mass rho vol / 20 (1 + delta) for tets; rho vol / 12
  // (1 + delta) for triangles with vol the signed doublearea
  const double denom = dim_ == 3 ? 20.0 : 12.0;
  const double rho = material_->config()->density;
  rowptr_.assign(nfree_ + 1, 0);
  col_.clear();
  slot_ptr_.assign(1, 0);
  slots_.clear();
  mass_.clear();
  diag_block_.assign(nfree_, -1);
  struct Ent { int col, e, b; };
  std::vector<Ent> ents;
  for (int i = 0; i < nfree_; ++i) {
    ents.clear();
    for (int p = inc_ptr_[i]; p < inc_ptr_[i + 1]; ++p) {
      int e = inc_[p] >> 2;
      for (int b = 0; b < N; ++b) {
        int fb = free_map_[T_[N * e + b]];
        if (fb >= 0) ents.push_back({fb, e, b});
      }
    }
    std::stable_sort(ents.begin(), ents.end(), [](const Ent& x, const Ent& y) {
      if (x.col != y.col) return x.col < y.col;
      if (x.e != y.e) return x.e < y.e;
      return x.b < y.b;
    });
    size_t k = 0;
    while (k < ents.size()) {
      int c = ents[k].col;
      double m = 0;
      if (c == i) diag_block_[i] = (int)col_.size();
      while (k < ents.size() && ents[k].col == c) {
        int e = ents[k].e;
        // local index of row vertex i in tet e
        int a = 0;
        while (free_map_[T_[N * e + a]] != i) ++a;
        slots_.push_back((e << 4) | (a << 2) | ents[k].b);
        m += rho * vols_[e] / denom * (a == ents[k].b ? 2.0 : 1.0);
        ++k;
      }
      col_.push_back(c);
      mass_.push_back(m);
      slot_ptr_.push_back((int)slots_.size());
    }
    rowptr_[i + 1] = (int)col_.size();
    if (diag_block_[i] < 0)
      throw std::runtime_error("Mesh: free vertex not referenced by any tetrahedron (singular system)");
  }
  nnzb_ = (int)col_.size();
  // blocks j >= i and the positions of their transposes (j, i): the pattern
  // is the symmetric vertex adjacency, and the slots of (j, i) are those of
  // (i, j) transposed, in the same (tet, local pair) order
  upper_.clear();
  tpos_.clear();
  std::vector<int> upper_row;
  for (int i = 0; i < nfree_; ++i)
    for (int p = rowptr_[i]; p < rowptr_[i + 1]; ++p) {
      const int j = col_[p];
      if (j < i) continue;
      const auto b = col_.begin() + rowptr_[j], e = col_.begin() + rowptr_[j + 1];
      const auto it = std::lower_bound(b, e, i);
      if (it == e || *it != i) throw std::runtime_error("Mesh: block pattern is not symmetric");
      upper_.push_back(p);
      tpos_.push_back((int)(it - col_.begin()));
      upper_row.push_back(i);
    }
  // Traversal order of the assembly gathers: rows by the Z-order (Morton)
  // code of their vertex's rest position. A tet's record is read by the rows
  // of its four corners; in row order those rows lie up to a cross-section
  // of the mesh apart, so the record has left the L2 before its last read,
  // while in Z-order most corners of a tet are visited close together. Only
  // which lane sums which block changes, not the order of any sum.
  asm_rows_.resize(nfree_);
  {
    std::vector<int> vert(nfree_);
    for (int v = 0; v < nV_; ++v)
      if (free_map_[v] >= 0) vert[free_map_[v]] = v;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int f = 0; f < nfree_; ++f)
      for (int c = 0; c < 3; ++c) {
        lo[c] = std::min(lo[c], V_[3 * (size_t)vert[f] + c]);
        hi[c] = std::max(hi[c], V_[3 * (size_t)vert[f] + c]);
      }
    constexpr int kBits = 21;
    std::vector<unsigned long long> code(nfree_);
    for (int f = 0; f < nfree_; ++f) {
      unsigned long long k = 0;
      for (int c = 0; c < 3; ++c) {
        const double ext = hi[c] - lo[c];
        const double u = ext > 0 ? (V_[3 * (size_t)vert[f] + c] - lo[c]) / ext : 0.0;
        const auto q = (unsigned long long)std::min<double>(
            std::max(0.0, u * (double)((1 << kBits) - 1)), (double)((1 << kBits) - 1));
        for (int bit = 0; bit < kBits; ++bit) k |= ((q >> bit) & 1ULL) << (3 * bit + c);
      }
      code[f] = k;
    }
    for (int f = 0; f < nfree_; ++f) asm_rows_[f] = f;
    std::stable_sort(asm_rows_.begin(), asm_rows_.end(),
                     [&](int a, int b) { return code[a] < code[b]; });
    std::vector<int> rank(nfree_);
    for (int r = 0; r < nfree_; ++r) rank[asm_rows_[r]] = r;
    std::vector<int> idx(upper_.size());
    for (size_t t = 0; t < idx.size(); ++t) idx[t] = (int)t;
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int a, int b) { return rank[upper_row[a]] < rank[upper_row[b]]; });
    std::vector<int> u2(idx.size()), t2(idx.size());
    for (size_t t = 0; t < idx.size(); ++t) {
      u2[t] = upper_[idx[t]];
      t2[t] = tpos_[idx[t]];
    }
    upper_.swap(u2);
    tpos_.swap(t2);
  }
}

bool Mesh::bc_moves() const {
  const auto t = bc_config_.type;
  return t == BC_STRETCH || t == BC_TWIST || t == BC_BEND;
}

// StretchBC / TwistBC / BendBC::step: group i moves with (-1)^i times the
// configured velocity: along the axis (stretch), or rotating by
// (-1)^i velocity pi dt about the x axis through the bounding-box center
// (twist; about z in 2D) or about the z axis through the group's last
// vertex (bend). Rotations act on the current positions.
void Mesh::update_bcs(double dt) {
  if (!bc_moves()) return;
  const auto t = bc_config_.type;
  for (size_t gi = 0; gi < bc_groups_.size(); ++gi) {
    const double sgn = gi % 2 == 0 ? 1.0 : -1.0;
    if (t == BC_STRETCH) {
      const double dv = sgn * bc_config_.velocity * dt;
      for (int j : bc_groups_[gi]) V_[3 * j + bc_config_.axis] += dv;
      continue;
    }
    const double a = sgn * bc_config_.velocity * M_PI * dt;
    const double c = std::cos(a), sn = std::sin(a);
    const double* o = t == BC_TWIST ? bc_center_ : bc_group_center_[gi].data();
    // rotation plane: (y, z) for the 3D twist, (x, y) otherwise
    const int p0 = (t == BC_TWIST && dim_ == 3) ? 1 : 0;
    const int p1 = p0 + 1;
    for (int j : bc_groups_[gi]) {
      const double u = V_[3 * j + p0] - o[p0], w = V_[3 * j + p1] - o[p1];
      V_[3 * j + p0] = c * u - sn * w + o[p0];
      V_[3 * j + p1] = sn * u + c * w + o[p1];
    }
  }
}

std::vector<double> Mesh::pinned_positions() const {
  std::vector<double> out(3 * pinned_.size());
  for (size_t k = 0; k < pinned_.size(); ++k)
    for (int d = 0; d < 3; ++d) out[3 * k + d] = V_[3 * pinned_[k] + d];
  return out;
}

mfemhip::TetMeshDev Mesh::dev() const {
  mfemhip::TetMeshDev m;
  m.nV = nV_;
  m.nT = nT_;
  m.nfree = nfree_;
  m.dim = dim_;
  m.T = d_T.data();
  m.dX = d_dX.data();
  m.vol = d_vol.data();
  m.mu = d_mu.data();
  m.la = d_la.data();
  m.mat = d_mat.data();
  m.free_map = d_free_map.data();
  m.density = material_->config()->density;
  return m;
}

}  // namespace mfem
