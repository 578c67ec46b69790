// Tetrahedral mesh: setup-time quantities of the reference's Mesh /
// TetrahedralMesh / Assembler (src/mesh/mesh.cpp:44-95, tet_mesh.cpp:16-80,
// src/utils/sparse_utils.cpp:19-138), built once on the host and uploaded.
//
// Device layout (structure of arrays, stride nT): T as int4, dX as 12 rows of
// dphi_k/dX_c, vol / mu / la / material id per tet; free_map per vertex.
// Assembly maps: BSR(3x3) pattern over free vertices; for every nonzero block
// its sorted (tet, local pair) slot list and its consistent-mass coefficient;
// for every free vertex its sorted (tet, corner) incidence list.
#pragma once
#include <array>
#include <memory>
#include <vector>

#include "../kernels.h"
#include "config.h"
#include "device.h"
#include "material_model.h"

namespace mfem {

struct Element {
  explicit Element(std::shared_ptr<MaterialModel> m) : material_(std::move(m)) {}
  std::shared_ptr<MaterialModel> material_;
};

class Mesh {
 public:
  // V: nV x 3 row-major, T: nT x 4 row-major.
  Mesh(const std::vector<double>& V, const std::vector<int>& T, const std::vector<int>& material_ids,
       const std::vector<std::shared_ptr<MaterialModel>>& materials);
  Mesh(const std::vector<double>& V, const std::vector<int>& T,
       std::shared_ptr<MaterialModel> material);
  // Per-element Lame parameters and model ids given directly (one material
  // per tet, as the reference's material_ids constructor with one model per
  // id); `material` supplies the density (reference tet_mesh.cpp:36-41).
  // dim 2: triangles (T holds 3 corners per element) in the z = 0 plane, V
  // still given as nV x 3 with z = 0 (reference Tri2DMesh, tri2d_mesh.cpp).
  Mesh(const std::vector<double>& V, const std::vector<int>& T, std::shared_ptr<MaterialModel> material,
       const std::vector<double>& mu, const std::vector<double>& la, const std::vector<int>& mat,
       int dim = 3);

  // BCs, volumes, shape gradients, assembly maps, upload (reference
  // Mesh::init, mesh.cpp:44-79). `fixed_override`, when non-empty, replaces
  // the boundary-condition pin set (one flag per vertex).
  void init(hipStream_t st, const std::vector<int>& fixed_override = {});

  mfemhip::TetMeshDev dev() const;

  // Scripted boundary motion (reference boundary_conditions/{stretch,twist,
  // bend}_boundary_condition.h, applied after every timestep by
  // Mesh::update_bcs, mesh.h:130-132): moves the pinned rows of V_.
  bool bc_moves() const;
  void update_bcs(double dt);
  // pinned vertex ids (ascending) and their current positions (3 each)
  const std::vector<int>& pinned() const { return pinned_; }
  std::vector<double> pinned_positions() const;
  int dim() const { return dim_; }
  int nodes() const { return nodes_; }  // corners per element
  int nV() const { return nV_; }
  int nT() const { return nT_; }
  int nfree() const { return nfree_; }
  int nnzb() const { return nnzb_; }
  int nupper() const { return (int)upper_.size(); }  // blocks j >= i of the pattern

  // host state (reference Mesh public members)
  std::vector<double> V_, Vref_, Vinit_, initial_velocity_;
  std::vector<int> T_;
  std::vector<int> is_fixed_, free_map_;
  BoundaryConditionConfig bc_config_;
  std::vector<Element> elements_;
  std::shared_ptr<MaterialModel> material_;
  std::vector<double> vols_, dX_;  // host copies (dX_[e*12 + 3k + c])
  std::vector<double> elem_mu_, elem_la_;  // per-element parameters (when given directly)
  std::vector<int> elem_mat_;

  // device state
  DeviceArray<int4> d_T;
  DeviceArray<double> d_dX, d_vol, d_mu, d_la;
  DeviceArray<int> d_mat, d_free_map;
  DeviceArray<int> d_rowptr, d_col, d_slot_ptr, d_slots, d_diag_block, d_inc_ptr, d_inc;
  DeviceArray<int> d_upper, d_tpos;  // upper blocks and the positions of their transposes
  // free rows in the order the assembly gathers visit them (Z-order of the
  // rest positions; the upper blocks are listed in the same row order)
  DeviceArray<int> d_asm_rows;
  DeviceArray<double> d_mass;
  std::vector<int> rowptr_, col_;  // host copy of the BSR pattern
  // material model id of every element (after init)
  const std::vector<int>& element_types() const { return elem_type_; }
  // consistent-mass coefficient per block, and the diagonal block of each row
  const std::vector<double>& mass() const { return mass_; }
  const std::vector<int>& diag_block() const { return diag_block_; }

 private:
  void build_maps();
  int dim_ = 3, nodes_ = 4;
  int nV_ = 0, nT_ = 0, nfree_ = 0, nnzb_ = 0;
  std::vector<int> slot_ptr_, slots_, diag_block_, inc_ptr_, inc_;
  std::vector<int> upper_, tpos_;  // pattern positions of the blocks j >= i; of their transposes
  std::vector<int> asm_rows_;      // assembly visiting order of the free rows
  std::vector<double> mass_;
  std::vector<int> elem_type_;
  std::vector<int> pinned_;
  std::vector<std::vector<int>> bc_groups_;   // min and max end groups along the axis
  double bc_center_[3] = {0, 0, 0};           // twist: bounding-box center
  std::vector<std::array<double, 3>> bc_group_center_;  // bend: rotation center per group
};

// Pinned set of the reference's boundary conditions
// (boundary_condition.cpp:5-21, fixed_boundary_conditions.h:75-166,
// bend/stretch/twist_boundary_condition.h).
std::vector<int> boundary_fixed(const BoundaryConditionConfig& bc, const std::vector<double>& V,
                                int dim);

}  // namespace mfem
