// Simulation configuration: the enums, defaults and Young's-modulus
// conversion of the reference (src/config.h:8-142), plus this backend's
// device options (precision mode, PCG polling cadence).
#pragma once
#include <set>
#include <string>

namespace mfem {

// reference src/config.h:13-16
constexpr void Enu_to_lame(double E, double nu, double& lambda, double& mu) {
  mu = E / (2.0 * (1.0 + nu));
  lambda = (E * nu) / ((1.0 + nu) * (1.0 - 2.0 * nu));
}

enum VariableType {
  VAR_DISPLACEMENT,
  VAR_STRETCH,
  VAR_COLLISION,
  VAR_FRICTION,
  VAR_MIXED_STRETCH,
  VAR_MIXED_COLLISION
};

enum OptimizerType {
  OPTIMIZER_ALM,
  OPTIMIZER_ADMM,
  OPTIMIZER_SQP,
  OPTIMIZER_SQP_PD,
  OPTIMIZER_NEWTON,
  OPTIMIZER_SQP_BENDING,
};

enum TimeIntegratorType { TI_BDF1, TI_BDF2, TI_BDF3, TI_BDF4, TI_BDF5, TI_BDF6 };

enum LinearSolverType {
  SOLVER_EIGEN_LLT,
  SOLVER_EIGEN_LDLT,
  SOLVER_EIGEN_LU,
  SOLVER_CHOLMOD,
  SOLVER_AFFINE_PCG,
  SOLVER_EIGEN_CG_DIAG,
  SOLVER_EIGEN_CG_IC,
  SOLVER_EIGEN_GS,
  SOLVER_MINRES_ID,
  SOLVER_MINRES_ADMM,
  SOLVER_ADMM,
  SOLVER_SUBSPACE,
  SOLVER_AMGCL
};

enum MaterialModelType {
  MATERIAL_SNH,
  MATERIAL_NH,
  MATERIAL_COROT,
  MATERIAL_FCR,
  MATERIAL_ARAP,
  MATERIAL_FUNG
};

enum BCScriptType {
  BC_NULL,
  BC_SCALEF,
  BC_HANG,
  BC_HANGENDS,
  BC_STRETCH,
  BC_SQUASH,
  BC_STRETCHNSQUASH,
  BC_BEND,
  BC_TWIST,
  BC_TWISTNSTRETCH,
  BC_TWISTNSNS,
  BC_TWISTNSNS_OLD,
  BC_RUBBERBANDPULL,
  BC_ONEPOINT,
  BC_RANDOM,
  BC_FALL,
};

// Element math / matrix precision of the device path.
enum class Precision {
  FP64,   // every kernel in FP64
  MIXED,  // FP32 PSD projection, Hessian inverse, stiffness blocks, K, SpMV
          // and inner CG; FP64 polar fit, stretch, gradients, residuals,
          // energies, dot products and an FP64 defect-correction loop
};

// reference src/config.h:91-124 (defaults kept)
struct SimConfig {
  double h = 0.034;
  float ext[3] = {0., -9.8f, 0.};
  bool show_timing = true;
  bool show_data = true;
  bool save_substeps = false;
  int outer_steps = 5;
  int ls_iters = 20;
  double mu = 0.5;
  double espv = 1e-3;
  double kappa = 10.0;
  double max_kappa = 1e6;
  double constraint_tol = 1e-2;
  bool enable_ccd = false;
  double dhat = 1e-2;
  double update_zone_tol = 1e-1;
  double newton_tol = 1e-10;
  double ls_tol = 1e-4;
  OptimizerType optimizer = OPTIMIZER_NEWTON;
  int max_iterative_solver_iters = 500;
  double itr_tol = 1e-4;
  LinearSolverType solver_type = SOLVER_EIGEN_LLT;
  TimeIntegratorType ti_type = TI_BDF1;
  std::set<VariableType> variables = {};
  std::set<VariableType> mixed_variables = {VAR_MIXED_STRETCH, VAR_MIXED_COLLISION};
  // device backend options
  Precision precision = Precision::FP64;
  int pcg_poll = 16;         // CG iterations launched between host polls (the
                             // Jacobi CG keeps a quarter as many queued ahead)
  double inner_tol = 1e-4;   // relative tolerance of each FP32 inner CG solve (MIXED)
  int max_refinements = 50;  // defect-correction sweeps (MIXED)
  int spmv_variant = -1;     // BSR SpMV shape; -1: autotune on the first solve
};

// reference src/config.h:127-135
struct MaterialConfig {
  MaterialModelType material_model = MATERIAL_SNH;
  double ym = 1e6;
  double pr = 0.45;
  double mu = ym / (2.0 * (1.0 + pr));
  double la = (ym * pr) / ((1.0 + pr) * (1.0 - 2.0 * pr));
  double density = 1000.0;
  double thickness = 1e-3;
};

// reference src/config.h:137-142
struct BoundaryConditionConfig {
  BCScriptType type = BCScriptType::BC_NULL;
  double ratio = 0.1;
  int axis = 0;
  double velocity = 0.1;
};

}  // namespace mfem
