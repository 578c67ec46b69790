// ORACLE — test infrastructure only. Not part of the product library.
//
// FP64 CPU restatement of the reference's per-Newton-step mixed-FEM path
// (tytrusty/mfem). Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline may load it, and only as the checker; the product library never
// links or calls it.
#pragma once
#include <cstdint>

#ifdef __cplusplus
#include "omath.h"
namespace orc {
enum { MAT_SNH = 0, MAT_NH = 1, MAT_COROT = 2, MAT_FCR = 3, MAT_ARAP = 4, MAT_FUNG = 5 };
int newton_procrustes(M3& R, const M3& A, const M3& B, bool compute_gradients,
                      Mat<9, 9>* dRdF, double tol, int max_iter, int* iters);
int polar_svd3(M3& R, Vec<6>& s, const M3& F, bool compute_gradients,
               Mat<6, 9>* dsdF, int* iters);
void material_s6(int type, double mu, double la, const Vec<6>& s, bool psd_fix,
                 double* E, Vec<6>* g, Mat<6, 6>* H);
int polar2d(Mat<2, 2>& R, Vec<3>& s, const Mat<2, 2>& F, bool compute_gradients,
            Mat<3, 4>* dsdF);
bool material_s3(int type, double mu, double la, const Vec<3>& s, bool psd_fix, double* E,
                 Vec<3>* g, Mat<3, 3>* H);
inline bool material_has_s3(int type) {
  return type == MAT_SNH || type == MAT_ARAP || type == MAT_COROT || type == MAT_FCR;
}
void snh_f9(double mu, double la, const Vec<9>& f, double* E, Vec<9>* g,
            Mat<9, 9>* H);
void snh_f4(double mu, double la, const Vec<4>& f, double* E, Vec<4>* g,
            Mat<4, 4>* H);
void nh_f9(double mu, double la, const Vec<9>& f, double* E, Vec<9>* g, Mat<9, 9>* H);
bool material_f9(int type, double mu, double la, const Vec<9>& f, double* E, Vec<9>* g,
                 Mat<9, 9>* H);
inline bool material_has_f9(int type) {
  return type == MAT_SNH || type == MAT_NH || type == MAT_FUNG;
}
}  // namespace orc
extern "C" {
#endif

// Simulation parameters (mirrors the reference's SimConfig + scene objects).
typedef struct {
  int dim;            // 2 or 3
  int nV, nT;
  const double* V;    // nV*dim, row per vertex
  const int* T;       // nT*(dim+1)
  const double* mu;   // nT per-element Lame mu
  const double* la;   // nT per-element Lame lambda
  const int* mat;     // nT material type (MAT_*)
  const int* fixed;   // nV, 1 = pinned
  const double* v0;   // nV*dim initial velocity, or NULL
  double density;
  double dt;
  double gravity[3];
  int bdf_order;      // 1..6
  int mixed;          // 1: "mixed-stretch", 0: "stretch"
  int solver;         // 0: dense LLT ("eigen-llt"), 1: Jacobi CG ("eigen-pcg-diag"),
                      // 2: MINRES on the KKT system ("minres-indefinite-block") with the
                      //    GPU form of its block preconditioner, 3: the same MINRES with
                      //    the reference's exact factorizations (dense; small meshes),
                      // 4: CG with the incomplete-Cholesky preconditioner of the
                      //    device ("eigen-pcg-IC": multicolour block IC(0))
  double itr_tol;
  int max_iters;
  int max_newton;
  double newton_tol;
  int ls_iters;
  int threads;        // OpenMP threads (<= 0: runtime default)
  // boundary-condition script (reference config.h:70-87 numbering): 1 scale
  // (start positions scaled 1.5x about the box center), 4 stretch, 7 bend,
  // 8 twist (end groups moved after every timestep); others: static pins
  int bc_type, bc_axis;
  double bc_ratio, bc_velocity;
  // start positions (nV*dim), or NULL for V: the reference's initial_state
  // x_path (simulation_state.cpp:263-285), which replaces Mesh::Vinit_ while
  // the rest shape stays Vref_
  const double* x0;
} oracle_params;

typedef struct {
  int newton_iters;
  int pcg_iters_total;
  int ls_fail;
  double last_decrement;
  double last_energy;
  double last_alpha;
  double last_res;
} oracle_stats;

// element-level entry points
int oracle_newton_procrustes(double* R, const double* A, const double* B,
                             int compute_gradients, double* dRdF, double tol,
                             int max_iter, int* iters);
int oracle_polar_svd3(double* R, double* s, const double* F, int grad,
                      double* dsdF, int* iters);
void oracle_material_s6(int type, double mu, double la, const double* s,
                        int psd_fix, double* E, double* g, double* H);
void oracle_snh_f9(double mu, double la, const double* F, double* E,
                   double* g, double* H);
void oracle_snh_f4(double mu, double la, const double* F, double* E,
                   double* g, double* H);
// 2D polar fit (F, R column-major 2x2; dsdF 3x4 column-major)
int oracle_polar2d(double* R, double* s, const double* F, int grad, double* dsdF);
// 3-vector (2D mixed) API of material `type` (0 when the model has none)
int oracle_material_s3(int type, double mu, double la, const double* s, int psd_fix, double* E,
                       double* g, double* H);
// 9-vector API of material `type` (0 when the model does not implement it)
int oracle_material_f9(int type, double mu, double la, const double* F, double* E, double* g,
                       double* H);
void oracle_psd_fix(int n, double* A, double tol);
void oracle_bc_fixed(int type, const double* V, int nV, int dim, int axis,
                     double ratio, int* fixed);
void oracle_tet_geometry(const double* V, const int* T, int nT, double* vols,
                         double* dX);

// simulation-level entry points
void* oracle_sim_create(const oracle_params* p);
void oracle_sim_destroy(void* sim);
int oracle_sim_step(void* sim);              // one timestep, 0 on success
int oracle_sim_update_system(void* sim);     // only update_system() of step
int oracle_sim_solve(void* sim);             // only substep() of step
void oracle_sim_stats(void* sim, oracle_stats* st);
void oracle_sim_get_positions(void* sim, double* out);  // nV*dim
// Fetch an internal array by name into out (size from oracle_sim_size).
long oracle_sim_size(void* sim, const char* name);
int oracle_sim_get(void* sim, const char* name, double* out);
int oracle_sim_get_int(void* sim, const char* name, int* out);
double oracle_sim_energy(void* sim, double alpha);
// Indefinite (KKT) system at the current state, unknowns [dx; ds; lambda]
// (n + 12 nT, 6-vectors element-major). oracle_kkt_precond rebuilds the
// preconditioner from the current state (exact: dense factorizations).
long oracle_kkt_size(void* sim);
int oracle_kkt_rhs(void* sim, double* out);
int oracle_kkt_product(void* sim, const double* w, double* y);
int oracle_kkt_precond(void* sim, int exact, const double* b, double* z);
// The MINRES iteration on a dense symmetric A with a dense SPD preconditioner
// P (z = P r), both row-major N x N. Returns the estimated relative residual
// (negative if P is not positive definite).
int oracle_sim_history(void* sim, double* out);
int oracle_ic_factor_dense(int n, int bs, const int* rowptr, const int* colind, const double* vals,
                           double* L, double* scale, double* shift, int* color, int* attempts);
int oracle_ic_pcg(int n, int bs, const int* rowptr, const int* colind, const double* vals,
                  const double* b, double* x, double tol, int max_iters, double* err);
int oracle_jacobi_pcg(int n, const int* rowptr, const int* colind, const double* vals,
                      const double* b, double* x, double tol, int max_iters, double* err);
int oracle_envelope_solve(int n, int bs, const int* rowptr, const int* colind, const double* vals,
                          const double* b, double* x, int* perm);
double oracle_minres_dense(int N, const double* A, const double* P, const double* b, double* x,
                           int max_iters, double tol, int* iters);

#ifdef __cplusplus
}
#endif
