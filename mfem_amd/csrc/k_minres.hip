// Indefinite (KKT) Newton system and MINRES ("minres-indefinite-block").
//
// Unknowns [dx (3 nfree); ds (ns x nT); lambda (ns x nT)], ns = 6 on
// tetrahedra and 3 on triangles, the element parts stored structure-of-arrays
// (component i of element e at i * nT + e) so that they alias the mixed
// variable's ds and lambda arrays. On triangles the z rows of dx hold only
// the mass, which never couples to the rest, and stay exactly zero.
//
//   K = [[PMP, 0, Gx], [0, Hloc, C], [Gx^T, C, 0]]
//   Gx = -PJW diag(dSdF), C = vol Sym, Hloc = vol Hpsi (Hpsi = h^2 Hessian)
//
// is applied matrix-free as BlockMatrix::scaleAndAddTo does (reference
// src/linear_solvers/block_matrix.h:74-125 with Displacement::product_hessian,
// displacement.h:59-63, and MixedStretch::product_*, mixed_stretch.cpp:373-434).
// In this build's notation dsdF is ds/dF (6 x 9) and dSdF = dsdF^T Sym, so the
// lambda rows read -vol Sym dsdF Jloc dx + vol Sym ds and the x rows gather
// -vol Jloc^T dsdF^T Sym lambda per tet.
//
// The preconditioner is diag(Chebyshev(PMP), Hloc^-1, S_e^-1), S_e the element
// blocks of the Schur complement Gx^T Mlump^-1 Gx + C Hloc^-1 C, and MINRES
// follows Eigen 3.4's preconditioned Lanczos/Givens recurrence
// (unsupported/Eigen/src/IterativeSolvers/MINRES.h). One iteration is
//   k_mr_tet, k_mr_vertex   p/x update of the previous iteration; v_new = K w - beta v_old
//   k_mr_fin_alpha          alpha = v_new . w
//   k_mr_prec_tet, _vtx0    v_new -= alpha v; w_new = P v_new (element blocks,
//                           first Chebyshev step of the mass block)
//   k_mr_cheb x 3           Chebyshev steps of the mass block
//   k_mr_fin_beta           beta_new, Givens rotation, stopping test
// with every scalar in one MinresCtrl block on the device. Lanczos vectors are
// stored unnormalised: v_k = V_k / beta_k, w_k = W_k / beta_k. Every sum is a
// per-workgroup partial reduced in index order by a one-workgroup kernel.
#include <hip/hip_runtime.h>
#include "device_math.h"
#include "kernels.h"
#include "reduce.h"

namespace mfemhip {

namespace {

constexpr int kTetBS = 128;
constexpr int kVtxBS = 256;
constexpr int kFinBS = 1024;
constexpr int kGather = 8;  // indices / values in flight per gather batch

// Per-dimension element shape: corners, entries of s (Voigt, off-diagonals
// weighted 2 by Sym), consistent-mass divisor (rho vol / 20 on a tet, / 12 on
// a triangle), packed index of a symmetric ns x ns matrix (i <= j).
template <int DIM>
struct Elem {
  static constexpr int NC = DIM + 1, NS = DIM == 3 ? 6 : 3, NF = DIM * DIM;
  static constexpr double kMassDiv = DIM == 3 ? 20.0 : 12.0;
  __device__ static double symw(int i) { return i < DIM ? 1.0 : 2.0; }
  __device__ static int sym(int i, int j) {
    if (i > j) { const int t = i; i = j; j = t; }
    return DIM == 3 ? sym6(i, j) : sym3(i, j);
  }
};

// XCD-contiguous work split: workgroup b runs on XCD b % 8 (round-robin
// dispatch), and XCD x owns items [n x / 8, n (x + 1) / 8), so that the
// vertices and tets one XCD touches, and the lines their gathers read, stay
// in that XCD's L2. Returns n for lanes past the end of their range.
__device__ __forceinline__ int xcd_index(int n) {
  const int xcd = blockIdx.x & 7, lb = blockIdx.x >> 3;
  const long long start = (long long)n * xcd / 8, end = (long long)n * (xcd + 1) / 8;
  const long long i = start + (long long)lb * blockDim.x + threadIdx.x;
  return i < end ? (int)i : n;
}

inline int blocks_for(int n, int bs) {
  if (n <= 0) return 0;
  const int per_xcd = (n + 7) / 8;
  return 8 * ((per_xcd + bs - 1) / bs);
}

// Sum of parts[0 .. n) in index order (strided by lane, then the block tree).
__device__ __forceinline__ double fin_sum(const double* parts, int n, double* sm) {
  return reduce_array<kFinBS, 0>(parts, n, 1, 0, sm);
}

// ---------------------------------------------------------------- setup
// S_e = Sym (vol^2 B Lambda B^T + vol Hpsi^-1) Sym with B = dsdF Jloc and
// Lambda = 1 / Mlump on free corners (0 on pinned ones), inverted by Cholesky
// into Sinv (packed symmetric, stride nT). A block that is not positive
// definite becomes NaN, which the MINRES scalars report.
template <typename Real, int DIM>
__global__ __launch_bounds__(kTetBS) void k_kkt_setup(TetMeshDev m, const Real* __restrict__ dsdF,
                                                      const Real* __restrict__ Hinv,
                                                      const double* __restrict__ mlump_inv,
                                                      double* __restrict__ Sinv) {
  using E = Elem<DIM>;
  constexpr int NC = E::NC, NS = E::NS;
  const int e = xcd_index(m.nT);
  if (e >= m.nT) return;
  const int nT = m.nT;
  const double vol = m.vol[e];
  const int4 t = m.T[e];
  const int v[4] = {t.x, t.y, t.z, t.w};
  double lam[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    const int f = m.free_map[v[k]];
    lam[k] = f >= 0 ? mlump_inv[f] : 0.0;
  }
  // G(c, c') = sum_k Lambda_k dX(k, c) dX(k, c')
  double dx[3 * NC];
#pragma unroll
  for (int k = 0; k < NC; ++k)
#pragma unroll
    for (int c = 0; c < DIM; ++c) dx[3 * k + c] = m.dX[(3 * k + c) * nT + e];
  double G[DIM * DIM];
#pragma unroll
  for (int c = 0; c < DIM; ++c)
#pragma unroll
    for (int c2 = 0; c2 < DIM; ++c2) {
      double a = 0;
#pragma unroll
      for (int k = 0; k < NC; ++k) a += lam[k] * dx[3 * k + c] * dx[3 * k + c2];
      G[c + DIM * c2] = a;
    }
  // B Lambda B^T = dsdF (I (x) G) dsdF^T, one row r of F at a time
  double A[NS * NS];
#pragma unroll
  for (int i = 0; i < NS * NS; ++i) A[i] = 0.0;
#pragma unroll 1
  for (int r = 0; r < DIM; ++r) {
    double J[NS * DIM];  // J[i + NS c] = dsdF(i, r + DIM c)
#pragma unroll
    for (int c = 0; c < DIM; ++c)
#pragma unroll
      for (int i = 0; i < NS; ++i) J[i + NS * c] = (double)dsdF[(i + NS * (r + DIM * c)) * nT + e];
    double T[NS * DIM];  // T = J G
#pragma unroll
    for (int c = 0; c < DIM; ++c)
#pragma unroll
      for (int i = 0; i < NS; ++i) {
        double a = 0;
#pragma unroll
        for (int c2 = 0; c2 < DIM; ++c2) a += J[i + NS * c2] * G[c2 + DIM * c];
        T[i + NS * c] = a;
      }
#pragma unroll
    for (int i = 0; i < NS; ++i)
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        double a = 0;
#pragma unroll
        for (int c = 0; c < DIM; ++c) a += T[i + NS * c] * J[j + NS * c];
        A[i + NS * j] += a;
      }
  }
#pragma unroll
  for (int i = 0; i < NS; ++i)
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      const double h = (double)Hinv[E::sym(i, j) * nT + e];
      A[i + NS * j] = E::symw(i) * (vol * vol * A[i + NS * j] + vol * h) * E::symw(j);
    }
  // Cholesky A = L L^T (lower, in place), then Sinv = L^-T L^-1
  bool ok = true;
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    double d = A[j + NS * j];
#pragma unroll
    for (int k = 0; k < j; ++k) d -= A[j + NS * k] * A[j + NS * k];
    ok = ok && d > 0.0;
    d = sqrt(fmax(d, 0.0));
    A[j + NS * j] = d;
    const double id = 1.0 / d;
#pragma unroll
    for (int i = j + 1; i < NS; ++i) {
      double s = A[i + NS * j];
#pragma unroll
      for (int k = 0; k < j; ++k) s -= A[i + NS * k] * A[j + NS * k];
      A[i + NS * j] = s * id;
    }
  }
  double Li[NS * NS];  // L^-1 (lower)
#pragma unroll
  for (int i = 0; i < NS * NS; ++i) Li[i] = 0.0;
#pragma unroll
  for (int c = 0; c < NS; ++c) {
    Li[c + NS * c] = 1.0 / A[c + NS * c];
#pragma unroll
    for (int i = c + 1; i < NS; ++i) {
      double s = 0;
#pragma unroll
      for (int k = c; k < i; ++k) s -= A[i + NS * k] * Li[k + NS * c];
      Li[i + NS * c] = s / A[i + NS * i];
    }
  }
#pragma unroll
  for (int i = 0; i < NS; ++i)
#pragma unroll
    for (int j = i; j < NS; ++j) {
      double s = 0;
#pragma unroll
      for (int k = j; k < NS; ++k) s += Li[k + NS * i] * Li[k + NS * j];
      Sinv[E::sym(i, j) * nT + e] = ok ? s : NAN;
    }
}

// ---------------------------------------------------------------- rhs
// b = [-PM (x_full - x_tilde - h^2 g); -vol g; vol Sym (S - s)]
// (linear_system.h:67-86): the inertia part per tet into fe (gathered per
// free vertex by k_kkt_rhs_vtx), the element parts straight into b.
template <int DIM>
__global__ __launch_bounds__(kTetBS) void k_kkt_rhs_tet(TetMeshDev m, const double* __restrict__ x,
                                                        Inertia in, const double* __restrict__ g,
                                                        const double* __restrict__ S,
                                                        const double* __restrict__ s,
                                                        double* __restrict__ fe, KktVec b,
                                                        double* __restrict__ parts) {
  __shared__ double sm[kTetBS / kWave];
  const int e = xcd_index(m.nT);
  double acc = 0.0;
  if (e < m.nT) {
    using E = Elem<DIM>;
    constexpr int NC = E::NC;
    const int nT = m.nT;
    const double vol = m.vol[e];
    const int4 t = m.T[e];
    const int v[4] = {t.x, t.y, t.z, t.w};
    double d[3 * NC], dsum[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < NC; ++k)
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const double dd = x[3 * v[k] + r] - in.x_tilde[3 * v[k] + r] - in.h2g[r];
        d[3 * k + r] = dd;
        dsum[r] += dd;
      }
    const double mc = m.density * vol / E::kMassDiv;
    double* fo = fe + (size_t)e * 12;
#pragma unroll
    for (int j = 0; j < 3 * NC; ++j) fo[j] = -mc * (d[j] + dsum[j % 3]);
#pragma unroll
    for (int i = 0; i < E::NS; ++i) {
      const double bs = -vol * g[i * nT + e];
      const double bl = vol * E::symw(i) * (S[i * nT + e] - s[i * nT + e]);
      b.s[i * nT + e] = bs;
      b.l[i * nT + e] = bl;
      acc += bs * bs + bl * bl;
    }
  }
  block_partial<kTetBS, 0>(acc, parts, sm);
}

__global__ __launch_bounds__(kVtxBS) void k_kkt_rhs_vtx(int nfree, const int* __restrict__ inc_ptr,
                                                        const int* __restrict__ inc,
                                                        const double* __restrict__ fe, KktVec b,
                                                        double* __restrict__ parts) {
  __shared__ double sm[kVtxBS / kWave];
  const int i = xcd_index(nfree);
  double acc = 0.0;
  if (i < nfree) {
    double a[3] = {0, 0, 0};
    const int p1 = inc_ptr[i + 1];
    for (int p = inc_ptr[i]; p < p1; p += kGather) {  // batched gather, sums in order
      int ek[kGather];
      double f[kGather][3];
#pragma unroll
      for (int u = 0; u < kGather; ++u) ek[u] = p + u < p1 ? inc[p + u] : -1;
#pragma unroll
      for (int u = 0; u < kGather; ++u)
        if (ek[u] >= 0) {
          const double* src = fe + (size_t)(ek[u] >> 2) * 12 + 3 * (ek[u] & 3);
#pragma unroll
          for (int r = 0; r < 3; ++r) f[u][r] = src[r];
        }
#pragma unroll
      for (int u = 0; u < kGather; ++u)
        if (ek[u] >= 0)
#pragma unroll
          for (int r = 0; r < 3; ++r) a[r] += f[u][r];
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      b.x[3 * i + r] = a[r];
      acc += a[r] * a[r];
    }
  }
  block_partial<kVtxBS, 0>(acc, parts, sm);
}

// ---------------------------------------------------------------- iteration
// Update of the previous iteration, p_{k-1} = (w_{k-1} - r2 p_{k-2} -
// r3 p_{k-3}) / r1 and x += beta_one c eta p_{k-1}, on one entry.
__device__ __forceinline__ void pending_update(const MinresCtrl& c, long i, const double* wo,
                                               const double* pa, const double* pb, double* pn,
                                               double* x) {
  const double p = (wo[i] / c.beta_prev - c.r2 * pa[i] - c.r3 * pb[i]) / c.r1;
  pn[i] = p;
  x[i] += c.coef * p;
}

struct MrVecs {
  KktVec wc;  // W_k (scale beta_k)
  KktVec vo;  // in: V_{k-1} (scale beta_{k-1}); out: V_{k+1} before alpha
  KktVec wo;  // W_{k-1} (scale beta_{k-1})
  KktVec pa;  // p_{k-2}
  KktVec pb;  // p_{k-3}
  KktVec pn;  // out: p_{k-1} (the slot of p_{k-4})
  KktVec x;   // solution
};

template <typename Real, int DIM>
__global__ __launch_bounds__(kTetBS) void k_mr_tet(TetMeshDev m, const MinresCtrl* __restrict__ ctrl,
                                                   const Real* __restrict__ dsdF,
                                                   const Real* __restrict__ Hpsi, MrVecs V,
                                                   double* __restrict__ fe,
                                                   double* __restrict__ parts) {
  using E = Elem<DIM>;
  constexpr int NC = E::NC, NS = E::NS, NF = E::NF;
  constexpr int kFePitch = 13;  // x-row contributions staged per tet (odd pitch: LDS banks)
  __shared__ double sm[kTetBS / kWave];
  __shared__ double fe_stage[kTetBS * kFePitch];
  const MinresCtrl c = *ctrl;
  const int e = xcd_index(m.nT);
  const int nT = m.nT;
  if (c.pending && e < nT) {
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      pending_update(c, (long)i * nT + e, V.wo.s, V.pa.s, V.pb.s, V.pn.s, V.x.s);
      pending_update(c, (long)i * nT + e, V.wo.l, V.pa.l, V.pb.l, V.pn.l, V.x.l);
    }
  }
  if (c.done) return;
  double acc = 0.0;
  if (e < nT) {
    const double vol = m.vol[e];
    const double ib = 1.0 / c.beta;
    const double ratio = c.beta / c.beta_prev;
    double ws[NS], wl[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      ws[i] = V.wc.s[i * nT + e] * ib;
      wl[i] = V.wc.l[i * nT + e] * ib;
    }
    // Jloc w_x over the free corners
    const int4 t = m.T[e];
    const int v[4] = {t.x, t.y, t.z, t.w};
    double Fw[NF];
#pragma unroll
    for (int i = 0; i < NF; ++i) Fw[i] = 0.0;
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const int f = m.free_map[v[k]];
      double wk[DIM];
#pragma unroll
      for (int r = 0; r < DIM; ++r) wk[r] = f >= 0 ? V.wc.x[3 * f + r] * ib : 0.0;
#pragma unroll
      for (int cc = 0; cc < DIM; ++cc) {
        const double d = m.dX[(3 * k + cc) * nT + e];
#pragma unroll
        for (int r = 0; r < DIM; ++r) Fw[r + DIM * cc] += wk[r] * d;
      }
    }
    // u = dsdF Fw and tv = dsdF^T (Sym w_l), one pass over dsdF
    double u[NS], tv[NF];
#pragma unroll
    for (int i = 0; i < NF; ++i) tv[i] = 0.0;
#pragma unroll
    for (int i = 0; i < NS; ++i) u[i] = 0.0;
#pragma unroll
    for (int mm = 0; mm < NF; ++mm)
#pragma unroll
      for (int i = 0; i < NS; ++i) {
        const double js = (double)dsdF[(i + NS * mm) * nT + e];
        u[i] += js * Fw[mm];
        tv[mm] += js * E::symw(i) * wl[i];
      }
    // x rows of this element: -vol Jloc^T tv (the z rows of a triangle: 0),
    // staged in LDS and written below as contiguous runs
    double* fo = fe_stage + threadIdx.x * kFePitch;
#pragma unroll
    for (int k = 0; k < NC; ++k)
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        double a = 0;
        if (r < DIM)
#pragma unroll
          for (int cc = 0; cc < DIM; ++cc) a += m.dX[(3 * k + cc) * nT + e] * tv[r + DIM * cc];
        fo[3 * k + r] = -vol * a;
      }
    // s rows: Hloc w_s + vol Sym w_l ; lambda rows: vol Sym (w_s - u)
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      double a = 0;
#pragma unroll
      for (int j = 0; j < NS; ++j) a += (double)Hpsi[E::sym(i, j) * nT + e] * ws[j];
      const double ys = vol * a + vol * E::symw(i) * wl[i];
      const double yl = vol * E::symw(i) * (ws[i] - u[i]);
      const long is = (long)i * nT + e;
      // v_new = K w - beta v_old, v_old = V_{k-1} / beta_{k-1}
      const double vs = ys - ratio * V.vo.s[is];
      const double vl = yl - ratio * V.vo.l[is];
      V.vo.s[is] = vs;
      V.vo.l[is] = vl;
      acc += vs * ws[i] + vl * wl[i];
    }
  }
  // the staged x-row contributions of this wave's live tets, which are a
  // contiguous prefix of the wave (xcd_index hands out consecutive items)
  {
    const int lane = threadIdx.x & (kWave - 1);
    const int nlive = __popcll(__ballot(e < nT));
    const int e0 = __shfl(e, 0);
    const double* ws = fe_stage + (threadIdx.x - lane) * kFePitch;
    __syncthreads();
    for (int idx = lane; idx < nlive * 3 * NC; idx += kWave) {
      const int t = idx / (3 * NC), j = idx - t * (3 * NC);
      fe[(size_t)(e0 + t) * 12 + j] = ws[t * kFePitch + j];
    }
  }
  block_partial<kTetBS, 0>(acc, parts, sm);
}

__global__ __launch_bounds__(kVtxBS) void k_mr_vertex(int nfree, const int* __restrict__ rowptr,
                                                      const int* __restrict__ col,
                                                      const double* __restrict__ mass,
                                                      const int* __restrict__ inc_ptr,
                                                      const int* __restrict__ inc,
                                                      const double* __restrict__ fe,
                                                      const MinresCtrl* __restrict__ ctrl, MrVecs V,
                                                      double* __restrict__ parts) {
  __shared__ double sm[kVtxBS / kWave];
  const MinresCtrl c = *ctrl;
  const int i = xcd_index(nfree);
  if (c.pending && i < nfree) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
      pending_update(c, 3L * i + r, V.wo.x, V.pa.x, V.pb.x, V.pn.x, V.x.x);
  }
  if (c.done) return;
  double acc = 0.0;
  if (i < nfree) {
    const double ib = 1.0 / c.beta;
    const double ratio = c.beta / c.beta_prev;
    // both gathers in batches of kGather: indices, then values, in flight
    // together (the sums keep their order)
    double y[3] = {0, 0, 0};
    const int b1 = rowptr[i + 1];
    for (int b = rowptr[i]; b < b1; b += kGather) {
      int j[kGather];
      double mm[kGather], w[kGather][3];
#pragma unroll
      for (int u = 0; u < kGather; ++u) {
        j[u] = b + u < b1 ? col[b + u] : -1;
        mm[u] = b + u < b1 ? mass[b + u] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kGather; ++u)
        if (j[u] >= 0)
#pragma unroll
          for (int r = 0; r < 3; ++r) w[u][r] = V.wc.x[3 * j[u] + r];
#pragma unroll
      for (int u = 0; u < kGather; ++u)
        if (j[u] >= 0)
#pragma unroll
          for (int r = 0; r < 3; ++r) y[r] += mm[u] * (w[u][r] * ib);
    }
    double z[3] = {0, 0, 0};
    const int p1 = inc_ptr[i + 1];
    for (int p = inc_ptr[i]; p < p1; p += kGather) {
      int ek[kGather];
      double f[kGather][3];
#pragma unroll
      for (int u = 0; u < kGather; ++u) ek[u] = p + u < p1 ? inc[p + u] : -1;
#pragma unroll
      for (int u = 0; u < kGather; ++u)
        if (ek[u] >= 0) {
          const double* src = fe + (size_t)(ek[u] >> 2) * 12 + 3 * (ek[u] & 3);
#pragma unroll
          for (int r = 0; r < 3; ++r) f[u][r] = src[r];
        }
#pragma unroll
      for (int u = 0; u < kGather; ++u)
        if (ek[u] >= 0)
#pragma unroll
          for (int r = 0; r < 3; ++r) z[r] += f[u][r];
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const long ir = 3L * i + r;
      const double vn = (y[r] + z[r]) - ratio * V.vo.x[ir];
      V.vo.x[ir] = vn;
      acc += vn * (V.wc.x[ir] * ib);
    }
  }
  block_partial<kVtxBS, 0>(acc, parts, sm);
}

// alpha = v_new . w; the pending update has been applied by now.
__global__ __launch_bounds__(kFinBS) void k_mr_fin_alpha(MinresCtrl* ctrl, const double* parts,
                                                         int nparts) {
  __shared__ double sm[kFinBS / kWave];
  const bool done = ctrl->done;
  __syncthreads();
  if (threadIdx.x == 0) ctrl->pending = 0;
  if (done) return;
  const double a = fin_sum(parts, nparts, sm);
  if (threadIdx.x == 0) ctrl->alpha = a;
}

// v_new -= alpha v_k (init: alpha = 0, v_k absent); w_new = P v_new on the
// element parts: Hloc^-1 = Hpsi^-1 / vol on s, S_e^-1 on lambda.
template <typename Real, int DIM>
__global__ __launch_bounds__(kTetBS) void k_mr_prec_tet(TetMeshDev m, const MinresCtrl* __restrict__ ctrl,
                                                        int init, const Real* __restrict__ Hinv,
                                                        const double* __restrict__ Sinv, KktVec vn,
                                                        KktVec vc, KktVec wn,
                                                        double* __restrict__ parts) {
  using E = Elem<DIM>;
  constexpr int NS = E::NS;
  __shared__ double sm[kTetBS / kWave];
  const MinresCtrl c = *ctrl;
  if (!init && c.done) return;
  const int e = xcd_index(m.nT);
  const int nT = m.nT;
  double acc = 0.0;
  if (e < nT) {
    const double ab = init ? 0.0 : c.alpha / c.beta;
    const double vol = m.vol[e];
    double vs[NS], vl[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      const long is = (long)i * nT + e;
      vs[i] = init ? vn.s[is] : vn.s[is] - ab * vc.s[is];
      vl[i] = init ? vn.l[is] : vn.l[is] - ab * vc.l[is];
      if (!init) {
        vn.s[is] = vs[i];
        vn.l[is] = vl[i];
      }
    }
    const double iv = 1.0 / vol;
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      double a = 0, b = 0;
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        const int k = E::sym(i, j);
        a += (double)Hinv[k * nT + e] * vs[j];
        b += Sinv[k * nT + e] * vl[j];
      }
      const long is = (long)i * nT + e;
      const double zs = a * iv;
      wn.s[is] = zs;
      wn.l[is] = b;
      acc += vs[i] * zs + vl[i] * b;
    }
  }
  block_partial<kTetBS, 0>(acc, parts, sm);
}

// x part: v_new -= alpha v_k, then the first Chebyshev step of the mass block:
// r = v_new, d = D^-1 r / theta, z = d.
__global__ __launch_bounds__(kVtxBS) void k_mr_prec_vtx0(int nfree, const MinresCtrl* __restrict__ ctrl,
                                                         int init, const double* __restrict__ dinv,
                                                         double theta, KktVec vn, KktVec vc,
                                                         KktVec wn, double* __restrict__ cr,
                                                         double* __restrict__ cd) {
  const MinresCtrl c = *ctrl;
  if (!init && c.done) return;
  const int i = xcd_index(nfree);
  if (i >= nfree) return;
  const double ab = init ? 0.0 : c.alpha / c.beta;
  const double di = dinv[i];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const long ir = 3L * i + r;
    double v = vn.x[ir];
    if (!init) {
      v -= ab * vc.x[ir];
      vn.x[ir] = v;
    }
    const double d = di * v / theta;
    cr[ir] = v;
    cd[ir] = d;
    wn.x[ir] = d;
  }
}

// One Chebyshev step: r -= PMP d_in; d_out = c1 d_in + c2 D^-1 r; z += d_out.
// The last step also writes the partials of v_new . z.
__global__ __launch_bounds__(kVtxBS) void k_mr_cheb(int nfree, const int* __restrict__ rowptr,
                                                    const int* __restrict__ col,
                                                    const double* __restrict__ mass,
                                                    const double* __restrict__ dinv,
                                                    const MinresCtrl* __restrict__ ctrl, double c1,
                                                    double c2, const double* __restrict__ d_in,
                                                    double* __restrict__ d_out,
                                                    double* __restrict__ cr, KktVec vn, KktVec wn,
                                                    int last, int init,
                                                    double* __restrict__ parts) {
  __shared__ double sm[kVtxBS / kWave];
  if (!init && ctrl->done) return;
  const int i = xcd_index(nfree);
  double acc = 0.0;
  if (i < nfree) {
    double y[3] = {0, 0, 0};
    const int b1 = rowptr[i + 1];
    for (int b = rowptr[i]; b < b1; b += kGather) {  // batched gather, sums in order
      int j[kGather];
      double mm[kGather], w[kGather][3];
#pragma unroll
      for (int u = 0; u < kGather; ++u) {
        j[u] = b + u < b1 ? col[b + u] : -1;
        mm[u] = b + u < b1 ? mass[b + u] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kGather; ++u)
        if (j[u] >= 0)
#pragma unroll
          for (int r = 0; r < 3; ++r) w[u][r] = d_in[3 * j[u] + r];
#pragma unroll
      for (int u = 0; u < kGather; ++u)
        if (j[u] >= 0)
#pragma unroll
          for (int r = 0; r < 3; ++r) y[r] += mm[u] * w[u][r];
    }
    const double di = dinv[i];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const long ir = 3L * i + r;
      const double rr = cr[ir] - y[r];
      cr[ir] = rr;
      const double d = c1 * d_in[ir] + c2 * di * rr;
      d_out[ir] = d;
      const double z = wn.x[ir] + d;
      wn.x[ir] = z;
      acc += vn.x[ir] * z;
    }
  }
  if (last) block_partial<kVtxBS, 0>(acc, parts, sm);
}

// beta_new, the Givens rotation and the stopping test of one iteration.
__global__ __launch_bounds__(kFinBS) void k_mr_fin_beta(MinresCtrl* ctrl, const double* parts,
                                                        int nparts) {
  __shared__ double sm[kFinBS / kWave];
  if (ctrl->done) return;
  const double b2 = fin_sum(parts, nparts, sm);
  if (threadIdx.x != 0) return;
  MinresCtrl c = *ctrl;
  if (!(b2 >= 0.0)) {  // preconditioner not positive definite (or NaN)
    c.fail = 1;
    c.done = 1;
    *ctrl = c;
    return;
  }
  const double bn = sqrt(b2);
  const double beta = c.beta, alpha = c.alpha;
  const double r2 = c.s * alpha + c.c * c.c_old * beta;
  const double r3 = c.s_old * beta;
  const double r1_hat = c.c * alpha - c.c_old * c.s * beta;
  const double r1 = sqrt(r1_hat * r1_hat + bn * bn);
  c.c_old = c.c;
  c.s_old = c.s;
  c.c = r1_hat / r1;
  c.s = bn / r1;
  c.r1 = r1;
  c.r2 = r2;
  c.r3 = r3;
  c.coef = c.beta_one * c.c * c.eta;
  c.pending = 1;
  c.res2 *= c.s * c.s;
  if (c.res2 < c.threshold2) {
    c.done = 1;
  } else {
    c.eta = -c.s * c.eta;
    c.iter += 1;
    if (c.iter >= c.max_iters) c.done = 1;
  }
  c.beta_prev = beta;
  c.beta = bn;
  *ctrl = c;
}

// Control block for a bare product (beta = 1, nothing pending, not stopped).
__global__ void k_mr_ctrl_unit(MinresCtrl* ctrl) {
  MinresCtrl c{};
  c.beta = c.beta_prev = c.beta_one = c.r1 = 1.0;
  *ctrl = c;
}

// Scalars after the right-hand side and its preconditioned image.
__global__ __launch_bounds__(kFinBS) void k_mr_fin_init(MinresCtrl* ctrl, const double* parts_b,
                                                        const double* parts_w, int nparts,
                                                        double tol, int max_iters) {
  __shared__ double sm[kFinBS / kWave];
  const double rhs2 = fin_sum(parts_b, nparts, sm);
  const double b2 = fin_sum(parts_w, nparts, sm);
  if (threadIdx.x != 0) return;
  MinresCtrl c{};
  c.rhs2 = rhs2;
  c.threshold2 = tol * tol * rhs2;
  c.res2 = rhs2;
  c.max_iters = max_iters;
  c.c = c.c_old = 1.0;
  c.s = c.s_old = 0.0;
  c.eta = 1.0;
  c.beta_prev = 1.0;
  c.r1 = 1.0;
  if (rhs2 == 0.0 || max_iters <= 0) {
    c.done = 1;
    c.beta = c.beta_one = 1.0;
  } else if (!(b2 >= 0.0)) {
    c.done = 1;
    c.fail = 1;
    c.beta = c.beta_one = 1.0;
  } else {
    c.beta = c.beta_one = sqrt(b2);
  }
  *ctrl = c;
}

}  // namespace

int kkt_tet_blocks(int nT) { return blocks_for(nT, kTetBS); }
int kkt_vertex_blocks(int nfree) { return blocks_for(nfree, kVtxBS); }

template <typename Real>
void launch_kkt_setup(const KktSystem<Real>& S, hipStream_t st) {
  if (S.m.nT == 0) return;
  if (S.m.dim == 2)
    hipLaunchKernelGGL((k_kkt_setup<Real, 2>), dim3(kkt_tet_blocks(S.m.nT)), dim3(kTetBS), 0, st,
                       S.m, S.dsdF, S.Hinv, S.mlump_inv, S.Sinv);
  else
    hipLaunchKernelGGL((k_kkt_setup<Real, 3>), dim3(kkt_tet_blocks(S.m.nT)), dim3(kTetBS), 0, st,
                       S.m, S.dsdF, S.Hinv, S.mlump_inv, S.Sinv);
}

// Chebyshev constants for the mass block over [1/2, 5/2]: theta 1.5,
// delta 1, sigma = theta / delta; rho_0 = 1 / sigma, rho_j = 1 / (2 sigma -
// rho_{j-1}); step j uses c1 = rho_j rho_{j-1}, c2 = 2 rho_j / delta.
namespace {
constexpr double kChebTheta = 1.5, kChebDelta = 1.0;

template <typename Real>
void launch_precondition(const KktSystem<Real>& S, int init, KktVec vn, KktVec vc, KktVec wn,
                         hipStream_t st) {
  const int nbT = kkt_tet_blocks(S.m.nT), nbV = kkt_vertex_blocks(S.m.nfree);
  if (nbT > 0)
    if (S.m.dim == 2)
      hipLaunchKernelGGL((k_mr_prec_tet<Real, 2>), dim3(nbT), dim3(kTetBS), 0, st, S.m, S.ctrl,
                         init, S.Hinv, S.Sinv, vn, vc, wn, S.parts[1]);
    else
      hipLaunchKernelGGL((k_mr_prec_tet<Real, 3>), dim3(nbT), dim3(kTetBS), 0, st, S.m, S.ctrl,
                         init, S.Hinv, S.Sinv, vn, vc, wn, S.parts[1]);
  if (nbV == 0) return;
  hipLaunchKernelGGL(k_mr_prec_vtx0, dim3(nbV), dim3(kVtxBS), 0, st, S.m.nfree, S.ctrl, init,
                     S.mdinv, kChebTheta, vn, vc, wn, S.cheb_r, S.cheb_d[0]);
  const double sigma = kChebTheta / kChebDelta;
  double rho = 1.0 / sigma;
  for (int j = 1; j < kMassChebSteps; ++j) {
    const double rho1 = 1.0 / (2.0 * sigma - rho);
    const int last = j == kMassChebSteps - 1;
    hipLaunchKernelGGL(k_mr_cheb, dim3(nbV), dim3(kVtxBS), 0, st, S.m.nfree, S.rowptr, S.col,
                       S.mass, S.mdinv, S.ctrl, rho1 * rho, 2.0 * rho1 / kChebDelta,
                       S.cheb_d[(j - 1) & 1], S.cheb_d[j & 1], S.cheb_r, vn, wn, last, init,
                       S.parts[1] + nbT);
    rho = rho1;
  }
}
}  // namespace

template <typename Real>
void launch_minres_init(const KktSystem<Real>& S, const double* x_full, const Inertia& in,
                        const double* g, const double* Smat, const double* s, KktVec b, KktVec w,
                        double tol, int max_iters, hipStream_t st) {
  const int nbT = kkt_tet_blocks(S.m.nT), nbV = kkt_vertex_blocks(S.m.nfree);
  if (nbT > 0)
    if (S.m.dim == 2)
      hipLaunchKernelGGL((k_kkt_rhs_tet<2>), dim3(nbT), dim3(kTetBS), 0, st, S.m, x_full, in, g,
                         Smat, s, S.fe, b, S.parts[0]);
    else
      hipLaunchKernelGGL((k_kkt_rhs_tet<3>), dim3(nbT), dim3(kTetBS), 0, st, S.m, x_full, in, g,
                         Smat, s, S.fe, b, S.parts[0]);
  if (nbV > 0)
    hipLaunchKernelGGL(k_kkt_rhs_vtx, dim3(nbV), dim3(kVtxBS), 0, st, S.m.nfree, S.inc_ptr,
                       S.inc, S.fe, b, S.parts[0] + nbT);
  launch_precondition(S, 1, b, b, w, st);
  hipLaunchKernelGGL(k_mr_fin_init, dim3(1), dim3(kFinBS), 0, st, S.ctrl, S.parts[0],
                     S.parts[1], nbT + nbV, tol, max_iters);
}

template <typename Real>
void launch_minres_iteration(const KktSystem<Real>& S, const MinresBuffers& B, long k,
                             hipStream_t st, hipEvent_t* tet_ev) {
  const int nbT = kkt_tet_blocks(S.m.nT), nbV = kkt_vertex_blocks(S.m.nfree);
  auto mod3 = [](long j) { return (int)(((j % 3) + 3) % 3); };
  MrVecs V;
  V.wc = B.w[k & 1];
  V.wo = B.w[(k + 1) & 1];
  V.vo = B.v[(k + 1) & 1];
  V.pa = B.p[mod3(k - 2)];
  V.pb = B.p[mod3(k - 3)];
  V.pn = B.p[mod3(k - 1)];  // p_{k-1} replaces p_{k-4}
  V.x = B.x;
  if (tet_ev) (void)hipEventRecord(tet_ev[0], st);
  if (nbT > 0)
    if (S.m.dim == 2)
      hipLaunchKernelGGL((k_mr_tet<Real, 2>), dim3(nbT), dim3(kTetBS), 0, st, S.m, S.ctrl, S.dsdF,
                         S.Hpsi, V, S.fe, S.parts[0]);
    else
      hipLaunchKernelGGL((k_mr_tet<Real, 3>), dim3(nbT), dim3(kTetBS), 0, st, S.m, S.ctrl, S.dsdF,
                         S.Hpsi, V, S.fe, S.parts[0]);
  if (tet_ev) (void)hipEventRecord(tet_ev[1], st);
  if (nbV > 0)
    hipLaunchKernelGGL(k_mr_vertex, dim3(nbV), dim3(kVtxBS), 0, st, S.m.nfree, S.rowptr, S.col,
                       S.mass, S.inc_ptr, S.inc, S.fe, S.ctrl, V, S.parts[0] + nbT);
  hipLaunchKernelGGL(k_mr_fin_alpha, dim3(1), dim3(kFinBS), 0, st, S.ctrl, S.parts[0], nbT + nbV);
  // v_new (in V_{k-1}'s buffer) -= alpha v_k; w_new into W_{k-1}'s buffer
  launch_precondition(S, 0, B.v[(k + 1) & 1], B.v[k & 1], B.w[(k + 1) & 1], st);
  hipLaunchKernelGGL(k_mr_fin_beta, dim3(1), dim3(kFinBS), 0, st, S.ctrl, S.parts[1], nbT + nbV);
}

template <typename Real>
void launch_kkt_apply(const KktSystem<Real>& S, int which, KktVec in, KktVec out, KktVec scratch,
                      hipStream_t st) {
  const int nbT = kkt_tet_blocks(S.m.nT), nbV = kkt_vertex_blocks(S.m.nfree);
  if (which == 1) {  // out = P in
    launch_kkt_setup(S, st);
    launch_precondition(S, 1, in, in, out, st);
    return;
  }
  // out = K in - 1 * 0: the product kernels with beta = 1 and v_old = 0
  hipLaunchKernelGGL(k_mr_ctrl_unit, dim3(1), dim3(1), 0, st, S.ctrl);
  (void)hipMemsetAsync(out.x, 0, 3 * (size_t)S.m.nfree * sizeof(double), st);
  (void)hipMemsetAsync(out.s, 0, 6 * (size_t)S.m.nT * sizeof(double), st);
  (void)hipMemsetAsync(out.l, 0, 6 * (size_t)S.m.nT * sizeof(double), st);
  MrVecs V;
  V.wc = in;
  V.wo = in;
  V.vo = out;
  V.pa = V.pb = V.pn = V.x = scratch;
  if (nbT > 0)
    if (S.m.dim == 2)
      hipLaunchKernelGGL((k_mr_tet<Real, 2>), dim3(nbT), dim3(kTetBS), 0, st, S.m, S.ctrl, S.dsdF,
                         S.Hpsi, V, S.fe, S.parts[0]);
    else
      hipLaunchKernelGGL((k_mr_tet<Real, 3>), dim3(nbT), dim3(kTetBS), 0, st, S.m, S.ctrl, S.dsdF,
                         S.Hpsi, V, S.fe, S.parts[0]);
  if (nbV > 0)
    hipLaunchKernelGGL(k_mr_vertex, dim3(nbV), dim3(kVtxBS), 0, st, S.m.nfree, S.rowptr, S.col,
                       S.mass, S.inc_ptr, S.inc, S.fe, S.ctrl, V, S.parts[0] + nbT);
}

#define MFEM_MINRES_INST(Real)                                                                  \
  template void launch_kkt_setup<Real>(const KktSystem<Real>&, hipStream_t);                    \
  template void launch_minres_init<Real>(const KktSystem<Real>&, const double*, const Inertia&, \
                                         const double*, const double*, const double*, KktVec,    \
                                         KktVec, double, int, hipStream_t);                      \
  template void launch_minres_iteration<Real>(const KktSystem<Real>&, const MinresBuffers&, long, \
                                              hipStream_t, hipEvent_t*);                         \
  template void launch_kkt_apply<Real>(const KktSystem<Real>&, int, KktVec, KktVec, KktVec,      \
                                       hipStream_t);
MFEM_MINRES_INST(float)
MFEM_MINRES_INST(double)

}  // namespace mfemhip
