// ORACLE — test infrastructure only. Not part of the product library.
//
// Small fixed-size dense linear algebra used by the FP64 CPU restatement of
// the reference's mixed-FEM Newton step. Storage is column-major, matching
// the Eigen fixed-size types the reference uses (so vec(F) of a 3x3 F is
// [F00,F10,F20,F01,...], as in reference src/svd/newton_procrustes.h:69-85
// and tests/test_dsvd.cpp:226).
#pragma once
#include <cmath>
#include <cstring>
#include <algorithm>

namespace orc {

template <int R, int C>
struct Mat {
  double a[R * C];
  Mat() { std::memset(a, 0, sizeof(a)); }
  double& operator()(int i, int j) { return a[i + R * j]; }
  double operator()(int i, int j) const { return a[i + R * j]; }
  double& operator[](int i) { return a[i]; }
  double operator[](int i) const { return a[i]; }
  static Mat Identity() {
    Mat m;
    for (int i = 0; i < (R < C ? R : C); ++i) m(i, i) = 1.0;
    return m;
  }
  static Mat Zero() { return Mat(); }
  Mat<C, R> T() const {
    Mat<C, R> t;
    for (int i = 0; i < R; ++i)
      for (int j = 0; j < C; ++j) t(j, i) = (*this)(i, j);
    return t;
  }
  Mat& operator+=(const Mat& o) {
    for (int i = 0; i < R * C; ++i) a[i] += o.a[i];
    return *this;
  }
  Mat& operator-=(const Mat& o) {
    for (int i = 0; i < R * C; ++i) a[i] -= o.a[i];
    return *this;
  }
  Mat& operator*=(double s) {
    for (int i = 0; i < R * C; ++i) a[i] *= s;
    return *this;
  }
  double trace() const {
    double t = 0;
    for (int i = 0; i < (R < C ? R : C); ++i) t += (*this)(i, i);
    return t;
  }
  double squaredNorm() const {
    double t = 0;
    for (int i = 0; i < R * C; ++i) t += a[i] * a[i];
    return t;
  }
  double norm() const { return std::sqrt(squaredNorm()); }
  double dot(const Mat& o) const {
    double t = 0;
    for (int i = 0; i < R * C; ++i) t += a[i] * o.a[i];
    return t;
  }
  template <int R2, int C2>
  Mat<R2, C2> block(int r0, int c0) const {
    Mat<R2, C2> b;
    for (int i = 0; i < R2; ++i)
      for (int j = 0; j < C2; ++j) b(i, j) = (*this)(r0 + i, c0 + j);
    return b;
  }
};

template <int R, int C>
inline Mat<R, C> operator+(Mat<R, C> x, const Mat<R, C>& y) { return x += y; }
template <int R, int C>
inline Mat<R, C> operator-(Mat<R, C> x, const Mat<R, C>& y) { return x -= y; }
template <int R, int C>
inline Mat<R, C> operator*(double s, Mat<R, C> x) { return x *= s; }
template <int R, int C>
inline Mat<R, C> operator-(Mat<R, C> x) { return x *= -1.0; }

template <int R, int K, int C>
inline Mat<R, C> operator*(const Mat<R, K>& x, const Mat<K, C>& y) {
  Mat<R, C> z;
  for (int j = 0; j < C; ++j)
    for (int k = 0; k < K; ++k) {
      double ykj = y(k, j);
      for (int i = 0; i < R; ++i) z(i, j) += x(i, k) * ykj;
    }
  return z;
}

template <int N>
using Vec = Mat<N, 1>;
using M3 = Mat<3, 3>;
using V3 = Vec<3>;

template <int N>
inline Mat<N, N> diag(const double* d) {
  Mat<N, N> m;
  for (int i = 0; i < N; ++i) m(i, i) = d[i];
  return m;
}

// Column-major vectorisation of an R x C matrix.
template <int R, int C>
inline Vec<R * C> flatten(const Mat<R, C>& m) {
  Vec<R * C> v;
  for (int i = 0; i < R * C; ++i) v[i] = m.a[i];
  return v;
}

// vec(A X) = flatten_multiply(A) vec(X)   ->  I (x) A
template <int D>
inline Mat<D * D, D * D> flatten_multiply(const Mat<D, D>& A) {
  Mat<D * D, D * D> K;
  for (int blk = 0; blk < D; ++blk)
    for (int i = 0; i < D; ++i)
      for (int j = 0; j < D; ++j) K(blk * D + i, blk * D + j) = A(i, j);
  return K;
}

// vec(X B) = flatten_multiply_right(B) vec(X)   ->  B^T (x) I
template <int D>
inline Mat<D * D, D * D> flatten_multiply_right(const Mat<D, D>& B) {
  Mat<D * D, D * D> K;
  for (int bi = 0; bi < D; ++bi)
    for (int bj = 0; bj < D; ++bj)
      for (int i = 0; i < D; ++i) K(bi * D + i, bj * D + i) = B(bj, bi);
  return K;
}

inline double det3(const M3& m) {
  return m(0, 0) * (m(1, 1) * m(2, 2) - m(1, 2) * m(2, 1)) -
         m(0, 1) * (m(1, 0) * m(2, 2) - m(1, 2) * m(2, 0)) +
         m(0, 2) * (m(1, 0) * m(2, 1) - m(1, 1) * m(2, 0));
}

inline M3 inv3(const M3& m) {
  M3 r;
  double d = det3(m);
  r(0, 0) = (m(1, 1) * m(2, 2) - m(1, 2) * m(2, 1)) / d;
  r(0, 1) = (m(0, 2) * m(2, 1) - m(0, 1) * m(2, 2)) / d;
  r(0, 2) = (m(0, 1) * m(1, 2) - m(0, 2) * m(1, 1)) / d;
  r(1, 0) = (m(1, 2) * m(2, 0) - m(1, 0) * m(2, 2)) / d;
  r(1, 1) = (m(0, 0) * m(2, 2) - m(0, 2) * m(2, 0)) / d;
  r(1, 2) = (m(0, 2) * m(1, 0) - m(0, 0) * m(1, 2)) / d;
  r(2, 0) = (m(1, 0) * m(2, 1) - m(1, 1) * m(2, 0)) / d;
  r(2, 1) = (m(0, 1) * m(2, 0) - m(0, 0) * m(2, 1)) / d;
  r(2, 2) = (m(0, 0) * m(1, 1) - m(0, 1) * m(1, 0)) / d;
  return r;
}

// Householder QR with column pivoting (Golub & Van Loan, Alg. 5.4.1),
// used where the reference calls Eigen's colPivHouseholderQr().solve()
// (reference src/svd/newton_procrustes.h:141,186). Solves A X = B.
template <int N, int NR>
inline Mat<N, NR> colpiv_qr_solve(Mat<N, N> A, Mat<N, NR> B) {
  int perm[N];
  for (int j = 0; j < N; ++j) perm[j] = j;
  double diagR[N];
  for (int k = 0; k < N; ++k) {
    // pivot: column with largest remaining norm
    int p = k;
    double best = -1;
    for (int j = k; j < N; ++j) {
      double s = 0;
      for (int i = k; i < N; ++i) s += A(i, j) * A(i, j);
      if (s > best) { best = s; p = j; }
    }
    if (p != k) {
      for (int i = 0; i < N; ++i) std::swap(A(i, k), A(i, p));
      std::swap(perm[k], perm[p]);
    }
    // Householder reflector for column k rows k..N-1
    double alpha = 0;
    for (int i = k; i < N; ++i) alpha += A(i, k) * A(i, k);
    alpha = std::sqrt(alpha);
    if (alpha == 0.0) { diagR[k] = 0.0; continue; }
    if (A(k, k) > 0) alpha = -alpha;
    double v[N];
    for (int i = 0; i < N; ++i) v[i] = 0;
    v[k] = A(k, k) - alpha;
    for (int i = k + 1; i < N; ++i) v[i] = A(i, k);
    double vv = 0;
    for (int i = k; i < N; ++i) vv += v[i] * v[i];
    if (vv > 0) {
      for (int j = k; j < N; ++j) {
        double s = 0;
        for (int i = k; i < N; ++i) s += v[i] * A(i, j);
        s = 2.0 * s / vv;
        for (int i = k; i < N; ++i) A(i, j) -= s * v[i];
      }
      for (int j = 0; j < NR; ++j) {
        double s = 0;
        for (int i = k; i < N; ++i) s += v[i] * B(i, j);
        s = 2.0 * s / vv;
        for (int i = k; i < N; ++i) B(i, j) -= s * v[i];
      }
    }
    diagR[k] = A(k, k);
  }
  // back substitution R y = Q^T B, then undo the column permutation
  Mat<N, NR> Y;
  for (int j = 0; j < NR; ++j) {
    for (int i = N - 1; i >= 0; --i) {
      double s = B(i, j);
      for (int k = i + 1; k < N; ++k) s -= A(i, k) * Y(k, j);
      Y(i, j) = (diagR[i] != 0.0) ? s / A(i, i) : 0.0;
    }
  }
  Mat<N, NR> X;
  for (int j = 0; j < NR; ++j)
    for (int i = 0; i < N; ++i) X(perm[i], j) = Y(i, j);
  return X;
}

// Inverse by LU with partial pivoting (Eigen's default for fixed-size
// inverse() above 4x4 is PartialPivLU; reference src/variables/
// mixed_stretch.cpp:162).
template <int N>
inline Mat<N, N> lu_inverse(Mat<N, N> A) {
  int piv[N];
  for (int i = 0; i < N; ++i) piv[i] = i;
  for (int k = 0; k < N; ++k) {
    int p = k;
    double best = std::fabs(A(k, k));
    for (int i = k + 1; i < N; ++i)
      if (std::fabs(A(i, k)) > best) { best = std::fabs(A(i, k)); p = i; }
    if (p != k) {
      for (int j = 0; j < N; ++j) std::swap(A(k, j), A(p, j));
      std::swap(piv[k], piv[p]);
    }
    double d = A(k, k);
    for (int i = k + 1; i < N; ++i) {
      A(i, k) /= d;
      double l = A(i, k);
      for (int j = k + 1; j < N; ++j) A(i, j) -= l * A(k, j);
    }
  }
  Mat<N, N> X;
  for (int c = 0; c < N; ++c) {
    double y[N];
    for (int i = 0; i < N; ++i) {
      double s = (piv[i] == c) ? 1.0 : 0.0;
      for (int k = 0; k < i; ++k) s -= A(i, k) * y[k];
      y[i] = s;
    }
    for (int i = N - 1; i >= 0; --i) {
      double s = y[i];
      for (int k = i + 1; k < N; ++k) s -= A(i, k) * X(k, c);
      X(i, c) = s / A(i, i);
    }
  }
  return X;
}

// Cyclic Jacobi eigen-decomposition of a symmetric matrix:
// A = Q diag(w) Q^T. Published method (Golub & Van Loan, Alg. 8.5.3).
template <int N>
inline void sym_eig(const Mat<N, N>& Ain, double* w, Mat<N, N>& Q) {
  Mat<N, N> A = Ain;
  Q = Mat<N, N>::Identity();
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0, tot = 0;
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) {
        tot += A(i, j) * A(i, j);
        if (i != j) off += A(i, j) * A(i, j);
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < N - 1; ++p)
      for (int q = p + 1; q < N; ++q) {
        double apq = A(p, q);
        if (apq == 0.0) continue;
        double theta = (A(q, q) - A(p, p)) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) /
                   (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < N; ++k) {
          double akp = A(k, p), akq = A(k, q);
          A(k, p) = c * akp - s * akq;
          A(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < N; ++k) {
          double apk = A(p, k), aqk = A(q, k);
          A(p, k) = c * apk - s * aqk;
          A(q, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < N; ++k) {
          double qkp = Q(k, p), qkq = Q(k, q);
          Q(k, p) = c * qkp - s * qkq;
          Q(k, q) = s * qkp + c * qkq;
        }
      }
  }
  for (int i = 0; i < N; ++i) w[i] = A(i, i);
}

// PSD projection used at the reference's sim::simple_psd_fix call sites
// (reference src/energies/stable_neohookean.cpp:92,232,275). Bartels is not
// vendored, so the rule is stated here: eigenvalues below tol are raised to
// tol; a matrix whose eigenvalues are all >= tol is returned unchanged.
template <int N>
inline void simple_psd_fix(Mat<N, N>& A, double tol = 1e-6) {
  double w[N];
  Mat<N, N> Q;
  sym_eig<N>(A, w, Q);
  bool any = false;
  for (int i = 0; i < N; ++i)
    if (w[i] < tol) { w[i] = tol; any = true; }
  if (!any) return;
  Mat<N, N> R;
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      double s = 0;
      for (int k = 0; k < N; ++k) s += Q(i, k) * w[k] * Q(j, k);
      R(i, j) = s;
    }
  A = R;
}

}  // namespace orc
